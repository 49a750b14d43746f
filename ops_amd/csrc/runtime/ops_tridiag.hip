// Batched tridiagonal solves on HBM-resident dats (ops_tridMultiDimBatch[_Inc]).
//
// Semantics follow the reference shim (ops/c/src/tridiag/ops_tridiag.cpp:78-249): the
// solve box starts at `range` (loop-index coordinates) from the dat's index-0 pointer,
// arrays share the padded strides of a->size, and the result overwrites d (or is
// added to u). The arithmetic is the textbook Thomas algorithm:
//   c'_0 = c_0/b_0, d'_0 = d_0/b_0,
//   m = 1/(b_i - a_i c'_{i-1}), c'_i = c_i m, d'_i = (d_i - a_i d'_{i-1}) m,
//   x_{n-1} = d'_{n-1}, x_i = d'_i - c'_i x_{i+1}.
//
// Calls are deferred: consecutive calls with the same solve dimension, the same kind
// (overwrite or increment) and type, and no dat written by one and touched by another,
// are solved together by one launch per 16 systems. A batch of uneven blocks (the ADI
// sweeps of many blocks) then fills the GPU, where one block's lines alone would occupy a
// fraction of it. Every other runtime operation that launches work, copies or
// synchronises flushes the pending batch first (ops_trid_flush).
//
// Two algorithms. The partitioned solve (trid_part, below; the default for lines of at most
// 256 rows) keeps every row in registers and moves 40 bytes per FP64 point. The sequential
// sweeps (OPS_TRID_ALGO=thomas, and lines longer than 256 rows) are laid out as follows.
// Layout on MI355X: one lane per line. For y- and z-lines consecutive lanes own
// consecutive x positions, so every row access of a wavefront is one contiguous segment.
// x-lines are contiguous per line, so a one-wavefront workgroup stages a 64-line x 16-row
// tile of a/b/c/d through LDS with coalesced loads before each lane sweeps its line
// (systems are padded to whole 64-line workgroups). The forward-sweep temporaries c', d'
// live in a scratch array laid out row-major over lines (row i of all lines contiguous),
// so they are written and re-read coalesced.
#include <hip/hip_runtime.h>

#include <cstdint>

#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ops_internal.h"
#include "ops_tridiag.h"

#define HIPCHECK(x) ops_hip_check((int)(x), #x, __FILE__, __LINE__)

namespace {

constexpr int MAXSYS = 16;  // systems per launch (kernel-argument size)
constexpr int XL = 64;      // x-lines: lines per workgroup (rows per LDS tile: XR)

template <typename T> struct Sys {
  T *a, *b, *c, *d, *u;  // box origins
  long long s1, s2;      // element strides of dimensions 1 and 2 (dimension 0: 1)
  int n[3];              // box extents
  long long line0;       // first global line (x-lines: padded to whole workgroups)
  long long nlines;
  long long scr;         // scratch offset (elements) of this system's c' (d' follows)
};

template <typename T> struct Batch {
  Sys<T> sys[MAXSYS];
  int nsys;
  int solvedim;
  unsigned group;  // partitioned solve: workgroups per XCD run (trid_group)
  int rev;         // partitioned solve: groups in reverse order (every other launch)
};

// line l of system S: its origin (elements from the box origin) and stride
template <typename T>
__device__ __forceinline__ long long line_origin(const Sys<T> &S, int sd, long long l) {
  if (sd == 0) {
    const long long p = l % S.n[1], q = l / S.n[1];
    return p * S.s1 + q * S.s2;
  } else if (sd == 1) {
    const long long p = l % S.n[0], q = l / S.n[0];
    return p + q * S.s2;
  }
  const long long p = l % S.n[0], q = l / S.n[0];
  return p + q * S.s1;
}

// y/z-line partitioned solves: workgroups per row of x positions (16 lines each, starting
// at an aligned x that may lie up to 15 before the row's first position)
__host__ __device__ inline long long trid_row_groups(int n0) { return (n0 + 30) / 16; }

// the system holding global line g: the count of later systems whose first line is at or
// before g (line0 increases with the system index). Every comparison reads its own line0,
// so the loads are issued together rather than one after another as a search would.
template <typename T> __device__ __forceinline__ int find_sys(const Batch<T> &B, long long g) {
  int k = 0;
#pragma unroll
  for (int i = 1; i < MAXSYS; i++) k += (i < B.nsys && g >= B.sys[i].line0) ? 1 : 0;
  return k;
}

// rows per chunk of the lane-serial sweeps: a chunk's loads are all issued before its
// arithmetic, so a lane waits for memory once per chunk rather than once per row (the
// recurrence itself is unchanged, row by row)
constexpr int TU = 8;

// y- and z-lines: one lane per line
template <typename T, bool INC>
__global__ __launch_bounds__(256) void thomas_strided(const Batch<T> B, T *__restrict__ scratch) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = find_sys(B, g);
  const Sys<T> &S = B.sys[k];
  const long long l = g - S.line0;
  if (l < 0 || l >= S.nlines) return;
  const int sd = B.solvedim;
  const long long o = line_origin(S, sd, l);
  const long long st = sd == 0 ? 1 : (sd == 1 ? S.s1 : S.s2);
  const int n = S.n[sd];
  const long long L = S.nlines;
  T *cp = scratch + S.scr, *dp = cp + (long long)n * L;
  const T *__restrict__ a = S.a, *__restrict__ b = S.b, *__restrict__ c = S.c;
  T *d = S.d, *u = S.u;
  T cprev = c[o] / b[o];
  T dprev = d[o] / b[o];
  cp[l] = cprev;
  dp[l] = dprev;
  for (int i0 = 1; i0 < n; i0 += TU) {
    T ra[TU], rb[TU], rc[TU], rd[TU];
#pragma unroll
    for (int j = 0; j < TU; j++)
      if (i0 + j < n) {
        const long long kk = o + (long long)(i0 + j) * st;
        ra[j] = a[kk];
        rb[j] = b[kk];
        rc[j] = c[kk];
        rd[j] = d[kk];
      }
#pragma unroll
    for (int j = 0; j < TU; j++)
      if (i0 + j < n) {
        const T m = T(1) / (rb[j] - ra[j] * cprev);
        cprev = rc[j] * m;
        dprev = (rd[j] - ra[j] * dprev) * m;
        cp[(long long)(i0 + j) * L + l] = cprev;
        dp[(long long)(i0 + j) * L + l] = dprev;
      }
  }
  T x = dprev;
  long long kk = o + (long long)(n - 1) * st;
  if (INC) u[kk] += x; else d[kk] = x;
  for (int i1 = n - 2; i1 >= 0; i1 -= TU) {
    // rows i1, i1 - 1, ..., i1 - TU + 1
    T rc[TU], rd[TU], ru[TU];
#pragma unroll
    for (int j = 0; j < TU; j++)
      if (i1 - j >= 0) {
        const long long i = i1 - j;
        rc[j] = cp[i * L + l];
        rd[j] = dp[i * L + l];
        if (INC) ru[j] = u[o + i * st];
      }
#pragma unroll
    for (int j = 0; j < TU; j++)
      if (i1 - j >= 0) {
        x = rd[j] - rc[j] * x;
        const long long kj = o + (long long)(i1 - j) * st;
        if (INC) u[kj] = ru[j] + x; else d[kj] = x;
      }
  }
}

// x-lines: 64 lines of one system per workgroup, rows staged through LDS in chunks of XR
// (16, 8 or 4: taller tiles coalesce better, shorter ones leave LDS for more workgroups
// per CU; the autotuner picks)
template <typename T, bool INC, int XR>
__global__ __launch_bounds__(64) void thomas_x(const Batch<T> B, T *__restrict__ scratch) {
  __shared__ T sa[XL][XR + 1], sb[XL][XR + 1], sc[XL][XR + 1], sd[XL][XR + 1];
  const int lane = threadIdx.x;
  const long long g0 = (long long)blockIdx.x * XL;  // first global line of the workgroup
  const int k = find_sys(B, g0);
  const Sys<T> &S = B.sys[k];
  const long long l0 = g0 - S.line0;
  if (l0 >= S.nlines) return;  // padding workgroup past the last system (uniform)
  const int n = S.n[0];
  const long long L = S.nlines;
  const long long myl = l0 + lane;
  const bool own = myl < L;
  T *cp = scratch + S.scr, *dp = cp + (long long)n * L;
  const T *a = S.a, *b = S.b, *c = S.c;
  T *d = S.d, *u = S.u;
  T cprev = T(0), dprev = T(0);
  for (int r0 = 0; r0 < n; r0 += XR) {
    const int rows = (n - r0) < XR ? (n - r0) : XR;
    // cooperative coalesced load: thread -> (line, row) with row fastest
    for (int e = lane; e < XL * XR; e += XL) {
      const int li = e / XR, ri = e % XR;
      const long long ll = l0 + li;
      if (ll < L && ri < rows) {
        const long long kk = line_origin(S, 0, ll) + r0 + ri;
        sa[li][ri] = a[kk];
        sb[li][ri] = b[kk];
        sc[li][ri] = c[kk];
        sd[li][ri] = d[kk];
      }
    }
    __syncthreads();
    if (own) {
      for (int ri = 0; ri < rows; ri++) {
        const int i = r0 + ri;
        if (i == 0) {
          cprev = sc[lane][ri] / sb[lane][ri];
          dprev = sd[lane][ri] / sb[lane][ri];
        } else {
          const T m = T(1) / (sb[lane][ri] - sa[lane][ri] * cprev);
          cprev = sc[lane][ri] * m;
          dprev = (sd[lane][ri] - sa[lane][ri] * dprev) * m;
        }
        cp[i * L + myl] = cprev;
        dp[i * L + myl] = dprev;
      }
    }
    __syncthreads();
  }
  // backward sweep: each lane walks its line; results are staged back through LDS so the
  // global stores are coalesced as well
  T x = T(0);
  const int nchunks = (n + XR - 1) / XR;
  for (int ci = nchunks - 1; ci >= 0; ci--) {
    const int r0 = ci * XR;
    const int rows = (n - r0) < XR ? (n - r0) : XR;
    if (own) {
      T rc[XR], rd[XR];
#pragma unroll
      for (int ri = 0; ri < XR; ri++)
        if (ri < rows) {
          rc[ri] = cp[(long long)(r0 + ri) * L + myl];
          rd[ri] = dp[(long long)(r0 + ri) * L + myl];
        }
#pragma unroll
      for (int ri = XR - 1; ri >= 0; ri--)
        if (ri < rows) {
          if (r0 + ri == n - 1) x = rd[ri];
          else x = rd[ri] - rc[ri] * x;
          sd[lane][ri] = x;
        }
    }
    __syncthreads();
    for (int e = lane; e < XL * XR; e += XL) {
      const int li = e / XR, ri = e % XR;
      const long long ll = l0 + li;
      if (ll < L && ri < rows) {
        const long long kk = line_origin(S, 0, ll) + r0 + ri;
        if (INC) u[kk] += sd[li][ri]; else d[kk] = sd[li][ri];
      }
    }
    __syncthreads();
  }
}

// c', d' scratch. A graph captured earlier may still hold an outgrown buffer, so it is
// retired, not freed, until ops_exit.
// Partitioned solve (no scratch; the default). A workgroup of 256 lanes takes 16 lines of one
// system; each line is cut into 16 partitions of m = ceil(n/16) consecutive rows, one lane
// per partition, the lane's rows held in registers. Lane index: line fastest, so for y-
// and z-lines a row of 16 lines is one contiguous 128-byte segment.
// Within a partition (local rows 0..c-1, c >= 2) the rows are reduced to two equations in
// the partition's first and last unknowns x_F, x_E and the neighbours' x_E(p-1), x_F(p+1):
//   forward, j = 1..c-1 (row j in terms of x_F and x_{j+1}):
//     r = 1/(b_j - a_j cc_{j-1}) (cc_0 = 0, aa_0 = -1),  dd_j = r (d_j - a_j dd_{j-1}),
//     aa_j = -r a_j aa_{j-1},  cc_j = r c_j
//     (the last row then reads aa x_F + x_E + cc x_F(p+1) = dd);
//   backward, j = c-3..1 (row j in terms of x_F and x_E):
//     dd_j -= cc_j dd_{j+1},  aa_j -= cc_j aa_{j+1},  cc_j = -cc_j cc_{j+1};
//   first row (x_1 = dd_1 - aa_1 x_F - cc_1 x_E; for c = 2, x_1 = x_E):
//     r = 1/(b_0 - c_0 aa_1):  (r a_0) x_E(p-1) + x_F + (-r c_0 cc_1) x_E = r (d_0 - c_0 dd_1).
// A one-row partition is its own reduced row (divided by b). The reduced rows, in order
// F_0, E_0, F_1, ..., form a tridiagonal system of 32 slots per line (slots of empty
// partitions, which only follow the last row, are identity rows; the first row's
// sub-diagonal and the last row's super-diagonal are dropped, as the sequential sweep
// ignores a_0 and c_{n-1}). It is solved by parallel cyclic reduction across the line's 16
// lanes, two slots per lane, five steps through LDS (Hockney): with row i's neighbours
// i -+ s (identity beyond the ends), k1 = a_i / b_{i-s}, k2 = c_i / b_{i+s},
//   a_i' = -(a_{i-s} k1), c_i' = -(c_{i+s} k2), b_i' = b_i - c_{i-s} k1 - a_{i+s} k2,
//   d_i' = d_i - d_{i-s} k1 - d_{i+s} k2,  s = 1, 2, 4, 8, 16;  then x_i = d_i / b_i.
// Every interior row is then x_j = dd_j - aa_j x_F - cc_j x_E. Traffic: a, b, c, d read
// once and the result written once (40 B per FP64 point; 48 B with the increment), where
// the sequential sweep also writes and re-reads c', d'. The result differs from the
// sequential sweep's in the last bits (the eliminations are ordered differently).
constexpr int PL = 16;  // lines per workgroup
// Line groups of consecutive workgroups run on the same XCD in runs of nb (workgroup b
// runs on XCD b % 8; within each span of 8 nb workgroups, XCD x takes groups nb x ..
// nb x + nb - 1). Neighbouring groups of 16 y- or z-lines share the cache lines that
// straddle them (a row of a group is 128-byte aligned only where the row pitch allows),
// and on one XCD the second reader finds them in that XCD's L2; the XCDs still work on
// neighbouring spans at any moment. A partial last span keeps its order.
// With rev set the groups are taken in reverse order (launches alternate, so each starts on
// the lines the previous one touched last, which the memory-side cache still holds;
// neighbouring groups stay on one XCD).
__device__ __forceinline__ unsigned trid_group(unsigned b, unsigned n, unsigned nb, int rev) {
  const unsigned S = 8 * nb;
  unsigned g = b;
  if (nb > 1 && b < n - n % S) {
    const unsigned q = b / S, r = b % S;
    g = q * S + (r % 8) * nb + r / 8;
  }
  return rev ? n - 1 - g : g;
}
constexpr int PP = 16;  // partitions (lanes) per line
constexpr int PM = 16;  // rows per partition at most: n <= PP * PM

// x-lines (XT): a line's rows are contiguous, so the 16 lines of the workgroup are staged
// through an LDS tile, one array at a time, with coalesced row loads (and the results
// stored back the same way); the lanes then read their partitions' rows from the tile.
constexpr int PXP = PP * PM + 1;  // tile pitch (odd: the 16 lanes reading one row of 16 lines
                                  // hit 16 different banks)

// Three 256-lane workgroups per CU (three wavefronts per SIMD): the x-line variant needs ~240
// registers per lane unconstrained (two wavefronts per SIMD) and spills a few rows to
// scratch at 168; the extra wavefront in flight hides more HBM latency than the spills cost
// (ADI's 48-block step 13.9 -> 12.4 ms; at four wavefronts, 128 registers, the spills
// dominate: 19 ms). The y/z variants fit in 166 registers without spilling.
template <typename T, bool INC, bool XT, bool PCRR>
__global__ __launch_bounds__(256, 3) void trid_part(const Batch<T> B) {
  // reduced rows (sub-diagonal, diagonal, super-diagonal, rhs), two buffers for the PCR
  // steps; for x-lines they share the memory of the row tile (never live at the same time)
  constexpr size_t RED_BYTES = 2 * sizeof(T) * PL * 2 * PP * 4;
  constexpr size_t TILE_BYTES = XT ? sizeof(T) * PL * PXP : 0;
  __shared__ __align__(16) char smem[RED_BYTES > TILE_BYTES ? RED_BYTES : TILE_BYTES];
  T(*red)[PL][2 * PP][4] = (T(*)[PL][2 * PP][4])smem;
  T(*tile)[PXP] = (T(*)[PXP])smem;
  const int tid = threadIdx.x, li = tid % PL, p = tid / PL;
  const long long g0 = (long long)trid_group(blockIdx.x, gridDim.x, B.group, B.rev) * PL;
  const int k = find_sys(B, g0);
  const Sys<T> &S = B.sys[k];
  const int sd = XT ? 0 : B.solvedim;
  const int n = S.n[sd];
  const int m = (n + PP - 1) / PP;
  int c = n - p * m;
  c = c < 0 ? 0 : (c > m ? m : c);
  const long long st = sd == 0 ? 1 : (sd == 1 ? S.s1 : S.s2);
  bool live;
  long long o = 0;
  if (XT) {
    const long long l = g0 - S.line0 + li;
    live = l < S.nlines;
    if (live) o = line_origin(S, 0, l);
  } else {
    // y- and z-lines: the workgroup takes 16 consecutive x positions of one row of the box
    // (a z-plane row for y-lines, a y-row for z-lines) starting at a 128-byte boundary of
    // the arrays (they share one layout), so that a row of its 16 lines is one whole cache
    // line rather than parts of two that neighbouring workgroups fetch again; lanes outside
    // the box idle, and trid_row_groups workgroups per row cover any alignment
    const long long G = trid_row_groups(S.n[0]);
    const long long w = (g0 - S.line0) / PL;
    const long long r = w / G;
    const long long ro = r * (sd == 1 ? S.s2 : S.s1);
    const int mis = (int)(((uintptr_t)(S.d + ro) / sizeof(T)) % PL);
    const long long x = PL * (w - r * G) - mis + li;
    live = x >= 0 && x < S.n[0];
    if (live) o = ro + x;
  }
  if (live) o += (long long)p * m * st;
  const T *__restrict__ a = S.a, *__restrict__ b = S.b, *__restrict__ cc_in = S.c;
  T *d = S.d, *u = S.u;
  const bool act = live && c > 0;
  T ra[PM], rb[PM], rc[PM], rd[PM];
  // x-lines: the line origins of the workgroup, kept in LDS rather than in 32 registers
  // per lane (the kernel's register count sets how many wavefronts a SIMD holds)
  __shared__ long long lo[XT ? PL : 1];
  if (XT) {
    // coalesced loads of the 16 lines' rows into the tile, one array at a time
    // thread r moves row r of every line (n <= 256): the 16 loads of an array are all in
    // flight before the first LDS write
    const long long nl = S.nlines - (g0 - S.line0);  // live lines of the workgroup
    if (tid < PL) lo[tid] = tid < nl ? line_origin(S, 0, g0 - S.line0 + tid) : 0;
    __syncthreads();
    // software-pipelined: the next array's loads are in flight while the lanes read the
    // current one from the tile
    T v[PL];
#pragma unroll
    for (int q = 0; q < PL; q++)
      if (q < nl && tid < n) v[q] = a[lo[q] + tid];
#pragma unroll
    for (int arr = 0; arr < 4; arr++) {
#pragma unroll
      for (int q = 0; q < PL; q++)
        if (q < nl && tid < n) tile[q][tid] = v[q];
      __syncthreads();
      if (arr < 3) {
        const T *src = arr == 0 ? b : (arr == 1 ? cc_in : (const T *)d);
#pragma unroll
        for (int q = 0; q < PL; q++)
          if (q < nl && tid < n) v[q] = src[lo[q] + tid];
      }
      T *dst = arr == 0 ? ra : (arr == 1 ? rb : (arr == 2 ? rc : rd));
#pragma unroll
      for (int j = 0; j < PM; j++)
        if (act && j < c) dst[j] = tile[li][p * m + j];
      __syncthreads();
    }
  } else if (act) {
#pragma unroll
    for (int j = 0; j < PM; j++)
      if (j < c) {
        const long long kk = o + (long long)j * st;
        ra[j] = a[kk];
        rb[j] = b[kk];
        rc[j] = cc_in[kk];
        rd[j] = d[kk];
      }
  }
  if (act) {
    const T a0 = ra[0], b0 = rb[0], c0 = rc[0], d0 = rd[0];
    // forward: rows 1..c-1 in terms of x_F (aa) and the next row (cc)
    T cprev = T(0), dprev = T(0), aprev = T(-1);
#pragma unroll
    for (int j = 1; j < PM; j++)
      if (j < c) {
        const T r = T(1) / (rb[j] - ra[j] * cprev);
        rd[j] = (rd[j] - ra[j] * dprev) * r;
        ra[j] = -(ra[j] * aprev) * r;
        rc[j] = rc[j] * r;
        cprev = rc[j];
        dprev = rd[j];
        aprev = ra[j];
      }
    // backward: rows c-3..1 in terms of x_F and x_E
#pragma unroll
    for (int j = PM - 3; j >= 1; j--)
      if (j <= c - 3) {
        rd[j] = rd[j] - rc[j] * rd[j + 1];
        ra[j] = ra[j] - rc[j] * ra[j + 1];
        rc[j] = -rc[j] * rc[j + 1];
      }
    T fa, fc, fd;
    if (c == 1) {
      const T r = T(1) / b0;
      fa = a0 * r;
      fc = c0 * r;
      fd = d0 * r;
    } else {
      // x_1 = dd_1 - aa_1 x_F - cc_1 x_E (for c = 2 the next row is x_E itself)
      const T d1 = c == 2 ? T(0) : rd[1], a1 = c == 2 ? T(0) : ra[1], e1 = c == 2 ? T(-1) : rc[1];
      const T r = T(1) / (b0 - c0 * a1);
      fa = a0 * r;
      fc = -(c0 * e1) * r;
      fd = (d0 - c0 * d1) * r;
    }
    ra[0] = fa;
    rc[0] = fc;
    rd[0] = fd;
  }
  // this lane's two reduced rows (slot 2p: first row, 2p + 1: last row), identity rows for
  // what the partition does not have
  const bool lastp = (p + 1) * m >= n;  // the partition holding row n - 1
  T sa[2], sb[2], sc[2], sdd[2];
#pragma unroll
  for (int e = 0; e < 2; e++) {
    sa[e] = T(0);
    sb[e] = T(1);
    sc[e] = T(0);
    sdd[e] = T(0);
  }
  if (act) {
    sa[0] = p == 0 ? T(0) : ra[0];
    sc[0] = (lastp && c == 1) ? T(0) : rc[0];
    sdd[0] = rd[0];
    if (c >= 2) {
#pragma unroll
      for (int j = 1; j < PM; j++)
        if (j == c - 1) {
          sa[1] = ra[j];
          sc[1] = lastp ? T(0) : rc[j];
          sdd[1] = rd[j];
        }
    }
  }
  // lines of at most 16 rows (one row per partition): slot i is row i, so that the chain
  // of reduced rows has no identity rows between real ones; the rows move to their slots'
  // owners through LDS
  const bool one = m == 1;
  if (one) {
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int sl = e == 0 ? p : p + PP;  // lane p's row (slot p), or an identity slot
      const bool mine = e == 0;
      red[0][li][sl][0] = mine ? sa[0] : T(0);
      red[0][li][sl][1] = mine ? sb[0] : T(1);
      red[0][li][sl][2] = mine ? sc[0] : T(0);
      red[0][li][sl][3] = mine ? sdd[0] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 2; e++) {
      sa[e] = red[0][li][2 * p + e][0];
      sb[e] = red[0][li][2 * p + e][1];
      sc[e] = red[0][li][2 * p + e][2];
      sdd[e] = red[0][li][2 * p + e][3];
    }
  }
  T xo[2];  // the solution of this lane's two slots (of its row, for one-row lines)
  if (PCRR) {
  // parallel cyclic reduction over the 32 slots of the line
  int buf = one ? 1 : 0;
#pragma unroll
  for (int s = 1; s < 2 * PP; s *= 2) {
#pragma unroll
    for (int e = 0; e < 2; e++) {
      red[buf][li][2 * p + e][0] = sa[e];
      red[buf][li][2 * p + e][1] = sb[e];
      red[buf][li][2 * p + e][2] = sc[e];
      red[buf][li][2 * p + e][3] = sdd[e];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int i = 2 * p + e;
      T am = T(0), bm = T(1), cm = T(0), dm = T(0), ap = T(0), bp = T(1), cp = T(0), dp = T(0);
      if (i - s >= 0) {
        am = red[buf][li][i - s][0];
        bm = red[buf][li][i - s][1];
        cm = red[buf][li][i - s][2];
        dm = red[buf][li][i - s][3];
      }
      if (i + s < 2 * PP) {
        ap = red[buf][li][i + s][0];
        bp = red[buf][li][i + s][1];
        cp = red[buf][li][i + s][2];
        dp = red[buf][li][i + s][3];
      }
      const T k1 = sa[e] / bm, k2 = sc[e] / bp;
      sa[e] = -(am * k1);
      sc[e] = -(cp * k2);
      sb[e] = sb[e] - cm * k1 - ap * k2;
      sdd[e] = sdd[e] - dm * k1 - dp * k2;
    }
    buf ^= 1;
  }
  xo[0] = sdd[0] / sb[0];
  xo[1] = sdd[1] / sb[1];
  if (one) {
    // lane p's row sits in slot p, owned by lane p / 2
#pragma unroll
    for (int e = 0; e < 2; e++) red[buf][li][2 * p + e][3] = xo[e];
    __syncthreads();
    xo[0] = red[buf][li][p][3];
  }
  } else {
    // the reduced rows through one lane per line: the sequential sweep over the 32 slots
    // (unit diagonal; identity rows only follow the last real row)
    if (!one) {
#pragma unroll
      for (int e = 0; e < 2; e++) {
        red[0][li][2 * p + e][0] = sa[e];
        red[0][li][2 * p + e][2] = sc[e];
        red[0][li][2 * p + e][3] = sdd[e];
      }
    }
    __syncthreads();
    if (p == 0) {
      T cp = T(0), dp = T(0);
      for (int sl = 0; sl < 2 * PP; sl++) {
        const T al = red[0][li][sl][0], ga = red[0][li][sl][2], de = red[0][li][sl][3];
        if (sl == 0) {
          cp = ga;
          dp = de;
        } else {
          const T r = T(1) / (T(1) - al * cp);
          cp = ga * r;
          dp = (de - al * dp) * r;
        }
        red[0][li][sl][2] = cp;
        red[0][li][sl][3] = dp;
      }
      T x = T(0);
      for (int sl = 2 * PP - 1; sl >= 0; sl--) {
        x = sl == 2 * PP - 1 ? red[0][li][sl][3] : red[0][li][sl][3] - red[0][li][sl][2] * x;
        red[1][li][sl][3] = x;
      }
    }
    __syncthreads();
    xo[0] = red[1][li][one ? p : 2 * p][3];
    xo[1] = red[1][li][2 * p + 1][3];
  }
  if (XT) __syncthreads();  // the tile's memory is reused below
  if (act) {
    const T xf = xo[0];
    const T xe = c >= 2 ? xo[1] : xf;
#pragma unroll
    for (int j = 0; j < PM; j++)
      if (j < c) {
        const T x = j == 0 ? xf : (j == c - 1 ? xe : rd[j] - ra[j] * xf - rc[j] * xe);
        if (XT) {
          tile[li][p * m + j] = x;
        } else {
          const long long kk = o + (long long)j * st;
          if (INC) u[kk] += x; else d[kk] = x;
        }
      }
  }
  if (XT) {
    // the solution back through the tile: coalesced row stores
    __syncthreads();
    const long long nl = S.nlines - (g0 - S.line0);
#pragma unroll
    for (int q = 0; q < PL; q++)
      if (q < nl && tid < n) {
        if (INC) u[lo[q] + tid] += tile[q][tid]; else d[lo[q] + tid] = tile[q][tid];
      }
  }
}

struct Scratch {
  char *p = nullptr;
  size_t cap = 0;
  std::vector<char *> retired;
};
Scratch g_scratch;

void *sweep_scratch(size_t bytes) {
  if (bytes > g_scratch.cap) {
    if (g_scratch.p) g_scratch.retired.push_back(g_scratch.p);
    g_scratch.p = nullptr;
    g_scratch.cap = 0;
    HIPCHECK(hipMalloc((void **)&g_scratch.p, bytes));
    g_scratch.cap = bytes;
  }
  return g_scratch.p;
}

// one deferred call
struct Call {
  int ndim, solvedim, inc, is_double;
  int range[6];
  ops_dat a, b, c, d, u;
};

struct Pending {
  std::vector<Call> calls;
  bool flushing = false;
};
Pending g_pending;

template <typename T> T *box_origin(ops_dat x, long long off) {
  return x == nullptr ? nullptr : (T *)(x->data_d + x->base_offset) + off;
}

// OPS_TRID_ALGO=thomas selects the sequential sweeps everywhere (results bitwise equal to
// a sequential Thomas solve of every line); the default is the partitioned solve for lines
// of at most 256 rows, the sweeps for longer ones
bool use_thomas() {
  const char *e = getenv("OPS_TRID_ALGO");
  return e != nullptr && !strcmp(e, "thomas");
}

// launches the systems calls[first, first + count) (same solvedim, kind and type).
// fixed != nullptr: the allocation-free form, sequential sweeps with c', d' in `fixed`
// (fixed_bytes of them), no autotuning
template <typename T, bool INC>
void launch_chunk(const Call *calls, size_t first, size_t count, T *fixed = nullptr,
                  size_t fixed_bytes = 0) {
  Batch<T> B;
  B.nsys = 0;
  B.solvedim = calls[first].solvedim;
  // runs of 4 neighbouring line groups per XCD: ADI's 48-block step 12.2 -> 11.8 ms (runs of
  // 2 and 8 measure the same; one XCD band per launch, ops_xcd_remap, 12.5 ms)
  B.group = 4;
  static unsigned launches = 0;
  B.rev = (int)(launches++ & 1u);
  const int sd = B.solvedim;
  bool part = fixed == nullptr && !use_thomas();
  for (size_t k = first; k < first + count; k++) {
    const Call &C = calls[k];
    if (C.range[2 * sd + 1] - C.range[2 * sd] > PP * PM) part = false;
  }
  // lines of each system are padded to whole workgroups of the kernel that solves them
  const long long pad = part ? PL : (sd == 0 ? XL : 1);
  long long line = 0, scr = 0;
  for (size_t k = first; k < first + count; k++) {
    const Call &C = calls[k];
    Sys<T> S;
    long long s[3] = {1, C.a->size[0], (long long)C.a->size[0] * C.a->size[1]};
    long long off = 0;
    bool empty = false;
    for (int i = 0; i < 3; i++) S.n[i] = 1;
    for (int i = 0; i < C.ndim; i++) {
      S.n[i] = C.range[2 * i + 1] - C.range[2 * i];
      if (S.n[i] <= 0) empty = true;
      off += (long long)C.range[2 * i] * s[i];
    }
    if (empty) continue;
    S.a = box_origin<T>(C.a, off);
    S.b = box_origin<T>(C.b, off);
    S.c = box_origin<T>(C.c, off);
    S.d = box_origin<T>(C.d, off);
    S.u = box_origin<T>(C.u, off);
    S.s1 = s[1];
    S.s2 = s[2];
    const long long total = (long long)S.n[0] * S.n[1] * S.n[2];
    S.nlines = total / S.n[sd];
    S.line0 = line;
    S.scr = scr;
    if (part && sd != 0)  // rows of aligned 16-line groups (trid_part)
      line += (S.nlines / S.n[0]) * trid_row_groups(S.n[0]) * PL;
    else
      line += ((S.nlines + pad - 1) / pad) * pad;
    if (!part) scr += 2 * total;
    B.sys[B.nsys++] = S;
  }
  if (B.nsys == 0) return;
  hipStream_t st = (hipStream_t)ops_hip_get_stream_raw();
  double pts = 0.0;  // system points of the launch (autotuning compares time per point)
  for (int k = 0; k < B.nsys; k++) pts += (double)B.sys[k].nlines * B.sys[k].n[sd];
  if (part) {
    const dim3 grid((unsigned)(line / PL)), blk(PL * PP);
    if (sd == 0) {
      // x-lines through the LDS tile (steady state 0.91 ms per ADI launch of 65 M points,
      // against 1.18 ms with each lane loading its own rows directly; the direct variant
      // computes the same operations and stays for reference)
      hipLaunchKernelGGL((trid_part<T, INC, true, false>), grid, blk, 0, st, B);
    } else {
      hipLaunchKernelGGL((trid_part<T, INC, false, true>), grid, blk, 0, st, B);
    }
    ops_hip_check_launch("ops_tridMultiDimBatch");
    return;
  }
  if (fixed != nullptr && sizeof(T) * (size_t)scr > fixed_bytes)
    ops_throw(OPS_INVALID_ARGUMENT,
              "ops_tridMultiDimBatch_noalloc: the solve needs " +
                  std::to_string(sizeof(T) * (size_t)scr) + " bytes of c', d' scratch, more than the " +
                  std::to_string(fixed_bytes) + " bytes of the static scratch");
  T *scr_p = fixed != nullptr ? fixed : (T *)sweep_scratch(sizeof(T) * (size_t)scr);
  if (sd == 0 && fixed != nullptr) {
    hipLaunchKernelGGL((thomas_x<T, INC, 8>), dim3((unsigned)(line / XL)), dim3(XL), 0, st, B, scr_p);
  } else if (sd == 0) {
    // candidates: LDS tiles of 16, 8 or 4 rows. One lane per line without LDS (the y/z
    // kernel with unit stride, the loads of a wavefront spread over 64 lines) measured
    // 0.049 ns per point on ADI's x solves against 0.026 for 8-row tiles
    static const int kid = ops_hip_register_kernel("ops_tridiag_x");
    const int c = ops_tune_next(kid, 3, 2);
    const dim3 grid((unsigned)(line / XL));
    if (c >= 0 && ops_kernel_table()[kid].tune_state == 1) ops_tune_begin(kid, c, pts);
    const bool measured = c >= 0 && ops_kernel_table()[kid].tune_state == 1;
    if (c == 1)
      hipLaunchKernelGGL((thomas_x<T, INC, 8>), grid, dim3(XL), 0, st, B, scr_p);
    else if (c == 2)
      hipLaunchKernelGGL((thomas_x<T, INC, 4>), grid, dim3(XL), 0, st, B, scr_p);
    else
      hipLaunchKernelGGL((thomas_x<T, INC, 16>), grid, dim3(XL), 0, st, B, scr_p);
    if (measured) ops_tune_end(kid, 3, 2);
  } else
    hipLaunchKernelGGL((thomas_strided<T, INC>), dim3((unsigned)((line + 255) / 256)), dim3(256), 0,
                       st, B, scr_p);
  ops_hip_check_launch("ops_tridMultiDimBatch");
}

bool writes(const Call &c, ops_dat x) { return x != nullptr && (c.inc ? c.u == x : c.d == x); }
bool touches(const Call &c, ops_dat x) {
  return x != nullptr && (c.a == x || c.b == x || c.c == x || c.d == x || c.u == x);
}

// a new call joins the pending batch when nothing one writes is touched by the other
bool compatible(const Call &n) {
  for (const Call &p : g_pending.calls) {
    if (p.solvedim != n.solvedim || p.inc != n.inc || p.is_double != n.is_double) return false;
    ops_dat nd[5] = {n.a, n.b, n.c, n.d, n.u};
    for (ops_dat x : nd)
      if (writes(p, x)) return false;
    ops_dat pd[5] = {p.a, p.b, p.c, p.d, p.u};
    for (ops_dat x : pd)
      if (writes(n, x)) return false;
  }
  return true;
}

void check_args(int ndim, int solvedim, ops_dat a, ops_dat b, ops_dat c, ops_dat d, ops_dat u) {
  if (ndim < 1 || ndim > 3 || solvedim < 0 || solvedim >= ndim)
    ops_throw(OPS_INVALID_ARGUMENT, "Tridsolver error: ndim must be 1..3 and solvedim < ndim");
  for (int i = 0; i < ndim; i++) {
    if (a->size[i] != b->size[i] || b->size[i] != c->size[i] || c->size[i] != d->size[i] ||
        (u && u->size[i] != d->size[i]))
      ops_throw(OPS_RUNTIME_ERROR,
                "Tridsolver error: the a,b,c,d datasets all need to be the same size");
  }
  if (strcmp(a->type, b->type) || strcmp(b->type, c->type) || strcmp(c->type, d->type) ||
      (u && strcmp(u->type, d->type)))
    ops_throw(OPS_RUNTIME_ERROR,
              "Tridsolver error: the a,b,c,d datasets must all be of the same type");
  if (strcmp(a->type, "double") && strcmp(a->type, "float"))
    ops_throw(OPS_RUNTIME_ERROR,
              "Tridsolver error: unsupported type. Dataset type must either be \"double\" or "
              "\"float\"");
  ops_dat all[5] = {a, b, c, d, u};
  for (ops_dat x : all)
    if (x) ops_put_data(x);
}

void enqueue(int ndim, int solvedim, const int *range, ops_dat a, ops_dat b, ops_dat c, ops_dat d,
             ops_dat u, int inc) {
  check_args(ndim, solvedim, a, b, c, d, u);
  Call C;
  C.ndim = ndim;
  C.solvedim = solvedim;
  C.inc = inc;
  C.is_double = !strcmp(a->type, "double");
  for (int i = 0; i < 6; i++) C.range[i] = i < 2 * ndim ? range[i] : 0;
  C.a = a;
  C.b = b;
  C.c = c;
  C.d = d;
  C.u = u;
  // a loop batch issued before this call runs before it
  ops_loop_batch_flush();
  if (!compatible(C)) ops_trid_flush();
  g_pending.calls.push_back(C);
  if (inc) u->dirty_hd = 2;
  else d->dirty_hd = 2;
}

}  // namespace

void ops_trid_flush() {
  // a pending loop batch was issued before any pending tridiagonal call (a loop's launch
  // flushes the tridiagonal queue first), so it runs first
  ops_loop_batch_flush();
  if (g_pending.calls.empty() || g_pending.flushing) return;
  g_pending.flushing = true;
  std::vector<Call> calls;
  calls.swap(g_pending.calls);
  try {
    const Call &c0 = calls[0];
    for (size_t first = 0; first < calls.size(); first += MAXSYS) {
      const size_t count = calls.size() - first < (size_t)MAXSYS ? calls.size() - first : MAXSYS;
      if (c0.is_double) {
        if (c0.inc) launch_chunk<double, true>(calls.data(), first, count);
        else launch_chunk<double, false>(calls.data(), first, count);
      } else {
        if (c0.inc) launch_chunk<float, true>(calls.data(), first, count);
        else launch_chunk<float, false>(calls.data(), first, count);
      }
    }
  } catch (...) {
    g_pending.flushing = false;
    throw;
  }
  g_pending.flushing = false;
}

void ops_trid_discard() { g_pending.calls.clear(); }

void ops_trid_release() {
  for (char *p : g_scratch.retired) (void)hipFree(p);
  g_scratch.retired.clear();
  if (g_scratch.p) (void)hipFree(g_scratch.p);
  g_scratch.p = nullptr;
  g_scratch.cap = 0;
}

// ---------------------------------------------------------------- allocation-free form
// c', d' of ops_tridMultiDimBatch_noalloc live in this file-scope device array, one copy per
// GPU, whose device address is taken once per device when the runtime initialises it
// (ops_trid_device_init); host code uses that address, never the array's name.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
void *g_noalloc_scratch[MAX_DEVICES] = {};
}  // namespace

void ops_trid_device_init(int device) {
  if (device < 0 || device >= MAX_DEVICES)
    ops_throw(OPS_RUNTIME_CONFIGURATION_ERROR, "ops_tridiag: device index beyond the supported count");
  if (g_noalloc_scratch[device] != nullptr) return;
  void *p = nullptr;
  HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_noalloc_scratch[device] = p;
}

extern "C" void ops_tridMultiDimBatch_noalloc(int ndim, int solvedim, int *range, ops_dat a,
                                              ops_dat b, ops_dat c, ops_dat d, ops_dat u,
                                              ops_tridsolver_params *ctx) {
  (void)ctx;
  check_args(ndim, solvedim, a, b, c, d, u);
  // work issued before this call (a pending loop batch, deferred solves) runs first
  ops_trid_flush();
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES || g_noalloc_scratch[dev] == nullptr)
    ops_throw(OPS_RUNTIME_ERROR, "ops_tridMultiDimBatch_noalloc: called before ops_init on this device");
  Call C;
  C.ndim = ndim;
  C.solvedim = solvedim;
  C.inc = u != nullptr;
  C.is_double = !strcmp(a->type, "double");
  for (int i = 0; i < 6; i++) C.range[i] = i < 2 * ndim ? range[i] : 0;
  C.a = a;
  C.b = b;
  C.c = c;
  C.d = d;
  C.u = u;
  void *scr = g_noalloc_scratch[dev];
  if (C.is_double) {
    if (C.inc) launch_chunk<double, true>(&C, 0, 1, (double *)scr, NOALLOC_SCRATCH_BYTES);
    else launch_chunk<double, false>(&C, 0, 1, (double *)scr, NOALLOC_SCRATCH_BYTES);
  } else {
    if (C.inc) launch_chunk<float, true>(&C, 0, 1, (float *)scr, NOALLOC_SCRATCH_BYTES);
    else launch_chunk<float, false>(&C, 0, 1, (float *)scr, NOALLOC_SCRATCH_BYTES);
  }
  if (C.inc) u->dirty_hd = 2;
  else d->dirty_hd = 2;
}

extern "C" void ops_tridMultiDimBatch(int ndim, int solvedim, int *range, ops_dat a, ops_dat b,
                                      ops_dat c, ops_dat d, ops_tridsolver_params *ctx) {
  (void)ctx;
  enqueue(ndim, solvedim, range, a, b, c, d, nullptr, 0);
}

extern "C" void ops_tridMultiDimBatch_Inc(int ndim, int solvedim, int *range, ops_dat a,
                                          ops_dat b, ops_dat c, ops_dat d, ops_dat u,
                                          ops_tridsolver_params *ctx) {
  (void)ctx;
  enqueue(ndim, solvedim, range, a, b, c, d, u, 1);
}
