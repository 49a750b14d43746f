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
// Layout on MI355X: one lane per line. For y- and z-lines consecutive lanes own
// consecutive x positions, so every row access of a wavefront is one contiguous segment.
// x-lines are contiguous per line, so a one-wavefront workgroup stages a 64-line x 16-row
// tile of a/b/c/d through LDS with coalesced loads before each lane sweeps its line
// (systems are padded to whole 64-line workgroups). The forward-sweep temporaries c', d'
// live in a scratch array laid out row-major over lines (row i of all lines contiguous),
// so they are written and re-read coalesced.
#include <hip/hip_runtime.h>

#include <cstring>
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

template <typename T> __device__ __forceinline__ int find_sys(const Batch<T> &B, long long g) {
  int k = 0;
  while (k + 1 < B.nsys && g >= B.sys[k + 1].line0) k++;
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
struct Scratch {
  char *p = nullptr;
  size_t cap = 0;
  std::vector<char *> retired;
};
Scratch g_scratch;

void *scratch(size_t bytes) {
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

// launches the systems calls[first, first + count) (same solvedim, kind and type)
template <typename T, bool INC>
void launch_chunk(const std::vector<Call> &calls, size_t first, size_t count) {
  Batch<T> B;
  B.nsys = 0;
  B.solvedim = calls[first].solvedim;
  const int sd = B.solvedim;
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
    line += (sd == 0) ? ((S.nlines + XL - 1) / XL) * XL : S.nlines;
    scr += 2 * total;
    B.sys[B.nsys++] = S;
  }
  if (B.nsys == 0) return;
  double pts = 0.0;  // system points of the launch (autotuning compares time per point)
  for (int k = 0; k < B.nsys; k++) pts += (double)B.sys[k].nlines * B.sys[k].n[sd];
  T *scr_p = (T *)scratch(sizeof(T) * (size_t)scr);
  hipStream_t st = (hipStream_t)ops_hip_get_stream_raw();
  if (sd == 0) {
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
        if (c0.inc) launch_chunk<double, true>(calls, first, count);
        else launch_chunk<double, false>(calls, first, count);
      } else {
        if (c0.inc) launch_chunk<float, true>(calls, first, count);
        else launch_chunk<float, false>(calls, first, count);
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
