// TeaLeaf implicit step: FP64 flexible preconditioned CG whose preconditioner is one
// FP32 multigrid V-cycle (no reference counterpart; interface in tealeaf_native.h).
//
// Operator on every level: (A x)_i = m x_i + sum over the four faces f of c_f (x_i - x_j),
// with x = 0 outside the mesh (a boundary face only adds to the diagonal). Level 0 is the
// TeaLeaf operator (m = 1, c = rx*Kx, ry*Ky). Level l+1 aggregates 2 x 2 cells: with the
// piecewise-constant prolongation P, A_{l+1} = P^T A_l P keeps the same form, with mass
// 4 m and each coarse face coefficient the sum of the two fine faces it covers.
// Smoother: red-black Gauss-Seidel, red then black before the coarse correction, black
// then red after it, so the V-cycle is a symmetric operator; the coarse correction is
// scaled by 1.5 (aggregation under-corrects smooth modes). The coarsest level gets a
// fixed number of symmetric sweeps.
//
// Arrays on level l have pitch s = n_l + 2 and one halo layer: cx[i] is the face to the
// left of cell i (the face right of the last cell sits in the halo column), cy[i] the
// face below it. x and b keep zero halos.
//
// The outer PCG runs in FP64 and stops on the reference's rule, sqrt(r.r) below
// eps * initial_residual. beta is the flexible (Polak-Ribiere) form, so an FP32
// preconditioner that is only nearly linear does not spoil the recurrences. Scalars stay
// on the device; an iteration is captured once per parity as a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ops_hip_kernel.h"
#include "ops_hip_rt.h"
#include "tealeaf_native.h"

// V-cycle arithmetic: FP32 (tea_native_mgcg), or FP64 when built with TEA_MG_FP64
// (tea_native_mgcg64; the same preconditioner in double, to measure what FP32 storage buys)
#ifdef TEA_MG_FP64
typedef double vreal;
#define TEA_MGCG_ENTRY tea_native_mgcg64
#else
typedef float vreal;
#define TEA_MGCG_ENTRY tea_native_mgcg
#endif

#define CK(x) ops_hip_check((int)(x), #x, __FILE__, __LINE__)

namespace {

constexpr int BX = 64, BY = 4, NT = BX * BY;
constexpr int FT = 1024;
constexpr int POLL = 8;          // replays of the two-iteration graph between polls
constexpr int COARSEST = 16;     // stop coarsening at this size (or at an odd size)
constexpr int COARSE_SWEEPS = 8; // symmetric sweep pairs on the coarsest level
constexpr vreal ALPHA_C = 1.5;  // coarse-correction scaling

struct Level {
  int n, s;      // cells per side, pitch
  vreal m;       // mass term
  vreal *cx, *cy, *x, *b;
};

struct Rec {
  double rz, rzp, pw, alpha, beta, rr, err, thresh;
  double dinv, dgam;  // correction direction: 1 / d^T A d, and its coefficient d^T r / d^T A d
  int stopped, iters, pad[10];
};

__device__ __forceinline__ unsigned lin_block() {
  return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

// tile of this workgroup: the XCD remap, each XCD's run of tiles walked from its end when
// rev is set (launches alternate, so each starts on what the previous one left in the
// memory-side cache); partials are indexed by tile, so sums do not depend on the direction
__device__ __forceinline__ unsigned tile_id(int rev) {
  return ops_xcd_remap_dir(lin_block(), gridDim.x * gridDim.y * gridDim.z, rev);
}

// 2-D cell of this thread (1-based interior coordinates); false outside
__device__ __forceinline__ bool cell2(int nxc, int nyc, int &x, int &y, int rev = 0) {
  unsigned bx, by, bz;
  ops_block_coords_dir(bx, by, bz, rev);
  x = 1 + (int)(bx * BX + threadIdx.x);
  y = 1 + (int)(by * BY + threadIdx.y);
  return x <= nxc && y <= nyc;
}

dim3 grid2(int nxc, int nyc) {
  return dim3((unsigned)((nxc + BX - 1) / BX), (unsigned)((nyc + BY - 1) / BY));
}

template <typename T> __device__ __forceinline__ void write_partial(T v, T *part, int rev) {
  __shared__ T lds[NT / OPS_WAVE];
  T s = ops_block_reduce<ops_red_inc<T>>(v, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[tile_id(rev)] = s;
}

// fixed-order sum of n partials by one FT-lane workgroup (valid in thread 0)
__device__ __forceinline__ double fold(const double *__restrict__ part, unsigned n) {
  __shared__ double lds[FT / OPS_WAVE];
  double acc[8];
#pragma unroll
  for (int u = 0; u < 8; u++) acc[u] = 0.0;
  for (unsigned base = 0; base < n; base += 8 * FT) {
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const unsigned j = base + u * FT + threadIdx.x;
      if (j < n) acc[u] += part[j];
    }
  }
  const double a = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  return ops_block_reduce<ops_red_inc<double>>(a, lds);
}

// ---------------------------------------------------------------- level set-up

// level 0 coefficients from the FP64 conduction coefficients
__global__ __launch_bounds__(NT) void k_fine_coef(const double *__restrict__ Kx,
                                                  const double *__restrict__ Ky, vreal *cx, vreal *cy,
                                                  int n, double rx, double ry) {
  int x, y;
  // faces: x and y up to n + 1 (the far boundary faces sit in the halo column / row)
  if (!cell2(n + 1, n + 1, x, y)) return;
  const long i = x + (long)y * (n + 2);
  cx[i] = (y <= n) ? (vreal)(rx * Kx[i]) : vreal(0);
  cy[i] = (x <= n) ? (vreal)(ry * Ky[i]) : vreal(0);
}

// coarse faces: the sum of the two fine faces each covers
__global__ __launch_bounds__(NT) void k_coarsen(const vreal *__restrict__ fcx, const vreal *__restrict__ fcy,
                                                int fs, vreal *ccx, vreal *ccy, int nc) {
  int X, Y;
  if (!cell2(nc + 1, nc + 1, X, Y)) return;
  const int cs = nc + 2;
  const long ic = X + (long)Y * cs;
  const int fx = 2 * X - 1, fy = 2 * Y - 1;  // first fine child (1-based)
  ccx[ic] = (Y <= nc) ? fcx[fx + (long)fy * fs] + fcx[fx + (long)(fy + 1) * fs] : vreal(0);
  ccy[ic] = (X <= nc) ? fcy[fx + (long)fy * fs] + fcy[fx + 1 + (long)fy * fs] : vreal(0);
}

// ---------------------------------------------------------------- V-cycle kernels

// One red-black Gauss-Seidel half sweep over the cells of colour `color` ((x + y) % 2).
// Work item (t, y) owns the pair of cells 2t+1, 2t+2 of row y; with init set the level's x
// is taken as zero (the sweep writes b / diag on its colour and zero on the other).
__device__ __forceinline__ void smooth_pair(const Level &L, int color, int init, int t, int y) {
  const int x0 = 2 * t + 1;
  if (x0 > L.n) return;
  const int x = ((x0 + y) % 2 == color) ? x0 : x0 + 1;  // the cell of this colour
  const int xo = (x == x0) ? x0 + 1 : x0;               // the other cell of the pair
  const int s = L.s;
  if (x <= L.n) {
    const long i = x + (long)y * s;
    const vreal c0 = L.cx[i], c1 = L.cx[i + 1], c2 = L.cy[i], c3 = L.cy[i + s];
    const vreal diag = L.m + ((c0 + c1) + (c2 + c3));
    vreal v = L.b[i];
    if (!init) v += (c0 * L.x[i - 1] + c1 * L.x[i + 1]) + (c2 * L.x[i - s] + c3 * L.x[i + s]);
    L.x[i] = v / diag;
  }
  if (init && xo <= L.n) L.x[xo + (long)y * s] = vreal(0);
}

// b_coarse at coarse cell (X, Y) = sum over its 2 x 2 children of (b - A x)
__device__ __forceinline__ void restrict_cell(const Level &F, const Level &C, int X, int Y) {
  const int s = F.s;
  vreal acc = vreal(0);
#pragma unroll
  for (int dy = 0; dy < 2; dy++)
#pragma unroll
    for (int dx = 0; dx < 2; dx++) {
      const long i = (2 * X - 1 + dx) + (long)(2 * Y - 1 + dy) * s;
      const vreal c0 = F.cx[i], c1 = F.cx[i + 1], c2 = F.cy[i], c3 = F.cy[i + s];
      const vreal diag = F.m + ((c0 + c1) + (c2 + c3));
      const vreal ax = diag * F.x[i] - ((c0 * F.x[i - 1] + c1 * F.x[i + 1]) + (c2 * F.x[i - s] + c3 * F.x[i + s]));
      acc += F.b[i] - ax;
    }
  C.b[X + (long)Y * C.s] = acc;
}

// x_fine += ALPHA_C * x_coarse(parent) at fine cell (x, y)
__device__ __forceinline__ void prolong_cell(const Level &F, const Level &C, int x, int y) {
  F.x[x + (long)y * F.s] += ALPHA_C * C.x[(x + 1) / 2 + (long)((y + 1) / 2) * C.s];
}

__global__ __launch_bounds__(NT) void k_restrict(Level F, Level C, const Rec *S, int rev) {
  if (S->stopped) return;
  int X, Y;
  if (!cell2(C.n, C.n, X, Y, rev)) return;
  restrict_cell(F, C, X, Y);
}

// Both half sweeps of one Gauss-Seidel step in one launch: colour `first`, then the other
// colour, over a tile of SW x SH cells per workgroup, from the iterate xin into L.x (a
// different array: a neighbouring workgroup may still read xin's cells that this one
// updates). Phase one computes colour `first` on the tile and a one-cell ring around it
// (the ring's cells are also computed by the neighbouring workgroups, with the same inputs
// and arithmetic, so the values agree) and keeps them in LDS; phase two computes the other
// colour of the tile from those LDS values. The result is the two separate half sweeps'
// (with init, xin is not read: the first colour starts from zero and the other colour's old
// values are never used), from one launch instead of two.
constexpr int SW = 64, SH = 16;
// With xc set (the up-sweep), the iterate is xin plus the coarse correction prolonged on
// the fly: x_j = xin_j + ALPHA_C * xc(parent of j) for interior j, so the prolonged field is
// neither written nor read back (only the other colour's values of it are ever used: a
// Gauss-Seidel update does not read its own cell).
// Two kernels do this step with the same arithmetic per cell: k_smooth2 (a lane per cell of
// the phase's colour) and k_smooth2p (a lane per pair of cells, for the widest levels).

// the value of first-colour cell (x, y) of level L in a smoothing step from the iterate xin
// (init: from zero), the neighbours prolonged-corrected when xc is set
__device__ __forceinline__ vreal first_value(const Level &L, const vreal *__restrict__ xin,
                                             const vreal *__restrict__ xc, int cs, int init, int x, int y) {
  const int s = L.s, n = L.n;
  const long i = x + (long)y * s;
  const vreal c0 = L.cx[i], c1 = L.cx[i + 1], c2 = L.cy[i], c3 = L.cy[i + s];
  const vreal diag = L.m + ((c0 + c1) + (c2 + c3));
  vreal v = L.b[i];
  if (!init) {
    vreal xw = xin[i - 1], xe = xin[i + 1], xs = xin[i - s], xn = xin[i + s];
    if (xc) {
      // the neighbours' prolonged values (halo cells get no correction)
      if (x > 1) xw = xw + ALPHA_C * xc[x / 2 + (long)((y + 1) / 2) * cs];
      if (x < n) xe = xe + ALPHA_C * xc[(x + 2) / 2 + (long)((y + 1) / 2) * cs];
      if (y > 1) xs = xs + ALPHA_C * xc[(x + 1) / 2 + (long)(y / 2) * cs];
      if (y < n) xn = xn + ALPHA_C * xc[(x + 1) / 2 + (long)((y + 2) / 2) * cs];
    }
    v += (c0 * xw + c1 * xe) + (c2 * xs + c3 * xn);
  }
  return v / diag;
}

__global__ __launch_bounds__(NT) void k_smooth2(Level L, const vreal *__restrict__ xin, const vreal *__restrict__ xc,
                                                int cs, int first, int init, const Rec *S, int rev) {
  if (S->stopped) return;
  __shared__ vreal R[SH + 2][SW + 2];
  unsigned bx, by, bz;
  ops_block_coords_dir(bx, by, bz, rev);
  const int x0 = 1 + (int)bx * SW, y0 = 1 + (int)by * SH;
  const int tid = threadIdx.x + BX * threadIdx.y;
  const int s = L.s, n = L.n;
  // lanes take the cells of the phase's own colour only: in window row ly those at lx of
  // parity (first + x0 + y0 + ly) (the first colour, phase 1; 33 of the ring window's 66) and
  // lx - 1 of that parity (the second, phase 2; 32 of the tile's 64). Entries of R at the
  // second colour are never read.
  constexpr int H1 = (SW + 2) / 2, H2 = SW / 2;
  const int par0 = first + x0 + y0;
  for (int e = tid; e < (SH + 2) * H1; e += NT) {
    const int ly = e / H1, lx = 2 * (e - ly * H1) + ((par0 + ly) & 1);
    const int x = x0 - 1 + lx, y = y0 - 1 + ly;
    vreal r = vreal(0);
    if (x >= 1 && x <= n && y >= 1 && y <= n) {
      r = first_value(L, xin, xc, cs, init, x, y);
      if (lx >= 1 && lx <= SW && ly >= 1 && ly <= SH) L.x[x + (long)y * s] = r;
    }
    R[ly][lx] = r;
  }
  __syncthreads();
  for (int e = tid; e < SH * H2; e += NT) {
    const int ly = 1 + e / H2, lx = 1 + 2 * (e - (ly - 1) * H2) + ((par0 + ly) & 1);
    const int x = x0 - 1 + lx, y = y0 - 1 + ly;
    if (x > n || y > n) continue;
    const long i = x + (long)y * s;
    const vreal c0 = L.cx[i], c1 = L.cx[i + 1], c2 = L.cy[i], c3 = L.cy[i + s];
    const vreal diag = L.m + ((c0 + c1) + (c2 + c3));
    vreal v = L.b[i];
    v += (c0 * R[ly][lx - 1] + c1 * R[ly][lx + 1]) + (c2 * R[ly - 1][lx] + c3 * R[ly + 1][lx]);
    L.x[i] = v / diag;
  }
}

// The same step with lanes taking horizontal pairs of cells, one of each colour (the
// window's pairs start at even x, so with an even pitch a pair is one aligned 2-vector), 33
// pairs per window row: phase one loads each array once with pair loads, computes the pair's
// first-colour cell and keeps the other cell's inputs in registers, so phase two reads LDS
// alone and stores the tile's pairs as one 2-vector. Pairs with a cell outside the mesh (and
// every pair when the pitch is odd) take the per-cell path. Used for levels wider than
// pair_min cells: at 4096^2 its launches take 78 and 114 us against 101 and 122 us; at
// 2048^2 it is no faster, and on smaller levels its longer per-workgroup chain makes the
// latency-bound launches slower.
typedef vreal vpair __attribute__((ext_vector_type(2)));
constexpr int NPAIR = (SW + 2) / 2, PPT = ((SH + 2) * NPAIR + NT - 1) / NT;

__global__ __launch_bounds__(NT) void k_smooth2p(Level L, const vreal *__restrict__ xin, const vreal *__restrict__ xc,
                                                int cs, int first, int init, const Rec *S, int rev) {
  if (S->stopped) return;
  __shared__ vreal R[SH + 2][SW + 2];
  unsigned bx, by, bz;
  ops_block_coords_dir(bx, by, bz, rev);
  const int x0 = 1 + (int)bx * SW, y0 = 1 + (int)by * SH;
  const int tid = threadIdx.x + BX * threadIdx.y;
  const int s = L.s, n = L.n;
  // per pair of the lane: the second-colour cell's b and coefficients, the first colour's value
  vreal sb[PPT], sc0[PPT], sc1[PPT], sc2[PPT], sc3[PPT], rf[PPT];
  bool sok[PPT];  // the second-colour cell is in the mesh
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    const int e = tid + k * NT;
    sok[k] = false;
    rf[k] = vreal(0);
    if (e >= (SH + 2) * NPAIR) continue;
    const int ly = e / NPAIR, lx = 2 * (e - ly * NPAIR);
    const int x = x0 - 1 + lx, y = y0 - 1 + ly;  // cells (x, y) and (x + 1, y); x is even
    const int fo = ((x + y) & 1) == first ? 0 : 1;  // which of the two is the first colour
    const bool yin = y >= 1 && y <= n;
    if (yin && x >= 1 && x + 1 <= n && !(s & 1)) {
      const long i = x + (long)y * s;
      const vpair bb = *(const vpair *)(L.b + i), cxa = *(const vpair *)(L.cx + i);
      const vpair cya = *(const vpair *)(L.cy + i), cyb = *(const vpair *)(L.cy + i + s);
      const vreal cx2 = L.cx[i + 2];
      // the left cell's coefficients {cx[i], cx[i + 1], cy[i], cy[i + s]}, the right's shifted by one
      const vreal l0 = cxa.x, l1 = cxa.y, l2 = cya.x, l3 = cyb.x;
      const vreal r0 = cxa.y, r1 = cx2, r2 = cya.y, r3 = cyb.y;
      const vreal c0 = fo ? r0 : l0, c1 = fo ? r1 : l1, c2 = fo ? r2 : l2, c3 = fo ? r3 : l3;
      const vreal diag = L.m + ((c0 + c1) + (c2 + c3));
      vreal v = fo ? bb.y : bb.x;
      if (!init) {
        const int xf = x + fo;
        const vpair xm = *(const vpair *)(xin + i);
        const vpair xsv = *(const vpair *)(xin + i - s), xnv = *(const vpair *)(xin + i + s);
        vreal xw = fo ? xm.x : xin[i - 1], xe = fo ? xin[i + 2] : xm.y;
        vreal xs = fo ? xsv.y : xsv.x, xn = fo ? xnv.y : xnv.x;
        if (xc) {
          if (xf > 1) xw = xw + ALPHA_C * xc[xf / 2 + (long)((y + 1) / 2) * cs];
          if (xf < n) xe = xe + ALPHA_C * xc[(xf + 2) / 2 + (long)((y + 1) / 2) * cs];
          if (y > 1) xs = xs + ALPHA_C * xc[(xf + 1) / 2 + (long)(y / 2) * cs];
          if (y < n) xn = xn + ALPHA_C * xc[(xf + 1) / 2 + (long)((y + 2) / 2) * cs];
        }
        v += (c0 * xw + c1 * xe) + (c2 * xs + c3 * xn);
      }
      rf[k] = v / diag;
      sb[k] = fo ? bb.x : bb.y;
      sc0[k] = fo ? l0 : r0;
      sc1[k] = fo ? l1 : r1;
      sc2[k] = fo ? l2 : r2;
      sc3[k] = fo ? l3 : r3;
      sok[k] = true;
    } else if (yin) {
      // a pair at the mesh edge (or an odd pitch): cell by cell
      const int xf = x + fo, xs2 = x + 1 - fo;
      if (xf >= 1 && xf <= n) rf[k] = first_value(L, xin, xc, cs, init, xf, y);
      if (xs2 >= 1 && xs2 <= n) {
        const long i = xs2 + (long)y * s;
        sb[k] = L.b[i];
        sc0[k] = L.cx[i];
        sc1[k] = L.cx[i + 1];
        sc2[k] = L.cy[i];
        sc3[k] = L.cy[i + s];
        sok[k] = true;
      }
    }
    R[ly][lx + fo] = rf[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PPT; k++) {
    const int e = tid + k * NT;
    if (e >= (SH + 2) * NPAIR) continue;
    const int ly = e / NPAIR, lx = 2 * (e - ly * NPAIR);
    if (ly < 1 || ly > SH) continue;
    const int x = x0 - 1 + lx, y = y0 - 1 + ly;
    const int fo = ((x + y) & 1) == first ? 0 : 1, lf = lx + fo, ls = lx + 1 - fo;
    const bool fin = lf >= 1 && lf <= SW && x + fo >= 1 && x + fo <= n && y <= n;  // tile cell in the mesh
    const bool sin = sok[k] && ls >= 1 && ls <= SW;
    vreal rs = vreal(0);
    if (sin) {
      const vreal diag = L.m + ((sc0[k] + sc1[k]) + (sc2[k] + sc3[k]));
      vreal v = sb[k];
      v += (sc0[k] * R[ly][ls - 1] + sc1[k] * R[ly][ls + 1]) + (sc2[k] * R[ly - 1][ls] + sc3[k] * R[ly + 1][ls]);
      rs = v / diag;
    }
    const long i = x + (long)y * s;
    if (fin && sin && !(s & 1)) {
      vpair o;
      o.x = fo ? rs : rf[k];
      o.y = fo ? rf[k] : rs;
      *(vpair *)(L.x + i) = o;
    } else {
      if (fin) L.x[i + fo] = rf[k];
      if (sin) L.x[i + 1 - fo] = rs;
    }
  }
}

// The levels of at most FOLD_N cells per side, the coarsest included, in one workgroup of
// CT threads: the same passes in the same order as the per-level launches (each pass a
// workgroup-stride loop over its cells, a barrier between passes), so the result is the
// same; one launch replaces the 44 latency-bound launches of those levels (4096^2: levels
// 64, 32 and 16; the solve 0.607 -> 0.554 s; folding the 128^2 level too measured 0.572 s).
constexpr int FOLD_N = 64;
constexpr int CT = 1024;
constexpr int MAXFOLD = 8;
struct Folded {
  Level lv[MAXFOLD];
  int n;  // levels, lv[n - 1] the coarsest
};

__device__ __forceinline__ void smooth_all(const Level &L, int color, int init) {
  const int P = (L.n + 1) / 2, items = P * L.n;
  for (int e = threadIdx.x; e < items; e += CT) smooth_pair(L, color, init, e % P, 1 + e / P);
  __syncthreads();
}

__global__ __launch_bounds__(CT) void k_coarse(Folded f, const Rec *S) {
  if (S->stopped) return;
  for (int l = 0; l + 1 < f.n; l++) {
    smooth_all(f.lv[l], 0, 1);
    smooth_all(f.lv[l], 1, 0);
    const int nc = f.lv[l + 1].n;
    for (int e = threadIdx.x; e < nc * nc; e += CT) restrict_cell(f.lv[l], f.lv[l + 1], 1 + e % nc, 1 + e / nc);
    __syncthreads();
  }
  const Level &c = f.lv[f.n - 1];
  for (int k = 0; k < COARSE_SWEEPS; k++) {
    smooth_all(c, 0, k == 0 ? 1 : 0);
    smooth_all(c, 1, 0);
  }
  for (int k = 0; k < COARSE_SWEEPS; k++) {
    smooth_all(c, 1, 0);
    smooth_all(c, 0, 0);
  }
  for (int l = f.n - 2; l >= 0; l--) {
    const int nf = f.lv[l].n;
    for (int e = threadIdx.x; e < nf * nf; e += CT) prolong_cell(f.lv[l], f.lv[l + 1], 1 + e % nf, 1 + e / nf);
    __syncthreads();
    smooth_all(f.lv[l], 1, 0);
    smooth_all(f.lv[l], 0, 0);
  }
}

// ---------------------------------------------------------------- outer FP64 PCG kernels

struct Fine {
  int n, s;
  double rx, ry;
  int rev = 0;  // tile walk direction of the launch
};

// The outer kernels' lanes take FR cells each, one column of the workgroup's tile of BX x
// (BY FR) cells, rows BY apart (row by BY FR + ty + BY r for r < FR), so a wavefront's loads
// stay contiguous along x; the tiles (and partials, one per tile) are FR times fewer than a
// one-cell-per-lane grid's, which keeps the single-workgroup folds short.
constexpr int FR = 4;
dim3 gridf(int n) { return dim3((unsigned)((n + BX - 1) / BX), (unsigned)((n + BY * FR - 1) / (BY * FR))); }

struct FCol {
  int x, y0;  // the lane's column and first row (1-based interior coordinates)
};
__device__ __forceinline__ FCol fcol(const Fine &g) {
  unsigned bx, by, bz;
  ops_block_coords_dir(bx, by, bz, g.rev);
  return FCol{1 + (int)(bx * BX + threadIdx.x), 1 + (int)(by * BY * FR + threadIdx.y)};
}
// cell r of the lane: false outside the mesh
__device__ __forceinline__ bool fcell(const Fine &g, const FCol &c, int r, long &i) {
  const int y = c.y0 + BY * r;
  i = c.x + (long)y * g.s;
  return c.x <= g.n && y <= g.n;
}

// start: b32 = (vreal) r, r.r partials
__global__ __launch_bounds__(NT) void k_start(const double *__restrict__ r, vreal *__restrict__ b32, Fine g,
                                              double *part) {
  const FCol c = fcol(g);
  double v = 0.0;
#pragma unroll
  for (int q = 0; q < FR; q++) {
    long i;
    if (!fcell(g, c, q, i)) continue;
    const double ri = r[i];
    b32[i] = (vreal)ri;
    v += ri * ri;
  }
  write_partial(v, part, g.rev);
}

__global__ __launch_bounds__(FT) void k_fold_start(const double *part, unsigned n, Rec *S, double thresh) {
  const double rr = fold(part, n);
  if (threadIdx.x == 0) {
    S->rr = rr;
    S->err = sqrt(fabs(rr));
    S->thresh = thresh;
    S->rz = 0.0;
    S->stopped = 0;
    S->iters = 0;
  }
}

// partials of r.z and r.z_prev (z = V-cycle output, z_prev the previous one)
__global__ __launch_bounds__(NT) void k_dots(const double *__restrict__ r, const vreal *__restrict__ z,
                                             const vreal *__restrict__ zp, Fine g, double *part,
                                             const Rec *S) {
  if (S->stopped) return;
  const FCol c = fcol(g);
  double a = 0.0, b = 0.0;
#pragma unroll
  for (int q = 0; q < FR; q++) {
    long i;
    if (!fcell(g, c, q, i)) continue;
    const double ri = r[i];
    a += ri * (double)z[i];
    b += ri * (double)zp[i];
  }
  __shared__ double lds[NT / OPS_WAVE];
  const double sa = ops_block_reduce<ops_red_inc<double>>(a, lds);
  const double sb = ops_block_reduce<ops_red_inc<double>>(b, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    part[2 * tile_id(g.rev)] = sa;
    part[2 * tile_id(g.rev) + 1] = sb;
  }
}

// beta = (r.z - r.z_prev) / rz_old (flexible form; zero on the first iteration)
__global__ __launch_bounds__(FT) void k_fold_dots(const double *part, unsigned n, Rec *S) {
  if (S->stopped) return;
  __shared__ double lds[FT / OPS_WAVE];
  double a0 = 0.0, b0 = 0.0, a1 = 0.0, b1 = 0.0;
  unsigned j = threadIdx.x;
  for (; j + FT < n; j += 2 * FT) {
    a0 += part[2 * j];
    b0 += part[2 * j + 1];
    a1 += part[2 * (j + FT)];
    b1 += part[2 * (j + FT) + 1];
  }
  if (j < n) {
    a0 += part[2 * j];
    b0 += part[2 * j + 1];
  }
  const double rz = ops_block_reduce<ops_red_inc<double>>(a0 + a1, lds);
  const double rzp = ops_block_reduce<ops_red_inc<double>>(b0 + b1, lds);
  if (threadIdx.x == 0) {
    S->beta = (S->iters == 0) ? 0.0 : (rz - rzp) / S->rz;
    S->rz = rz;
    S->rzp = rzp;
  }
}

// p_out = z + beta p_in at the five stencil points, w = A p_out (FP64), p.w partials
__global__ __launch_bounds__(NT) void k_pw(const double *__restrict__ Kx, const double *__restrict__ Ky,
                                           const vreal *__restrict__ z, const double *__restrict__ pin,
                                           double *__restrict__ pout, double *__restrict__ w, Fine g,
                                           double *part, const Rec *S) {
  if (S->stopped) return;
  const double beta = S->beta;
  const FCol c = fcol(g);
  double v = 0.0;
#pragma unroll
  for (int q = 0; q < FR; q++) {
    long i;
    if (!fcell(g, c, q, i)) continue;
    const int s = g.s;
    const double pc = (double)z[i] + beta * pin[i];
    const double pe = (double)z[i + 1] + beta * pin[i + 1];
    const double pwst = (double)z[i - 1] + beta * pin[i - 1];
    const double pn = (double)z[i + s] + beta * pin[i + s];
    const double ps = (double)z[i - s] + beta * pin[i - s];
    const double kx0 = Kx[i], kx1 = Kx[i + 1], ky0 = Ky[i], ky1 = Ky[i + s];
    const double wi = (1.0 + g.ry * (ky1 + ky0) + g.rx * (kx1 + kx0)) * pc - g.ry * (ky1 * pn + ky0 * ps) -
                      g.rx * (kx1 * pe + kx0 * pwst);
    pout[i] = pc;
    w[i] = wi;
    v += wi * pc;
  }
  write_partial(v, part, g.rev);
}

__global__ __launch_bounds__(FT) void k_fold_pw(const double *part, unsigned n, Rec *S) {
  if (S->stopped) return;
  const double pw = fold(part, n);
  if (threadIdx.x == 0) {
    S->pw = pw;
    S->alpha = S->rz / pw;
  }
}

// u += alpha p, r -= alpha w, b32 = (vreal) r for the next V-cycle, r.r partials
__global__ __launch_bounds__(NT) void k_ur(double *__restrict__ u, double *__restrict__ r,
                                           const double *__restrict__ p, const double *__restrict__ w,
                                           vreal *__restrict__ b32, Fine g, double *part, const Rec *S) {
  if (S->stopped) return;
  const double alpha = S->alpha;
  const FCol c = fcol(g);
  double v = 0.0;
#pragma unroll
  for (int q = 0; q < FR; q++) {
    long i;
    if (!fcell(g, c, q, i)) continue;
    u[i] = u[i] + alpha * p[i];
    const double ri = r[i] - alpha * w[i];
    r[i] = ri;
    b32[i] = (vreal)ri;
    v += ri * ri;
  }
  write_partial(v, part, g.rev);
}

__global__ __launch_bounds__(FT) void k_fold_rr(const double *part, unsigned n, Rec *S) {
  if (S->stopped) return;
  const double rr = fold(part, n);
  if (threadIdx.x == 0) {
    S->rr = rr;
    S->iters = S->iters + 1;
    S->err = sqrt(fabs(rr));
    if (S->err < S->thresh) S->stopped = 1;
  }
}

// ---------------------------------------------------------------- preconditioner correction
// z = V(r) + d (d^T r) / (d^T A d) with d the given direction less its mean: the V-cycle
// plus the A-orthogonal projection onto d, a symmetric positive semi-definite term, so the
// preconditioner stays symmetric positive definite (a rank-one coarse space).

// (d - m)^T A (d - m) partials (once per solve)
__global__ __launch_bounds__(NT) void k_defl_dad(const double *__restrict__ Kx, const double *__restrict__ Ky,
                                                 const double *__restrict__ d, double m, Fine g, double *part) {
  const FCol c = fcol(g);
  double v = 0.0;
#pragma unroll
  for (int q = 0; q < FR; q++) {
    long i;
    if (!fcell(g, c, q, i)) continue;
    const int s = g.s, n = g.n;
    const int x = (int)(i % s), y = (int)(i / s);
    // d outside the mesh is zero (not d - m): the operator's boundary faces see x = 0
    auto dv = [&](long j, bool in) { return in ? d[j] - m : 0.0; };
    const double dc = d[i] - m;
    const double de = dv(i + 1, x < n), dw = dv(i - 1, x > 1), dn = dv(i + s, y < n), ds = dv(i - s, y > 1);
    const double kx0 = Kx[i], kx1 = Kx[i + 1], ky0 = Ky[i], ky1 = Ky[i + s];
    const double ad = (1.0 + g.ry * (ky1 + ky0) + g.rx * (kx1 + kx0)) * dc - g.ry * (ky1 * dn + ky0 * ds) -
                      g.rx * (kx1 * de + kx0 * dw);
    v += dc * ad;
  }
  write_partial(v, part, g.rev);
}

__global__ __launch_bounds__(FT) void k_fold_dad(const double *part, unsigned n, Rec *S) {
  const double dad = fold(part, n);
  if (threadIdx.x == 0) S->dinv = dad > 0.0 ? 1.0 / dad : 0.0;
}

// (d - m)^T r partials
__global__ __launch_bounds__(NT) void k_defl_dot(const double *__restrict__ d, double m,
                                                 const double *__restrict__ r, Fine g, double *part,
                                                 const Rec *S) {
  if (S->stopped) return;
  const FCol c = fcol(g);
  double v = 0.0;
#pragma unroll
  for (int q = 0; q < FR; q++) {
    long i;
    if (fcell(g, c, q, i)) v += (d[i] - m) * r[i];
  }
  write_partial(v, part, g.rev);
}

__global__ __launch_bounds__(FT) void k_fold_defl(const double *part, unsigned n, Rec *S) {
  if (S->stopped) return;
  const double dr = fold(part, n);
  if (threadIdx.x == 0) S->dgam = dr * S->dinv;
}

// z += dgam (d - m)
__global__ __launch_bounds__(NT) void k_defl_add(const double *__restrict__ d, double m, vreal *__restrict__ z,
                                                 Fine g, const Rec *S) {
  if (S->stopped) return;
  const FCol c = fcol(g);
#pragma unroll
  for (int q = 0; q < FR; q++) {
    long i;
    if (fcell(g, c, q, i)) z[i] = (vreal)((double)z[i] + S->dgam * (d[i] - m));
  }
}

// ---------------------------------------------------------------- host side

struct MG {
  int n = 0;
  std::vector<Level> lv;
  std::vector<vreal *> xb;  // second iterate array of each level above the fold (k_smooth2)
  void *mem = nullptr;
  size_t cap = 0;
  Rec *rec = nullptr, *host = nullptr;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_n = -1;
  const void *graph_key[11] = {};
  int pair_min = 2049;  // smoothing levels of at least this many cells per side use k_smooth2p
};
MG G;

void drop_graphs() {
  for (auto &e : G.graph)
    if (e) {
      (void)hipGraphExecDestroy(e);
      e = nullptr;
    }
  G.graph_n = -1;
}

char *carve(char *&p, size_t bytes) {
  char *r = p;
  p += (bytes + 255) & ~(size_t)255;
  return r;
}

// V-cycle on the levels; level 0's b must hold the right-hand side, x0 receives z. Levels
// of more than FOLD_N cells per side are swept by one launch per pass; the rest (the
// coarsest included) by one k_coarse launch.
void vcycle(hipStream_t st, vreal *x0, int &dir) {
  const dim3 blk(BX, BY);
  const int L = (int)G.lv.size();
  std::vector<Level> lv = G.lv;
  lv[0].x = x0;
  int l0 = 0;  // the first folded level
  while (l0 < L - 1 && lv[l0].n > FOLD_N) l0++;
  if (L - l0 > MAXFOLD) throw OPSException(OPS_INTERNAL_ERROR, "tealeaf multigrid: too many coarse levels");
  auto s2grid = [](const Level &l) {
    return dim3((unsigned)((l.n + SW - 1) / SW), (unsigned)((l.n + SH - 1) / SH));
  };
  // levels above the fold: the down-sweep's iterate in the level's second array xb[l], the
  // up-sweep's result in its x (x0 on level 0)
  for (int l = 0; l < l0; l++) {
    Level a = lv[l];
    a.x = G.xb[l];
    if (a.n >= G.pair_min)
      hipLaunchKernelGGL(k_smooth2p, s2grid(a), blk, 0, st, a, (const vreal *)nullptr, (const vreal *)nullptr, 0, 0,
                         1, G.rec, dir++ & 1);
    else
      hipLaunchKernelGGL(k_smooth2, s2grid(a), blk, 0, st, a, (const vreal *)nullptr, (const vreal *)nullptr, 0, 0,
                         1, G.rec, dir++ & 1);
    hipLaunchKernelGGL(k_restrict, grid2(lv[l + 1].n, lv[l + 1].n), blk, 0, st, a, lv[l + 1], G.rec, dir++ & 1);
  }
  Folded f;
  f.n = L - l0;
  for (int k = 0; k < f.n; k++) f.lv[k] = lv[l0 + k];
  hipLaunchKernelGGL(k_coarse, dim3(1), dim3(CT), 0, st, f, G.rec);
  // up-sweep: the coarse correction prolonged inside the smoothing (k_smooth2 with xc)
  for (int l = l0 - 1; l >= 0; l--) {
    if (lv[l].n >= G.pair_min)
      hipLaunchKernelGGL(k_smooth2p, s2grid(lv[l]), blk, 0, st, lv[l], (const vreal *)G.xb[l],
                         (const vreal *)lv[l + 1].x, lv[l + 1].s, 1, 0, G.rec, dir++ & 1);
    else
      hipLaunchKernelGGL(k_smooth2, s2grid(lv[l]), blk, 0, st, lv[l], (const vreal *)G.xb[l],
                         (const vreal *)lv[l + 1].x, lv[l + 1].s, 1, 0, G.rec, dir++ & 1);
  }
}

}  // namespace

void TEA_MGCG_ENTRY(ops_dat u, ops_dat u0, ops_dat r, ops_dat p, ops_dat w, ops_dat Kx, ops_dat Ky,
                     const tea_native_solve_params *P, tea_native_result *R) {
  (void)u0;
  if (P->nx != P->ny)
    throw OPSException(OPS_NOT_IMPLEMENTED, "tealeaf multigrid solver: square meshes only");
  ops_dat all[] = {u, u0, r, p, w, Kx, Ky};
  for (ops_dat d : all) {
    ops_memspace ms = OPS_DEVICE;
    ops_dat_get_raw_pointer(d, 0, nullptr, &ms);  // uploads a host-newer copy
    ops_dat_release_raw_data(d, 0, OPS_RW);
  }
  hipStream_t st = (hipStream_t)ops_hip_get_stream();
  // OPS_TEA_PAIR_MIN: the narrowest level smoothed by the pair kernel (default 2049; the
  // kernels give the same values, so this moves time, not results)
  const char *pm = getenv("OPS_TEA_PAIR_MIN");
  G.pair_min = pm && atoi(pm) > 0 ? atoi(pm) : 2049;
  const int n = P->nx;
  const Fine g{n, n + 2, P->rx, P->ry};
  const dim3 blk(BX, BY), gf = gridf(n);
  const unsigned nb = gf.x * gf.y;
  // levels
  std::vector<int> sizes{n};
  while (sizes.back() > COARSEST && sizes.back() % 2 == 0) sizes.push_back(sizes.back() / 2);
  size_t bytes = 4 * (((size_t)nb * 2 * sizeof(double) + 255) & ~(size_t)255);
  const size_t fine64 = (((size_t)(n + 2) * (n + 2) * sizeof(double)) + 255) & ~(size_t)255;
  bytes += fine64;  // second p buffer
  for (int s : sizes)
    bytes += (s > FOLD_N ? 5 : 4) * ((((size_t)(s + 2) * (s + 2) * sizeof(vreal)) + 255) & ~(size_t)255);
  bytes += (((size_t)(n + 2) * (n + 2) * sizeof(vreal)) + 255) & ~(size_t)255;  // second z buffer
  if (bytes > G.cap || G.n != n) {
    ops_hip_device_sync();
    drop_graphs();
    if (G.mem) CK(hipFree(G.mem));
    CK(hipMalloc(&G.mem, bytes));
    G.cap = bytes;
    G.n = n;
  }
  if (!G.rec) {
    CK(hipMalloc((void **)&G.rec, sizeof(Rec)));
    CK(hipHostMalloc((void **)&G.host, sizeof(Rec), hipHostMallocDefault));
  }
  char *q = (char *)G.mem;
  double *part = (double *)carve(q, (size_t)nb * 2 * sizeof(double));
  double *pb = (double *)carve(q, fine64);
  G.lv.clear();
  G.xb.clear();
  vreal mass = 1;
  for (int s : sizes) {
    const size_t lb = (size_t)(s + 2) * (s + 2) * sizeof(vreal);
    Level L{s, s + 2, mass, (vreal *)carve(q, lb), (vreal *)carve(q, lb), (vreal *)carve(q, lb),
            (vreal *)carve(q, lb)};
    G.lv.push_back(L);
    G.xb.push_back(s > FOLD_N ? (vreal *)carve(q, lb) : nullptr);
    mass *= 4.f;
  }
  vreal *z2 = (vreal *)carve(q, (size_t)(n + 2) * (n + 2) * sizeof(vreal));
  vreal *zbuf[2] = {G.lv[0].x, z2};
  // zero every array once (halos stay zero), then build the coefficients
  CK(hipMemsetAsync(G.mem, 0, bytes, st));
  double *u64 = (double *)u->data_d, *r64 = (double *)r->data_d, *pa = (double *)p->data_d,
         *w64 = (double *)w->data_d;
  const double *kx = (const double *)Kx->data_d, *ky = (const double *)Ky->data_d;
  hipLaunchKernelGGL(k_fine_coef, grid2(n + 1, n + 1), blk, 0, st, kx, ky, G.lv[0].cx, G.lv[0].cy, n,
                     P->rx, P->ry);
  for (size_t l = 1; l < G.lv.size(); l++)
    hipLaunchKernelGGL(k_coarsen, grid2(G.lv[l].n + 1, G.lv[l].n + 1), blk, 0, st, G.lv[l - 1].cx,
                       G.lv[l - 1].cy, G.lv[l - 1].s, G.lv[l].cx, G.lv[l].cy, G.lv[l].n);
  // start: r holds u0 - A u; b32 = r
  const double thresh = P->eps * P->initial_residual;
  hipLaunchKernelGGL(k_start, gf, blk, 0, st, r64, G.lv[0].b, g, part);
  hipLaunchKernelGGL(k_fold_start, dim3(1), dim3(FT), 0, st, part, nb, G.rec, thresh);
  const double *defl = P->deflation;
  const double dmean = P->deflation_mean;
  if (defl) {
    hipLaunchKernelGGL(k_defl_dad, gf, blk, 0, st, kx, ky, defl, dmean, g, part);
    hipLaunchKernelGGL(k_fold_dad, dim3(1), dim3(FT), 0, st, part, nb, G.rec);
  }
  ops_hip_check_launch("tealeaf multigrid set-up");
  // one PCG iteration with z in zbuf[par], the previous z in zbuf[1-par], p written to
  // the buffer of parity par
  auto iteration = [&](int par) {
    vreal *z = zbuf[par], *zp = zbuf[1 - par];
    const double *pin = par ? pa : pb;
    double *pout = par ? pb : pa;
    int dir = 0;  // alternating tile walks, launch by launch
    vcycle(st, z, dir);
    Fine gd = g;
    if (defl) {
      hipLaunchKernelGGL(k_defl_dot, gf, blk, 0, st, defl, dmean, r64, g, part, G.rec);
      hipLaunchKernelGGL(k_fold_defl, dim3(1), dim3(FT), 0, st, part, nb, G.rec);
      hipLaunchKernelGGL(k_defl_add, gf, blk, 0, st, defl, dmean, z, g, G.rec);
    }
    gd.rev = dir++ & 1;
    hipLaunchKernelGGL(k_dots, gf, blk, 0, st, r64, z, zp, gd, part, G.rec);
    hipLaunchKernelGGL(k_fold_dots, dim3(1), dim3(FT), 0, st, part, nb, G.rec);
    gd.rev = dir++ & 1;
    hipLaunchKernelGGL(k_pw, gf, blk, 0, st, kx, ky, z, pin, pout, w64, gd, part, G.rec);
    hipLaunchKernelGGL(k_fold_pw, dim3(1), dim3(FT), 0, st, part, nb, G.rec);
    gd.rev = dir++ & 1;
    hipLaunchKernelGGL(k_ur, gf, blk, 0, st, u64, r64, pout, w64, G.lv[0].b, gd, part, G.rec);
    hipLaunchKernelGGL(k_fold_rr, dim3(1), dim3(FT), 0, st, part, nb, G.rec);
  };
  // capture the two parities as graphs (rebuilt when the arrays move)
  // the captured iteration bakes in the arrays, the size and the correction direction with
  // its mean
  double dm_bits = dmean;
  long dm_key;
  memcpy(&dm_key, &dm_bits, sizeof dm_key);
  const void *key[11] = {G.mem, u64, r64, pa, w64, kx, ky, (const void *)(long)n, defl, (const void *)dm_key,
                         (const void *)(long)G.pair_min};
  bool same = G.graph[0] != nullptr;
  for (int k = 0; k < 11 && same; k++) same = key[k] == G.graph_key[k];
  // OPS_TEA_GRAPH=0: launch every iteration's kernels eagerly (the same kernels and order;
  // for tools that do not follow graph launches)
  const char *ge = getenv("OPS_TEA_GRAPH");
  const bool use_graph = !(ge && atoi(ge) == 0);
  if (use_graph && !same) {
    drop_graphs();
    for (int par = 0; par < 2; par++) {
      hipGraph_t gr;
      CK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      iteration(par);
      CK(hipStreamEndCapture(st, &gr));
      CK(hipGraphInstantiate(&G.graph[par], gr, nullptr, nullptr, 0));
      CK(hipGraphDestroy(gr));
    }
    for (int k = 0; k < 11; k++) G.graph_key[k] = key[k];
  }
  // iterations 1, 2, ...: odd iterations use parity 1 (p from pa into pb)
  int issued = 0;
  bool pending = false;
  hipEvent_t ev;
  CK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  while (issued < P->max_iters) {
    for (int k = 0; k < POLL && issued < P->max_iters; k++) {
      if (use_graph)
        CK(hipGraphLaunch(G.graph[(issued + 1) & 1], st));
      else
        iteration((issued + 1) & 1);
      issued++;
    }
    if (pending) {
      CK(hipEventSynchronize(ev));
      if (G.host->stopped) break;
    }
    CK(hipMemcpyAsync(G.host, G.rec, sizeof(Rec), hipMemcpyDeviceToHost, st));
    CK(hipEventRecord(ev, st));
    pending = true;
  }
  CK(hipMemcpyAsync(G.host, G.rec, sizeof(Rec), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  CK(hipEventDestroy(ev));
  const Rec h = *G.host;
  // odd iteration counts leave the last p in the second buffer
  if (h.iters & 1)
    CK(hipMemcpyAsync(pa, pb, (size_t)(n + 2) * (n + 2) * sizeof(double), hipMemcpyDeviceToDevice, st));
  R->iterations = h.iters;
  R->outer = (int)G.lv.size();
  R->error = h.err;
  u->dirty_hd = 2;
  r->dirty_hd = 2;
  p->dirty_hd = 2;
  w->dirty_hd = 2;
}
