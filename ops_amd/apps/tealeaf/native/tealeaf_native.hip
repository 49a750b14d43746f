// Device-resident CG solvers for the TeaLeaf implicit step on gfx950 (interface and
// algorithms described in tealeaf_native.h).
//
// Mesh arrays carry one halo layer: pitch sx = nx + 2, interior cells at array indices
// 1..nx, 1..ny. The operator is the reference's calc_w stencil evaluated in the same
// operation order:
//   (1 + ry*(Ky(0,1)+Ky) + rx*(Kx(1,0)+Kx)) * p - ry*(Ky(0,1)*p(0,1) + Ky*p(0,-1))
//                                              - rx*(Kx(1,0)*p(1,0) + Kx*p(-1,0))
//
// One CG iteration k is two launches on a capped grid of G workgroups:
//   W_k : p_k = beta p_{k-1} + r (evaluated at the five stencil points, so p ping-pongs
//         between two buffers), w = A p_k, partial sums of p_k.w
//   UR_k: u += alpha p_k, r -= alpha w, partial sums of r.r
// Every workgroup folds the previous launch's G partials itself, in the same fixed order,
// so all workgroups agree on alpha, beta and the stop test without atomics, device-scope
// fences or extra launches. r.r partials alternate between two arrays (W_k needs r.r of
// iterations k-1 and k-2). Workgroup 0 records the state (stop iteration, residual) for
// the host, which polls it every POLL iterations while the next batch is already queued;
// launches issued after the stop return at once.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "ops_hip_kernel.h"
#include "ops_hip_rt.h"
#include "tealeaf_native.h"

#define CK(x) ops_hip_check((int)(x), #x, __FILE__, __LINE__)

namespace {

constexpr int BX = 64, BY = 4, NT = BX * BY;
constexpr int POLL = 32;
constexpr int GMAX = OPS_HIP_RED_BLOCKS;

struct State {
  double err;  // sqrt(|r.r|) of the last folded residual
  int stop;    // iteration whose W launch saw convergence (INT_MAX while running)
  int iters;   // completed iterations (UR launches that ran)
  int pad[12];
};

struct Geo {
  int nx, ny, sx;
  unsigned tx, ntiles;  // tiles of BX x BY cells
};

template <typename T> struct Parts {
  T *rr[2];  // r.r partials, by iteration parity
  T *pw;     // p.w partials
};

// sum of n partials, in the same order in every workgroup; the value is returned to
// every thread
template <typename T> __device__ __forceinline__ T fold(const T *__restrict__ part, unsigned n) {
  __shared__ T lds[NT / OPS_WAVE];
  __shared__ T bcast;
  const int tid = threadIdx.x + BX * threadIdx.y;
  T a0 = T(0), a1 = T(0);
  unsigned i = tid;
  for (; i + NT < n; i += 2 * NT) {
    a0 += part[i];
    a1 += part[i + NT];
  }
  if (i < n) a0 += part[i];
  T s = ops_block_reduce<ops_red_inc<T>>(a0 + a1, lds);
  if (tid == 0) bcast = s;
  __syncthreads();
  s = bcast;
  __syncthreads();
  return s;
}

template <typename T> __device__ __forceinline__ void write_partial(T v, T *part) {
  __shared__ T lds[NT / OPS_WAVE];
  T s = ops_block_reduce<ops_red_inc<T>>(v, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) part[blockIdx.x] = s;
}

template <typename T, typename K>
__device__ __forceinline__ T apply_A(const K *Kx, const K *Ky, const T *p, long i, int sx, T rx, T ry) {
  const T kx0 = (T)Kx[i], kx1 = (T)Kx[i + 1], ky0 = (T)Ky[i], ky1 = (T)Ky[i + sx];
  return (T(1.0) + ry * (ky1 + ky0) + rx * (kx1 + kx0)) * p[i] - ry * (ky1 * p[i + sx] + ky0 * p[i - sx]) -
         rx * (kx1 * p[i + 1] + kx0 * p[i - 1]);
}

// array index of this thread's cell in tile t (false outside the interior)
__device__ __forceinline__ bool tile_cell(const Geo &g, unsigned t, long &i) {
  const int x = 1 + (int)((t % g.tx) * BX + threadIdx.x);
  const int y = 1 + (int)((t / g.tx) * BY + threadIdx.y);
  i = x + (long)y * g.sx;
  return x <= g.nx && y <= g.ny;
}

__device__ __forceinline__ bool lead() {
  return blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0;
}

// ---------------------------------------------------------------- CG kernels (T = double or float)

// p = r, r.r partials into rr[0]
template <typename T>
__global__ __launch_bounds__(NT) void k_cg_start(const T *__restrict__ r, T *__restrict__ p, Geo g,
                                                 Parts<T> P, State *S) {
  unsigned t, t1, dt;
  ops_tile_span(g.ntiles, t, t1, dt);
  T v = T(0);
  for (; t < t1; t += dt) {
    long i;
    if (tile_cell(g, t, i)) {
      const T ri = r[i];
      p[i] = ri;
      v += ri * ri;
    }
  }
  write_partial(v, P.rr[0]);
  if (lead()) {
    S->stop = INT_MAX;
    S->iters = 0;
  }
}

// W_k: p_out = beta p_in + r at the five stencil points, w = A p_out, p_out.w partials.
// The stop test on r.r of iteration k-1 runs here (from k = 2): the iteration that
// converged still updates p, as the reference's loop does.
template <typename T, typename K>
__global__ __launch_bounds__(NT) void k_cg_w(const K *__restrict__ Kx, const K *__restrict__ Ky,
                                             const T *__restrict__ pin, T *__restrict__ pout,
                                             const T *__restrict__ r, T *__restrict__ w, Geo g, T rx,
                                             T ry, Parts<T> P, State *S, int k, double thresh) {
  if (S->stop < k) return;
  const unsigned G = gridDim.x;
  T beta = T(0);
  bool conv = false;
  if (k >= 2) {
    const T rr1 = fold(P.rr[(k - 1) & 1], G), rr2 = fold(P.rr[k & 1], G);
    beta = rr1 / rr2;
    const double err = sqrt(fabs((double)rr1));
    conv = err < thresh;
    if (lead()) {
      S->err = err;
      if (conv) S->stop = k;
    }
  }
  const T one = T(1.0);
  const int sx = g.sx;
  unsigned t, t1, dt;
  ops_tile_span(g.ntiles, t, t1, dt);
  T v = T(0);
  for (; t < t1; t += dt) {
    long i;
    if (!tile_cell(g, t, i)) continue;
    const T pc = beta * pin[i] + one * r[i];
    pout[i] = pc;
    if (conv) continue;
    // neighbours of p_k from p_{k-1} and r (zero halos give zero neighbours at the rim)
    const T pe = beta * pin[i + 1] + one * r[i + 1];
    const T pwst = beta * pin[i - 1] + one * r[i - 1];
    const T pn = beta * pin[i + sx] + one * r[i + sx];
    const T ps = beta * pin[i - sx] + one * r[i - sx];
    const T kx0 = (T)Kx[i], kx1 = (T)Kx[i + 1], ky0 = (T)Ky[i], ky1 = (T)Ky[i + sx];
    const T wi = (one + ry * (ky1 + ky0) + rx * (kx1 + kx0)) * pc - ry * (ky1 * pn + ky0 * ps) -
                 rx * (kx1 * pe + kx0 * pwst);
    w[i] = wi;
    v += wi * pc;
  }
  if (!conv) write_partial(v, P.pw);
}

// UR_k: alpha = rr_{k-1} / pw_k; u += alpha p; r -= alpha w; r.r partials into rr[k & 1]
template <typename T, typename U>
__global__ __launch_bounds__(NT) void k_cg_ur(U *__restrict__ u, T *__restrict__ r, const T *__restrict__ p,
                                              const T *__restrict__ w, Geo g, Parts<T> P, State *S,
                                              int k) {
  if (S->stop <= k) return;
  const unsigned G = gridDim.x;
  const T rr = fold(P.rr[(k - 1) & 1], G);
  const T alpha = rr / fold(P.pw, G);
  unsigned t, t1, dt;
  ops_tile_span(g.ntiles, t, t1, dt);
  T v = T(0);
  for (; t < t1; t += dt) {
    long i;
    if (!tile_cell(g, t, i)) continue;
    u[i] = u[i] + (U)(alpha * p[i]);
    const T ri = r[i] - alpha * w[i];
    r[i] = ri;
    v += ri * ri;
  }
  write_partial(v, P.rr[k & 1]);
  if (lead()) S->iters = k;
}

// residual of iteration `last` when no W launch has folded it
template <typename T>
__global__ __launch_bounds__(NT) void k_cg_final(Parts<T> P, unsigned G, State *S, int last) {
  if (S->stop != INT_MAX) return;
  const T rr = fold(P.rr[last & 1], G);
  if (threadIdx.x == 0 && threadIdx.y == 0) S->err = sqrt(fabs((double)rr));
}

// ---------------------------------------------------------------- refinement kernels

__global__ __launch_bounds__(NT) void k_to_f32(const double *__restrict__ a, float *__restrict__ b,
                                               long n) {
  long i = (long)blockIdx.x * NT + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// r32 = r64 * scale, p32 = r32, d32 = 0, r32.r32 partials into rr[0]
__global__ __launch_bounds__(NT) void k_inner_start(const double *__restrict__ r64, float *__restrict__ r32,
                                                    float *__restrict__ p32, float *__restrict__ d32,
                                                    Geo g, double scale, Parts<float> P, State *S) {
  unsigned t, t1, dt;
  ops_tile_span(g.ntiles, t, t1, dt);
  float v = 0.f;
  for (; t < t1; t += dt) {
    long i;
    if (!tile_cell(g, t, i)) continue;
    const float ri = (float)(r64[i] * scale);
    r32[i] = ri;
    p32[i] = ri;
    d32[i] = 0.f;
    v += ri * ri;
  }
  write_partial(v, P.rr[0]);
  if (lead()) {
    S->stop = INT_MAX;
    S->iters = 0;
  }
}

// u64 += d32 / scale
__global__ __launch_bounds__(NT) void k_add_corr(double *__restrict__ u, const float *__restrict__ d32, Geo g,
                                                 double inv_scale) {
  unsigned t, t1, dt;
  ops_tile_span(g.ntiles, t, t1, dt);
  for (; t < t1; t += dt) {
    long i;
    if (tile_cell(g, t, i)) u[i] = u[i] + (double)d32[i] * inv_scale;
  }
}

// r = u0 - A u (the reference's residual kernel) in FP64, r.r partials
__global__ __launch_bounds__(NT) void k_residual64(const double *__restrict__ Kx, const double *__restrict__ Ky,
                                                   const double *__restrict__ u, const double *__restrict__ u0,
                                                   double *__restrict__ r, Geo g, double rx, double ry,
                                                   double *part) {
  unsigned t, t1, dt;
  ops_tile_span(g.ntiles, t, t1, dt);
  double v = 0.0;
  for (; t < t1; t += dt) {
    long i;
    if (!tile_cell(g, t, i)) continue;
    const double smvp = apply_A<double, double>(Kx, Ky, u, i, g.sx, rx, ry);
    const double ri = u0[i] - smvp;
    r[i] = ri;
    v += ri * ri;
  }
  write_partial(v, part);
}

__global__ __launch_bounds__(NT) void k_fold_to(const double *part, unsigned n, double *out) {
  const double s = fold(part, n);
  if (threadIdx.x == 0 && threadIdx.y == 0) *out = s;
}

// ---------------------------------------------------------------- host side

struct Work {
  void *buf = nullptr;
  size_t cap = 0;
  State *state = nullptr;  // device records: [0] FP64 CG, [1] FP32 inner CG
  double *dres = nullptr;  // refinement residual r.r
  State *pinned = nullptr; // host copies of the records, then r.r
  hipEvent_t ev = nullptr;
};
Work W;

void *work(size_t bytes) {
  if (bytes > W.cap) {
    ops_hip_device_sync();
    if (W.buf) CK(hipFree(W.buf));
    CK(hipMalloc(&W.buf, bytes));
    W.cap = bytes;
  }
  if (!W.state) {
    CK(hipMalloc((void **)&W.state, 4096));
    W.dres = (double *)((char *)W.state + 1024);
    CK(hipHostMalloc((void **)&W.pinned, 4096, hipHostMallocDefault));
    CK(hipEventCreateWithFlags(&W.ev, hipEventDisableTiming));
  }
  return W.buf;
}

void sync_device_copy(ops_dat d) {
  ops_memspace ms = OPS_DEVICE;
  ops_dat_get_raw_pointer(d, 0, nullptr, &ms);  // uploads a host-newer copy
  ops_dat_release_raw_data(d, 0, OPS_RW);
}

// array position 0 of a dat declared with base 1 and halo depth 1
template <typename T> T *ix0(ops_dat d) { return (T *)(d->data_d); }

// Runs CG iterations k = 1..max_iters until a W launch sees convergence and returns the
// host copy of the record. *last_p is the iteration whose W launch wrote the final p.
template <typename T, typename K, typename U>
State cg_loop(const K *Kx, const K *Ky, U *u, T *r, T *pa, T *pb, T *w, const Geo &g, T rx, T ry,
              Parts<T> P, State *S, int max_iters, double thresh, unsigned G, hipStream_t st,
              int *last_p) {
  dim3 blk(BX, BY), grid(G);
  State *host = W.pinned + (S - W.state);
  int k = 1;
  bool pending = false;
  State h{};
  h.stop = INT_MAX;
  while (k <= max_iters) {
    const int end = (k + POLL - 1 < max_iters) ? k + POLL - 1 : max_iters;
    for (; k <= end; k++) {
      const T *pin = (k & 1) ? pa : pb;
      T *pout = (k & 1) ? pb : pa;
      hipLaunchKernelGGL((k_cg_w<T, K>), grid, blk, 0, st, Kx, Ky, pin, pout, r, w, g, rx, ry, P, S,
                         k, thresh);
      hipLaunchKernelGGL((k_cg_ur<T, U>), grid, blk, 0, st, u, r, pout, w, g, P, S, k);
    }
    ops_hip_check_launch("tealeaf cg");
    // the previous batch's record is in host memory once its event has completed
    if (pending) {
      CK(hipEventSynchronize(W.ev));
      h = *host;
      if (h.stop != INT_MAX) break;
    }
    CK(hipMemcpyAsync(host, S, sizeof(State), hipMemcpyDeviceToHost, st));
    CK(hipEventRecord(W.ev, st));
    pending = true;
  }
  if (h.stop == INT_MAX && k == max_iters + 1) {
    // one more W launch tests the last issued iteration; if it did not converge either,
    // the final kernel folds its residual
    const T *pin = (k & 1) ? pa : pb;
    T *pout = (k & 1) ? pb : pa;
    hipLaunchKernelGGL((k_cg_w<T, K>), grid, blk, 0, st, Kx, Ky, pin, pout, r, w, g, rx, ry, P, S, k,
                       thresh);
    hipLaunchKernelGGL((k_cg_final<T>), dim3(1), blk, 0, st, P, G, S, k - 1);
    ops_hip_check_launch("tealeaf cg final");
  }
  CK(hipMemcpyAsync(host, S, sizeof(State), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  h = *host;
  *last_p = (h.stop != INT_MAX) ? h.stop : max_iters + 1;
  return h;
}

}  // namespace

void tea_native_cg(ops_dat u, ops_dat u0, ops_dat r, ops_dat p, ops_dat w, ops_dat z, ops_dat Kx,
                   ops_dat Ky, const tea_native_solve_params *P, tea_native_result *R) {
  (void)z;
  hipStream_t st = (hipStream_t)ops_hip_get_stream();
  ops_dat all[] = {u, u0, r, p, w, Kx, Ky};
  for (ops_dat d : all) sync_device_copy(d);
  Geo g;
  g.nx = P->nx;
  g.ny = P->ny;
  g.sx = P->nx + 2;
  g.tx = (unsigned)((P->nx + BX - 1) / BX);
  g.ntiles = g.tx * (unsigned)((P->ny + BY - 1) / BY);
  const unsigned G = g.ntiles < (unsigned)GMAX ? g.ntiles : (unsigned)GMAX;
  dim3 blk(BX, BY), grid(G);
  const long n = (long)(P->nx + 2) * (P->ny + 2);
  double *u64 = ix0<double>(u), *u064 = ix0<double>(u0), *r64 = ix0<double>(r);
  double *p64 = ix0<double>(p), *w64 = ix0<double>(w);
  const double *kx = ix0<double>(Kx), *ky = ix0<double>(Ky);
  const double thresh = P->eps * P->initial_residual;
  // scratch: 4 partial arrays, then a second FP64 p buffer (solver 1) or the FP32 arrays
  // Kx, Ky, r, pa, pb, w, d (solver 2)
  const size_t part_bytes = ((size_t)GMAX * sizeof(double) + 255) & ~(size_t)255;
  const size_t f64 = ((size_t)n * sizeof(double) + 255) & ~(size_t)255;
  const size_t f32 = ((size_t)n * sizeof(float) + 255) & ~(size_t)255;
  char *base = (char *)work(4 * part_bytes + (P->solver == 2 ? 7 * f32 : f64));
  char *arr = base + 4 * part_bytes;
  State *S64 = W.state, *S32 = W.state + 1;
  if (P->solver == 1) {
    Parts<double> Pd{{(double *)base, (double *)(base + part_bytes)}, (double *)(base + 2 * part_bytes)};
    double *pb = (double *)arr;
    CK(hipMemsetAsync(pb, 0, f64, st));  // zero halos of the second p buffer
    hipLaunchKernelGGL((k_cg_start<double>), grid, blk, 0, st, r64, p64, g, Pd, S64);
    int last_p = 0;
    State h = cg_loop<double, double, double>(kx, ky, u64, r64, p64, pb, w64, g, P->rx, P->ry, Pd,
                                              S64, P->max_iters, thresh, G, st, &last_p);
    // odd iterations write the second buffer: leave the final p in the dat
    if (last_p & 1)
      CK(hipMemcpyAsync(p64, pb, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
    R->iterations = h.iters;
    R->outer = 1;
    R->error = h.err;
  } else {
    Parts<float> Pf{{(float *)base, (float *)(base + part_bytes)}, (float *)(base + 2 * part_bytes)};
    double *part64 = (double *)(base + 3 * part_bytes);
    float *kx32 = (float *)arr, *ky32 = (float *)(arr + f32), *r32 = (float *)(arr + 2 * f32);
    float *pa32 = (float *)(arr + 3 * f32), *pb32 = (float *)(arr + 4 * f32);
    float *w32 = (float *)(arr + 5 * f32), *d32 = (float *)(arr + 6 * f32);
    const unsigned lin = (unsigned)((n + NT - 1) / NT);
    hipLaunchKernelGGL(k_to_f32, dim3(lin), dim3(NT), 0, st, kx, kx32, n);
    hipLaunchKernelGGL(k_to_f32, dim3(lin), dim3(NT), 0, st, ky, ky32, n);
    // halos of the FP32 work arrays (r, pa, pb, w, d) must be zero: the operator reads them
    CK(hipMemsetAsync(r32, 0, 5 * f32, st));
    const char *tau_env = getenv("OPS_TEA_MIXED_TAU");
    const double tau = tau_env ? atof(tau_env) : 1e-4;
    int total = 0, outer = 0;
    double rr = P->initial_residual * P->initial_residual, err = P->initial_residual;
    // the caller's r holds u0 - A u in FP64
    while (err >= thresh && total < P->max_iters) {
      outer++;
      // scale the correction equation so the FP32 residual is O(1)
      const double scale = 1.0 / sqrt(rr);
      double inner_thresh = tau;  // relative to the scaled initial residual (= 1)
      const double floor_rel = 0.5 * thresh * scale;
      if (floor_rel > inner_thresh) inner_thresh = floor_rel;
      hipLaunchKernelGGL(k_inner_start, grid, blk, 0, st, r64, r32, pa32, d32, g, scale, Pf, S32);
      int last_p = 0;
      State h = cg_loop<float, float, float>(kx32, ky32, d32, r32, pa32, pb32, w32, g, (float)P->rx,
                                             (float)P->ry, Pf, S32, P->max_iters - total,
                                             inner_thresh, G, st, &last_p);
      total += h.iters;
      hipLaunchKernelGGL(k_add_corr, grid, blk, 0, st, u64, d32, g, 1.0 / scale);
      hipLaunchKernelGGL(k_residual64, grid, blk, 0, st, kx, ky, u64, u064, r64, g, P->rx, P->ry,
                         part64);
      hipLaunchKernelGGL(k_fold_to, dim3(1), blk, 0, st, part64, G, W.dres);
      ops_hip_check_launch("tealeaf refinement");
      double *rr_host = (double *)(W.pinned + 2);
      CK(hipMemcpyAsync(rr_host, W.dres, sizeof(double), hipMemcpyDeviceToHost, st));
      CK(hipStreamSynchronize(st));
      rr = *rr_host;
      err = sqrt(fabs(rr));
      if (h.iters == 0) break;  // no progress possible
    }
    R->iterations = total;
    R->outer = outer;
    R->error = err;
  }
  // the solver wrote u, r, p and w on the device
  u->dirty_hd = 2;
  r->dirty_hd = 2;
  p->dirty_hd = 2;
  w->dirty_hd = 2;
}
