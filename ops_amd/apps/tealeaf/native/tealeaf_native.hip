// Device-resident CG solvers for the TeaLeaf implicit step on gfx950 (interface and
// algorithms described in tealeaf_native.h).
//
// Mesh arrays carry one halo layer: pitch sx = nx + 2, interior cells at array indices
// 1..nx, 1..ny. The operator is the reference's calc_w stencil evaluated in the same
// operation order:
//   (1 + ry*(Ky(0,1)+Ky) + rx*(Kx(1,0)+Kx)) * p - ry*(Ky(0,1)*p(0,1) + Ky*p(0,-1))
//                                              - rx*(Kx(1,0)*p(1,0) + Kx*p(-1,0))
// Scalars (rro, pw, alpha, beta, the stop flag) live in a small device record updated by
// the last workgroup of each reduction, so an iteration issues three launches and no host
// synchronisation; the host polls the stop flag every POLL iterations, and kernels queued
// after convergence return at once.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "ops_hip_kernel.h"
#include "ops_hip_rt.h"
#include "tealeaf_native.h"

#define CK(x) ops_hip_check((int)(x), #x, __FILE__, __LINE__)

namespace {

constexpr int BX = 64, BY = 4, NT = BX * BY;
constexpr int POLL = 16;

template <typename T> struct Scal {
  T rro, pw, alpha, rrn, beta, rr0;
  double err, thresh;
  int done, done_p, iters, pad;
};

struct Geo {
  int nx, ny, sx;
};

__device__ __forceinline__ bool cell(const Geo &g, int &i) {
  unsigned bx, by, bz;
  ops_block_coords(bx, by, bz);
  const int x = 1 + (int)(bx * BX + threadIdx.x);
  const int y = 1 + (int)(by * BY + threadIdx.y);
  i = x + y * g.sx;
  return x <= g.nx && y <= g.ny;
}

// workgroup sum, then grid sum by the last workgroup (fixed order); returns true in
// thread 0 of the last workgroup, with the total in *tot
template <typename T>
__device__ __forceinline__ bool grid_sum(T v, T *part, unsigned *count, T *tot) {
  __shared__ T lds[NT / OPS_WAVE];
  __shared__ int last;
  const int tid = threadIdx.x + BX * threadIdx.y;
  const unsigned nb = gridDim.x * gridDim.y;
  const unsigned b = blockIdx.x + gridDim.x * blockIdx.y;
  T s = ops_block_reduce<ops_red_inc<T>>(v, lds);
  if (tid == 0) {
    part[b] = s;
    __threadfence();
    last = (atomicAdd(count, 1u) == nb - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  T acc = T(0);
  for (unsigned k = tid; k < nb; k += NT) acc += part[k];
  T t = ops_block_reduce<ops_red_inc<T>>(acc, lds);
  if (tid == 0) {
    *count = 0u;
    *tot = t;
    return true;
  }
  return false;
}

template <typename T, typename K>
__device__ __forceinline__ T apply_A(const K *Kx, const K *Ky, const T *p, int i, int sx, T rx, T ry) {
  const T kx0 = (T)Kx[i], kx1 = (T)Kx[i + 1], ky0 = (T)Ky[i], ky1 = (T)Ky[i + sx];
  return (T(1.0) + ry * (ky1 + ky0) + rx * (kx1 + kx0)) * p[i] - ry * (ky1 * p[i + sx] + ky0 * p[i - sx]) -
         rx * (kx1 * p[i + 1] + kx0 * p[i - 1]);
}

// ---------------------------------------------------------------- CG kernels (T = double or float)

// p = r, rro = r.r ; resets the record
template <typename T>
__global__ __launch_bounds__(NT) void k_cg_start(const T *__restrict__ r, T *__restrict__ p, Geo g,
                                                 Scal<T> *S, T *part, unsigned *cnt, double thresh) {
  int i;
  T v = T(0);
  if (cell(g, i)) {
    T ri = r[i];
    p[i] = ri;
    v = ri * ri;
  }
  T tot;
  if (grid_sum(v, part, cnt, &tot)) {
    S->rro = tot;
    S->rr0 = tot;
    S->done = 0;
    S->done_p = 0;
    S->iters = 0;
    S->thresh = thresh;
    S->err = sqrt(fabs((double)tot));
  }
}

// w = A p, pw = p.w, alpha = rro / pw
template <typename T, typename K>
__global__ __launch_bounds__(NT) void k_cg_w(const K *__restrict__ Kx, const K *__restrict__ Ky,
                                             const T *__restrict__ p, T *__restrict__ w, Geo g, T rx,
                                             T ry, Scal<T> *S, T *part, unsigned *cnt) {
  if (S->done) return;
  int i;
  T v = T(0);
  if (cell(g, i)) {
    T wi = apply_A<T, K>(Kx, Ky, p, i, g.sx, rx, ry);
    w[i] = wi;
    v = wi * p[i];
  }
  T tot;
  if (grid_sum(v, part, cnt, &tot)) {
    S->pw = tot;
    S->alpha = S->rro / tot;
  }
}

// u += alpha p ; r -= alpha w ; rrn = r.r ; beta = rrn / rro ; stop test on sqrt(|rrn|)
template <typename T, typename U>
__global__ __launch_bounds__(NT) void k_cg_ur(U *__restrict__ u, T *__restrict__ r, const T *__restrict__ p,
                                              const T *__restrict__ w, Geo g, Scal<T> *S, T *part,
                                              unsigned *cnt) {
  if (S->done) return;
  const T alpha = S->alpha;
  int i;
  T v = T(0);
  if (cell(g, i)) {
    u[i] = u[i] + (U)(alpha * p[i]);
    T ri = r[i] - alpha * w[i];
    r[i] = ri;
    v = ri * ri;
  }
  T tot;
  if (grid_sum(v, part, cnt, &tot)) {
    S->rrn = tot;
    S->beta = tot / S->rro;
    S->rro = tot;
    S->iters = S->iters + 1;
    S->err = sqrt(fabs((double)tot));
    if (S->err < S->thresh) S->done = 1;
  }
}

// p = beta p + r (the reference's axpby with alpha = beta, beta = 1)
template <typename T>
__global__ __launch_bounds__(NT) void k_cg_p(T *__restrict__ p, const T *__restrict__ r, Geo g,
                                             Scal<T> *S) {
  if (S->done_p) return;
  const T beta = S->beta, one = T(1.0);
  int i;
  if (cell(g, i)) p[i] = beta * p[i] + one * r[i];
  // the iteration that converged still updates p, later ones do not
  if (threadIdx.x == 0 && threadIdx.y == 0 && blockIdx.x == 0 && blockIdx.y == 0 && S->done)
    S->done_p = 1;
}

// ---------------------------------------------------------------- refinement kernels

__global__ __launch_bounds__(NT) void k_to_f32(const double *__restrict__ a, float *__restrict__ b,
                                               long n) {
  long i = (long)blockIdx.x * NT + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// r32 = r64 * scale, p32 = r32, d32 = 0, rr = r32.r32
__global__ __launch_bounds__(NT) void k_inner_start(const double *__restrict__ r64, float *__restrict__ r32,
                                                    float *__restrict__ p32, float *__restrict__ d32,
                                                    Geo g, double scale, Scal<float> *S, float *part,
                                                    unsigned *cnt, double thresh) {
  int i;
  float v = 0.f;
  if (cell(g, i)) {
    float ri = (float)(r64[i] * scale);
    r32[i] = ri;
    p32[i] = ri;
    d32[i] = 0.f;
    v = ri * ri;
  }
  float tot;
  if (grid_sum(v, part, cnt, &tot)) {
    S->rro = tot;
    S->rr0 = tot;
    S->done = 0;
    S->done_p = 0;
    S->iters = 0;
    S->thresh = thresh;
    S->err = sqrt((double)tot);
  }
}

// u64 += d32 / scale
__global__ __launch_bounds__(NT) void k_add_corr(double *__restrict__ u, const float *__restrict__ d32, Geo g,
                                                 double inv_scale) {
  int i;
  if (cell(g, i)) u[i] = u[i] + (double)d32[i] * inv_scale;
}

// r = u0 - A u (the reference's residual kernel), rr = r.r in FP64
__global__ __launch_bounds__(NT) void k_residual64(const double *__restrict__ Kx, const double *__restrict__ Ky,
                                                   const double *__restrict__ u, const double *__restrict__ u0,
                                                   double *__restrict__ r, Geo g, double rx, double ry,
                                                   double *out, double *part, unsigned *cnt) {
  int i;
  double v = 0.0;
  if (cell(g, i)) {
    double smvp = apply_A<double, double>(Kx, Ky, u, i, g.sx, rx, ry);
    double ri = u0[i] - smvp;
    r[i] = ri;
    v = ri * ri;
  }
  double tot;
  if (grid_sum(v, part, cnt, &tot)) *out = tot;
}

// ---------------------------------------------------------------- host side

struct Work {
  void *buf = nullptr;
  size_t cap = 0;
  void *scal = nullptr;     // Scal<double> + Scal<float> + doubles
  unsigned *cnt = nullptr;  // reduction counters
  void *pinned = nullptr;
};
Work W;

void *work(size_t bytes) {
  if (bytes > W.cap) {
    ops_hip_device_sync();
    if (W.buf) CK(hipFree(W.buf));
    CK(hipMalloc(&W.buf, bytes));
    W.cap = bytes;
  }
  if (!W.scal) {
    CK(hipMalloc(&W.scal, 4096));
    CK(hipMalloc((void **)&W.cnt, 64 * sizeof(unsigned)));
    CK(hipMemset(W.cnt, 0, 64 * sizeof(unsigned)));
    CK(hipHostMalloc(&W.pinned, 4096, hipHostMallocDefault));
  }
  return W.buf;
}

char *dev(ops_dat d) {
  ops_memspace ms = OPS_DEVICE;
  char *p = ops_dat_get_raw_pointer(d, 0, nullptr, &ms);
  ops_dat_release_raw_data(d, 0, OPS_RW);
  // the raw pointer addresses loop index 0; the halo-inclusive array starts one row and
  // one column earlier (base 1, d_m -1 gives index 0 at array position 0)
  return p;
}

template <typename T> T *ix0(ops_dat d) {
  // array position 0 of a dat declared with base 1 and halo depth 1
  return (T *)(d->data_d);
}

template <typename T> Scal<T> read_scal(Scal<T> *S, hipStream_t st) {
  CK(hipMemcpyAsync(W.pinned, S, sizeof(Scal<T>), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  return *(Scal<T> *)W.pinned;
}

// runs CG iterations until the record's stop flag is set or max_iters are issued
template <typename T, typename K, typename U>
int cg_loop(const K *Kx, const K *Ky, U *u, T *r, T *p, T *w, Geo g, T rx, T ry, Scal<T> *S, T *part,
            unsigned *cnt, int max_iters, dim3 grid, dim3 blk, hipStream_t st) {
  int issued = 0;
  while (issued < max_iters) {
    int batch = max_iters - issued < POLL ? max_iters - issued : POLL;
    for (int k = 0; k < batch; k++) {
      hipLaunchKernelGGL((k_cg_w<T, K>), grid, blk, 0, st, Kx, Ky, p, w, g, rx, ry, S, part, cnt);
      hipLaunchKernelGGL((k_cg_ur<T, U>), grid, blk, 0, st, u, r, p, w, g, S, part, cnt + 1);
      hipLaunchKernelGGL((k_cg_p<T>), grid, blk, 0, st, p, r, g, S);
    }
    ops_hip_check_launch("tealeaf cg");
    issued += batch;
    Scal<T> h = read_scal(S, st);
    if (h.done) break;
  }
  return issued;
}

}  // namespace

void tea_native_cg(ops_dat u, ops_dat u0, ops_dat r, ops_dat p, ops_dat w, ops_dat z, ops_dat Kx,
                   ops_dat Ky, const tea_native_solve_params *P, tea_native_result *R) {
  (void)z;
  hipStream_t st = (hipStream_t)ops_hip_get_stream();
  // make sure every dat's device copy is current
  ops_dat all[] = {u, u0, r, p, w, Kx, Ky};
  for (ops_dat d : all) dev(d);
  Geo g{P->nx, P->ny, P->nx + 2};
  const long n = (long)(P->nx + 2) * (P->ny + 2);
  dim3 blk(BX, BY), grid((P->nx + BX - 1) / BX, (P->ny + BY - 1) / BY);
  const unsigned nb = grid.x * grid.y;
  double *u64 = ix0<double>(u), *u064 = ix0<double>(u0), *r64 = ix0<double>(r);
  double *p64 = ix0<double>(p), *w64 = ix0<double>(w);
  const double *kx = ix0<double>(Kx), *ky = ix0<double>(Ky);
  const double thresh = P->eps * P->initial_residual;
  // scratch: partials (nb doubles), FP32 arrays (Kx, Ky, r, p, w, d)
  size_t part_bytes = ((size_t)nb * sizeof(double) + 255) & ~(size_t)255;
  size_t f32 = ((size_t)n * sizeof(float) + 255) & ~(size_t)255;
  char *base = (char *)work(part_bytes + (P->solver == 2 ? 6 * f32 : 0));
  Scal<double> *S64 = (Scal<double> *)W.scal;
  Scal<float> *S32 = (Scal<float> *)((char *)W.scal + 512);
  double *dres = (double *)((char *)W.scal + 1024);
  if (P->solver == 1) {
    double *part = (double *)base;
    hipLaunchKernelGGL((k_cg_start<double>), grid, blk, 0, st, r64, p64, g, S64, part, W.cnt, thresh);
    int issued = cg_loop<double, double, double>(kx, ky, u64, r64, p64, w64, g, P->rx, P->ry, S64,
                                                 part, W.cnt, P->max_iters, grid, blk, st);
    Scal<double> h = read_scal(S64, st);
    (void)issued;
    R->iterations = h.iters;
    R->outer = 1;
    R->error = h.err;
  } else {
    float *part32 = (float *)base;
    double *part64 = (double *)base;
    float *kx32 = (float *)(base + part_bytes), *ky32 = kx32 + f32 / sizeof(float);
    float *r32 = ky32 + f32 / sizeof(float), *p32 = r32 + f32 / sizeof(float);
    float *w32 = p32 + f32 / sizeof(float), *d32 = w32 + f32 / sizeof(float);
    unsigned lin = (unsigned)((n + NT - 1) / NT);
    hipLaunchKernelGGL(k_to_f32, dim3(lin), dim3(NT), 0, st, kx, kx32, n);
    hipLaunchKernelGGL(k_to_f32, dim3(lin), dim3(NT), 0, st, ky, ky32, n);
    // halos of the FP32 work arrays must be zero (the operator reads them at the rim)
    CK(hipMemsetAsync(r32, 0, 4 * f32, st));
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
      double floor_rel = 0.5 * thresh * scale;
      if (floor_rel > inner_thresh) inner_thresh = floor_rel;
      hipLaunchKernelGGL(k_inner_start, grid, blk, 0, st, r64, r32, p32, d32, g, scale, S32, part32,
                         W.cnt + 2, inner_thresh);
      int issued = cg_loop<float, float, float>(kx32, ky32, d32, r32, p32, w32, g, (float)P->rx,
                                                (float)P->ry, S32, part32, W.cnt + 3,
                                                P->max_iters - total, grid, blk, st);
      Scal<float> h = read_scal(S32, st);
      total += h.iters > 0 ? h.iters : issued;
      hipLaunchKernelGGL(k_add_corr, grid, blk, 0, st, u64, d32, g, 1.0 / scale);
      hipLaunchKernelGGL(k_residual64, grid, blk, 0, st, kx, ky, u64, u064, r64, g, P->rx, P->ry,
                         dres, part64, W.cnt + 5);
      ops_hip_check_launch("tealeaf refinement");
      CK(hipMemcpyAsync(W.pinned, dres, sizeof(double), hipMemcpyDeviceToHost, st));
      CK(hipStreamSynchronize(st));
      rr = *(double *)W.pinned;
      err = sqrt(fabs(rr));
      if (h.iters == 0) break;  // no progress possible
    }
    R->iterations = total;
    R->outer = outer;
    R->error = err;
  }
  // the solver wrote u and r on the device
  u->dirty_hd = 2;
  r->dirty_hd = 2;
  p->dirty_hd = 2;
  w->dirty_hd = 2;
}
