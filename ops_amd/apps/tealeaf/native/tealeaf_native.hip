// Device-resident CG solvers for the TeaLeaf implicit step on gfx950 (interface and
// algorithms described in tealeaf_native.h).
//
// Mesh arrays carry one halo layer: pitch sx = nx + 2, interior cells at array indices
// 1..nx, 1..ny. The operator is the reference's calc_w stencil evaluated in the same
// operation order:
//   (1 + ry*(Ky(0,1)+Ky) + rx*(Kx(1,0)+Kx)) * p - ry*(Ky(0,1)*p(0,1) + Ky*p(0,-1))
//                                              - rx*(Kx(1,0)*p(1,0) + Kx*p(-1,0))
//
// One CG iteration k is four launches, each streaming kernel on a full grid:
//   W_k     p_k = beta p_{k-1} + r, evaluated at the five stencil points (p ping-pongs
//           between two buffers), w = A p_k, per-workgroup partials of p_k.w
//   PW_k    one workgroup folds the partials: alpha = rro / pw
//   UR_k    u += alpha p_k, r -= alpha w, per-workgroup partials of r.r
//   RR_k    one workgroup folds them: beta = rrn / rro, the stop test
// Scalars live in a device record, so there is no per-iteration host round trip and no
// atomics or device-scope fences. The host polls the record every POLL iterations while
// the next batch is already queued; launches issued after the stop return at once.
// When the pitch is even each thread owns two neighbouring cells (8- or 16-byte
// accesses: 4-byte lanes reach far less HBM bandwidth), otherwise one.
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
constexpr int FT = 1024;  // lanes of the folding workgroup

template <typename T> struct Rec {
  T rro, pw, alpha, beta;
  double err, thresh;
  double maxrr;  // r.r at the last residual replacement (mixed precision)
  double rr64;   // FP64 r.r of the last residual replacement
  int stop;      // iteration whose r.r met the threshold (INT_MAX while running)
  int iters;     // completed iterations
  int pad[8];
};

struct Geo {
  int nx, ny, sx;
};

template <typename T> struct Vec2;
template <> struct Vec2<double> { typedef double v __attribute__((ext_vector_type(2))); };
template <> struct Vec2<float> { typedef float v __attribute__((ext_vector_type(2))); };

// V consecutive values of a row starting at index i (i even when V == 2)
template <int V, typename T> struct Row {
  T v[V];
  __device__ __forceinline__ void load(const T *__restrict__ a, long i) {
    if constexpr (V == 2) {
      const typename Vec2<T>::v q = *(const typename Vec2<T>::v *)(a + i);
      v[0] = q.x;
      v[1] = q.y;
    } else {
      v[0] = a[i];
    }
  }
};

template <int V, typename T>
__device__ __forceinline__ void store_masked(T *__restrict__ a, long i, const T (&v)[V],
                                             const bool (&ok)[V]) {
  if constexpr (V == 2) {
    if (ok[0] && ok[1]) {
      typename Vec2<T>::v q;
      q.x = v[0];
      q.y = v[1];
      *(typename Vec2<T>::v *)(a + i) = q;
      return;
    }
  }
#pragma unroll
  for (int c = 0; c < V; c++)
    if (ok[c]) a[i + c] = v[c];
}

// first cell x0 and row index i of this thread's cells; false when the row is outside
template <int V>
__device__ __forceinline__ bool cells(const Geo &g, long &i, int &x0, bool (&ok)[V]) {
  unsigned bx, by, bz;
  ops_block_coords(bx, by, bz);
  const int t = (int)(bx * BX + threadIdx.x);
  x0 = (V == 2) ? 2 * t : 1 + t;
  const int y = 1 + (int)(by * BY + threadIdx.y);
  i = x0 + (long)y * g.sx;
#pragma unroll
  for (int c = 0; c < V; c++) ok[c] = x0 + c >= 1 && x0 + c <= g.nx;
  return y <= g.ny && x0 < g.sx;
}

template <int V> dim3 grid_of(const Geo &g) {
  const int lanes_x = (V == 2) ? g.sx / 2 : g.nx;
  return dim3((unsigned)((lanes_x + BX - 1) / BX), (unsigned)((g.ny + BY - 1) / BY));
}

template <typename T> __device__ __forceinline__ void write_partial(T v, T *part) {
  __shared__ T lds[NT / OPS_WAVE];
  T s = ops_block_reduce<ops_red_inc<T>>(v, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0)
    part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = s;
}

// sum of n partials by one FT-lane workgroup: eight independent lane-strided
// accumulators, then the workgroup tree; fixed order, valid in thread 0
template <typename T> __device__ __forceinline__ T fold(const T *__restrict__ part, unsigned n) {
  __shared__ T lds[FT / OPS_WAVE];
  T acc[8];
#pragma unroll
  for (int u = 0; u < 8; u++) acc[u] = T(0);
  for (unsigned base = 0; base < n; base += 8 * FT) {
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const unsigned j = base + u * FT + threadIdx.x;
      if (j < n) acc[u] += part[j];
    }
  }
  T a = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  return ops_block_reduce<ops_red_inc<T>>(a, lds);
}

template <typename T, typename K>
__device__ __forceinline__ T apply_A(const K *Kx, const K *Ky, const T *p, long i, int sx, T rx, T ry) {
  const T kx0 = (T)Kx[i], kx1 = (T)Kx[i + 1], ky0 = (T)Ky[i], ky1 = (T)Ky[i + sx];
  return (T(1.0) + ry * (ky1 + ky0) + rx * (kx1 + kx0)) * p[i] - ry * (ky1 * p[i + sx] + ky0 * p[i - sx]) -
         rx * (kx1 * p[i + 1] + kx0 * p[i - 1]);
}

// ---------------------------------------------------------------- CG kernels

// p = r; r.r partials
template <int V, typename T>
__global__ __launch_bounds__(NT) void k_start(const T *__restrict__ r, T *__restrict__ p, Geo g,
                                              T *part) {
  long i;
  int x0;
  bool ok[V];
  T v = T(0);
  if (cells<V>(g, i, x0, ok)) {
    Row<V, T> rr;
    rr.load(r, i);
    store_masked<V, T>(p, i, rr.v, ok);
#pragma unroll
    for (int c = 0; c < V; c++)
      if (ok[c]) v += rr.v[c] * rr.v[c];
  }
  write_partial(v, part);
}

// record after the start: rro = r.r
template <typename T>
__global__ __launch_bounds__(FT) void k_fold_start(const T *part, unsigned n, Rec<T> *S,
                                                   double thresh) {
  const T s = fold(part, n);
  if (threadIdx.x == 0) {
    S->rro = s;
    S->err = sqrt(fabs((double)s));
    S->thresh = thresh;
    S->maxrr = (double)s;
    S->stop = INT_MAX;
    S->iters = 0;
  }
}

// W_k. A stop at iteration k-1 makes this the reference's last p update (p only);
// p_only forces that form after the last issued iteration.
template <int V, typename T, typename K>
__global__ __launch_bounds__(NT) void k_w(const K *__restrict__ Kx, const K *__restrict__ Ky,
                                          const T *__restrict__ pin, T *__restrict__ pout,
                                          const T *__restrict__ r, T *__restrict__ w, Geo g, T rx,
                                          T ry, const Rec<T> *S, T *part, int k, int p_only) {
  const int stop = S->stop;
  if (stop < k - 1) return;
  const bool ponly = p_only || stop == k - 1;
  const T beta = k >= 2 ? S->beta : T(0), one = T(1.0);
  long i;
  int x0;
  bool ok[V];
  T v = T(0);
  if (cells<V>(g, i, x0, ok)) {
    const int sx = g.sx;
    Row<V, T> a, b;
    a.load(pin, i);
    b.load(r, i);
    T pc[V];
#pragma unroll
    for (int c = 0; c < V; c++) pc[c] = beta * a.v[c] + one * b.v[c];
    store_masked<V, T>(pout, i, pc, ok);
    if (!ponly) {
      // neighbours of p_k from p_{k-1} and r (zero halos give zero neighbours at the rim)
      const T pwst = ok[0] ? beta * pin[i - 1] + one * r[i - 1] : T(0);
      const T peas = ok[V - 1] ? beta * pin[i + V] + one * r[i + V] : T(0);
      Row<V, T> pn, rn, ps, rs;
      pn.load(pin, i + sx);
      rn.load(r, i + sx);
      ps.load(pin, i - sx);
      rs.load(r, i - sx);
      Row<V, K> kxk, kyk, kynk;
      kxk.load(Kx, i);
      kyk.load(Ky, i);
      kynk.load(Ky, i + sx);
      const T kxe = ok[V - 1] ? (T)Kx[i + V] : T(0);
      T wv[V];
#pragma unroll
      for (int c = 0; c < V; c++) {
        const T pnc = beta * pn.v[c] + one * rn.v[c];
        const T psc = beta * ps.v[c] + one * rs.v[c];
        const T pe = (c + 1 < V) ? pc[c + 1 < V ? c + 1 : c] : peas;
        const T pw = (c > 0) ? pc[c > 0 ? c - 1 : 0] : pwst;
        const T kx0 = (T)kxk.v[c];
        const T kx1 = (c + 1 < V) ? (T)kxk.v[c + 1 < V ? c + 1 : c] : kxe;
        const T ky0 = (T)kyk.v[c], ky1 = (T)kynk.v[c];
        wv[c] = (one + ry * (ky1 + ky0) + rx * (kx1 + kx0)) * pc[c] - ry * (ky1 * pnc + ky0 * psc) -
                rx * (kx1 * pe + kx0 * pw);
        if (ok[c]) v += wv[c] * pc[c];
      }
      store_masked<V, T>(w, i, wv, ok);
    }
  }
  if (!ponly) write_partial(v, part);
}

template <typename T>
__global__ __launch_bounds__(FT) void k_fold_pw(const T *part, unsigned n, Rec<T> *S, int k) {
  if (S->stop < k) return;
  const T pw = fold(part, n);
  if (threadIdx.x == 0) {
    S->pw = pw;
    S->alpha = S->rro / pw;
  }
}

// UR_k: u += alpha p; r -= alpha w; r.r partials
template <int V, typename T, typename U>
__global__ __launch_bounds__(NT) void k_ur(U *__restrict__ u, T *__restrict__ r, const T *__restrict__ p,
                                           const T *__restrict__ w, Geo g, const Rec<T> *S, T *part,
                                           int k) {
  if (S->stop < k) return;
  const T alpha = S->alpha;
  long i;
  int x0;
  bool ok[V];
  T v = T(0);
  if (cells<V>(g, i, x0, ok)) {
    Row<V, U> uu;
    Row<V, T> pp, rr, ww;
    uu.load(u, i);
    pp.load(p, i);
    rr.load(r, i);
    ww.load(w, i);
    U un[V];
    T rn[V];
#pragma unroll
    for (int c = 0; c < V; c++) {
      un[c] = uu.v[c] + (U)(alpha * pp.v[c]);
      rn[c] = rr.v[c] - alpha * ww.v[c];
      if (ok[c]) v += rn[c] * rn[c];
    }
    store_masked<V, U>(u, i, un, ok);
    store_masked<V, T>(r, i, rn, ok);
  }
  write_partial(v, part);
}

template <typename T>
__global__ __launch_bounds__(FT) void k_fold_rr(const T *part, unsigned n, Rec<T> *S, int k) {
  if (S->stop < k) return;
  const T rrn = fold(part, n);
  if (threadIdx.x == 0) {
    S->beta = rrn / S->rro;
    S->rro = rrn;
    S->iters = k;
    const double err = sqrt(fabs((double)rrn));
    S->err = err;
    if (err < S->thresh) S->stop = k;
  }
}

// ---------------------------------------------------------------- mixed-precision kernels
// FP32 CG iterations with reliable updates (residual replacement): the FP32 residual is
// replaced by the FP64 residual u0 - A u whenever the iterated residual has dropped by
// delta since the last replacement, and when it meets the threshold. The search direction
// is kept, so the Krylov process continues instead of restarting. The solution increment
// accumulates in FP32 (x32 += alpha p) and is folded into the FP64 solution after every
// batch of POLL iterations: an FP32 sum over many iterations would round each small
// increment against a large total, an error that A amplifies by up to its largest
// eigenvalue; over one batch it stays far below the replacement factor.

__global__ __launch_bounds__(NT) void k_to_f32(const double *__restrict__ a, float *__restrict__ b,
                                               long n) {
  const long i = (long)blockIdx.x * NT + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// u64 += x32; x32 = 0
template <int V>
__global__ __launch_bounds__(NT) void k_fold_increment(double *__restrict__ u, float *__restrict__ x32,
                                                       Geo g) {
  long i;
  int x0;
  bool ok[V];
  if (cells<V>(g, i, x0, ok)) {
    Row<V, double> uu;
    Row<V, float> xx;
    uu.load(u, i);
    xx.load(x32, i);
    double un[V];
    float z[V];
#pragma unroll
    for (int c = 0; c < V; c++) {
      un[c] = uu.v[c] + (double)xx.v[c];
      z[c] = 0.f;
    }
    store_masked<V, double>(u, i, un, ok);
    store_masked<V, float>(x32, i, z, ok);
  }
}

// r64 = u0 - A u (the reference's residual kernel) in FP64, r32 = (float)r64, r.r partials
__global__ __launch_bounds__(NT) void k_replace_residual(const double *__restrict__ Kx,
                                                         const double *__restrict__ Ky,
                                                         const double *__restrict__ u,
                                                         const double *__restrict__ u0,
                                                         double *__restrict__ r, float *__restrict__ r32,
                                                         Geo g, double rx, double ry, double *part) {
  long i;
  int x0;
  bool ok[1];
  double v = 0.0;
  if (cells<1>(g, i, x0, ok) && ok[0]) {
    const double smvp = apply_A<double, double>(Kx, Ky, u, i, g.sx, rx, ry);
    const double ri = u0[i] - smvp;
    r[i] = ri;
    r32[i] = (float)ri;
    v = ri * ri;
  }
  write_partial(v, part);
}

// after a replacement: rr = FP64 r.r; the next beta uses it in place of the iterated r.r
__global__ __launch_bounds__(FT) void k_fold_replace(const double *part, unsigned n, Rec<float> *S) {
  const double rr = fold(part, n);
  if (threadIdx.x == 0) {
    const double rro = (double)S->rro;
    S->beta = (float)((double)S->beta * (rr / rro));
    S->rro = (float)rr;
    S->rr64 = rr;
    S->maxrr = rr;
    S->err = sqrt(fabs(rr));
    S->stop = INT_MAX;
  }
}

// r32 = r64, p32 = r32, r32.r32 partials
__global__ __launch_bounds__(NT) void k_mixed_start(const double *__restrict__ r64, float *__restrict__ r32,
                                                    float *__restrict__ p32, Geo g, float *part) {
  long i;
  int x0;
  bool ok[1];
  float v = 0.f;
  if (cells<1>(g, i, x0, ok) && ok[0]) {
    const float ri = (float)r64[i];
    r32[i] = ri;
    p32[i] = ri;
    v = ri * ri;
  }
  write_partial(v, part);
}

// ---------------------------------------------------------------- host side

struct Work {
  void *buf = nullptr;
  size_t cap = 0;
  char *rec = nullptr;     // device records: FP64 CG, FP32 inner CG, refinement r.r
  char *pinned = nullptr;  // host copies
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
  if (!W.rec) {
    CK(hipMalloc((void **)&W.rec, 4096));
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

// CG iterations k = 1..max_iters until the record's stop flag is set; returns the host
// copy of the record and, in *last_p, the iteration whose W launch wrote the final p
template <int V, typename T, typename K, typename U>
Rec<T> cg_loop(const K *Kx, const K *Ky, U *u, T *r, T *pa, T *pb, T *w, const Geo &g, T rx, T ry,
               Rec<T> *S, Rec<T> *host, T *part, int max_iters, hipStream_t st, int *last_p) {
  const dim3 blk(BX, BY), grid = grid_of<V>(g);
  const unsigned nb = grid.x * grid.y;
  int k = 1;
  bool pending = false;
  Rec<T> h{};
  h.stop = INT_MAX;
  auto issue_w = [&](int kk, int p_only) {
    const T *pin = (kk & 1) ? pa : pb;
    T *pout = (kk & 1) ? pb : pa;
    hipLaunchKernelGGL((k_w<V, T, K>), grid, blk, 0, st, Kx, Ky, pin, pout, r, w, g, rx, ry, S, part,
                       kk, p_only);
  };
  while (k <= max_iters) {
    const int end = (k + POLL - 1 < max_iters) ? k + POLL - 1 : max_iters;
    for (; k <= end; k++) {
      issue_w(k, 0);
      hipLaunchKernelGGL((k_fold_pw<T>), dim3(1), dim3(FT), 0, st, part, nb, S, k);
      hipLaunchKernelGGL((k_ur<V, T, U>), grid, blk, 0, st, u, r, (k & 1) ? pb : pa, w, g, S, part, k);
      hipLaunchKernelGGL((k_fold_rr<T>), dim3(1), dim3(FT), 0, st, part, nb, S, k);
    }
    ops_hip_check_launch("tealeaf cg");
    // the previous batch's record is in host memory once its event has completed
    if (pending) {
      CK(hipEventSynchronize(W.ev));
      h = *host;
      if (h.stop != INT_MAX) break;
    }
    CK(hipMemcpyAsync(host, S, sizeof(Rec<T>), hipMemcpyDeviceToHost, st));
    CK(hipEventRecord(W.ev, st));
    pending = true;
  }
  // the last iteration's p update (W returns at once when an earlier stop made it)
  if (h.stop == INT_MAX && k == max_iters + 1) issue_w(k, 1);
  ops_hip_check_launch("tealeaf cg");
  CK(hipMemcpyAsync(host, S, sizeof(Rec<T>), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  h = *host;
  *last_p = (h.stop != INT_MAX) ? h.stop + 1 : max_iters + 1;
  return h;
}

template <int V>
void solve(ops_dat u, ops_dat u0, ops_dat r, ops_dat p, ops_dat w, ops_dat Kx, ops_dat Ky,
           const tea_native_solve_params *P, tea_native_result *R, const Geo &g) {
  hipStream_t st = (hipStream_t)ops_hip_get_stream();
  const dim3 blk(BX, BY), grid = grid_of<V>(g), grid1 = grid_of<1>(g);
  const unsigned nb = grid.x * grid.y, nb1 = grid1.x * grid1.y;
  const long n = (long)g.sx * (g.ny + 2);
  double *u64 = ix0<double>(u), *u064 = ix0<double>(u0), *r64 = ix0<double>(r);
  double *p64 = ix0<double>(p), *w64 = ix0<double>(w);
  const double *kx = ix0<double>(Kx), *ky = ix0<double>(Ky);
  const double thresh = P->eps * P->initial_residual;
  // scratch: partials, then a second FP64 p buffer (solver 1) or the FP32 arrays
  // Kx, Ky, r, pa, pb, w, x (solver 2)
  const unsigned nbmax = nb > nb1 ? nb : nb1;
  const size_t part_bytes = ((size_t)nbmax * sizeof(double) + 255) & ~(size_t)255;
  const size_t f64 = ((size_t)n * sizeof(double) + 255) & ~(size_t)255;
  const size_t f32 = ((size_t)n * sizeof(float) + 255) & ~(size_t)255;
  char *base = (char *)work(2 * part_bytes + (P->solver == 2 ? 7 * f32 : f64));
  char *arr = base + 2 * part_bytes;
  Rec<double> *S64 = (Rec<double> *)W.rec, *H64 = (Rec<double> *)W.pinned;
  Rec<float> *S32 = (Rec<float> *)(W.rec + 256), *H32 = (Rec<float> *)(W.pinned + 256);
  double *dres = (double *)(W.rec + 512), *hres = (double *)(W.pinned + 512);
  if (P->solver == 1) {
    double *part = (double *)base;
    double *pb = (double *)arr;
    CK(hipMemsetAsync(pb, 0, f64, st));  // zero halos of the second p buffer
    hipLaunchKernelGGL((k_start<V, double>), grid, blk, 0, st, r64, p64, g, part);
    hipLaunchKernelGGL((k_fold_start<double>), dim3(1), dim3(FT), 0, st, part, nb, S64, thresh);
    int last_p = 0;
    Rec<double> h = cg_loop<V, double, double, double>(kx, ky, u64, r64, p64, pb, w64, g, P->rx, P->ry,
                                                       S64, H64, part, P->max_iters, st, &last_p);
    // odd iterations write the second buffer: leave the final p in the dat
    if (last_p & 1)
      CK(hipMemcpyAsync(p64, pb, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
    R->iterations = h.iters;
    R->outer = 1;
    R->error = h.err;
    return;
  }
  float *part32 = (float *)base;
  double *part64 = (double *)(base + part_bytes);
  float *kx32 = (float *)arr, *ky32 = (float *)(arr + f32), *r32 = (float *)(arr + 2 * f32);
  float *pa32 = (float *)(arr + 3 * f32), *pb32 = (float *)(arr + 4 * f32);
  float *w32 = (float *)(arr + 5 * f32), *x32 = (float *)(arr + 6 * f32);
  const unsigned lin = (unsigned)((n + NT - 1) / NT);
  hipLaunchKernelGGL(k_to_f32, dim3(lin), dim3(NT), 0, st, kx, kx32, n);
  hipLaunchKernelGGL(k_to_f32, dim3(lin), dim3(NT), 0, st, ky, ky32, n);
  // FP32 work arrays (r, pa, pb, w, x) start at zero: the operator reads their halos
  CK(hipMemsetAsync(r32, 0, 5 * f32, st));
  const char *delta_env = getenv("OPS_TEA_RELIABLE_DELTA");
  const double delta = delta_env ? atof(delta_env) : 0.1;
  const double delta2 = delta * delta;
  // the caller's r holds u0 - A u in FP64
  hipLaunchKernelGGL(k_mixed_start, grid1, blk, 0, st, r64, r32, pa32, g, part32);
  hipLaunchKernelGGL((k_fold_start<float>), dim3(1), dim3(FT), 0, st, part32, nb1, S32, thresh);
  const dim3 gridv = grid_of<V>(g);
  auto replace = [&]() {
    hipLaunchKernelGGL((k_fold_increment<V>), gridv, blk, 0, st, u64, x32, g);
    hipLaunchKernelGGL(k_replace_residual, grid1, blk, 0, st, kx, ky, u64, u064, r64, r32, g, P->rx,
                       P->ry, part64);
    hipLaunchKernelGGL(k_fold_replace, dim3(1), dim3(FT), 0, st, part64, nb1, S32);
    ops_hip_check_launch("tealeaf residual replacement");
    CK(hipMemcpyAsync(H32, S32, sizeof(Rec<float>), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    return *H32;
  };
  const float rx32 = (float)P->rx, ry32 = (float)P->ry;
  int k = 1, replacements = 0, done_iters = 0;
  bool pending = false, converged = false, delta_on = true;
  double err = P->initial_residual;
  while (k <= P->max_iters) {
    const int end = (k + POLL - 1 < P->max_iters) ? k + POLL - 1 : P->max_iters;
    for (; k <= end; k++) {
      const float *pin = (k & 1) ? pa32 : pb32;
      float *pout = (k & 1) ? pb32 : pa32;
      hipLaunchKernelGGL((k_w<V, float, float>), gridv, blk, 0, st, kx32, ky32, pin, pout, r32, w32, g,
                         rx32, ry32, S32, part32, k, 0);
      hipLaunchKernelGGL((k_fold_pw<float>), dim3(1), dim3(FT), 0, st, part32, nb, S32, k);
      hipLaunchKernelGGL((k_ur<V, float, float>), gridv, blk, 0, st, x32, r32, pout, w32, g, S32,
                         part32, k);
      hipLaunchKernelGGL((k_fold_rr<float>), dim3(1), dim3(FT), 0, st, part32, nb, S32, k);
    }
    hipLaunchKernelGGL((k_fold_increment<V>), gridv, blk, 0, st, u64, x32, g);
    ops_hip_check_launch("tealeaf mixed cg");
    bool need = false;
    if (pending) {
      // record of the previous batch: replace when the iterated residual fell by delta
      // since the last replacement, or met the threshold
      CK(hipEventSynchronize(W.ev));
      const Rec<float> h = *H32;
      need = h.stop != INT_MAX || (delta_on && (double)h.rro < delta2 * h.maxrr);
    }
    if (need) {
      CK(hipStreamSynchronize(st));
      CK(hipMemcpyAsync(H32, S32, sizeof(Rec<float>), hipMemcpyDeviceToHost, st));
      CK(hipStreamSynchronize(st));
      const bool candidate = H32->stop != INT_MAX;  // the iterated residual met the threshold
      const int kc = candidate ? H32->stop : H32->iters;  // state after iteration kc
      const Rec<float> h = replace();
      replacements++;
      done_iters = kc;
      const double prev = err;
      err = h.err;
      // the stop follows the reference's rule on the iterated residual; the replacement
      // made before it leaves u and the reported error on the FP64 residual
      if (err < thresh || candidate) {
        converged = true;
        break;
      }
      // a replacement that no longer halves the FP64 residual has reached the attainable
      // accuracy (rounding of the FP64 residual itself, of order eps64 |A| |u|): further
      // replacements would only reset the iterated residual, so the solve runs on to the
      // stop rule
      if (err >= 0.5 * prev) delta_on = false;
      k = kc + 1;  // resume the Krylov process after iteration kc
      pending = false;
      continue;
    }
    CK(hipMemcpyAsync(H32, S32, sizeof(Rec<float>), hipMemcpyDeviceToHost, st));
    CK(hipEventRecord(W.ev, st));
    pending = true;
  }
  if (!converged) {
    // iteration limit: report the FP64 residual
    CK(hipStreamSynchronize(st));
    CK(hipMemcpyAsync(H32, S32, sizeof(Rec<float>), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    done_iters = H32->stop != INT_MAX ? H32->stop : H32->iters;
    const Rec<float> h = replace();
    replacements++;
    err = h.err;
  }
  (void)dres;
  (void)hres;
  R->iterations = done_iters;
  R->outer = replacements;
  R->error = err;
}

}  // namespace

void tea_native_cg(ops_dat u, ops_dat u0, ops_dat r, ops_dat p, ops_dat w, ops_dat z, ops_dat Kx,
                   ops_dat Ky, const tea_native_solve_params *P, tea_native_result *R) {
  (void)z;
  ops_dat all[] = {u, u0, r, p, w, Kx, Ky};
  for (ops_dat d : all) sync_device_copy(d);
  const Geo g{P->nx, P->ny, P->nx + 2};
  if (g.sx % 2 == 0)
    solve<2>(u, u0, r, p, w, Kx, Ky, P, R, g);
  else
    solve<1>(u, u0, r, p, w, Kx, Ky, P, R, g);
  // the solver wrote u, r, p and w on the device
  u->dirty_hd = 2;
  r->dirty_hd = 2;
  p->dirty_hd = 2;
  w->dirty_hd = 2;
}
