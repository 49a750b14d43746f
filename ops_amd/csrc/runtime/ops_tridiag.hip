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
// Layout on MI355X: one lane per line. For y- and z-lines consecutive lanes own
// consecutive x positions, so every row access of a wavefront is one contiguous 512-byte
// segment. x-lines are contiguous per line, so a one-wavefront workgroup stages a 64-line x 16-row
// tile of a/b/c/d through LDS with coalesced loads before each lane sweeps its line.
// The forward-sweep temporaries c', d' live in a scratch array laid out row-major over
// lines (row i of all lines contiguous), so they are written and re-read coalesced.
#include <hip/hip_runtime.h>

#include <cstring>

#include "ops_internal.h"
#include "ops_tridiag.h"

#define HIPCHECK(x) ops_hip_check((int)(x), #x, __FILE__, __LINE__)

namespace {

struct TridGeom {
  long long s[3];   // element strides of the padded arrays per dimension
  int n[3];         // extents of the solve box
  int solvedim;
  long long nlines;
};

// origin of line `l` (elements, relative to the box origin) and the stride along it
__device__ __forceinline__ long long line_origin(const TridGeom &g, long long l) {
  if (g.solvedim == 0) {
    long long p = l % g.n[1], q = l / g.n[1];
    return p * g.s[1] + q * g.s[2];
  } else if (g.solvedim == 1) {
    long long p = l % g.n[0], q = l / g.n[0];
    return p + q * g.s[2];
  } else {
    long long p = l % g.n[0], q = l / g.n[0];
    return p + q * g.s[1];
  }
}

template <typename T, bool INC>
__global__ __launch_bounds__(256) void thomas_strided(const T *__restrict__ a,
                                                      const T *__restrict__ b,
                                                      const T *__restrict__ c, T *__restrict__ d,
                                                      T *__restrict__ u, T *__restrict__ cp,
                                                      T *__restrict__ dp, TridGeom g) {
  long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= g.nlines) return;
  const long long o = line_origin(g, l);
  const long long st = g.s[g.solvedim];
  const int n = g.n[g.solvedim];
  const long long L = g.nlines;
  T cprev = c[o] / b[o];
  T dprev = d[o] / b[o];
  cp[l] = cprev;
  dp[l] = dprev;
  for (int i = 1; i < n; i++) {
    long long k = o + i * st;
    T m = T(1) / (b[k] - a[k] * cprev);
    cprev = c[k] * m;
    dprev = (d[k] - a[k] * dprev) * m;
    cp[i * L + l] = cprev;
    dp[i * L + l] = dprev;
  }
  T x = dprev;
  long long k = o + (long long)(n - 1) * st;
  if (INC) u[k] += x; else d[k] = x;
  for (int i = n - 2; i >= 0; i--) {
    x = dp[i * L + l] - cp[i * L + l] * x;
    k = o + i * st;
    if (INC) u[k] += x; else d[k] = x;
  }
}

// x-lines: 64 lines per workgroup, rows staged through LDS in chunks of XR.
constexpr int XL = 64, XR = 16;  // one wavefront per workgroup, 34.8 KB LDS (double)

template <typename T, bool INC>
__global__ __launch_bounds__(64) void thomas_x(const T *__restrict__ a,
                                                const T *__restrict__ b,
                                                const T *__restrict__ c, T *__restrict__ d,
                                                T *__restrict__ u, T *__restrict__ cp,
                                                T *__restrict__ dp, TridGeom g) {
  __shared__ T sa[XL][XR + 1], sb[XL][XR + 1], sc[XL][XR + 1], sd[XL][XR + 1];
  const int tid = threadIdx.x;  // 64 threads, lane tid sweeps line l0 + tid
  const long long l0 = (long long)blockIdx.x * XL;
  const int n = g.n[0];
  const long long L = g.nlines;
  const int lane = tid;
  const long long myl = l0 + lane;
  const bool own = myl < L;
  T cprev = T(0), dprev = T(0);
  for (int r0 = 0; r0 < n; r0 += XR) {
    const int rows = (n - r0) < XR ? (n - r0) : XR;
    // cooperative coalesced load: thread -> (line, row) with row fastest
    for (int e = tid; e < XL * XR; e += XL) {
      int li = e / XR, ri = e % XR;
      long long ll = l0 + li;
      if (ll < L && ri < rows) {
        long long k = line_origin(g, ll) + r0 + ri;
        sa[li][ri] = a[k];
        sb[li][ri] = b[k];
        sc[li][ri] = c[k];
        sd[li][ri] = d[k];
      }
    }
    __syncthreads();
    if (own) {
      for (int ri = 0; ri < rows; ri++) {
        int i = r0 + ri;
        if (i == 0) {
          cprev = sc[lane][ri] / sb[lane][ri];
          dprev = sd[lane][ri] / sb[lane][ri];
        } else {
          T m = T(1) / (sb[lane][ri] - sa[lane][ri] * cprev);
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
  for (int cidx = nchunks - 1; cidx >= 0; cidx--) {
    const int r0 = cidx * XR;
    const int rows = (n - r0) < XR ? (n - r0) : XR;
    if (own) {
      for (int ri = rows - 1; ri >= 0; ri--) {
        int i = r0 + ri;
        if (i == n - 1) x = dp[i * L + myl];
        else x = dp[i * L + myl] - cp[i * L + myl] * x;
        sd[lane][ri] = x;
      }
    }
    __syncthreads();
    for (int e = tid; e < XL * XR; e += XL) {
      int li = e / XR, ri = e % XR;
      long long ll = l0 + li;
      if (ll < L && ri < rows) {
        long long k = line_origin(g, ll) + r0 + ri;
        if (INC) u[k] += sd[li][ri]; else d[k] = sd[li][ri];
      }
    }
    __syncthreads();
  }
}

struct Scratch {
  char *p = nullptr;
  size_t cap = 0;
};
Scratch g_scratch;

void *scratch(size_t bytes) {
  if (bytes > g_scratch.cap) {
    ops_dev_sync();
    if (g_scratch.p) (void)hipFree(g_scratch.p);
    HIPCHECK(hipMalloc((void **)&g_scratch.p, bytes));
    g_scratch.cap = bytes;
  }
  return g_scratch.p;
}

template <typename T, bool INC>
void solve(int ndim, int solvedim, const int *range, ops_dat a, ops_dat b, ops_dat c, ops_dat d,
           ops_dat u) {
  TridGeom g;
  long long off = 0;
  g.s[0] = 1;
  g.s[1] = a->size[0];
  g.s[2] = (long long)a->size[0] * a->size[1];
  for (int i = 0; i < 3; i++) g.n[i] = 1;
  for (int i = 0; i < ndim; i++) {
    g.n[i] = range[2 * i + 1] - range[2 * i];
    if (g.n[i] <= 0) return;
    off += (long long)range[2 * i] * g.s[i];
  }
  g.solvedim = solvedim;
  long long total = (long long)g.n[0] * g.n[1] * g.n[2];
  g.nlines = total / g.n[solvedim];
  // index-0 pointers (raw-pointer semantics: base_offset) plus the box origin
  auto at = [&](ops_dat x) -> T * {
    if (x == nullptr) return nullptr;
    return (T *)(x->data_d + x->base_offset) + off;
  };
  T *cp = (T *)scratch(sizeof(T) * 2 * (size_t)total);
  T *dp = cp + total;
  hipStream_t s = (hipStream_t)ops_hip_get_stream();
  if (solvedim == 0) {
    unsigned blocks = (unsigned)((g.nlines + XL - 1) / XL);
    hipLaunchKernelGGL((thomas_x<T, INC>), dim3(blocks), dim3(XL), 0, s, at(a), at(b), at(c),
                       at(d), at(u), cp, dp, g);
  } else {
    unsigned blocks = (unsigned)((g.nlines + 255) / 256);
    hipLaunchKernelGGL((thomas_strided<T, INC>), dim3(blocks), dim3(256), 0, s, at(a), at(b),
                       at(c), at(d), at(u), cp, dp, g);
  }
  ops_hip_check_launch("ops_tridMultiDimBatch");
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
  ops_dat all[5] = {a, b, c, d, u};
  for (ops_dat x : all)
    if (x) ops_put_data(x);
}

}  // namespace

extern "C" void ops_tridMultiDimBatch(int ndim, int solvedim, int *range, ops_dat a, ops_dat b,
                                      ops_dat c, ops_dat d, ops_tridsolver_params *ctx) {
  (void)ctx;
  check_args(ndim, solvedim, a, b, c, d, nullptr);
  if (!strcmp(a->type, "double"))
    solve<double, false>(ndim, solvedim, range, a, b, c, d, nullptr);
  else if (!strcmp(a->type, "float"))
    solve<float, false>(ndim, solvedim, range, a, b, c, d, nullptr);
  else
    ops_throw(OPS_RUNTIME_ERROR,
              "Tridsolver error: unsupported type. Dataset type must either be \"double\" or "
              "\"float\"");
  d->dirty_hd = 2;
}

extern "C" void ops_tridMultiDimBatch_Inc(int ndim, int solvedim, int *range, ops_dat a,
                                          ops_dat b, ops_dat c, ops_dat d, ops_dat u,
                                          ops_tridsolver_params *ctx) {
  (void)ctx;
  check_args(ndim, solvedim, a, b, c, d, u);
  if (!strcmp(a->type, "double"))
    solve<double, true>(ndim, solvedim, range, a, b, c, d, u);
  else if (!strcmp(a->type, "float"))
    solve<float, true>(ndim, solvedim, range, a, b, c, d, u);
  else
    ops_throw(OPS_RUNTIME_ERROR,
              "Tridsolver error: unsupported type. Dataset type must either be \"double\" or "
              "\"float\"");
  u->dirty_hd = 2;
}
