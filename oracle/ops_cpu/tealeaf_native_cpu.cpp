// ORACLE — test infrastructure only. Host restatement of the TeaLeaf device-resident
// solvers (ops_amd/apps/tealeaf/tealeaf_native.h) for the oracle build of the
// application: the same recurrences (FP64 CG; FP32 CG with reliable updates, i.e. FP64
// residual replacement and solution accumulation), evaluated in plain loops. Dot
// products are summed in row order; the replacement test runs every iteration.
#include <math.h>
#include <stdlib.h>

#include <vector>

#include "ops_lib_core.h"
#include "tealeaf_native.h"

namespace {

template <typename T, typename K>
T apply_A(const K *Kx, const K *Ky, const T *p, long i, long sx, T rx, T ry) {
  const T kx0 = (T)Kx[i], kx1 = (T)Kx[i + 1], ky0 = (T)Ky[i], ky1 = (T)Ky[i + sx];
  return (T(1.0) + ry * (ky1 + ky0) + rx * (kx1 + kx0)) * p[i] - ry * (ky1 * p[i + sx] + ky0 * p[i - sx]) -
         rx * (kx1 * p[i + 1] + kx0 * p[i - 1]);
}

template <typename T, typename K, typename U>
int cg(const K *Kx, const K *Ky, U *u, T *r, T *p, T *w, int nx, int ny, T rx, T ry, double thresh,
       int max_iters, double *err_out) {
  const long sx = nx + 2;
  T rro = 0;
  for (int y = 1; y <= ny; y++)
    for (int x = 1; x <= nx; x++) {
      long i = x + y * sx;
      p[i] = r[i];
      rro += r[i] * r[i];
    }
  double err = sqrt(fabs((double)rro));
  int it = 0;
  while (it < max_iters && !(err < thresh)) {
    T pw = 0;
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++) {
        long i = x + y * sx;
        w[i] = apply_A<T, K>(Kx, Ky, p, i, sx, rx, ry);
        pw += w[i] * p[i];
      }
    T alpha = rro / pw, rrn = 0;
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++) {
        long i = x + y * sx;
        u[i] = u[i] + (U)(alpha * p[i]);
        r[i] = r[i] - alpha * w[i];
        rrn += r[i] * r[i];
      }
    T beta = rrn / rro;
    rro = rrn;
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++) {
        long i = x + y * sx;
        p[i] = beta * p[i] + T(1.0) * r[i];
      }
    it++;
    err = sqrt(fabs((double)rrn));
  }
  *err_out = err;
  return it;
}

}  // namespace

void tea_native_cg(ops_dat u, ops_dat u0, ops_dat r, ops_dat p, ops_dat w, ops_dat z, ops_dat Kx,
                   ops_dat Ky, const tea_native_solve_params *P, tea_native_result *R) {
  (void)z;
  const int nx = P->nx, ny = P->ny;
  const long sx = nx + 2, n = sx * (ny + 2);
  double *U = (double *)u->data, *U0 = (double *)u0->data, *Rr = (double *)r->data;
  double *Pp = (double *)p->data, *Ww = (double *)w->data;
  const double *kx = (const double *)Kx->data, *ky = (const double *)Ky->data;
  const double thresh = P->eps * P->initial_residual;
  if (P->solver == 1) {
    double err;
    R->iterations = cg<double, double, double>(kx, ky, U, Rr, Pp, Ww, nx, ny, P->rx, P->ry, thresh,
                                               P->max_iters, &err);
    R->outer = 1;
    R->error = err;
    return;
  }
  // FP32 CG with reliable updates: the residual is replaced by the FP64 one when the
  // iterated r.r fell by delta^2 since the last replacement and when it meets the
  // threshold (then the solve stops); the direction p is kept; the FP32 increment is
  // folded into the FP64 solution every 32 iterations and at each replacement.
  const char *delta_env = getenv("OPS_TEA_RELIABLE_DELTA");
  const double delta = delta_env ? atof(delta_env) : 0.1;
  std::vector<float> kx32(n), ky32(n), r32(n, 0.f), p32(n, 0.f), w32(n, 0.f), x32(n, 0.f);
  for (long i = 0; i < n; i++) {
    kx32[i] = (float)kx[i];
    ky32[i] = (float)ky[i];
  }
  float rro = 0.f;
  for (int y = 1; y <= ny; y++)
    for (int x = 1; x <= nx; x++) {
      long i = x + y * sx;
      r32[i] = (float)Rr[i];
      p32[i] = r32[i];
      rro += r32[i] * r32[i];
    }
  double maxrr = rro, err = P->initial_residual;
  const float rx = (float)P->rx, ry = (float)P->ry;
  int it = 0, outer = 0;
  bool delta_on = true;
  bool stalled = false;  // stopped by the rule (converged or at the iterated threshold)
  auto fold_increment = [&]() {
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++) {
        long i = x + y * sx;
        U[i] = U[i] + (double)x32[i];
        x32[i] = 0.f;
      }
  };
  while (it < P->max_iters) {
    float pw = 0.f;
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++) {
        long i = x + y * sx;
        w32[i] = apply_A<float, float>(kx32.data(), ky32.data(), p32.data(), i, sx, rx, ry);
        pw += w32[i] * p32[i];
      }
    const float alpha = rro / pw;
    float rrn = 0.f;
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++) {
        long i = x + y * sx;
        x32[i] = x32[i] + alpha * p32[i];
        r32[i] = r32[i] - alpha * w32[i];
        rrn += r32[i] * r32[i];
      }
    it++;
    float beta = rrn / rro;
    rro = rrn;
    const bool candidate = sqrt(fabs((double)rrn)) < thresh;
    const bool replace = candidate || (delta_on && (double)rrn < delta * delta * maxrr);
    if (it % 32 == 0 || replace) fold_increment();
    if (replace) {
      double rr = 0.0;
      for (int y = 1; y <= ny; y++)
        for (int x = 1; x <= nx; x++) {
          long i = x + y * sx;
          double ri = U0[i] - apply_A<double, double>(kx, ky, U, i, sx, P->rx, P->ry);
          Rr[i] = ri;
          r32[i] = (float)ri;
          rr += ri * ri;
        }
      outer++;
      beta = (float)((double)beta * (rr / (double)rro));
      rro = (float)rr;
      maxrr = rr;
      const double prev = err;
      err = sqrt(fabs(rr));
      if (err < thresh || candidate) {
        stalled = true;
        break;
      }
      if (err >= 0.5 * prev) delta_on = false;  // attainable accuracy reached
    }
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++) {
        long i = x + y * sx;
        p32[i] = beta * p32[i] + 1.0f * r32[i];
      }
  }
  int total = it;
  if (!stalled) {
    // iteration limit: fold the increment in and report the FP64 residual
    fold_increment();
    double rr = 0.0;
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++) {
        long i = x + y * sx;
        double ri = U0[i] - apply_A<double, double>(kx, ky, U, i, sx, P->rx, P->ry);
        Rr[i] = ri;
        rr += ri * ri;
      }
    outer++;
    err = sqrt(fabs(rr));
  }
  R->iterations = total;
  R->outer = outer;
  R->error = err;
}
