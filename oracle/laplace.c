/*
 * ORACLE — test infrastructure only. CPU restatement of the reference's Laplace2D
 * Jacobi path, used by tests/ and by bench.py's cpu_baseline leg as the checker; it is
 * never part of the product.
 *
 * Follows:
 *   apps/c/laplace2d_tutorial/step7/laplace2d.cpp:36-127   (dats with 1-cell halos,
 *       zero fill, boundary loops, the iterate: stencil + MAX reduce, copy)
 *   apps/c/laplace2d_tutorial/step7/laplace_kernels.h      (kernel arithmetic)
 *   ops/c/include/ops_seq_v2.h:55-62                        (index -> address)
 *
 * Arrays are (nx+2) x (ny+2) doubles, x fastest, halo included: element (i,j) of the
 * loop index space lives at [(j+1)*(nx+2) + (i+1)].
 *   mode 0: top halo row u(x) = sin(pi*x/(nx-1)), other halos 0
 *   mode 1: left halo column u = sin(pi*(j+1)/(ny+1)) for j = -1..ny, right column the
 *           same times exp(-pi)   (the tutorial's boundary loops)
 * The stencil sums A(1,0)+A(-1,0)+A(0,-1)+A(0,1) left to right and scales by 0.25, the
 * error is fmax of |Anew-A|; OpenMP splits rows (max is order independent, so results do
 * not depend on the thread count).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static void laplace_bc(double *u, int nx, int ny, int mode, double pi) {
  const int sx = nx + 2;
  if (mode == 0) {
    for (int x = 0; x < nx; x++) u[(ny + 1) * sx + (x + 1)] = sin(pi * x / (nx - 1));
  } else {
    for (int j = -1; j <= ny; j++) {
      u[(j + 1) * sx + 0] = sin(pi * (j + 1) / (ny + 1));
      u[(j + 1) * sx + (nx + 1)] = sin(pi * (j + 1) / (ny + 1)) * exp(-pi);
    }
  }
}

/* Runs `iters` Jacobi iterations; A and Anew must hold (nx+2)*(ny+2) doubles.
   Returns the error of the last iteration. */
double oracle_laplace(int nx, int ny, int iters, int mode, double *A, double *Anew,
                      int threads) {
  const double pi = 2.0 * asin(1.0);
  const int sx = nx + 2;
  const size_t n = (size_t)(nx + 2) * (size_t)(ny + 2);
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  memset(A, 0, n * sizeof(double));
  memset(Anew, 0, n * sizeof(double));
  laplace_bc(A, nx, ny, mode, pi);
  laplace_bc(Anew, nx, ny, mode, pi);
  double error = 0.0;
  for (int it = 0; it < iters; it++) {
    double err = -1.7976931348623157e308; /* -DBL_MAX: MAX handle initial value */
#pragma omp parallel for reduction(max : err) schedule(static)
    for (int j = 0; j < ny; j++) {
      const double *a = A + (size_t)(j + 1) * sx + 1;
      double *an = Anew + (size_t)(j + 1) * sx + 1;
      for (int i = 0; i < nx; i++) {
        double v = 0.25f * (a[i + 1] + a[i - 1] + a[i - sx] + a[i + sx]);
        an[i] = v;
        err = fmax(err, fabs(v - a[i]));
      }
    }
    error = err;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; j++) {
      memcpy(A + (size_t)(j + 1) * sx + 1, Anew + (size_t)(j + 1) * sx + 1,
             (size_t)nx * sizeof(double));
    }
  }
  return error;
}
