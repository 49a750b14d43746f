// Laplace2D Jacobi relaxation on an OPS block: the plain-method workload.
//
// Mirrors the structure of the reference tutorial (apps/c/laplace2d_tutorial/step7/
// laplace2d.cpp:14-143): two dats A and Anew with one-cell halos, a 5-point stencil loop
// with a MAX reduction of the update, ops_reduction_result each iteration, and a copy
// loop back. Two boundary setups:
//   mode 0  interior nx x ny, top halo row u(x) = sin(pi*x/(nx-1)), other halos 0
//   mode 1  the tutorial's: left halo column u = sin(pi*(j+1)/(ny+1)), right column the
//           same times exp(-pi), top and bottom 0
// Boundary values are computed on the host (glibc sin/exp) and uploaded with
// ops_dat_set_data_slab_memspace, so every run sees bit-identical boundary data.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

int imax, jmax;
double pi = 2.0 * asin(1.0);

#define OPS_2D
#include <ops_seq_v2.h>

#include "laplace_kernels.h"

namespace {
struct Laplace {
  ops_block block;
  ops_dat A, Anew;
  ops_stencil S00, S5;
  ops_reduction h_err;
  int iter = 0;
  double last_error = 0.0;
} L;

void upload_row(ops_dat d, int y, const std::vector<double> &v, int x0) {
  int range[4] = {x0, x0 + (int)v.size(), y, y + 1};
  ops_dat_set_data_slab_memspace(d, 0, (char *)v.data(), range, OPS_HOST);
}
void upload_col(ops_dat d, int x, const std::vector<double> &v, int y0) {
  int range[4] = {x, x + 1, y0, y0 + (int)v.size()};
  ops_dat_set_data_slab_memspace(d, 0, (char *)v.data(), range, OPS_HOST);
}

void set_boundaries(ops_dat d, int mode) {
  if (mode == 0) {
    std::vector<double> top(imax);
    for (int x = 0; x < imax; x++) top[x] = sin(pi * x / (imax - 1));
    upload_row(d, jmax, top, 0);
  } else {
    std::vector<double> left(jmax + 2), right(jmax + 2);
    for (int j = -1; j <= jmax; j++) {
      left[j + 1] = sin(pi * (j + 1) / (jmax + 1));
      right[j + 1] = sin(pi * (j + 1) / (jmax + 1)) * exp(-pi);
    }
    upload_col(d, -1, left, -1);
    upload_col(d, imax, right, -1);
  }
}
}  // namespace

extern "C" int laplace2d_init(int argc, const char **argv, int nx, int ny, int mode) {
  try {
    ops_init(argc, argv, 1);
    imax = nx;
    jmax = ny;
    L = Laplace();
    L.block = ops_decl_block(2, "my_grid");
    int size[] = {imax, jmax}, base[] = {0, 0}, d_m[] = {-1, -1}, d_p[] = {1, 1};
    double *none = NULL;
    L.A = ops_decl_dat(L.block, 1, size, base, d_m, d_p, none, "double", "A");
    L.Anew = ops_decl_dat(L.block, 1, size, base, d_m, d_p, none, "double", "Anew");
    int s00[] = {0, 0};
    L.S00 = ops_decl_stencil(2, 1, s00, "0,0");
    int s5[] = {0, 0, 1, 0, -1, 0, 0, 1, 0, -1};
    L.S5 = ops_decl_stencil(2, 5, s5, "5pt");
    L.h_err = ops_decl_reduction_handle(sizeof(double), "double", "error");
    ops_decl_const("imax", 1, "int", &imax);
    ops_decl_const("jmax", 1, "int", &jmax);
    ops_decl_const("pi", 1, "double", &pi);
    int full[] = {-1, imax + 1, -1, jmax + 1};
    ops_par_loop(set_zero, "set zero", L.block, 2, full,
                 ops_arg_dat(L.A, 1, L.S00, "double", OPS_WRITE));
    ops_par_loop(set_zero, "set zero", L.block, 2, full,
                 ops_arg_dat(L.Anew, 1, L.S00, "double", OPS_WRITE));
    if (mode == 2) {
      // tutorial boundaries evaluated on the device from the loop index
      int left[] = {-1, 0, -1, jmax + 1}, right[] = {imax, imax + 1, -1, jmax + 1};
      double one = 1.0, damp = exp(-pi);
      ops_dat ds[2] = {L.A, L.Anew};
      for (ops_dat d : ds) {
        ops_par_loop(sine_bndcon, "left bndcon", L.block, 2, left,
                     ops_arg_dat(d, 1, L.S00, "double", OPS_WRITE), ops_arg_idx(),
                     ops_arg_gbl(&one, 1, "double", OPS_READ));
        ops_par_loop(sine_bndcon, "right bndcon", L.block, 2, right,
                     ops_arg_dat(d, 1, L.S00, "double", OPS_WRITE), ops_arg_idx(),
                     ops_arg_gbl(&damp, 1, "double", OPS_READ));
      }
    } else {
      set_boundaries(L.A, mode);
      set_boundaries(L.Anew, mode);
    }
    L.iter = 0;
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "laplace2d_init: %s\n", e.what());
    return 1;
  }
}

// `n` Jacobi iterations; the error of the last one is returned in *err
extern "C" int laplace2d_iterate(int n, double *err) {
  try {
    int interior[] = {0, imax, 0, jmax};
    double error = 0.0;
    for (int it = 0; it < n; it++) {
      ops_par_loop(apply_stencil, "apply stencil", L.block, 2, interior,
                   ops_arg_dat(L.A, 1, L.S5, "double", OPS_READ),
                   ops_arg_dat(L.Anew, 1, L.S00, "double", OPS_WRITE),
                   ops_arg_reduce(L.h_err, 1, "double", OPS_MAX));
      ops_reduction_result(L.h_err, &error);
      ops_par_loop(copy, "copy", L.block, 2, interior,
                   ops_arg_dat(L.A, 1, L.S00, "double", OPS_WRITE),
                   ops_arg_dat(L.Anew, 1, L.S00, "double", OPS_READ));
      L.iter++;
    }
    L.last_error = error;
    if (err) *err = error;
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "laplace2d_iterate: %s\n", e.what());
    return 1;
  }
}

// copies the whole allocation of A (halos included, (nx+2)*(ny+2) doubles)
extern "C" int laplace2d_fetch_full(double *out) {
  try {
    int range[] = {-1, imax + 1, -1, jmax + 1};
    ops_dat_fetch_data_slab_memspace(L.A, 0, (char *)out, range, OPS_HOST);
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "laplace2d_fetch_full: %s\n", e.what());
    return 1;
  }
}

extern "C" int laplace2d_finalize(void) {
  try {
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "laplace2d_finalize: %s\n", e.what());
    return 1;
  }
}

#ifdef OPS_APP_MAIN
// command line: laplace2d [nx ny iters mode] — the tutorial run by default
int main(int argc, const char **argv) {
  int nx = argc > 1 ? atoi(argv[1]) : 4094, ny = argc > 2 ? atoi(argv[2]) : 4094;
  int iters = argc > 3 ? atoi(argv[3]) : 100, mode = argc > 4 ? atoi(argv[4]) : 1;
  if (laplace2d_init(argc, argv, nx, ny, mode)) return 1;
  double err = 0;
  if (laplace2d_iterate(iters, &err)) return 1;
  printf("%5d, %0.6f\n", iters, err);
  if (mode == 1 && nx == 4094 && ny == 4094 && iters == 100) {
    double diff = fabs((100.0 * (err / 2.421354960840227e-03)) - 100.0);
    printf("Total error is within %3.15E %% of the expected error\n", diff);
    printf(diff < 0.001 ? "This run is considered PASSED\n" : "This test is considered FAILED\n");
  }
  return laplace2d_finalize();
}
#endif
