// Multigrid transfer test (the reference's apps/c/mgrid/mgrid.cpp:52-257): dats on grids
// of strides 1, 2 and 4 of one block, restrict and prolong stencils of stride 2, periodic
// halos of depth 2, and self-checking kernels with MAX reductions. Sizes are scaled by
// `fac` (the reference runs fac = 1: a 24 x 24 fine grid).
#define OPS_2D
#include "ops_seq_v2.h"

#include "mgrid_kernels.h"

#include <cstdio>
#include <exception>

namespace {

// periodic halo group of a dat along x (dir 0) or y (dir 1), depth 2
ops_halo_group periodic(ops_dat d, const int *n, int dir) {
  int dirs[] = {1, 2};
  int iter[2], from[2], to[2];
  iter[dir] = 2;
  iter[1 - dir] = n[1 - dir] + 4;
  from[1 - dir] = to[1 - dir] = -2;
  from[dir] = 0;
  to[dir] = n[dir];
  ops_halo h1 = ops_decl_halo(d, d, iter, from, to, dirs, dirs);
  from[dir] = n[dir] - 2;
  to[dir] = -2;
  ops_halo h2 = ops_decl_halo(d, d, iter, from, to, dirs, dirs);
  ops_halo hs[] = {h1, h2};
  return ops_decl_halo_group(2, hs);
}

void fetch(ops_dat d, const int *n, double *out) {
  if (!out) return;
  int r[4] = {0, n[0], 0, n[1]};
  ops_dat_fetch_data_slab_memspace(d, 0, (char *)out, r, OPS_HOST);
}

}  // namespace

// errs[2]: the prolong and restrict checks (0 = pass); the outputs (packed interiors, x
// fastest, may be null): fine and middle grids after prolongation, middle and coarse grids
// after restriction
extern "C" int mgrid_run(int argc, const char **argv, int fac, int *errs, double *prolong_fine,
                         double *prolong_mid, double *restrict_mid, double *restrict_coarse) {
  try {
    ops_init(argc, argv, 1);
    ops_block grid0 = ops_decl_block(2, "grid0");
    int s00[] = {0, 0};
    int s00_m10_p10[] = {0, 0, -1, 0, 1, 0};
    int s5[] = {0, 0, -1, 0, 1, 0, 0, -1, 0, 1};
    ops_stencil S00 = ops_decl_stencil(2, 1, s00, "00");
    ops_stencil S5 = ops_decl_stencil(2, 5, s5, "5pt");
    int d_p[2] = {2, 2}, d_m[2] = {-2, -2}, base[2] = {0, 0};
    int fine[2] = {24 * fac, 24 * fac}, mid[2] = {12 * fac, 12 * fac}, coarse[2] = {6 * fac, 6 * fac};
    int st1[2] = {1, 1}, st2[2] = {2, 2}, st4[2] = {4, 4};
    ops_stencil RESTRICT = ops_decl_restrict_stencil(2, 3, s00_m10_p10, st2, "RESTRICT_00_M10_P10");
    ops_stencil PROLONG = ops_decl_prolong_stencil(2, 3, s00_m10_p10, st2, "PROLONG_00_M10_P10");
    double *none = nullptr;
    ops_dat data1 = ops_decl_dat(grid0, 1, coarse, base, d_m, d_p, st4, none, "double", "data1");
    ops_dat data0 = ops_decl_dat(grid0, 1, mid, base, d_m, d_p, st2, none, "double", "data0");
    ops_dat data5 = ops_decl_dat(grid0, 1, fine, base, d_m, d_p, st1, none, "double", "data5");
    ops_dat data6 = ops_decl_dat(grid0, 1, mid, base, d_m, d_p, st2, none, "double", "data6");
    ops_dat data3 = ops_decl_dat(grid0, 1, coarse, base, d_m, d_p, st4, none, "double", "data3");
    ops_reduction err = ops_decl_reduction_handle(sizeof(int), "int", "reduct_err");
    ops_halo_group fine_x = periodic(data5, fine, 0), fine_y = periodic(data5, fine, 1);
    ops_halo_group mid_x = periodic(data0, mid, 0), coarse_x = periodic(data1, coarse, 0);
    ops_partition("");

    int r_mid[] = {0, mid[0], 0, mid[1]};
    int r_fine[] = {0, fine[0], 0, fine[1]};
    int r_coarse[] = {0, coarse[0], 0, coarse[1]};
    // prolongation: coarse -> mid -> fine
    ops_par_loop(mgrid_populate, "mgrid_populate_kernel_1", grid0, 2, r_coarse,
                 ops_arg_dat(data1, 1, S00, "double", OPS_WRITE), ops_arg_idx(),
                 ops_arg_gbl(&coarse[0], 1, "int", OPS_READ));
    ops_halo_transfer(coarse_x);
    ops_par_loop(mgrid_prolong_kernel, "mgrid_prolong_kernel", grid0, 2, r_mid,
                 ops_arg_dat(data1, 1, PROLONG, "double", OPS_READ),
                 ops_arg_dat(data0, 1, S00, "double", OPS_WRITE));
    ops_halo_transfer(mid_x);
    ops_par_loop(mgrid_prolong_kernel, "mgrid_prolong_kernel", grid0, 2, r_fine,
                 ops_arg_dat(data0, 1, PROLONG, "double", OPS_READ),
                 ops_arg_dat(data5, 1, S00, "double", OPS_WRITE));
    ops_halo_transfer(fine_x);
    ops_halo_transfer(fine_y);
    ops_par_loop(mgrid_prolong_check, "prolong_check", grid0, 2, r_fine,
                 ops_arg_dat(data5, 1, S5, "double", OPS_READ), ops_arg_idx(),
                 ops_arg_reduce(err, 1, "int", OPS_MAX), ops_arg_gbl(&fine[0], 1, "int", OPS_READ),
                 ops_arg_gbl(&fine[1], 1, "int", OPS_READ));
    ops_reduction_result(err, &errs[0]);
    fetch(data5, fine, prolong_fine);
    fetch(data0, mid, prolong_mid);
    // restriction: fine -> mid -> coarse
    ops_par_loop(mgrid_populate, "mgrid_populate_kernel_3", grid0, 2, r_fine,
                 ops_arg_dat(data5, 1, S00, "double", OPS_WRITE), ops_arg_idx(),
                 ops_arg_gbl(&fine[0], 1, "int", OPS_READ));
    ops_par_loop(mgrid_restrict_kernel, "mgrid_restrict_kernel", grid0, 2, r_mid,
                 ops_arg_dat(data5, 1, RESTRICT, "double", OPS_READ),
                 ops_arg_dat(data6, 1, S00, "double", OPS_WRITE));
    ops_par_loop(mgrid_restrict_kernel, "mgrid_restrict_kernel", grid0, 2, r_coarse,
                 ops_arg_dat(data6, 1, RESTRICT, "double", OPS_READ),
                 ops_arg_dat(data3, 1, S00, "double", OPS_WRITE));
    ops_par_loop(mgrid_restrict_check, "restrict_check", grid0, 2, r_coarse,
                 ops_arg_dat(data3, 1, S00, "double", OPS_READ), ops_arg_idx(),
                 ops_arg_reduce(err, 1, "int", OPS_MAX), ops_arg_gbl(&fine[0], 1, "int", OPS_READ));
    ops_reduction_result(err, &errs[1]);
    fetch(data6, mid, restrict_mid);
    fetch(data3, coarse, restrict_coarse);
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "mgrid_run: %s\n", e.what());
    return 1;
  }
}
