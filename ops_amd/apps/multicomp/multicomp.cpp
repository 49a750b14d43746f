// Multi-component dats (the reference's apps/c/multiDim/multidim.cpp): a 2-component and a
// 3-component dat, a stencil loop reading neighbours of different components, a reduction
// over components, a periodic halo transfer of a multi-component dat, and packed fetches.
// Built twice: AoS (default) and SoA (OPS_SOA defined, ops_set_soa(1)).
#define OPS_2D
#include "ops_seq.h"

#include "multicomp_kernels.h"

#include <cstdio>
#include <exception>

// out: v (nx*ny*2, packed x fastest, components innermost), w (nx*ny*3), the stencil result
// (nx-2)*(ny-2) over the interior, then {sum, max}; v again after a periodic halo transfer
// along x, halos included ((nx+2)*ny*2); the stencil again from its API v1 kernel
// ((nx-2)*(ny-2))
extern "C" int multicomp_run(int argc, const char **argv, int nx, int ny, double *out) {
  try {
    ops_init(argc, argv, 1);
#ifdef OPS_SOA
    ops_set_soa(1);
#endif
    ops_block blk = ops_decl_block(2, "grid");
    int size[2] = {nx, ny}, base[2] = {0, 0}, dm[2] = {-1, -1}, dp[2] = {1, 1};
    double *none = nullptr;
    ops_dat v = ops_decl_dat(blk, 2, size, base, dm, dp, none, "double", "v");
    ops_dat w = ops_decl_dat(blk, 3, size, base, dm, dp, none, "double", "w");
    ops_dat o = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "out");
    ops_dat o1 = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "out_v1");
    int s00[] = {0, 0}, s4[] = {0, 0, 1, 0, 0, -1, -1, 0, 0, 1};
    ops_stencil S00 = ops_decl_stencil(2, 1, s00, "00");
    ops_stencil S4 = ops_decl_stencil(2, 5, s4, "00:10:0-1:-10:01");
    ops_reduction hs = ops_decl_reduction_handle(sizeof(double), "double", "sum");
    ops_reduction hm = ops_decl_reduction_handle(sizeof(double), "double", "max");
    int full[4] = {0, nx, 0, ny}, inner[4] = {1, nx - 1, 1, ny - 1};
    ops_par_loop(mc_set, "mc_set", blk, 2, full, ops_arg_dat(v, 2, S00, "double", OPS_WRITE),
                 ops_arg_dat(w, 3, S00, "double", OPS_WRITE), ops_arg_idx());
    ops_par_loop(mc_stencil, "mc_stencil", blk, 2, inner, ops_arg_dat(v, 2, S4, "double", OPS_READ),
                 ops_arg_dat(w, 3, S4, "double", OPS_READ), ops_arg_dat(o, 1, S00, "double", OPS_WRITE));
    ops_par_loop(mc_stencil_v1, "mc_stencil_v1", blk, 2, inner,
                 ops_arg_dat(v, 2, S4, "double", OPS_READ), ops_arg_dat(w, 3, S4, "double", OPS_READ),
                 ops_arg_dat(o1, 1, S00, "double", OPS_WRITE));
    ops_par_loop(mc_reduce, "mc_reduce", blk, 2, full, ops_arg_dat(v, 2, S00, "double", OPS_READ),
                 ops_arg_dat(w, 3, S00, "double", OPS_READ), ops_arg_reduce(hs, 1, "double", OPS_INC),
                 ops_arg_reduce(hm, 1, "double", OPS_MAX));
    double *p = out;
    ops_dat_fetch_data_slab_memspace(v, 0, (char *)p, full, OPS_HOST);
    p += (long)nx * ny * 2;
    ops_dat_fetch_data_slab_memspace(w, 0, (char *)p, full, OPS_HOST);
    p += (long)nx * ny * 3;
    ops_dat_fetch_data_slab_memspace(o, 0, (char *)p, inner, OPS_HOST);
    p += (long)(nx - 2) * (ny - 2);
    ops_reduction_result(hs, p);
    ops_reduction_result(hm, p + 1);
    p += 2;
    // periodic halo along x: column -1 <- column nx-1, column nx <- column 0
    int dirs[2] = {1, 2}, iter[2] = {1, ny}, from1[2] = {nx - 1, 0}, to1[2] = {-1, 0};
    int from2[2] = {0, 0}, to2[2] = {nx, 0};
    ops_halo h1 = ops_decl_halo(v, v, iter, from1, to1, dirs, dirs);
    ops_halo h2 = ops_decl_halo(v, v, iter, from2, to2, dirs, dirs);
    ops_halo hs2[] = {h1, h2};
    ops_halo_group g = ops_decl_halo_group(2, hs2);
    ops_halo_transfer(g);
    int wide[4] = {-1, nx + 1, 0, ny};
    ops_dat_fetch_data_slab_memspace(v, 0, (char *)p, wide, OPS_HOST);
    p += (long)(nx + 2) * ny * 2;
    ops_dat_fetch_data_slab_memspace(o1, 0, (char *)p, inner, OPS_HOST);
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "multicomp_run: %s\n", e.what());
    return 1;
  }
}
