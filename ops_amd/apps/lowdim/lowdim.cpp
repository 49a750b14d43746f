// Low-dimensional reductions (the reference's apps/c/lowdim_test/lowdim.cpp:53-472): a
// 3-D dat reduced with MAX, MIN and INC onto the three coordinate planes and three lines
// (dats of size 1 in the collapsed dimensions, strided stencils with stride 0 there), and
// the reduced dats read back over the 3-D range.
#define OPS_3D
#include "ops_seq_v2.h"

#include "lowdim_kernels.h"

#include <cstdio>
#include <exception>

namespace {

ops_dat lowdim(ops_block blk, int nx, int ny, int nz, const char *name) {
  int size[3] = {nx, ny, nz}, base[3] = {0, 0, 0};
  int hm[3] = {nx > 1 ? -1 : 0, ny > 1 ? -1 : 0, nz > 1 ? -1 : 0};
  int hp[3] = {nx > 1 ? 1 : 0, ny > 1 ? 1 : 0, nz > 1 ? 1 : 0};
  double *none = nullptr;
  return ops_decl_dat(blk, 1, size, base, hm, hp, none, "double", name);
}

double *fetch(ops_dat d, double *out) {
  int r[6] = {0, d->size[0] - d->d_p[0] + d->d_m[0], 0, d->size[1] - d->d_p[1] + d->d_m[1], 0,
              d->size[2] - d->d_p[2] + d->d_m[2]};
  ops_dat_fetch_data_slab_memspace(d, 0, (char *)out, r, OPS_HOST);
  return out + (long)r[1] * r[3] * r[5];
}

}  // namespace

// out (x fastest within each array): for MAX, MIN, INC, then INC run a second time (the
// reduction accumulates on top: same access as the last write), the planes XY, YZ, XZ and
// the lines X, Y, Z; then the XY plane of the sums plus the Z line, read over nx*ny*nz
extern "C" int lowdim_run(int argc, const char **argv, int nx, int ny, int nz, double *out) {
  try {
    ops_init(argc, argv, 1);
    ops_block blk = ops_decl_block(3, "block");
    ops_dat v = lowdim(blk, nx, ny, nz, "dat3D");
    ops_dat b = lowdim(blk, nx, ny, nz, "dat3D_bcast");
    ops_dat xy = lowdim(blk, nx, ny, 1, "dat2D_XY"), yz = lowdim(blk, 1, ny, nz, "dat2D_YZ");
    ops_dat xz = lowdim(blk, nx, 1, nz, "dat2D_XZ");
    ops_dat x = lowdim(blk, nx, 1, 1, "dat1D_X"), y = lowdim(blk, 1, ny, 1, "dat1D_Y");
    ops_dat z = lowdim(blk, 1, 1, nz, "dat1D_Z");
    int s000[] = {0, 0, 0};
    int sxy[] = {1, 1, 0}, syz[] = {0, 1, 1}, sxz[] = {1, 0, 1}, sx[] = {1, 0, 0}, sy[] = {0, 1, 0},
        sz[] = {0, 0, 1};
    ops_stencil S000 = ops_decl_stencil(3, 1, s000, "S3D_000");
    ops_stencil SXY = ops_decl_strided_stencil(3, 1, s000, sxy, "stride3D_xy");
    ops_stencil SYZ = ops_decl_strided_stencil(3, 1, s000, syz, "stride3D_yz");
    ops_stencil SXZ = ops_decl_strided_stencil(3, 1, s000, sxz, "stride3D_xz");
    ops_stencil SX = ops_decl_strided_stencil(3, 1, s000, sx, "stride3D_x");
    ops_stencil SY = ops_decl_strided_stencil(3, 1, s000, sy, "stride3D_y");
    ops_stencil SZ = ops_decl_strided_stencil(3, 1, s000, sz, "stride3D_z");
    ops_partition("");
    int r3[] = {0, nx, 0, ny, 0, nz};
    ops_par_loop(lowdim_set, "set3D", blk, 3, r3, ops_arg_dat(v, 1, S000, "double", OPS_WRITE), ops_arg_idx());
    double *o = out;
    auto all = [&]() {
      for (ops_dat d : {xy, yz, xz, x, y, z}) o = fetch(d, o);
    };
    ops_par_loop(lowdim_max, "reduct_max", blk, 3, r3, ops_arg_dat(v, 1, S000, "double", OPS_READ),
                 ops_arg_dat(xy, 1, SXY, "double", OPS_MAX), ops_arg_dat(yz, 1, SYZ, "double", OPS_MAX),
                 ops_arg_dat(xz, 1, SXZ, "double", OPS_MAX), ops_arg_dat(x, 1, SX, "double", OPS_MAX),
                 ops_arg_dat(y, 1, SY, "double", OPS_MAX), ops_arg_dat(z, 1, SZ, "double", OPS_MAX));
    all();
    ops_par_loop(lowdim_min, "reduct_min", blk, 3, r3, ops_arg_dat(v, 1, S000, "double", OPS_READ),
                 ops_arg_dat(xy, 1, SXY, "double", OPS_MIN), ops_arg_dat(yz, 1, SYZ, "double", OPS_MIN),
                 ops_arg_dat(xz, 1, SXZ, "double", OPS_MIN), ops_arg_dat(x, 1, SX, "double", OPS_MIN),
                 ops_arg_dat(y, 1, SY, "double", OPS_MIN), ops_arg_dat(z, 1, SZ, "double", OPS_MIN));
    all();
    for (int k = 0; k < 2; k++) {
      ops_par_loop(lowdim_inc, "reduct_inc", blk, 3, r3, ops_arg_dat(v, 1, S000, "double", OPS_READ),
                   ops_arg_dat(xy, 1, SXY, "double", OPS_INC), ops_arg_dat(yz, 1, SYZ, "double", OPS_INC),
                   ops_arg_dat(xz, 1, SXZ, "double", OPS_INC), ops_arg_dat(x, 1, SX, "double", OPS_INC),
                   ops_arg_dat(y, 1, SY, "double", OPS_INC), ops_arg_dat(z, 1, SZ, "double", OPS_INC));
      all();
    }
    ops_par_loop(lowdim_bcast, "bcast", blk, 3, r3, ops_arg_dat(xy, 1, SXY, "double", OPS_READ),
                 ops_arg_dat(z, 1, SZ, "double", OPS_READ), ops_arg_dat(b, 1, S000, "double", OPS_WRITE));
    fetch(b, o);
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "lowdim_run: %s\n", e.what());
    return 1;
  }
}
