// Generated by ops_amd.translator (target hip_gfx950) for application 'multicomp_soa'.
// Do not edit: regenerate from the application sources.
#define OPS_2D
#include <hip/hip_runtime.h>
#include <string.h>
#include <math.h>
#include "ops_lib_core.h"
#include "ops_hip_rt.h"
#include "ops_hip_kernel.h"

namespace ops_gen_multicomp_soa {

// ---- ops_decl_const symbols

// ---- user kernels, compiled for host and gfx950
#pragma clang force_cuda_host_device begin
// from multicomp_kernels.h
// Kernels of the multi-component dat test (the reference's apps/c/multiDim: dats of 2 and 3
// components addressed as ACC(component, x, y), AoS by default and SoA under OPS_SOA).
#ifndef MULTICOMP_KERNELS_H
#define MULTICOMP_KERNELS_H

void mc_set(ACC<double> &v, ACC<double> &w, const int *idx) {
  v(0, 0, 0) = (double)idx[0];
  v(1, 0, 0) = (double)(10 * idx[1]);
  for (int c = 0; c < 3; c++) w(c, 0, 0) = (double)(100 * c + idx[0] + 7 * idx[1]);
}

// neighbours across components
void mc_stencil(const ACC<double> &v, const ACC<double> &w, ACC<double> &out) {
  out(0, 0) = v(0, 1, 0) + v(1, 0, -1) + w(2, -1, 0) - w(0, 0, 1);
}

void mc_reduce(const ACC<double> &v, const ACC<double> &w, double *sum, double *mx) {
  *sum += v(0, 0, 0) + v(1, 0, 0);
  *mx = *mx > w(2, 0, 0) ? *mx : w(2, 0, 0);
}

#endif
#pragma clang force_cuda_host_device end

struct ops_kp_mc_set {
  int cnt0, cnt1, cnt2, st0, st1, st2;
  double *p0; int sx0, sy0, sz0, xd0, yd0, zd0, md0;
  double *p1; int sx1, sy1, sz1, xd1, yd1, zd1, md1;
};

__global__ __launch_bounds__(256) void ops_kern_mc_set(const ops_kp_mc_set p, double *__restrict__ q0, double *__restrict__ q1) {
  unsigned bx, by, bz;
  ops_block_coords(bx, by, bz);
  const int y = (int)(by * blockDim.y + threadIdx.y);
  const int z = (int)(bz * blockDim.z + threadIdx.z);
  const int x = (int)(bx * blockDim.x + threadIdx.x);
  if (x < p.cnt0 && y < p.cnt1 && z < p.cnt2) {
    ACC<double> a0(p.md0, p.xd0, p.yd0, q0 + (x * p.sx0 + y * p.sy0 + z * p.sz0));
    ACC<double> a1(p.md1, p.xd1, p.yd1, q1 + (x * p.sx1 + y * p.sy1 + z * p.sz1));
    int idx2[3] = {p.st0 + x, p.st1 + y, p.st2 + z};
    mc_set(a0, a1, idx2);
  }
}

__global__ __launch_bounds__(256) void ops_kern_mc_set_alias(const ops_kp_mc_set p) {
  unsigned bx, by, bz;
  ops_block_coords(bx, by, bz);
  const int y = (int)(by * blockDim.y + threadIdx.y);
  const int z = (int)(bz * blockDim.z + threadIdx.z);
  const int x = (int)(bx * blockDim.x + threadIdx.x);
  if (x < p.cnt0 && y < p.cnt1 && z < p.cnt2) {
    ACC<double> a0(p.md0, p.xd0, p.yd0, p.p0 + (x * p.sx0 + y * p.sy0 + z * p.sz0));
    ACC<double> a1(p.md1, p.xd1, p.yd1, p.p1 + (x * p.sx1 + y * p.sy1 + z * p.sz1));
    int idx2[3] = {p.st0 + x, p.st1 + y, p.st2 + z};
    mc_set(a0, a1, idx2);
  }
}

struct ops_kp_mc_stencil {
  int cnt0, cnt1, cnt2, st0, st1, st2;
  double *p0; int sx0, sy0, sz0, xd0, yd0, zd0, md0;
  double *p1; int sx1, sy1, sz1, xd1, yd1, zd1, md1;
  double *p2; int sx2, sy2, sz2, xd2, yd2, zd2, md2;
};

__global__ __launch_bounds__(256) void ops_kern_mc_stencil(const ops_kp_mc_stencil p, double *__restrict__ q0, double *__restrict__ q1, double *__restrict__ q2) {
  unsigned bx, by, bz;
  ops_block_coords(bx, by, bz);
  const int y = (int)(by * blockDim.y + threadIdx.y);
  const int z = (int)(bz * blockDim.z + threadIdx.z);
  const int x = (int)(bx * blockDim.x + threadIdx.x);
  if (x < p.cnt0 && y < p.cnt1 && z < p.cnt2) {
    const ACC<double> a0(p.md0, p.xd0, p.yd0, q0 + (x * p.sx0 + y * p.sy0 + z * p.sz0));
    const ACC<double> a1(p.md1, p.xd1, p.yd1, q1 + (x * p.sx1 + y * p.sy1 + z * p.sz1));
    ACC<double> a2(p.xd2, q2 + (x * p.sx2 + y * p.sy2 + z * p.sz2));
    mc_stencil(a0, a1, a2);
  }
}

__global__ __launch_bounds__(256) void ops_kern_mc_stencil_alias(const ops_kp_mc_stencil p) {
  unsigned bx, by, bz;
  ops_block_coords(bx, by, bz);
  const int y = (int)(by * blockDim.y + threadIdx.y);
  const int z = (int)(bz * blockDim.z + threadIdx.z);
  const int x = (int)(bx * blockDim.x + threadIdx.x);
  if (x < p.cnt0 && y < p.cnt1 && z < p.cnt2) {
    const ACC<double> a0(p.md0, p.xd0, p.yd0, p.p0 + (x * p.sx0 + y * p.sy0 + z * p.sz0));
    const ACC<double> a1(p.md1, p.xd1, p.yd1, p.p1 + (x * p.sx1 + y * p.sy1 + z * p.sz1));
    ACC<double> a2(p.xd2, p.p2 + (x * p.sx2 + y * p.sy2 + z * p.sz2));
    mc_stencil(a0, a1, a2);
  }
}

struct ops_kp_mc_reduce {
  int cnt0, cnt1, cnt2, st0, st1, st2;
  double *p0; int sx0, sy0, sz0, xd0, yd0, zd0, md0;
  double *p1; int sx1, sy1, sz1, xd1, yd1, zd1, md1;
  double *rp2;
  double *rp3;
};

__global__ __launch_bounds__(256) void ops_kern_mc_reduce(const ops_kp_mc_reduce p) {
  double r2[1];
  for (int d = 0; d < 1; d++) r2[d] = ops_red_inc<double>::identity();
  double r3[1];
  for (int d = 0; d < 1; d++) r3[d] = ops_red_max<double>::identity();
  unsigned bx, by, bz;
  ops_block_coords(bx, by, bz);
  const int y = (int)(by * blockDim.y + threadIdx.y);
  const int z = (int)(bz * blockDim.z + threadIdx.z);
  const int x = (int)(bx * blockDim.x + threadIdx.x);
  if (x < p.cnt0 && y < p.cnt1 && z < p.cnt2) {
    const ACC<double> a0(p.md0, p.xd0, p.yd0, p.p0 + (x * p.sx0 + y * p.sy0 + z * p.sz0));
    const ACC<double> a1(p.md1, p.xd1, p.yd1, p.p1 + (x * p.sx1 + y * p.sy1 + z * p.sz1));
    mc_reduce(a0, a1, r2, r3);
  }
  ops_reduce_partial<ops_red_inc<double>, double, 1>(r2, p.rp2);
  ops_reduce_partial<ops_red_max<double>, double, 1>(r3, p.rp3);
}

}  // namespace ops_gen_multicomp_soa

void ops_par_loop_mc_set(char const *name, ops_block block, int dim, int *range, ops_arg arg0, ops_arg arg1, ops_arg arg2) {
  static int kid = ops_hip_register_kernel("mc_set");
  ops_arg args[3] = {arg0, arg1, arg2};
  ops_hip_loop lp;
  lp.vx = 1;
  lp.prolong_mask = 0u;
  if (ops_hip_loop_begin(&lp, kid, name, block, dim, range, args, 3)) {
    ops_gen_multicomp_soa::ops_kp_mc_set p;
    p.cnt0 = lp.count[0]; p.cnt1 = lp.count[1]; p.cnt2 = lp.count[2];
    p.st0 = lp.start[0]; p.st1 = lp.start[1]; p.st2 = lp.start[2];
    p.p0 = (double *)lp.ptr[0]; p.sx0 = lp.step[0][0]; p.sy0 = lp.step[0][1]; p.sz0 = lp.step[0][2];
    p.xd0 = lp.xdim[0]; p.yd0 = lp.ydim[0]; p.zd0 = lp.zdim[0]; p.md0 = lp.mdim[0];
    p.p1 = (double *)lp.ptr[1]; p.sx1 = lp.step[1][0]; p.sy1 = lp.step[1][1]; p.sz1 = lp.step[1][2];
    p.xd1 = lp.xdim[1]; p.yd1 = lp.ydim[1]; p.zd1 = lp.zdim[1]; p.md1 = lp.mdim[1];
    if (lp.alias)
      hipLaunchKernelGGL(ops_gen_multicomp_soa::ops_kern_mc_set_alias, dim3(lp.grid[0], lp.grid[1], lp.grid[2]), dim3(lp.block[0], lp.block[1], lp.block[2]), 0, (hipStream_t)lp.stream, p);
    else
      hipLaunchKernelGGL(ops_gen_multicomp_soa::ops_kern_mc_set, dim3(lp.grid[0], lp.grid[1], lp.grid[2]), dim3(lp.block[0], lp.block[1], lp.block[2]), 0, (hipStream_t)lp.stream, p, p.p0, p.p1);
    ops_hip_check_launch(name);
  }
  ops_hip_loop_end(&lp, args, 3);
}
void ops_par_loop_mc_stencil(char const *name, ops_block block, int dim, int *range, ops_arg arg0, ops_arg arg1, ops_arg arg2) {
  static int kid = ops_hip_register_kernel("mc_stencil");
  ops_arg args[3] = {arg0, arg1, arg2};
  ops_hip_loop lp;
  lp.vx = 1;
  lp.prolong_mask = 0u;
  if (ops_hip_loop_begin(&lp, kid, name, block, dim, range, args, 3)) {
    ops_gen_multicomp_soa::ops_kp_mc_stencil p;
    p.cnt0 = lp.count[0]; p.cnt1 = lp.count[1]; p.cnt2 = lp.count[2];
    p.st0 = lp.start[0]; p.st1 = lp.start[1]; p.st2 = lp.start[2];
    p.p0 = (double *)lp.ptr[0]; p.sx0 = lp.step[0][0]; p.sy0 = lp.step[0][1]; p.sz0 = lp.step[0][2];
    p.xd0 = lp.xdim[0]; p.yd0 = lp.ydim[0]; p.zd0 = lp.zdim[0]; p.md0 = lp.mdim[0];
    p.p1 = (double *)lp.ptr[1]; p.sx1 = lp.step[1][0]; p.sy1 = lp.step[1][1]; p.sz1 = lp.step[1][2];
    p.xd1 = lp.xdim[1]; p.yd1 = lp.ydim[1]; p.zd1 = lp.zdim[1]; p.md1 = lp.mdim[1];
    p.p2 = (double *)lp.ptr[2]; p.sx2 = lp.step[2][0]; p.sy2 = lp.step[2][1]; p.sz2 = lp.step[2][2];
    p.xd2 = lp.xdim[2]; p.yd2 = lp.ydim[2]; p.zd2 = lp.zdim[2]; p.md2 = lp.mdim[2];
    if (lp.alias)
      hipLaunchKernelGGL(ops_gen_multicomp_soa::ops_kern_mc_stencil_alias, dim3(lp.grid[0], lp.grid[1], lp.grid[2]), dim3(lp.block[0], lp.block[1], lp.block[2]), 0, (hipStream_t)lp.stream, p);
    else
      hipLaunchKernelGGL(ops_gen_multicomp_soa::ops_kern_mc_stencil, dim3(lp.grid[0], lp.grid[1], lp.grid[2]), dim3(lp.block[0], lp.block[1], lp.block[2]), 0, (hipStream_t)lp.stream, p, p.p0, p.p1, p.p2);
    ops_hip_check_launch(name);
  }
  ops_hip_loop_end(&lp, args, 3);
}
void ops_par_loop_mc_reduce(char const *name, ops_block block, int dim, int *range, ops_arg arg0, ops_arg arg1, ops_arg arg2, ops_arg arg3) {
  static int kid = ops_hip_register_kernel("mc_reduce");
  ops_arg args[4] = {arg0, arg1, arg2, arg3};
  ops_hip_loop lp;
  lp.vx = 1;
  lp.prolong_mask = 0u;
  if (ops_hip_loop_begin(&lp, kid, name, block, dim, range, args, 4)) {
    ops_gen_multicomp_soa::ops_kp_mc_reduce p;
    p.cnt0 = lp.count[0]; p.cnt1 = lp.count[1]; p.cnt2 = lp.count[2];
    p.st0 = lp.start[0]; p.st1 = lp.start[1]; p.st2 = lp.start[2];
    p.p0 = (double *)lp.ptr[0]; p.sx0 = lp.step[0][0]; p.sy0 = lp.step[0][1]; p.sz0 = lp.step[0][2];
    p.xd0 = lp.xdim[0]; p.yd0 = lp.ydim[0]; p.zd0 = lp.zdim[0]; p.md0 = lp.mdim[0];
    p.p1 = (double *)lp.ptr[1]; p.sx1 = lp.step[1][0]; p.sy1 = lp.step[1][1]; p.sz1 = lp.step[1][2];
    p.xd1 = lp.xdim[1]; p.yd1 = lp.ydim[1]; p.zd1 = lp.zdim[1]; p.md1 = lp.mdim[1];
    p.rp2 = (double *)lp.red_part[2];
    p.rp3 = (double *)lp.red_part[3];
    hipLaunchKernelGGL(ops_gen_multicomp_soa::ops_kern_mc_reduce, dim3(lp.grid[0], lp.grid[1], lp.grid[2]), dim3(lp.block[0], lp.block[1], lp.block[2]), 0, (hipStream_t)lp.stream, p);
    ops_hip_check_launch(name);
  }
  ops_hip_loop_end(&lp, args, 4);
}

namespace {
struct ops_const_registration {
  ops_const_registration() {
  }
} ops_const_registration_instance;
}  // namespace
