/*
 * Device-side helpers shared by translator-generated gfx950 loop kernels and by the
 * runtime's own kernels: XCD-aware workgroup placement and wavefront/workgroup
 * reductions with an on-device second pass.
 *
 * Replaces the reference's ops_reduction_hip<> LDS tree (ops/c/include/
 * ops_hip_reduction.h:49-258) and its host-side final pass
 * (templates/cpp/cuda/loop_host.cpp.j2:720-744).
 */
#ifndef OPS_AMD_HIP_KERNEL_H
#define OPS_AMD_HIP_KERNEL_H

#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>
#include "ops_lib_core.h"

#define OPS_WAVE 64
#define OPS_XCDS 8

/*
 * Workgroups are dispatched round-robin over the 8 XCDs (linear id b runs on XCD b % 8).
 * Remap so that each XCD owns one contiguous run of tiles: vertically adjacent tiles of
 * a stencil then share that XCD's L2 instead of being fetched into several.
 * The map is a bijection on [0, n) for any n.
 */
__device__ __forceinline__ unsigned ops_xcd_remap(unsigned b, unsigned n) {
  unsigned q = n / OPS_XCDS, r = n % OPS_XCDS;
  unsigned x = b % OPS_XCDS, local = b / OPS_XCDS;
  unsigned first = (x < r) ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return first + local;
}

/* The XCD remap with each XCD's run of tiles walked from its end when rev is set: a launch
   that follows another in the opposite direction starts on the tiles that the previous one
   touched last, whose lines the 256 MB memory-side cache still holds. */
__device__ __forceinline__ unsigned ops_xcd_remap_dir(unsigned b, unsigned n, int rev) {
  unsigned q = n / OPS_XCDS, r = n % OPS_XCDS;
  unsigned x = b % OPS_XCDS, local = b / OPS_XCDS;
  unsigned first = (x < r) ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  unsigned len = q + (x < r ? 1u : 0u);
  return first + (rev ? len - 1 - local : local);
}

/* Tile coordinates of this workgroup after the XCD remap, walked forwards or backwards. */
__device__ __forceinline__ void ops_block_coords_dir(unsigned &bx, unsigned &by, unsigned &bz, int rev) {
  unsigned gx = gridDim.x, gy = gridDim.y;
  unsigned n = gx * gy * gridDim.z;
  unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  unsigned t = ops_xcd_remap_dir(b, n, rev);
  bx = t % gx;
  t /= gx;
  by = t % gy;
  bz = t / gy;
}

/* Tile coordinates of this workgroup after the XCD remap. */
__device__ __forceinline__ void ops_block_coords(unsigned &bx, unsigned &by, unsigned &bz) {
  unsigned gx = gridDim.x, gy = gridDim.y;
  unsigned n = gx * gy * gridDim.z;
  unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  unsigned t = ops_xcd_remap(b, n);
  bx = t % gx;
  t /= gx;
  by = t % gy;
  bz = t / gy;
}

/* Tile coordinates of tile vt of a loop walked grid-stride (loops with reductions run on a
   capped 1-D grid): the same XCD remap over all n tiles of the tile grid gx x gy x ... */
__device__ __forceinline__ void ops_tile_coords(unsigned vt, unsigned n, unsigned gx, unsigned gy,
                                                unsigned &bx, unsigned &by, unsigned &bz) {
  unsigned t = ops_xcd_remap(vt, n);
  bx = t % gx;
  t /= gx;
  by = t % gy;
  bz = t / gy;
}

/* ---- reduction operators ---- */
template <typename T> struct ops_red_inc {
  static __device__ __forceinline__ T identity() { return T(0); }
  static __device__ __forceinline__ T op(T a, T b) { return a + b; }
};
template <typename T> struct ops_lim;
template <> struct ops_lim<double> {
  static __device__ __forceinline__ double hi() { return DBL_MAX; }
  static __device__ __forceinline__ double lo() { return -DBL_MAX; }
};
template <> struct ops_lim<float> {
  static __device__ __forceinline__ float hi() { return FLT_MAX; }
  static __device__ __forceinline__ float lo() { return -FLT_MAX; }
};
template <> struct ops_lim<int> {
  static __device__ __forceinline__ int hi() { return INT_MAX; }
  static __device__ __forceinline__ int lo() { return INT_MIN; }
};
template <> struct ops_lim<long long> {
  static __device__ __forceinline__ long long hi() { return LLONG_MAX; }
  static __device__ __forceinline__ long long lo() { return LLONG_MIN; }
};
/* identities follow ops_arg_reduce's handle initialisation (ops_lib_core.cpp:840-877) */
template <typename T> struct ops_red_min {
  static __device__ __forceinline__ T identity() { return ops_lim<T>::hi(); }
  static __device__ __forceinline__ T op(T a, T b) { return b < a ? b : a; }
};
template <typename T> struct ops_red_max {
  static __device__ __forceinline__ T identity() { return ops_lim<T>::lo(); }
  static __device__ __forceinline__ T op(T a, T b) { return b > a ? b : a; }
};

/* Butterfly over the 64 lanes; every lane ends with the same value. The combine order
   is fixed, so the result does not depend on scheduling. */
template <typename OP, typename T> __device__ __forceinline__ T ops_wave_reduce(T v) {
#pragma unroll
  for (int m = OPS_WAVE / 2; m >= 1; m >>= 1) v = OP::op(v, __shfl_xor(v, m, OPS_WAVE));
  return v;
}

/* Workgroup reduction of one value (256 threads = 4 waves); result valid in thread 0. */
template <typename OP, typename T> __device__ __forceinline__ T ops_block_reduce(T v, T *lds) {
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int nthreads = blockDim.x * blockDim.y * blockDim.z;
  v = ops_wave_reduce<OP>(v);
  const int wave = tid / OPS_WAVE, lane = tid % OPS_WAVE;
  __syncthreads();
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  T r = OP::identity();
  if (tid == 0) {
    const int nw = (nthreads + OPS_WAVE - 1) / OPS_WAVE;
    r = lds[0];
    for (int w = 1; w < nw; w++) r = OP::op(r, lds[w]);
  }
  return r;
}

/*
 * Writes this workgroup's partial of one reduction argument of dimension DIM to
 * part[workgroup * DIM + d]. No atomics and no device-scope fences: a second kernel,
 * launched by the runtime on the same stream, folds the partials in a fixed order.
 */
template <typename OP, typename T, int DIM>
__device__ __forceinline__ void ops_reduce_partial(const T (&vals)[DIM], T *part) {
  __shared__ T lds[16];
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    T r = ops_block_reduce<OP>(vals[d], lds);
    if (tid == 0) part[(size_t)b * DIM + d] = r;
  }
}

/* by-value carrier for small ops_arg_gbl READ arguments */
template <typename T, int N> struct ops_gbl {
  T v[N];
};

#endif /* OPS_AMD_HIP_KERNEL_H */
