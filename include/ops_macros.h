/*
 * API v1 accessors: kernels that take plain pointers (`const double *a`) and address
 * them with OPS_ACC<n>(x, y, z) / OPS_ACC_MD<n>(d, x, y, z), where <n> is the position
 * of the argument in the ops_par_loop call (the reference's ops/c/include/ops_macros.h,
 * the non-debug definitions at :723, :928, :1133 and :1765-2058).
 *
 * The pointer handed to the kernel already points at the loop's current point; the
 * macros add the offset of a stencil point in elements, using the sizes of the dat
 * (size[0] as xdim, size[1] as ydim, size[2] as zdim) and its component count
 * (multi_d). Those per-argument sizes live in `ops_v1`:
 *   - in translated gfx950 kernels `ops_v1` is a by-value parameter that the
 *     translator appends to every kernel using these macros (it shadows the global);
 *   - in host builds it is the global below, which the host loop driver fills before
 *     each loop.
 * Multi-component dats follow the dat layout: components innermost (AoS) by default,
 * component planes (SoA) under OPS_SOA.
 */
#ifndef OPS_AMD_MACROS_H
#define OPS_AMD_MACROS_H

#define OPS_V1_MAXARGS 32

struct ops_v1_shape {
  int xd[OPS_V1_MAXARGS], yd[OPS_V1_MAXARGS], zd[OPS_V1_MAXARGS], md[OPS_V1_MAXARGS];
};

#ifdef __cplusplus
extern "C" {
#endif
extern struct ops_v1_shape ops_v1;
#ifdef __cplusplus
}
#endif

#if defined(OPS_1D)
#define OPS_V1_OFF(n, x) (x)
#ifdef OPS_SOA
#define OPS_V1_OFF_MD(n, d, x) ((x) + (d) * ops_v1.xd[n])
#else
#define OPS_V1_OFF_MD(n, d, x) ((x) * ops_v1.md[n] + (d))
#endif
#elif defined(OPS_2D)
#define OPS_V1_OFF(n, x, y) ((x) + ops_v1.xd[n] * (y))
#ifdef OPS_SOA
#define OPS_V1_OFF_MD(n, d, x, y) ((x) + ops_v1.xd[n] * (y) + (d) * ops_v1.xd[n] * ops_v1.yd[n])
#else
#define OPS_V1_OFF_MD(n, d, x, y) ((x) * ops_v1.md[n] + (d) + ops_v1.xd[n] * (y) * ops_v1.md[n])
#endif
#elif defined(OPS_3D)
#define OPS_V1_OFF(n, x, y, z) ((x) + ops_v1.xd[n] * (y) + ops_v1.xd[n] * ops_v1.yd[n] * (z))
#ifdef OPS_SOA
#define OPS_V1_OFF_MD(n, d, x, y, z)                                                       \
  ((x) + ops_v1.xd[n] * (y) + ops_v1.xd[n] * ops_v1.yd[n] * (z) +                           \
   (d) * ops_v1.xd[n] * ops_v1.yd[n] * ops_v1.zd[n])
#else
#define OPS_V1_OFF_MD(n, d, x, y, z)                                                       \
  ((x) * ops_v1.md[n] + (d) + ops_v1.xd[n] * (y) * ops_v1.md[n] +                           \
   ops_v1.xd[n] * ops_v1.yd[n] * (z) * ops_v1.md[n])
#endif
#endif

#define OPS_ACC0(...) OPS_V1_OFF(0, __VA_ARGS__)
#define OPS_ACC1(...) OPS_V1_OFF(1, __VA_ARGS__)
#define OPS_ACC2(...) OPS_V1_OFF(2, __VA_ARGS__)
#define OPS_ACC3(...) OPS_V1_OFF(3, __VA_ARGS__)
#define OPS_ACC4(...) OPS_V1_OFF(4, __VA_ARGS__)
#define OPS_ACC5(...) OPS_V1_OFF(5, __VA_ARGS__)
#define OPS_ACC6(...) OPS_V1_OFF(6, __VA_ARGS__)
#define OPS_ACC7(...) OPS_V1_OFF(7, __VA_ARGS__)
#define OPS_ACC8(...) OPS_V1_OFF(8, __VA_ARGS__)
#define OPS_ACC9(...) OPS_V1_OFF(9, __VA_ARGS__)
#define OPS_ACC10(...) OPS_V1_OFF(10, __VA_ARGS__)
#define OPS_ACC11(...) OPS_V1_OFF(11, __VA_ARGS__)
#define OPS_ACC12(...) OPS_V1_OFF(12, __VA_ARGS__)
#define OPS_ACC13(...) OPS_V1_OFF(13, __VA_ARGS__)
#define OPS_ACC14(...) OPS_V1_OFF(14, __VA_ARGS__)
#define OPS_ACC15(...) OPS_V1_OFF(15, __VA_ARGS__)
#define OPS_ACC16(...) OPS_V1_OFF(16, __VA_ARGS__)
#define OPS_ACC17(...) OPS_V1_OFF(17, __VA_ARGS__)
#define OPS_ACC18(...) OPS_V1_OFF(18, __VA_ARGS__)
#define OPS_ACC19(...) OPS_V1_OFF(19, __VA_ARGS__)
#define OPS_ACC20(...) OPS_V1_OFF(20, __VA_ARGS__)
#define OPS_ACC21(...) OPS_V1_OFF(21, __VA_ARGS__)
#define OPS_ACC22(...) OPS_V1_OFF(22, __VA_ARGS__)
#define OPS_ACC23(...) OPS_V1_OFF(23, __VA_ARGS__)
#define OPS_ACC24(...) OPS_V1_OFF(24, __VA_ARGS__)
#define OPS_ACC25(...) OPS_V1_OFF(25, __VA_ARGS__)
#define OPS_ACC26(...) OPS_V1_OFF(26, __VA_ARGS__)
#define OPS_ACC27(...) OPS_V1_OFF(27, __VA_ARGS__)
#define OPS_ACC28(...) OPS_V1_OFF(28, __VA_ARGS__)
#define OPS_ACC29(...) OPS_V1_OFF(29, __VA_ARGS__)
#define OPS_ACC30(...) OPS_V1_OFF(30, __VA_ARGS__)
#define OPS_ACC31(...) OPS_V1_OFF(31, __VA_ARGS__)

#define OPS_ACC_MD0(...) OPS_V1_OFF_MD(0, __VA_ARGS__)
#define OPS_ACC_MD1(...) OPS_V1_OFF_MD(1, __VA_ARGS__)
#define OPS_ACC_MD2(...) OPS_V1_OFF_MD(2, __VA_ARGS__)
#define OPS_ACC_MD3(...) OPS_V1_OFF_MD(3, __VA_ARGS__)
#define OPS_ACC_MD4(...) OPS_V1_OFF_MD(4, __VA_ARGS__)
#define OPS_ACC_MD5(...) OPS_V1_OFF_MD(5, __VA_ARGS__)
#define OPS_ACC_MD6(...) OPS_V1_OFF_MD(6, __VA_ARGS__)
#define OPS_ACC_MD7(...) OPS_V1_OFF_MD(7, __VA_ARGS__)
#define OPS_ACC_MD8(...) OPS_V1_OFF_MD(8, __VA_ARGS__)
#define OPS_ACC_MD9(...) OPS_V1_OFF_MD(9, __VA_ARGS__)
#define OPS_ACC_MD10(...) OPS_V1_OFF_MD(10, __VA_ARGS__)
#define OPS_ACC_MD11(...) OPS_V1_OFF_MD(11, __VA_ARGS__)
#define OPS_ACC_MD12(...) OPS_V1_OFF_MD(12, __VA_ARGS__)
#define OPS_ACC_MD13(...) OPS_V1_OFF_MD(13, __VA_ARGS__)
#define OPS_ACC_MD14(...) OPS_V1_OFF_MD(14, __VA_ARGS__)
#define OPS_ACC_MD15(...) OPS_V1_OFF_MD(15, __VA_ARGS__)
#define OPS_ACC_MD16(...) OPS_V1_OFF_MD(16, __VA_ARGS__)
#define OPS_ACC_MD17(...) OPS_V1_OFF_MD(17, __VA_ARGS__)
#define OPS_ACC_MD18(...) OPS_V1_OFF_MD(18, __VA_ARGS__)
#define OPS_ACC_MD19(...) OPS_V1_OFF_MD(19, __VA_ARGS__)
#define OPS_ACC_MD20(...) OPS_V1_OFF_MD(20, __VA_ARGS__)
#define OPS_ACC_MD21(...) OPS_V1_OFF_MD(21, __VA_ARGS__)
#define OPS_ACC_MD22(...) OPS_V1_OFF_MD(22, __VA_ARGS__)
#define OPS_ACC_MD23(...) OPS_V1_OFF_MD(23, __VA_ARGS__)
#define OPS_ACC_MD24(...) OPS_V1_OFF_MD(24, __VA_ARGS__)
#define OPS_ACC_MD25(...) OPS_V1_OFF_MD(25, __VA_ARGS__)
#define OPS_ACC_MD26(...) OPS_V1_OFF_MD(26, __VA_ARGS__)
#define OPS_ACC_MD27(...) OPS_V1_OFF_MD(27, __VA_ARGS__)
#define OPS_ACC_MD28(...) OPS_V1_OFF_MD(28, __VA_ARGS__)
#define OPS_ACC_MD29(...) OPS_V1_OFF_MD(29, __VA_ARGS__)
#define OPS_ACC_MD30(...) OPS_V1_OFF_MD(30, __VA_ARGS__)
#define OPS_ACC_MD31(...) OPS_V1_OFF_MD(31, __VA_ARGS__)

#endif
