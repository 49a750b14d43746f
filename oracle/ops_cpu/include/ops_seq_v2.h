/*
 * ORACLE — test infrastructure only. CPU execution of ops_par_loop for the oracle build
 * of the applications (oracle/ops_cpu.cpp supplies the rest of the API on host memory).
 *
 * Restates the reference's sequential loop semantics, ops/c/include/ops_seq_v2.h:55-335:
 *   - ACC views built at the loop start from address() (:55-62): per dimension
 *     elem * cumulative size * (start*stride - base - d_m);
 *   - every index of the half-open range is visited, x fastest;
 *   - ops_arg_gbl READ passes the user's pointer, ops_arg_reduce the handle's running
 *     value (:117-126), ops_arg_idx the loop index (:127-134);
 *   - writes mark the dat host-dirty (:299-301);
 *   - a dat argument is addressed whatever its opt flag (ops_seq.h:184-196 of the
 *     reference computes every dat argument's pointer), unless its stencil would leave
 *     the allocation.
 * Kernels written against API v1 (plain pointer parameters, ops/c/include/ops_seq.h:129-296)
 * receive the element pointer at the current point; the sizes behind their OPS_ACC<n>
 * macros (size[0..2] and the component count of each dat argument) are set in `ops_v1`
 * before the loop.
 * Multigrid stencils follow the reference's generated loops, which the sequential header
 * does not implement (templates/cpp/mpi_openmp/loop_host.cpp.j2:230-237): a restrict
 * argument is addressed at loop index * mgrid stride, a prolong argument at loop index /
 * mgrid stride (C division).
 * The outer loop (y in 2D, y*z in 3D) is split over OpenMP threads, as the reference's
 * mpi_openmp target does (templates/cpp/mpi_openmp/loop_host.cpp.j2:150-330); reductions
 * go to per-thread copies combined in thread order, and with one thread the kernel
 * accumulates straight into the handle exactly as the sequential path does.
 */
#ifndef ORACLE_OPS_SEQ_V2_H
#define ORACLE_OPS_SEQ_V2_H

#include <string.h>

#include <new>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "ops_lib_core.h"
#include "ops_macros.h"
#include "ops_tridiag.h"

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" void ops_cpu_before_loop(ops_arg *args, int nargs);
extern "C" void ops_cpu_check_bounds(int dim, const int *range, const ops_arg *args, int nargs);
extern "C" int ops_cpu_arg_in_bounds(int dim, const int *range, const ops_arg *a);
extern "C" void ops_cpu_after_loop(ops_arg *args, int nargs);
extern "C" int ops_cpu_soa(void);
extern "C" int ops_cpu_lowdim_reduction(const ops_arg *a);

namespace ops_cpu {

constexpr int MAXRED = 64;  // doubles of private reduction storage per argument

struct Ctx {
  int idx[OPS_MAX_DIM];
  double red[MAXRED];
  alignas(16) unsigned char acc[64];  // the ACC<T> view handed to the kernel (an lvalue)
};

template <typename P> struct Param;

// addressing of one dat argument at the loop start (ops_seq_v2.h:55-62 of the reference)
struct DatSt {
  char *base;
  long sx, sy, sz;
  int xd, yd, zd, md;
  int prolong, mg[3], st[3];
};

inline DatSt dat_prepare(const ops_arg &a, const int *start, const int *range, int ndim) {
  DatSt s{};
  if (a.argtype != OPS_ARG_DAT || a.dat == nullptr || a.stencil == nullptr) return s;
  if (!a.opt && !ops_cpu_arg_in_bounds(ndim, range, &a)) return s;
  ops_dat d = a.dat;
  int soa = ops_cpu_soa();
  long esz = soa ? d->type_size : d->elem_size;
  const int type = a.stencil->type;  // 0 regular, 1 prolong, 2 restrict
  int st[3] = {1, 1, 1};
  for (int i = 0; i < 3; i++) s.mg[i] = 1;
  for (int i = 0; i < ndim && i < 3; i++) {
    const int m = (type != 0 && i < a.stencil->dims) ? a.stencil->mgrid_stride[i] : 1;
    st[i] = a.stencil->stride[i] * (type == 2 ? m : 1);
    s.mg[i] = type == 1 ? m : 1;
    s.st[i] = start[i];
  }
  s.prolong = type == 1;
  long off = 0, cum = 1;
  for (int i = 0; i < ndim; i++) {
    const long first = type == 1 ? 0 : (long)start[i] * st[i];
    off += esz * cum * (first - d->base[i] - d->d_m[i]);
    cum *= d->size[i];
  }
  s.base = d->data + off;
  long comp = soa ? 1 : d->dim;
  s.sx = st[0] * comp;
  s.sy = ndim > 1 ? (long)st[1] * d->size[0] * comp : 0;
  s.sz = ndim > 2 ? (long)st[2] * d->size[0] * d->size[1] * comp : 0;
  s.xd = d->size[0];
  s.yd = d->size[1];
  s.zd = d->size[2];
  s.md = d->dim;
  return s;
}

// element pointer of a dat argument at loop-local index lc
template <typename T> inline T *dat_at(const DatSt &s, const int *lc) {
  if (!s.base) return nullptr;
  T *b = (T *)s.base;
  if (s.prolong)
    return b + ((s.st[0] + lc[0]) / s.mg[0]) * s.sx + ((s.st[1] + lc[1]) / s.mg[1]) * s.sy +
           ((s.st[2] + lc[2]) / s.mg[2]) * s.sz;
  return b + (lc[0] * s.sx + lc[1] * s.sy + lc[2] * s.sz);
}

template <typename T> struct Param<ACC<T>> {
  using St = DatSt;
  static St prepare(const ops_arg &a, const int *start, const int *range, int ndim) {
    return dat_prepare(a, start, range, ndim);
  }
  static ACC<T> &get(const St &s, const ops_arg &, Ctx &c, const int *lc, const int *, int) {
    static_assert(sizeof(ACC<T>) <= sizeof(c.acc), "ACC view does not fit its slot");
    T *p = dat_at<T>(s, lc);
#if defined(OPS_1D)
    return *(s.md == 1 ? new (c.acc) ACC<T>(p) : new (c.acc) ACC<T>(s.md, s.xd, p));
#elif defined(OPS_2D)
    return *(s.md == 1 ? new (c.acc) ACC<T>(s.xd, p) : new (c.acc) ACC<T>(s.md, s.xd, s.yd, p));
#else
    return *(s.md == 1 ? new (c.acc) ACC<T>(s.xd, s.yd, p)
                       : new (c.acc) ACC<T>(s.md, s.xd, s.yd, s.zd, p));
#endif
  }
};

// plain pointers: API v1 dats (the pointer at the current point, addressed with the
// OPS_ACC<n> macros; ops_seq.h:129-296 of the reference), gbl/reduce data and ops_arg_idx
template <typename T> struct Param<T *> {
  using St = DatSt;
  static St prepare(const ops_arg &a, const int *start, const int *range, int ndim) {
    return dat_prepare(a, start, range, ndim);
  }
  static T *get(const St &s, const ops_arg &a, Ctx &c, const int *lc, const int *start,
                int direct) {
    if (a.argtype == OPS_ARG_DAT) return dat_at<T>(s, lc);
    if (a.argtype == OPS_ARG_IDX) {
      for (int d = 0; d < 3; d++) c.idx[d] = start[d] + lc[d];
      return (T *)c.idx;
    }
    if (a.acc == OPS_READ) return (T *)a.data;
    if (direct) return (T *)((ops_reduction)a.data)->data;
    return (T *)c.red;
  }
};

template <typename P> using base_t = typename std::remove_cv<typename std::remove_reference<P>::type>::type;
template <typename P> struct Strip { using type = base_t<P>; };
template <typename T> struct Strip<const T *> { using type = T *; };
template <typename T> struct Strip<T *> { using type = T *; };

void red_init(const ops_arg &a, double *buf);
void red_combine(const ops_arg &a, const double *buf);

template <typename... P, size_t... I>
void run(std::index_sequence<I...>, void (*kernel)(P...), ops_block block, int dim, int *range,
         ops_arg *args) {
  constexpr int N = sizeof...(P);
  int start[3] = {0, 0, 0}, cnt[3] = {1, 1, 1};
  for (int d = 0; d < dim && d < 3; d++) {
    start[d] = range[2 * d];
    cnt[d] = range[2 * d + 1] - range[2 * d];
    if (cnt[d] <= 0) return;
  }
  (void)block;
  ops_cpu_check_bounds(dim, range, args, N);
  ops_cpu_before_loop(args, N);
  auto st = std::make_tuple(Param<typename Strip<P>::type>::prepare(args[I], start, range, dim)...);
  // sizes behind the OPS_ACC<n> macros of API v1 kernels (xdim<n> = size[0], ...)
  for (int i = 0; i < N && i < OPS_V1_MAXARGS; i++)
    if (args[i].argtype == OPS_ARG_DAT) {
      ops_v1.xd[i] = args[i].dat->size[0];
      ops_v1.yd[i] = args[i].dat->size[1];
      ops_v1.zd[i] = args[i].dat->size[2];
      ops_v1.md[i] = args[i].dat->dim;
    }
  bool has_red = false;
  for (int i = 0; i < N; i++)
    if (args[i].argtype == OPS_ARG_GBL && args[i].acc != OPS_READ) has_red = true;
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
  if ((long)cnt[1] * cnt[2] < nth) nth = 1;
#endif
  // low-dim reductions combine many points into one element: one thread keeps them in
  // loop order
  for (int i = 0; i < N; i++)
    if (ops_cpu_lowdim_reduction(&args[i])) nth = 1;
  const long outer = (long)cnt[1] * cnt[2];
  if (nth == 1) {
    Ctx c[N > 0 ? N : 1];
    for (long o = 0; o < outer; o++) {
      int lc[3] = {0, (int)(o % cnt[1]), (int)(o / cnt[1])};
      for (lc[0] = 0; lc[0] < cnt[0]; lc[0]++)
        kernel(Param<typename Strip<P>::type>::get(std::get<I>(st), args[I], c[I], lc, start,
                                                   1)...);
    }
  } else {
    std::vector<Ctx> ctx((size_t)nth * (N > 0 ? N : 1));
    if (has_red)
      for (int t = 0; t < nth; t++)
        for (int i = 0; i < N; i++)
          if (args[i].argtype == OPS_ARG_GBL && args[i].acc != OPS_READ)
            red_init(args[i], ctx[(size_t)t * N + i].red);
#pragma omp parallel num_threads(nth)
    {
      int t = 0;
#ifdef _OPENMP
      t = omp_get_thread_num();
#endif
      Ctx *c = &ctx[(size_t)t * (N > 0 ? N : 1)];
#pragma omp for schedule(static)
      for (long o = 0; o < outer; o++) {
        int lc[3] = {0, (int)(o % cnt[1]), (int)(o / cnt[1])};
        for (lc[0] = 0; lc[0] < cnt[0]; lc[0]++)
          kernel(Param<typename Strip<P>::type>::get(std::get<I>(st), args[I], c[I], lc, start,
                                                     0)...);
      }
    }
    if (has_red)
      for (int t = 0; t < nth; t++)
        for (int i = 0; i < N; i++)
          if (args[i].argtype == OPS_ARG_GBL && args[i].acc != OPS_READ)
            red_combine(args[i], ctx[(size_t)t * N + i].red);
  }
  ops_cpu_after_loop(args, N);
}

}  // namespace ops_cpu

template <typename... P, typename... A>
void ops_par_loop(void (*kernel)(P...), char const *name, ops_block block, int dim, int *range,
                  A... a) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel parameters and ops_args differ in number");
  (void)name;
  ops_arg args[sizeof...(A) > 0 ? sizeof...(A) : 1] = {a...};
  ops_cpu::run(std::index_sequence_for<P...>{}, kernel, block, dim, range, args);
}

#endif
