/*
 * OPS public API for the MI355X (gfx950) backend.
 *
 * Source-compatible with the reference's user API (ops/c/include/ops_lib_core.h:183-552
 * for the types, :581-1309 for the functions, :1356-2038 for ACC<T>): applications are
 * recompiled against this header and, after translation of their ops_par_loop calls,
 * linked against libops_hip.so.
 *
 * Every non-template entry point below is exported with C linkage ("extern C") so that
 * the shared library has a plain C ABI (pointers, ints, sizes). The C++ conveniences of
 * the reference (ops_decl_dat<T>, ops_arg_gbl<T>, ops_reduction_result<T>,
 * ops_decl_const<T>, ACC<T>, OPSException) are header-only wrappers over that ABI.
 * Errors are raised as OPSException with the reference's codes
 * (ops/c/include/ops_exceptions.h:46-53).
 */
#ifndef OPS_AMD_LIB_CORE_H
#define OPS_AMD_LIB_CORE_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <float.h>
#include <limits.h>
#include <math.h>

#ifdef __cplusplus
#include <exception>
#include <sstream>
#include <string>
#include <ostream>
#include <iostream>
#include <complex>
#endif

#if defined(__HIPCC__)
#define OPS_HD __host__ __device__
#else
#define OPS_HD
#endif

#ifndef OPS_ALIGNMENT
#define OPS_ALIGNMENT 64
#endif
#define MAX_DEPTH 15
#define OPS_MAX_DIM 5

/* access modes (ops_lib_core.h:80-85 in the reference) */
#define OPS_READ 0
#define OPS_WRITE 1
#define OPS_RW 2
#define OPS_INC 3
#define OPS_MIN 4
#define OPS_MAX 5

#define OPS_ARG_GBL 0
#define OPS_ARG_DAT 1
#define OPS_ARG_IDX 2

#define OPS_HOST 1
#define OPS_DEVICE 2

/* error codes (ops_exceptions.h:46-53) */
#define OPS_NOT_IMPLEMENTED 1
#define OPS_RUNTIME_CONFIGURATION_ERROR 2
#define OPS_INTERNAL_ERROR 3
#define OPS_RUNTIME_ERROR 4
#define OPS_HDF5_ERROR 5
#define OPS_INVALID_ARGUMENT 6

typedef int ops_access;
typedef int ops_arg_type;
typedef int ops_memspace;

#ifdef __cplusplus
typedef std::complex<double> complexd;
typedef std::complex<float> complexf;
#endif

/* reduction identities used by generated code (ops_lib_core.h:99-140) */
#define ZERO_double 0.0;
#define INFINITY_double DBL_MAX;
#define ZERO_float 0.0f;
#define INFINITY_float FLT_MAX;
#define ZERO_int 0;
#define INFINITY_int INT_MAX;
#define ZERO_ll 0;
#define INFINITY_ll LLONG_MAX;

#ifndef MIN
#define MIN(a, b) ((a < b) ? (a) : (b))
#endif
#ifndef MAX
#define MAX(a, b) ((a > b) ? (a) : (b))
#endif
#ifndef MIN3
#define MIN3(a, b, c) MIN(a, MIN(b, c))
#endif
#ifndef MAX3
#define MAX3(a, b, c) MAX(a, MAX(b, c))
#endif
#ifndef SIGN
#define SIGN(a, b) ((b < 0.0) ? (a * (-1)) : (a))
#endif

#ifdef __cplusplus
extern "C" {
#endif

#ifdef __cplusplus
}
/* The part of the reference's OPS_instance (ops/c/include/ops_instance.h:67-377) that
   applications use: the singleton created by ops_init, the diagnostics level (OPS_diags,
   ops_init's diags_level) and the structure-of-arrays flag (OPS_soa, as ops_set_soa) they
   read or set, and the power-counter hook (no counters on this backend). Each backend keeps
   the rest of its state in a type derived from this one. */
struct OPS_instance {
  static OPS_instance *getOPSInstance();
  int OPS_diags = 1;
  int OPS_soa = 0;
  void reset_power_counters() {}
};
extern "C" {
#else
typedef struct OPS_instance OPS_instance;
#endif

/* ---- handle types: same field names and meaning as the reference (ops_lib_core.h:183-520) ---- */

typedef struct ops_block_core {
  int index;
  int dims;
  char const *name;
  OPS_instance *instance;
} ops_block_core;
typedef ops_block_core *ops_block;

typedef struct ops_reduction_core {
  char *data;       /* host copy of the reduced value(s) */
  int size;         /* bytes */
  int initialized;  /* set by the first ops_arg_reduce after a result was read */
  int index;
  ops_access acc;   /* access the handle was last used with */
  char *type;
  char *name;
  OPS_instance *instance;
  char *data_d;     /* device-resident running value (this backend) */
  int dirty_d;      /* 1: device value is newer than data */
  char *data_m;     /* pinned host copy of data_d, written by the same kernel (this backend) */
  unsigned seq_expect; /* finisher launches of the handle outside graph capture */
  int seq_valid;    /* 1: the last finisher launch stores seq_expect after data_m */
} ops_reduction_core;
typedef ops_reduction_core *ops_reduction;

typedef struct ops_dat_core {
  int index;
  ops_block block;
  int dim;                  /* components per grid point */
  int type_size;            /* bytes per component */
  int elem_size;            /* type_size * dim */
  int size[OPS_MAX_DIM];    /* allocated extent per dimension, halos included */
  int base[OPS_MAX_DIM];
  int d_m[OPS_MAX_DIM];     /* negative-side halo depth (<= 0) */
  int d_p[OPS_MAX_DIM];     /* positive-side halo depth (>= 0) */
  int x_pad;
  char *data;               /* host mirror (materialised on demand) */
  char *data_d;             /* HBM allocation */
  char const *name;
  char const *type;
  int dirty_hd;             /* 1: host newer, 2: device newer, 0: in sync */
  int locked_hd;            /* memspace of an outstanding raw pointer, 0 if none */
  int user_managed;
  int is_hdf5;
  char const *hdf5_file;
  int e_dat;                /* 1 if some dimension has extent 1 (edge dat) */
  int edat_acc;             /* access of the last loop that wrote this edge dat: a low-dim
                               reduction (INC/MIN/MAX) with a different access first fills
                               it with that reduction's identity */
  size_t mem;               /* bytes allocated */
  size_t base_offset;       /* bytes from allocation start to loop index 0 */
  int stride[OPS_MAX_DIM];
} ops_dat_core;
typedef ops_dat_core *ops_dat;

typedef struct ops_stencil_core {
  int index;
  int dims;
  int points;
  char const *name;
  int *stencil;       /* points*dims offsets */
  int *stride;        /* per-dimension stride, 0 = dimension not iterated */
  int *mgrid_stride;
  int type;           /* 0 regular, 1 prolong, 2 restrict */
} ops_stencil_core;
typedef ops_stencil_core *ops_stencil;

typedef struct ops_arg {
  ops_dat dat;
  ops_stencil stencil;
  int dim;
  int elem_size;
  char *data;
  char *data_d;
  ops_access acc;
  ops_arg_type argtype;
  int opt;
} ops_arg;

typedef struct ops_halo_core {
  ops_dat from;
  ops_dat to;
  int iter_size[OPS_MAX_DIM];
  int from_base[OPS_MAX_DIM];
  int to_base[OPS_MAX_DIM];
  int from_dir[OPS_MAX_DIM];
  int to_dir[OPS_MAX_DIM];
  int index;
} ops_halo_core;
typedef ops_halo_core *ops_halo;

typedef struct ops_halo_group_core {
  int nhalos;
  ops_halo *halos;
  int index;
  OPS_instance *instance;
} ops_halo_group_core;
typedef ops_halo_group_core *ops_halo_group;

/* ---- C API (ops_lib_core.h:581-1297) ---- */

void ops_init(const int argc, const char *const argv[], const int diags_level);
void ops_exit(void);
void ops_set_soa(const int soa_val);
ops_block ops_decl_block(int dims, const char *name);
ops_dat ops_decl_dat_char(ops_block block, int dim, int *block_size, int *base, int *d_m,
                          int *d_p, int *stride, char *data, int type_size, char const *type,
                          char const *name);
void ops_free_dat(ops_dat dat);
ops_dat ops_dat_copy(ops_dat orig_dat);
void ops_dat_deep_copy(ops_dat target, ops_dat orig_dat);

ops_stencil ops_decl_stencil(int dims, int points, int *stencil, char const *name);
ops_stencil ops_decl_strided_stencil(int dims, int points, int *sten, int *stride,
                                     char const *name);
ops_stencil ops_decl_restrict_stencil(int dims, int points, int *sten, int *stride,
                                      char const *name);
ops_stencil ops_decl_prolong_stencil(int dims, int points, int *sten, int *stride,
                                     char const *name);

ops_arg ops_arg_dat(ops_dat dat, int dim, ops_stencil stencil, char const *type,
                    ops_access acc);
ops_arg ops_arg_dat_opt(ops_dat dat, int dim, ops_stencil stencil, char const *type,
                        ops_access acc, int flag);
ops_arg ops_arg_idx(void);
ops_arg ops_arg_reduce(ops_reduction handle, int dim, const char *type, ops_access acc);
ops_arg ops_arg_gbl_char(char *data, int dim, int size, ops_access acc);

ops_reduction ops_decl_reduction_handle(int size, const char *type, const char *name);
void ops_reduction_result_char(ops_reduction handle, int type_size, char *ptr);

ops_halo ops_decl_halo(ops_dat from, ops_dat to, int *iter_size, int *from_base,
                       int *to_base, int *from_dir, int *to_dir);
ops_halo_group ops_decl_halo_group(int nhalos, ops_halo *halos);
void ops_halo_transfer(ops_halo_group group);

void ops_decl_const_char(OPS_instance *instance, int dim, char const *type, int type_size,
                         char *data, char const *name);

void ops_printf(const char *format, ...);
void ops_fprintf(FILE *stream, const char *format, ...);
void ops_diagnostic_output(void);
void ops_timing_output_stdout(void);
void ops_timers_core(double *cpu, double *et);
void ops_timers(double *cpu, double *et);
void ops_print_dat_to_txtfile(ops_dat dat, const char *file_name);
void ops_get_data(ops_dat dat);
int ops_is_root(void);
void ops_partition(const char *routine);
void ops_execute(OPS_instance *instance);

int ops_dat_get_local_npartitions(ops_dat dat);
void ops_dat_get_extents(ops_dat dat, int part, int *disp, int *size);
size_t ops_dat_get_slab_extents(ops_dat dat, int part, int *disp, int *size, int *slab);
void ops_dat_get_raw_metadata(ops_dat dat, int part, int *disp, int *size, int *stride,
                              int *d_m, int *d_p);
char *ops_dat_get_raw_pointer(ops_dat dat, int part, ops_stencil stencil,
                              ops_memspace *memspace);
void ops_dat_release_raw_data(ops_dat dat, int part, ops_access acc);
void ops_dat_release_raw_data_memspace(ops_dat dat, int part, ops_access acc,
                                       ops_memspace *memspace);
void ops_dat_fetch_data(ops_dat dat, int part, char *data);
void ops_dat_fetch_data_memspace(ops_dat dat, int part, char *data, ops_memspace memspace);
void ops_dat_fetch_data_slab_memspace(ops_dat dat, int part, char *data, int *range,
                                      ops_memspace memspace);
void ops_dat_set_data(ops_dat dat, int part, char *data);
void ops_dat_set_data_memspace(ops_dat dat, int part, char *data, ops_memspace memspace);
void ops_dat_set_data_slab_memspace(ops_dat dat, int part, char *data, int *range,
                                    ops_memspace memspace);
int ops_dat_get_global_npartitions(ops_dat dat);

void ops_randomgen_init(unsigned int seed, int options);
void ops_fill_random_uniform(ops_dat dat);
void ops_fill_random_normal(ops_dat dat);
void ops_randomgen_exit(void);

/* Checkpointing (ops/c/include/ops_checkpointing.h:61-181). The reference saves its
   checkpoints as HDF5 files; built without HDF5 (ops/c/src/externlib/ops_checkpointing.cpp:
   1734-1777) every entry point is inert: init reports "not restarting", every loop runs,
   the fast-forward calls report "no checkpoint taken". No HDF5 library is available to this
   backend either, so it exports exactly that behaviour: applications that call these
   (apps/c/CloverLeaf/build_field.cpp:236) link and run unchanged. */
typedef enum {
  OPS_CHECKPOINT_INITPHASE = 1,
  OPS_CHECKPOINT_MANUAL_DATLIST = 2,
  OPS_CHECKPOINT_FASTFW = 4,
  OPS_CHECKPOINT_MANUAL = 8
} ops_checkpoint_options;
#ifndef __cplusplus
#include <stdbool.h>
#endif
bool ops_checkpointing_init(const char *filename, double interval, int options);
void ops_checkpointing_initphase_done(void);
bool ops_checkpointing_before(ops_arg *args, int nargs, int *range, int loop_id);
bool ops_checkpointing_name_before(ops_arg *args, int nargs, int *range, const char *s);
void ops_checkpointing_exit(OPS_instance *instance);
void ops_checkpointing_reduction(ops_reduction red);
void ops_checkpointing_manual_datlist(int ndats, ops_dat *datlist);
bool ops_checkpointing_fastfw(int nbytes, char *payload);
bool ops_checkpointing_manual_datlist_fastfw(int ndats, ops_dat *datlist, int nbytes,
                                             char *payload);
bool ops_checkpointing_manual_datlist_fastfw_trigger(int ndats, ops_dat *datlist, int nbytes,
                                                     char *payload);

/* backend-specific: last error text for C/ctypes callers that cannot catch C++ */
const char *ops_get_last_error(void);

#ifdef __cplusplus
} /* extern "C" */

/* ---- C++ conveniences (header-only, over the C ABI) ---- */

struct OPSException : public std::exception {
  int code;
  std::ostringstream msg;
  mutable std::string text;
  explicit OPSException(int c) : code(c) {}
  OPSException(int c, const char *m) : code(c) { msg << m; }
  OPSException(const OPSException &o) : std::exception(), code(o.code) { msg << o.msg.str(); }
  template <class T> OPSException &operator<<(const T &v) {
    msg << v;
    return *this;
  }
  const char *what() const noexcept override {
    text = msg.str();
    return text.c_str();
  }
};

/* type-name check used by the typed wrappers (ops_internal1.h:105-160 in the reference) */
template <class T> inline bool ops_type_ok(const char *t);
template <> inline bool ops_type_ok<double>(const char *t) {
  return !strcmp(t, "double") || !strcmp(t, "real(8)") || !strcmp(t, "double precision");
}
template <> inline bool ops_type_ok<float>(const char *t) {
  return !strcmp(t, "float") || !strcmp(t, "real") || !strcmp(t, "real(4)");
}
template <> inline bool ops_type_ok<int>(const char *t) {
  return !strcmp(t, "int") || !strcmp(t, "integer") || !strcmp(t, "integer(4)");
}
template <> inline bool ops_type_ok<long long>(const char *t) {
  return !strcmp(t, "long long") || !strcmp(t, "ll");
}
template <class T> inline bool ops_type_ok(const char *) { return true; } /* user types */

template <class T>
ops_dat ops_decl_dat(ops_block block, int data_size, int *block_size, int *base, int *d_m,
                     int *d_p, int *stride, T *data, char const *type, char const *name) {
  if (!ops_type_ok<T>(type)) {
    OPSException ex(OPS_HDF5_ERROR);
    ex << "Error: incorrect type specified for dataset " << name;
    throw ex;
  }
  return ops_decl_dat_char(block, data_size, block_size, base, d_m, d_p, stride, (char *)data,
                           (int)sizeof(T), type, name);
}

template <class T>
ops_dat ops_decl_dat(ops_block block, int data_size, int *block_size, int *base, int *d_m,
                     int *d_p, T *data, char const *type, char const *name) {
  int stride[OPS_MAX_DIM] = {1, 1, 1, 1, 1};
  return ops_decl_dat(block, data_size, block_size, base, d_m, d_p, stride, data, type, name);
}

template <class T> ops_arg ops_arg_gbl(T *data, int dim, char const *type, ops_access acc) {
  (void)type;
  return ops_arg_gbl_char((char *)data, dim, (int)sizeof(T), acc);
}

template <class T> void ops_reduction_result(ops_reduction handle, T *ptr) {
  if (!ops_type_ok<T>(handle->type)) {
    OPSException ex(OPS_HDF5_ERROR);
    ex << "Error: incorrect type specified for constant " << handle->name
       << " in ops_reduction_result";
    throw ex;
  }
  if (!handle->initialized) {
    OPSException ex(OPS_INVALID_ARGUMENT);
    ex << "Error: ops_reduction_result called for " << handle->name
       << " but the handle was not previously used in a reduction since the last "
          "ops_reduction_result call.";
    throw ex;
  }
  ops_reduction_result_char(handle, (int)sizeof(T), (char *)ptr);
}

template <class T> void ops_decl_const(char const *name, int dim, char const *type, T *data) {
  if (!ops_type_ok<T>(type)) {
    OPSException ex(OPS_HDF5_ERROR);
    ex << "Error: incorrect type specified for constant " << name << " in ops_decl_const";
    throw ex;
  }
  ops_decl_const_char(NULL, dim, type, (int)sizeof(T), (char *)data, name);
}

template <class T> void ops_update_const(char const *name, int dim, char const *type, T *data) {
  ops_decl_const(name, dim, type, data);
}

/* the C core cannot hand a std::ostream across the ABI; the table goes to stdout */
inline void ops_timing_output(std::ostream &stream) {
  (void)stream;
  ops_timing_output_stdout();
}

/*
 * ACC<T>: relative-offset accessor handed to user kernels (ops_lib_core.h:1356-2038).
 * The dimensionality is fixed by OPS_1D/OPS_2D/OPS_3D exactly as in the reference, and
 * multi-component dats are addressed AoS, ACC(d, x, y), unless OPS_SOA is defined.
 * The same class is compiled for the host and for gfx950.
 */
template <typename T> class ACC {
public:
#if defined(OPS_1D)
  OPS_HD ACC(T *p) : ptr(p) {}
  OPS_HD ACC(int mdim_, int sx, T *p) : mdim(mdim_), sizex(sx), ptr(p) {}
  OPS_HD const T &operator()(int x) const { return ptr[x]; }
  OPS_HD T &operator()(int x) { return ptr[x]; }
  OPS_HD const T &operator()(int d, int x) const { return ptr[off(d) + x * step()]; }
  OPS_HD T &operator()(int d, int x) { return ptr[off(d) + x * step()]; }
#elif defined(OPS_2D)
  OPS_HD ACC(int sx, T *p) : sizex(sx), ptr(p) {}
  OPS_HD ACC(int mdim_, int sx, int sy, T *p) : mdim(mdim_), sizex(sx), sizey(sy), ptr(p) {}
  OPS_HD const T &operator()(int x, int y) const { return ptr[x + y * sizex]; }
  OPS_HD T &operator()(int x, int y) { return ptr[x + y * sizex]; }
  OPS_HD const T &operator()(int d, int x, int y) const {
    return ptr[off(d) + (x + y * sizex) * step()];
  }
  OPS_HD T &operator()(int d, int x, int y) { return ptr[off(d) + (x + y * sizex) * step()]; }
#elif defined(OPS_3D)
  OPS_HD ACC(int sx, int sy, T *p) : sizex(sx), sizey(sy), ptr(p) {}
  OPS_HD ACC(int mdim_, int sx, int sy, int sz, T *p)
      : mdim(mdim_), sizex(sx), sizey(sy), sizez(sz), ptr(p) {}
  OPS_HD const T &operator()(int x, int y, int z) const {
    return ptr[x + sizex * (y + sizey * z)];
  }
  OPS_HD T &operator()(int x, int y, int z) { return ptr[x + sizex * (y + sizey * z)]; }
  OPS_HD const T &operator()(int d, int x, int y, int z) const {
    return ptr[off(d) + (x + sizex * (y + sizey * z)) * step()];
  }
  OPS_HD T &operator()(int d, int x, int y, int z) {
    return ptr[off(d) + (x + sizex * (y + sizey * z)) * step()];
  }
#endif
  // low-dimensional reductions (ops_lib_core.h:1396-1900 of the reference): many loop
  // points combine into one element of a dat with size 1 in the collapsed dimensions;
  // atomics on the device, plain updates on the host (the host backend runs such loops on
  // one thread)
#if defined(OPS_1D)
  OPS_HD void combine_max(int x, const T v) { cmax((*this)(x), v); }
  OPS_HD void combine_min(int x, const T v) { cmin((*this)(x), v); }
  OPS_HD void combine_inc(int x, const T v) { cinc((*this)(x), v); }
#elif defined(OPS_2D)
  OPS_HD void combine_max(int x, int y, const T v) { cmax((*this)(x, y), v); }
  OPS_HD void combine_min(int x, int y, const T v) { cmin((*this)(x, y), v); }
  OPS_HD void combine_inc(int x, int y, const T v) { cinc((*this)(x, y), v); }
#elif defined(OPS_3D)
  OPS_HD void combine_max(int x, int y, int z, const T v) { cmax((*this)(x, y, z), v); }
  OPS_HD void combine_min(int x, int y, int z, const T v) { cmin((*this)(x, y, z), v); }
  OPS_HD void combine_inc(int x, int y, int z, const T v) { cinc((*this)(x, y, z), v); }
#endif
  OPS_HD void next(int offset) { ptr += offset; }
  OPS_HD T *raw() const { return ptr; }

private:
#if defined(__HIP_DEVICE_COMPILE__)
  __device__ static void cmax(T &x, T v) { atomicMax(&x, v); }
  __device__ static void cmin(T &x, T v) { atomicMin(&x, v); }
  __device__ static void cinc(T &x, T v) { atomicAdd(&x, v); }
#else
  OPS_HD static void cmax(T &x, T v) { if (v > x) x = v; }
  OPS_HD static void cmin(T &x, T v) { if (v < x) x = v; }
  OPS_HD static void cinc(T &x, T v) { x += v; }
#endif
#ifdef OPS_SOA
  OPS_HD int step() const { return 1; }
  OPS_HD long off(int d) const { return (long)d * plane(); }
#else
  OPS_HD int step() const { return mdim; }
  OPS_HD long off(int d) const { return d; }
#endif
  OPS_HD long plane() const {
    long p = sizex;
#if defined(OPS_2D) || defined(OPS_3D)
    p *= sizey;
#endif
#if defined(OPS_3D)
    p *= sizez;
#endif
    return p;
  }
  int mdim = 1;
  int sizex = 1;
  int sizey = 1;
  int sizez = 1;
  T *ptr;
};

#endif /* __cplusplus */

/* the reference's core header brings these in as well (ops_util.h, ops_hdf5.h) */
#include "ops_util.h"
#include "ops_hdf5.h"

#endif /* OPS_AMD_LIB_CORE_H */
