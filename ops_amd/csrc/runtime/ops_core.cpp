// libops_hip.so host runtime: declarations, arguments, reduction handles, consts,
// timing, and the data-access API. Device work is delegated to ops_device.hip.
//
// Behaviour follows the reference's single-node device build:
//   declarations / metadata   ops/c/src/core/ops_lib_core.cpp:376-738
//   arguments / reductions    ops/c/src/core/ops_lib_core.cpp:840-1149
//   argv flags                ops/c/src/core/ops_lib_core.cpp:100-249
//   device dat allocation     ops/c/src/core/ops_device_singlenode_common.cpp:45-106
//   data access               ops/c/src/core/ops_dummy_singlenode.cpp:626-860,
//                             ops/c/src/externlib/ops_util.cpp:353-500
//   timing / bytes            ops/c/src/core/ops_lib_core.cpp:1468-1697
#include <sys/time.h>
#include <time.h>

#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ops_internal.h"
#include "ops_macros.h"

// host-side home of the API v1 accessor sizes: translated kernels receive their sizes as
// a kernel parameter, so this is only referenced by the host copies of user kernels
struct ops_v1_shape ops_v1;

static ops_runtime_instance *g_inst = nullptr;
static std::string g_last_error;
static std::mt19937 g_rng;

std::vector<ops_kernel_entry> &ops_kernel_table() {
  static std::vector<ops_kernel_entry> t;
  return t;
}
std::map<const void *, ops_const_entry> &ops_const_table() {
  static std::map<const void *, ops_const_entry> t;
  return t;
}

void ops_throw(int code, const std::string &msg) {
  g_last_error = msg;
  throw OPSException(code, msg.c_str());
}

extern "C" const char *ops_get_last_error(void) { return g_last_error.c_str(); }

// The reference addresses dats without a range check (ops_seq_v2.h:55-62); on the GPU an
// access outside the allocation is a silent corruption or a memory fault, so every loop
// checks, on the host, the extent its stencils reach in each dimension.
bool ops_loop_arg_in_bounds(const char *name, int dim, const int *start, const int *end,
                            const ops_arg &a, std::string *why) {
  if (a.argtype != OPS_ARG_DAT || a.dat == nullptr || a.stencil == nullptr) return true;
  const ops_stencil s = a.stencil;
  const ops_dat dat = a.dat;
  for (int d = 0; d < dim && d < dat->block->dims; d++) {
    int stride = d < s->dims ? s->stride[d] : 1;
    // multigrid stencils: restrict reads index * m, prolong index / m (C division)
    const int m = (s->type != 0 && d < s->dims) ? s->mgrid_stride[d] : 1;
    int pmin = 0, pmax = 0;
    if (d < s->dims)
      for (int p = 0; p < s->points; p++) {
        int v = s->stencil[p * s->dims + d];
        if (p == 0 || v < pmin) pmin = v;
        if (p == 0 || v > pmax) pmax = v;
      }
    long off = -(long)dat->base[d] - dat->d_m[d];
    long lo = (long)start[d] * stride + pmin + off;
    long hi = (long)(end[d] - 1) * stride + pmax + off;
    if (s->type == 2) {
      lo = (long)start[d] * stride * m + pmin + off;
      hi = (long)(end[d] - 1) * stride * m + pmax + off;
    } else if (s->type == 1) {
      lo = (long)(start[d] / m) * stride + pmin + off;
      hi = (long)((end[d] - 1) / m) * stride + pmax + off;
    }
    if (stride == 0) {
      lo = pmin + off;
      hi = pmax + off;
    }
    if (lo < 0 || hi >= dat->size[d]) {
      if (why) {
        char buf[512];
        snprintf(buf, sizeof buf,
                 "Error: ops_par_loop %s: dat %s is addressed at %ld..%ld in dimension %d, "
                 "outside its allocation 0..%d (size %d, d_m %d, d_p %d)",
                 name ? name : "", dat->name, lo, hi, d, dat->size[d] - 1, dat->size[d],
                 dat->d_m[d], dat->d_p[d]);
        *why = buf;
      }
      return false;
    }
  }
  return true;
}

void ops_check_loop_bounds(const char *name, int dim, const int *start, const int *end,
                           const ops_arg *args, int nargs) {
  std::string why;
  for (int i = 0; i < nargs; i++)
    if (args[i].opt && !ops_loop_arg_in_bounds(name, dim, start, end, args[i], &why))
      ops_throw(OPS_INVALID_ARGUMENT, why);
}

ops_runtime_instance *ops_instance_get(bool create) {
  if (g_inst == nullptr) {
    if (!create) ops_throw(OPS_RUNTIME_ERROR, "Error: OPS used before ops_init");
    g_inst = new ops_runtime_instance();
  }
  return g_inst;
}

OPS_instance *OPS_instance::getOPSInstance() { return ops_instance_get(); }

static char *copy_str(const char *s) {
  if (s == nullptr) return nullptr;
  size_t n = strlen(s);
  char *d = (char *)malloc(n + 1);
  memcpy(d, s, n + 1);
  return d;
}

// ---------------------------------------------------------------- init / exit

static void parse_arg(ops_runtime_instance *in, const char *a) {
  const char *p;
  if ((p = strstr(a, "OPS_BLOCK_SIZE_X="))) in->block_size_x = atoi(p + 17);
  if ((p = strstr(a, "OPS_BLOCK_SIZE_Y="))) in->block_size_y = atoi(p + 17);
  if ((p = strstr(a, "OPS_BLOCK_SIZE_Z="))) in->block_size_z = atoi(p + 17);
  if ((p = strstr(a, "-OPS_DIAGS="))) in->OPS_diags = atoi(p + 11);
  if ((p = strstr(a, "OPS_DEVICE="))) in->device_id = atoi(p + 11);
  if ((p = strstr(a, "OPS_DUMP_DIR="))) in->dump_dir = std::string(p + 13);
  if ((p = strstr(a, "OPS_HIP_TIMING="))) in->timing = atoi(p + 15);
  if ((p = strstr(a, "OPS_AUTOTUNE="))) g_autotune = atoi(p + 13);
  // other reference tokens (OPS_TILING, OPS_CACHE_SIZE, OPS_FORCE_DECOMP_*, checkpointing)
  // are accepted and ignored
}

extern "C" void ops_init(const int argc, const char *const argv[], const int diags_level) {
  if (g_inst != nullptr) ops_exit();
  ops_runtime_instance *in = ops_instance_get(true);
  in->OPS_diags = diags_level;
  const char *env = getenv("OPS_DUMP_DIR");
  if (env) in->dump_dir = env;
  // autotuning is per session: off unless this ops_init's environment or argv asks
  // (shapes already decided stay decided for later sessions that tune)
  g_autotune = 0;
  const char *aenv = getenv("OPS_AUTOTUNE");
  if (aenv) g_autotune = atoi(aenv);
  const char *tenv = getenv("OPS_HIP_TIMING");
  if (tenv) in->timing = atoi(tenv);
  for (int i = 1; i < argc; i++)
    if (argv && argv[i]) parse_arg(in, argv[i]);
  for (int d = 1; d <= OPS_MAX_DIM; d++) {
    int zero[OPS_MAX_DIM] = {0, 0, 0, 0, 0};
    in->internal0[d] = ops_decl_stencil(d, 1, zero, "OPS_internal_0");
  }
  ops_dev_init(in);
  ops_hip_set_timing(in->timing);
}

static void free_dat(ops_dat d) {
  if (d->data_d) ops_dev_free(d->data_d);
  if (!d->user_managed && d->data) free(d->data);
  free((void *)d->name);
  free((void *)d->type);
  delete d;
}

extern "C" void ops_exit(void) {
  if (g_inst == nullptr) return;
  ops_runtime_instance *in = g_inst;
  ops_dev_sync();
  if (!in->dump_dir.empty()) ops_dump_all_dats(in, in->dump_dir.c_str());
  for (ops_dat d : in->dats)
    if (d) free_dat(d);
  for (ops_reduction r : in->reductions) {
    if (r->data_d) ops_dev_free(r->data_d);
    ops_dev_host_free(r->data_m);
    free(r->data);
    free(r->name);
    free(r->type);
    free(r);
  }
  for (ops_stencil s : in->stencils) {
    free(s->stencil);
    free(s->stride);
    free(s->mgrid_stride);
    free((void *)s->name);
    free(s);
  }
  for (ops_halo h : in->halos) free(h);
  for (ops_halo_group g : in->groups) {
    free(g->halos);
    free(g);
  }
  for (ops_block b : in->blocks) {
    free((void *)b->name);
    free(b);
  }
  for (auto &kv : ops_const_table()) {
    if (kv.second.owned) ops_dev_free(kv.second.owned);
    kv.second.owned = nullptr;
    kv.second.owned_bytes = 0;
  }
  ops_dev_exit(in);
  delete in;
  g_inst = nullptr;
}

extern "C" void ops_set_soa(const int soa_val) { ops_instance_get()->OPS_soa = soa_val; }

extern "C" void ops_partition(const char *routine) { (void)routine; }
extern "C" int ops_is_root(void) { return 1; }
extern "C" void ops_execute(OPS_instance *instance) { (void)instance; }

// ---------------------------------------------------------------- declarations

extern "C" ops_block ops_decl_block(int dims, const char *name) {
  ops_runtime_instance *in = ops_instance_get();
  if (dims < 1 || dims > OPS_MAX_DIM)
    ops_throw(OPS_INVALID_ARGUMENT, "Error: ops_decl_block -- dims must be between 1 and 5");
  ops_block b = (ops_block)calloc(1, sizeof(ops_block_core));
  b->index = (int)in->blocks.size();
  b->dims = dims;
  b->name = copy_str(name);
  b->instance = in;
  in->blocks.push_back(b);
  return b;
}

// metadata rules of ops_dat_init_metadata_core (ops_lib_core.cpp:468-560)
static void init_dat_metadata(ops_dat dat, ops_block block, int dim, const int *block_size,
                              const int *base, const int *d_m, const int *d_p,
                              const int *stride, int type_size, const char *type,
                              const char *name) {
  if (dim <= 0) {
    ops_throw(OPS_INVALID_ARGUMENT,
              std::string("Error: ops_decl_dat -- negative/zero number of items per grid "
                          "point in dataset: ") + name);
  }
  dat->block = block;
  dat->dim = dim;
  dat->type_size = type_size;
  dat->elem_size = type_size * dim;
  dat->e_dat = 0;
  dat->edat_acc = OPS_READ;
  for (int n = 0; n < block->dims; n++) {
    if (block_size[n] != 1) {
      dat->size[n] = block_size[n] - d_m[n] + d_p[n];
    } else {
      dat->size[n] = 1;
      dat->e_dat = 1;
    }
  }
  for (int n = 0; n < block->dims; n++) dat->base[n] = (dat->size[n] == 1) ? 0 : base[n];
  for (int n = 0; n < block->dims; n++) {
    if (d_m[n] > 0)
      ops_throw(OPS_INVALID_ARGUMENT,
                std::string("Error: ops_decl_dat -- Non negative d_m during declaration of: ") +
                    name);
    if (d_p[n] < 0)
      ops_throw(OPS_INVALID_ARGUMENT,
                std::string("Error: ops_decl_dat -- Non positive d_p during declaration of: ") +
                    name);
    dat->d_m[n] = d_m[n];
    dat->d_p[n] = d_p[n];
    dat->stride[n] = stride ? stride[n] : 1;
  }
  for (int n = block->dims; n < OPS_MAX_DIM; n++) {
    dat->size[n] = 1;
    dat->base[n] = 0;
    dat->d_m[n] = 0;
    dat->d_p[n] = 0;
    dat->stride[n] = 1;
  }
  dat->x_pad = 0;
  dat->type = copy_str(type);
  dat->name = copy_str(name);
  dat->hdf5_file = "none";
  dat->is_hdf5 = 0;
}

static void compute_base_offset(ops_dat dat, int soa) {
  // bytes from the allocation start to loop index 0 (ops_host_singlenode.cpp:92-100)
  size_t off = 0, cum = 1;
  for (int i = 0; i < dat->block->dims; i++) {
    off += (size_t)(soa ? dat->type_size : dat->elem_size) * cum *
           (size_t)(-dat->base[i] - dat->d_m[i]);
    cum *= (size_t)dat->size[i];
  }
  dat->base_offset = off;
}

static size_t dat_bytes(ops_dat dat) {
  size_t b = (size_t)dat->elem_size;
  for (int i = 0; i < dat->block->dims; i++) b *= (size_t)dat->size[i];
  return b;
}

extern "C" ops_dat ops_decl_dat_char(ops_block block, int dim, int *block_size, int *base,
                                     int *d_m, int *d_p, int *stride, char *data,
                                     int type_size, char const *type, char const *name) {
  ops_runtime_instance *in = ops_instance_get();
  if (block == nullptr) ops_throw(OPS_INVALID_ARGUMENT, "Error: ops_decl_dat -- invalid block");
  ops_dat dat = new ops_dat_core();
  memset((void *)dat, 0, sizeof(ops_dat_core));
  dat->index = (int)in->dats.size();
  init_dat_metadata(dat, block, dim, block_size, base, d_m, d_p, stride, type_size, type, name);
  compute_base_offset(dat, in->OPS_soa);
  size_t bytes = dat_bytes(dat);
  dat->mem = bytes;
  // HBM allocation, zero-filled (ops_device_singlenode_common.cpp:84-104)
  dat->data_d = (char *)ops_dev_malloc(bytes);
  ops_dev_memset(dat->data_d, 0, bytes);
  if (data != nullptr) {
    // user-owned host buffer: the device copy is the working copy
    dat->data = data;
    dat->user_managed = 1;
    ops_dev_h2d(dat->data_d, data, bytes);
  } else {
    dat->data = nullptr;
    dat->user_managed = 0;
  }
  dat->dirty_hd = 0;
  in->dats.push_back(dat);
  return dat;
}

void ops_dat_host_alloc(ops_dat dat) {
  if (dat->data == nullptr) {
    dat->data = (char *)calloc(1, dat->mem);
    if (!dat->data) ops_throw(OPS_RUNTIME_ERROR, "Error: host allocation failed");
    dat->user_managed = 0;
    if (dat->dirty_hd == 0) dat->dirty_hd = 2;  // device holds the data
  }
}

extern "C" void ops_get_data(ops_dat dat) {
  ops_dat_host_alloc(dat);
  if (dat->dirty_hd == 2) {
    ops_dev_d2h(dat->data, dat->data_d, dat->mem);
    dat->dirty_hd = 0;
  }
}

void ops_put_data(ops_dat dat) {
  if (dat->dirty_hd == 1) {
    ops_dev_h2d(dat->data_d, dat->data, dat->mem);
    dat->dirty_hd = 0;
  }
}

extern "C" void ops_free_dat(ops_dat dat) {
  ops_runtime_instance *in = ops_instance_get();
  ops_dev_sync();
  for (auto &d : in->dats)
    if (d == dat) d = nullptr;
  free_dat(dat);
}

extern "C" void ops_dat_deep_copy(ops_dat target, ops_dat orig) {
  if (target->block && target->block->index != orig->block->index)
    ops_throw(OPS_INVALID_ARGUMENT,
              "Error: ops_dat_copy_metadata_core -- the target and source blocks are not "
              "equal.  Cannot copy ops_dats between two different blocks");
  if (target->mem != orig->mem) {
    if (target->data_d) ops_dev_free(target->data_d);
    if (target->data && !target->user_managed) free(target->data);
    target->data = nullptr;
    target->data_d = (char *)ops_dev_malloc(orig->mem);
  }
  const char *nm = target->name, *ty = target->type;
  int idx = target->index;
  char *dd = target->data_d, *hd = target->data;
  int um = target->user_managed;
  *target = *orig;
  target->index = idx;
  target->name = nm ? nm : copy_str(orig->name);
  target->type = ty ? ty : copy_str(orig->type);
  target->data_d = dd;
  target->data = hd;
  target->user_managed = um;
  target->locked_hd = 0;
  if (orig->dirty_hd == 1) ops_put_data(orig);
  ops_dev_d2d(target->data_d, orig->data_d, orig->mem);
  target->dirty_hd = 2;
}

extern "C" ops_dat ops_dat_copy(ops_dat orig) {
  ops_runtime_instance *in = ops_instance_get();
  ops_dat dat = new ops_dat_core();
  memset((void *)dat, 0, sizeof(ops_dat_core));
  dat->index = (int)in->dats.size();
  dat->block = orig->block;
  in->dats.push_back(dat);
  ops_dat_deep_copy(dat, orig);
  return dat;
}

static ops_stencil make_stencil(int dims, int points, const int *sten, const int *stride,
                                const char *name, int type) {
  ops_runtime_instance *in = ops_instance_get();
  ops_stencil s = (ops_stencil)calloc(1, sizeof(ops_stencil_core));
  s->index = (int)in->stencils.size();
  s->dims = dims;
  s->points = points;
  s->name = copy_str(name);
  s->stencil = (int *)malloc(sizeof(int) * (size_t)std::max(1, dims * points));
  memcpy(s->stencil, sten, sizeof(int) * (size_t)dims * points);
  s->stride = (int *)malloc(sizeof(int) * (size_t)dims);
  s->mgrid_stride = (int *)malloc(sizeof(int) * (size_t)dims);
  for (int i = 0; i < dims; i++) {
    s->stride[i] = stride ? stride[i] : 1;
    s->mgrid_stride[i] = 1;
  }
  s->type = type;
  in->stencils.push_back(s);
  return s;
}

extern "C" ops_stencil ops_decl_stencil(int dims, int points, int *sten, char const *name) {
  return make_stencil(dims, points, sten, nullptr, name, 0);
}
extern "C" ops_stencil ops_decl_strided_stencil(int dims, int points, int *sten, int *stride,
                                                char const *name) {
  return make_stencil(dims, points, sten, stride, name, 0);
}
extern "C" ops_stencil ops_decl_restrict_stencil(int dims, int points, int *sten, int *stride,
                                                 char const *name) {
  ops_stencil s = make_stencil(dims, points, sten, nullptr, name, 2);
  for (int i = 0; i < dims; i++) s->mgrid_stride[i] = stride[i];
  return s;
}
extern "C" ops_stencil ops_decl_prolong_stencil(int dims, int points, int *sten, int *stride,
                                                char const *name) {
  ops_stencil s = make_stencil(dims, points, sten, nullptr, name, 1);
  for (int i = 0; i < dims; i++) s->mgrid_stride[i] = stride[i];
  return s;
}

// ---------------------------------------------------------------- arguments

extern "C" ops_arg ops_arg_dat(ops_dat dat, int dim, ops_stencil stencil, char const *type,
                               ops_access acc) {
  (void)type;
  ops_arg a;
  memset(&a, 0, sizeof(a));
  a.argtype = OPS_ARG_DAT;
  a.dat = dat;
  a.stencil = stencil;
  if (acc == OPS_WRITE && stencil->points != 1)
    ops_throw(OPS_INVALID_ARGUMENT,
              "Error: OPS does not support OPS_WRITE arguments with a non (0,0,0) stencil due "
              "to potential race conditions");
  if (dat) {
    a.data = dat->data;
    a.data_d = dat->data_d;
  }
  a.acc = acc;
  a.opt = 1;
  a.dim = dim;
  return a;
}

extern "C" ops_arg ops_arg_dat_opt(ops_dat dat, int dim, ops_stencil stencil, char const *type,
                                   ops_access acc, int flag) {
  ops_arg a = ops_arg_dat(dat, dim, stencil, type, acc);
  a.opt = flag;
  return a;
}

extern "C" ops_arg ops_arg_idx(void) {
  ops_arg a;
  memset(&a, 0, sizeof(a));
  a.argtype = OPS_ARG_IDX;
  return a;
}

extern "C" ops_arg ops_arg_gbl_char(char *data, int dim, int size, ops_access acc) {
  ops_arg a;
  memset(&a, 0, sizeof(a));
  a.argtype = OPS_ARG_GBL;
  a.dim = dim;
  a.data = data;
  a.acc = acc;
  a.elem_size = size;
  return a;
}

template <typename T> static void red_fill(ops_reduction h, T v) {
  for (size_t i = 0; i < (size_t)h->size / sizeof(T); i++) ((T *)h->data)[i] = v;
}

// handle initialisation on first use after a read (ops_lib_core.cpp:840-963)
extern "C" ops_arg ops_arg_reduce(ops_reduction handle, int dim, const char *type,
                                  ops_access acc) {
  ops_arg a;
  memset(&a, 0, sizeof(a));
  a.argtype = OPS_ARG_GBL;
  a.dim = dim;
  a.data = (char *)handle;
  a.acc = acc;
  if (handle->initialized == 0) {
    handle->initialized = 1;
    handle->acc = acc;
    handle->dirty_d = 0;
    if (acc == OPS_INC) memset(handle->data, 0, (size_t)handle->size);
    bool is_d = !strcmp(type, "double") || !strcmp(type, "real(8)") ||
                !strcmp(type, "real(kind=8)") || !strcmp(type, "double precision");
    bool is_f = !strcmp(type, "float") || !strcmp(type, "real") || !strcmp(type, "real(4)") ||
                !strcmp(type, "real(kind=4)");
    bool is_i = !strcmp(type, "int") || !strcmp(type, "int(4)") || !strcmp(type, "integer") ||
                !strcmp(type, "integer(4)") || !strcmp(type, "integer(kind=4)");
    bool is_ll = !strcmp(type, "long long") || !strcmp(type, "ll");
    if (is_d) {
      if (acc == OPS_MIN) red_fill<double>(handle, DBL_MAX);
      if (acc == OPS_MAX) red_fill<double>(handle, -DBL_MAX);
    } else if (is_f) {
      if (acc == OPS_MIN) red_fill<float>(handle, FLT_MAX);
      if (acc == OPS_MAX) red_fill<float>(handle, -FLT_MAX);
    } else if (is_i) {
      if (acc == OPS_MIN) red_fill<int>(handle, INT_MAX);
      if (acc == OPS_MAX) red_fill<int>(handle, INT_MIN);
    } else if (is_ll) {
      if (acc == OPS_MIN) red_fill<long long>(handle, LLONG_MAX);
      if (acc == OPS_MAX) red_fill<long long>(handle, LLONG_MIN);
    } else {
      ops_throw(OPS_NOT_IMPLEMENTED,
                "Error, reduction type not recognised, please add in ops_lib_core.cpp");
    }
  } else if (handle->acc != acc) {
    ops_throw(OPS_INVALID_ARGUMENT, std::string("Error: ops_reduction handle ") + handle->name +
                                        " was already used with a different access type");
  }
  return a;
}

extern "C" ops_reduction ops_decl_reduction_handle(int size, const char *type,
                                                   const char *name) {
  ops_runtime_instance *in = ops_instance_get();
  ops_reduction r = (ops_reduction)calloc(1, sizeof(ops_reduction_core));
  r->size = size;
  r->data = (char *)calloc((size_t)size, 1);
  r->data_d = (char *)ops_dev_malloc((size_t)std::max(size, 8));
  // the finisher writes each result here too, so reading it back needs no copy launch
  // (and, on the next 64-byte line, the number of the launch that wrote it)
  r->data_m = (char *)ops_dev_host_alloc((((size_t)std::max(size, 8) + 63) & ~(size_t)63) + 64);
  r->name = copy_str(name);
  r->type = copy_str(type);
  r->instance = in;
  r->index = (int)in->reductions.size();
  in->reductions.push_back(r);
  return r;
}

extern "C" void ops_reduction_result_char(ops_reduction handle, int type_size, char *ptr) {
  (void)type_size;
  if (handle->dirty_d) {
    if (handle->data_m) {
      // the device value's last writer, the reduction finisher, also wrote the pinned copy
      ops_dev_red_wait(handle);
      memcpy(handle->data, handle->data_m, (size_t)handle->size);
    } else {
      ops_dev_d2h(handle->data, handle->data_d, (size_t)handle->size);
    }
    handle->dirty_d = 0;
  }
  memcpy(ptr, handle->data, (size_t)handle->size);
  handle->initialized = 0;
}

// ---------------------------------------------------------------- ops_util.h, ops_hdf5.h

static void *alloc_or_exit(void *p, size_t size, const char *what) {
  if (p == nullptr && size != 0) {
    fprintf(stderr, "%s: allocation of %zu bytes failed\n", what, size);
    exit(-1);
  }
  return p;
}
extern "C" void *xmalloc(size_t size) { return alloc_or_exit(malloc(size), size, "xmalloc"); }
extern "C" void *xrealloc(void *ptr, size_t size) {
  return alloc_or_exit(realloc(ptr, size), size, "xrealloc");
}
extern "C" void *xcalloc(size_t number, size_t size) {
  return alloc_or_exit(calloc(number, size), number * size, "xcalloc");
}

[[noreturn]] static void no_hdf5(const char *fn) {
  ops_throw(OPS_HDF5_ERROR, std::string(fn) + ": HDF5 I/O is not available (this build has no HDF5 library)");
}
extern "C" ops_dat ops_decl_dat_hdf5(ops_block, int, char const *, char const *, char const *) {
  no_hdf5("ops_decl_dat_hdf5");
}
extern "C" ops_block ops_decl_block_hdf5(int, const char *, char const *) { no_hdf5("ops_decl_block_hdf5"); }
extern "C" ops_stencil ops_decl_stencil_hdf5(int, int, const char *, char const *) {
  no_hdf5("ops_decl_stencil_hdf5");
}
extern "C" ops_stencil ops_decl_strided_stencil_hdf5(int, int, const char *, char const *) {
  no_hdf5("ops_decl_strided_stencil_hdf5");
}
extern "C" ops_halo ops_decl_halo_hdf5(ops_dat, ops_dat, char const *) { no_hdf5("ops_decl_halo_hdf5"); }
extern "C" void ops_fetch_dat_hdf5_file(ops_dat, char const *) { no_hdf5("ops_fetch_dat_hdf5_file"); }
extern "C" void ops_fetch_block_hdf5_file(ops_block, char const *) { no_hdf5("ops_fetch_block_hdf5_file"); }
extern "C" void ops_fetch_stencil_hdf5_file(ops_stencil, char const *) {
  no_hdf5("ops_fetch_stencil_hdf5_file");
}
extern "C" void ops_fetch_halo_hdf5_file(ops_halo, char const *) { no_hdf5("ops_fetch_halo_hdf5_file"); }
extern "C" void ops_read_dat_hdf5(ops_dat) { no_hdf5("ops_read_dat_hdf5"); }
extern "C" void ops_dump_to_hdf5(char const *) { no_hdf5("ops_dump_to_hdf5"); }
extern "C" void ops_write_const_hdf5(char const *, int, char const *, char *, char const *) {
  no_hdf5("ops_write_const_hdf5");
}
extern "C" void ops_get_const_hdf5(char const *, int, char const *, char *, char const *) {
  no_hdf5("ops_get_const_hdf5");
}

// ---------------------------------------------------------------- checkpointing
// The reference's behaviour when built without HDF5 (ops/c/src/externlib/
// ops_checkpointing.cpp:1734-1777): nothing is saved or restored, every loop runs.

extern "C" bool ops_checkpointing_init(const char *, double, int) { return false; }
extern "C" void ops_checkpointing_initphase_done(void) {}
extern "C" bool ops_checkpointing_before(ops_arg *, int, int *, int) { return true; }
extern "C" bool ops_checkpointing_name_before(ops_arg *, int, int *, const char *) {
  return true;
}
extern "C" void ops_checkpointing_exit(OPS_instance *) {}
// the reference stub completes the reduction (ops_execute_reduction); here a reduction's
// value is on the device until it is read, so this brings it to the host handle
extern "C" void ops_checkpointing_reduction(ops_reduction red) {
  if (red && red->dirty_d) {
    ops_dev_d2h(red->data, red->data_d, (size_t)red->size);
    red->dirty_d = 0;
  }
}
extern "C" void ops_checkpointing_manual_datlist(int, ops_dat *) {}
extern "C" bool ops_checkpointing_fastfw(int, char *) { return false; }
extern "C" bool ops_checkpointing_manual_datlist_fastfw(int, ops_dat *, int, char *) {
  return false;
}
extern "C" bool ops_checkpointing_manual_datlist_fastfw_trigger(int, ops_dat *, int, char *) {
  return false;
}

// ---------------------------------------------------------------- halos

// Every index a halo reads or writes must lie inside the dat's allocation: the window of
// dat dimension i starts at base[i] - d_m[i] - dat->base[i] and spans
// iter_size[|dir[i]| - 1] cells, whichever way it is walked.
static void check_halo_side(ops_dat dat, const int *base, const int *dir, const int *iter,
                            const char *which) {
  const int nd = dat->block->dims;
  for (int i = 0; i < nd; i++) {
    const int ax = abs(dir[i]) - 1;
    if (ax < 0 || ax >= nd)
      ops_throw(OPS_INVALID_ARGUMENT,
                std::string("Error: ops_decl_halo -- invalid direction for ") + which + " dat " +
                    dat->name);
    const long first = (long)base[i] - dat->d_m[i] - dat->base[i];
    const long last = first + iter[ax] - 1;
    if (iter[ax] < 0 || first < 0 || last >= dat->size[i]) {
      char buf[400];
      snprintf(buf, sizeof buf,
               "Error: ops_decl_halo -- the %s window of dat %s spans %ld..%ld in dimension %d, "
               "outside its allocation 0..%d",
               which, dat->name, first, last, i, dat->size[i] - 1);
      ops_throw(OPS_INVALID_ARGUMENT, buf);
    }
  }
}

extern "C" ops_halo ops_decl_halo(ops_dat from, ops_dat to, int *iter_size, int *from_base,
                                  int *to_base, int *from_dir, int *to_dir) {
  ops_runtime_instance *in = ops_instance_get();
  ops_halo h = (ops_halo)calloc(1, sizeof(ops_halo_core));
  h->index = (int)in->halos.size();
  h->from = from;
  h->to = to;
  int nd = from->block->dims;
  for (int i = 0; i < nd; i++) {
    h->iter_size[i] = iter_size[i];
    h->from_base[i] = from_base[i];
    h->to_base[i] = to_base[i];
    h->from_dir[i] = from_dir[i];
    h->to_dir[i] = to_dir[i];
  }
  for (int i = nd; i < OPS_MAX_DIM; i++) {
    h->iter_size[i] = 1;
    h->from_base[i] = 0;
    h->to_base[i] = 0;
    h->from_dir[i] = i + 1;
    h->to_dir[i] = i + 1;
  }
  check_halo_side(from, h->from_base, h->from_dir, h->iter_size, "from");
  check_halo_side(to, h->to_base, h->to_dir, h->iter_size, "to");
  in->halos.push_back(h);
  return h;
}

extern "C" ops_halo_group ops_decl_halo_group(int nhalos, ops_halo *halos) {
  ops_runtime_instance *in = ops_instance_get();
  ops_halo_group g = (ops_halo_group)calloc(1, sizeof(ops_halo_group_core));
  g->nhalos = nhalos;
  g->halos = (ops_halo *)calloc((size_t)std::max(1, nhalos), sizeof(ops_halo));
  memcpy(g->halos, halos, sizeof(ops_halo) * (size_t)nhalos);
  g->instance = in;
  g->index = (int)in->groups.size();
  in->groups.push_back(g);
  return g;
}

extern "C" void ops_halo_transfer(ops_halo_group group) {
  for (int h = 0; h < group->nhalos; h++) {
    ops_halo halo = group->halos[h];
    ops_put_data(halo->from);
    ops_put_data(halo->to);
    ops_dev_halo_copy(halo);
    halo->to->dirty_hd = 2;
  }
}

// ---------------------------------------------------------------- consts

// An application without translated kernels has no device home for its constants: its
// ops_decl_const calls land here and are ignored. Translated applications define their own
// ops_decl_const_char over their symbols (generated code), which calls ops_hip_const_set.
extern "C" void ops_decl_const_char(OPS_instance *instance, int dim, char const *type,
                                    int type_size, char *data, char const *name) {
  (void)instance;
  (void)type;
  (void)dim;
  (void)type_size;
  (void)data;
  (void)name;
}

extern "C" void ops_hip_const_set(const void *symbol, long nbytes, const char *name,
                                  const char *data, long bytes) {
  if (symbol == nullptr) return;
  ops_const_entry &e = ops_const_table()[symbol];
  e.symbol = symbol;
  e.nbytes = nbytes;
  if (e.dev == nullptr) e.dev = ops_dev_symbol_address(e.symbol);
  if (e.nbytes >= 0) {
    if (bytes > e.nbytes)
      ops_throw(OPS_INVALID_ARGUMENT, std::string("Error: ops_decl_const -- constant ") + name +
                                          " is larger than its device declaration");
    ops_dev_stage_h2d(e.dev, data, (size_t)bytes);
  } else {
    if (bytes > e.owned_bytes) {
      ops_dev_sync();
      if (e.owned) ops_dev_free(e.owned);
      e.owned = ops_dev_malloc((size_t)bytes);
      e.owned_bytes = bytes;
      ops_dev_stage_h2d(e.dev, &e.owned, sizeof(void *));
    }
    ops_dev_stage_h2d(e.owned, data, (size_t)bytes);
  }
}

// device address of a fixed-size ops_decl_const symbol, for kernels that update a constant
// on the device (a captured time step cannot upload it from the host)
extern "C" void *ops_hip_const_symbol_ptr(const void *symbol, long nbytes, const char *name) {
  if (symbol == nullptr || nbytes < 0)
    ops_throw(OPS_INVALID_ARGUMENT, std::string("ops_hip_const_device_ptr: no fixed-size "
                                                "device constant named ") + (name ? name : ""));
  ops_const_entry &e = ops_const_table()[symbol];
  e.symbol = symbol;
  e.nbytes = nbytes;
  if (e.dev == nullptr) e.dev = ops_dev_symbol_address(e.symbol);
  return e.dev;
}

// an application without translated kernels has no device constants
extern "C" void *ops_hip_const_device_ptr(const char *name) {
  return ops_hip_const_symbol_ptr(nullptr, 0, name);
}

extern "C" void *ops_hip_reduction_device_ptr(ops_reduction h) { return h->data_d; }

// drop a reduction handle's device value without reading it back: the next loop that
// reduces into the handle starts from that loop's own result, as after
// ops_reduction_result
extern "C" void ops_hip_reduction_discard(ops_reduction h) {
  h->dirty_d = 0;
  h->initialized = 0;
}

// ---------------------------------------------------------------- output / timers

extern "C" void ops_printf(const char *format, ...) {
  va_list ap;
  va_start(ap, format);
  vprintf(format, ap);
  va_end(ap);
}

extern "C" void ops_fprintf(FILE *stream, const char *format, ...) {
  va_list ap;
  va_start(ap, format);
  vfprintf(stream, format, ap);
  va_end(ap);
}

extern "C" void ops_timers_core(double *cpu, double *et) {
  struct timeval t;
  gettimeofday(&t, nullptr);
  *cpu = (double)clock() / CLOCKS_PER_SEC;
  *et = t.tv_sec + t.tv_usec * 1.0e-6;
}

extern "C" void ops_timers(double *cpu, double *et) {
  ops_dev_sync();
  ops_timers_core(cpu, et);
}

extern "C" void ops_diagnostic_output(void) {
  ops_runtime_instance *in = ops_instance_get();
  if (in->OPS_diags <= 1) return;
  printf("\n OPS diagnostic output\n --------------------\n");
  printf("\n block dimension\n -------------------\n");
  for (ops_block b : in->blocks) printf(" %-10s %dD\n", b->name, b->dims);
  printf("\n dats: name, type, size, halo(-), halo(+), base\n");
  double total = 0;
  for (ops_dat d : in->dats) {
    if (!d) continue;
    printf(" %-12s %-8s", d->name, d->type);
    for (int i = 0; i < d->block->dims; i++) printf(" %d", d->size[i]);
    printf("\n");
    total += (double)d->mem;
  }
  printf(" total device memory %.2f MB\n", total / 1e6);
}

extern "C" void ops_timing_output_stdout(void) {
  ops_dev_timing_flush();
  auto &kt = ops_kernel_table();
  printf("\n %-40s %8s %12s %12s %10s\n", "Name", "Count", "Time(s)", "Bytes(GB)", "BW(GB/s)");
  for (auto &k : kt) {
    if (k.calls == 0) continue;
    double t = k.total_ms * 1e-3;
    printf(" %-40s %8ld %12.6f %12.4f %10.1f\n", k.name.c_str(), k.calls, t, k.bytes / 1e9,
           t > 0 ? k.bytes / 1e9 / t : 0.0);
  }
}

// per-dimension reference print format (ops_lib_core.cpp:1277-1464, doubles as %3.10lf)
extern "C" void ops_print_dat_to_txtfile(ops_dat dat, const char *file_name) {
  ops_get_data(dat);
  FILE *fp = fopen(file_name, "a");
  if (!fp) ops_throw(OPS_RUNTIME_ERROR, std::string("can't open file ") + file_name);
  fprintf(fp, "ops_dat: %s\n", dat->name);
  fprintf(fp, "ops_dat dim: %d\n", dat->dim);
  fprintf(fp, "block Dims : %d ", dat->block->dims);
  for (int i = 0; i < dat->block->dims; i++) fprintf(fp, "[%d]", dat->size[i]);
  fprintf(fp, "\n");
  size_t n = dat->mem / (size_t)dat->type_size;
  size_t row = (size_t)dat->size[0] * dat->dim;
  for (size_t i = 0; i < n; i++) {
    if (!strcmp(dat->type, "double"))
      fprintf(fp, " %3.10lf", ((double *)dat->data)[i]);
    else if (!strcmp(dat->type, "float"))
      fprintf(fp, " %3.10f", ((float *)dat->data)[i]);
    else if (!strcmp(dat->type, "int"))
      fprintf(fp, " %d", ((int *)dat->data)[i]);
    if ((i + 1) % row == 0) fprintf(fp, "\n");
  }
  fclose(fp);
}

// full-precision dump of every dat (same format as the oracle's reference dumps)
void ops_dump_all_dats(ops_runtime_instance *in, const char *dir) {
  std::string idx = std::string(dir) + "/index.txt";
  FILE *fi = fopen(idx.c_str(), "w");
  if (!fi) return;
  for (ops_dat d : in->dats) {
    if (!d) continue;
    ops_get_data(d);
    std::string name(d->name);  // some applications pad names with blanks
    while (!name.empty() && name.back() == ' ') name.pop_back();
    fprintf(fi, "%s %d %d %d", name.c_str(), d->dim, d->type_size, d->block->dims);
    for (int i = 0; i < OPS_MAX_DIM; i++) fprintf(fi, " %d", d->size[i]);
    for (int i = 0; i < OPS_MAX_DIM; i++) fprintf(fi, " %d", d->d_m[i]);
    for (int i = 0; i < OPS_MAX_DIM; i++) fprintf(fi, " %d", d->d_p[i]);
    for (int i = 0; i < OPS_MAX_DIM; i++) fprintf(fi, " %d", d->base[i]);
    fprintf(fi, " %d %zu\n", d->x_pad, d->base_offset);
    std::string fn = std::string(dir) + "/" + name + ".bin";
    FILE *fb = fopen(fn.c_str(), "wb");
    if (fb) {
      fwrite(d->data, 1, d->mem, fb);
      fclose(fb);
    }
  }
  fclose(fi);
  // each reduction handle's last value read by ops_reduction_result: name, type, size and
  // its bytes in hex
  std::string rfn = std::string(dir) + "/reductions.txt";
  FILE *fr = fopen(rfn.c_str(), "w");
  if (!fr) return;
  for (ops_reduction r : in->reductions) {
    if (!r) continue;
    fprintf(fr, "%s %s %d", r->name, r->type, r->size);
    for (int b = 0; b < r->size; b++) fprintf(fr, " %02x", (unsigned char)r->data[b]);
    fprintf(fr, "\n");
  }
  fclose(fr);
}

// ---------------------------------------------------------------- data access

extern "C" int ops_dat_get_local_npartitions(ops_dat dat) {
  (void)dat;
  return 1;
}
extern "C" int ops_dat_get_global_npartitions(ops_dat dat) {
  (void)dat;
  return 1;
}

extern "C" void ops_dat_get_extents(ops_dat dat, int part, int *disp, int *size) {
  (void)part;
  if (disp)
    for (int d = 0; d < dat->block->dims; d++) disp[d] = 0;
  if (size)
    for (int d = 0; d < dat->block->dims; d++) size[d] = dat->size[d] + dat->d_m[d] - dat->d_p[d];
}

extern "C" size_t ops_dat_get_slab_extents(ops_dat dat, int part, int *disp, int *size,
                                           int *slab) {
  int sizel[OPS_MAX_DIM], displ[OPS_MAX_DIM];
  ops_dat_get_extents(dat, part, displ, sizel);
  size_t bytes = (size_t)dat->elem_size;
  for (int d = 0; d < dat->block->dims; d++) {
    if (slab[2 * d] < 0 || slab[2 * d + 1] > sizel[d]) {
      OPSException ex(OPS_RUNTIME_ERROR);
      ex << "Error: ops_dat_get_slab_extents() called on " << dat->name
         << " with slab ranges in dimension " << d << ": " << slab[2 * d] << "-"
         << slab[2 * d + 1] << " beyond data size 0-" << sizel[d];
      g_last_error = ex.what();
      throw ex;
    }
    disp[d] = slab[2 * d];
    size[d] = slab[2 * d + 1] - slab[2 * d];
    bytes *= (size_t)size[d];
  }
  return bytes;
}

extern "C" void ops_dat_get_raw_metadata(ops_dat dat, int part, int *disp, int *size,
                                         int *stride, int *d_m, int *d_p) {
  ops_dat_get_extents(dat, part, disp, size);
  if (stride)
    for (int d = 0; d < dat->block->dims; d++) stride[d] = dat->size[d];
  if (d_m)
    for (int d = 0; d < dat->block->dims; d++) d_m[d] = dat->d_m[d];
  if (d_p)
    for (int d = 0; d < dat->block->dims; d++) d_p[d] = dat->d_p[d];
}

// pointer to loop index 0 in the requested memory space (ops_dummy_singlenode.cpp:662-707)
extern "C" char *ops_dat_get_raw_pointer(ops_dat dat, int part, ops_stencil stencil,
                                         ops_memspace *memspace) {
  (void)part;
  (void)stencil;
  if (*memspace == OPS_HOST) {
    ops_get_data(dat);
  } else if (*memspace == OPS_DEVICE) {
    ops_put_data(dat);
  } else {
    if (dat->dirty_hd == 1)
      *memspace = OPS_HOST;
    else
      *memspace = OPS_DEVICE;
  }
  if (*memspace == OPS_DEVICE) ops_dev_sync();
  dat->locked_hd = *memspace;
  return (*memspace == OPS_HOST ? dat->data : dat->data_d) + dat->base_offset;
}

extern "C" void ops_dat_release_raw_data(ops_dat dat, int part, ops_access acc) {
  (void)part;
  if (dat->locked_hd == 0)
    ops_throw(OPS_RUNTIME_ERROR,
              std::string("Error: ops_dat_release_raw_data() called, but with no matching "
                          "ops_dat_get_raw_pointer() beforehand: ") + dat->name);
  if (acc != OPS_READ) dat->dirty_hd = dat->locked_hd;
  dat->locked_hd = 0;
}

extern "C" void ops_dat_release_raw_data_memspace(ops_dat dat, int part, ops_access acc,
                                                  ops_memspace *memspace) {
  (void)part;
  if (dat->locked_hd == 0)
    ops_throw(OPS_RUNTIME_ERROR,
              std::string("Error: ops_dat_release_raw_data_memspace() called, but with no "
                          "matching ops_dat_get_raw_pointer() beforehand: ") + dat->name);
  if (acc != OPS_READ) dat->dirty_hd = *memspace;
  dat->locked_hd = 0;
}

// fetch ranges: [0, extent) of the interior (ops_dummy_singlenode.cpp:735-760)
extern "C" void ops_dat_fetch_data_memspace(ops_dat dat, int part, char *data,
                                            ops_memspace memspace) {
  int disp[OPS_MAX_DIM], size[OPS_MAX_DIM], range[2 * OPS_MAX_DIM];
  ops_dat_get_extents(dat, part, disp, size);
  for (int d = 0; d < dat->block->dims; d++) {
    range[2 * d] = 0;
    range[2 * d + 1] = size[d];
  }
  ops_put_data(dat);
  ops_dev_slab_copy(dat, data, memspace == OPS_DEVICE, range, 0);
}

extern "C" void ops_dat_fetch_data(ops_dat dat, int part, char *data) {
  ops_dat_fetch_data_memspace(dat, part, data, OPS_HOST);
}

extern "C" void ops_dat_fetch_data_slab_memspace(ops_dat dat, int part, char *data, int *range,
                                                 ops_memspace memspace) {
  (void)part;
  ops_put_data(dat);
  ops_dev_slab_copy(dat, data, memspace == OPS_DEVICE, range, 0);
}

extern "C" void ops_dat_set_data_slab_memspace(ops_dat dat, int part, char *data, int *range,
                                               ops_memspace memspace) {
  (void)part;
  ops_put_data(dat);
  ops_dev_slab_copy(dat, data, memspace == OPS_DEVICE, range, 1);
  dat->dirty_hd = 2;
}

// set ranges: [d_m, size - d_p) per dimension, i.e. the buffer carries the halos of the
// negative side twice over (ops_dummy_singlenode.cpp:774-785)
extern "C" void ops_dat_set_data_memspace(ops_dat dat, int part, char *data,
                                          ops_memspace memspace) {
  int range[2 * OPS_MAX_DIM];
  for (int d = 0; d < dat->block->dims; d++) {
    range[2 * d] = dat->d_m[d];
    range[2 * d + 1] = dat->size[d] - dat->d_p[d];
  }
  ops_dat_set_data_slab_memspace(dat, part, data, range, memspace);
}

extern "C" void ops_dat_set_data(ops_dat dat, int part, char *data) {
  ops_dat_set_data_memspace(dat, part, data, OPS_HOST);
}

// ---------------------------------------------------------------- random numbers
// host generator, as the reference's host path (ops_lib_core.cpp:2394-2490): the same
// std::mt19937 stream fills the whole allocation, halos included, and is uploaded

extern "C" void ops_randomgen_init(unsigned int seed, int options) {
  (void)options;
  g_rng.seed(seed);
}

static void fill_random(ops_dat dat, int normal) {
  ops_dev_sync();
  ops_dat_host_alloc(dat);
  size_t n = dat->mem / (size_t)dat->type_size;
  if (!strcmp(dat->type, "double")) {
    if (normal) {
      std::normal_distribution<double> dist(0.0, 1.0);
      for (size_t i = 0; i < n; i++) ((double *)dat->data)[i] = dist(g_rng);
    } else {
      std::uniform_real_distribution<double> dist(0.0, 1.0);
      for (size_t i = 0; i < n; i++) ((double *)dat->data)[i] = dist(g_rng);
    }
  } else if (!strcmp(dat->type, "float")) {
    if (normal) {
      std::normal_distribution<float> dist(0.0f, 1.0f);
      for (size_t i = 0; i < n; i++) ((float *)dat->data)[i] = dist(g_rng);
    } else {
      std::uniform_real_distribution<float> dist(0.0f, 1.0f);
      for (size_t i = 0; i < n; i++) ((float *)dat->data)[i] = dist(g_rng);
    }
  } else if (!strcmp(dat->type, "int") && !normal) {
    std::uniform_int_distribution<int> dist(0, INT_MAX);
    for (size_t i = 0; i < n; i++) ((int *)dat->data)[i] = dist(g_rng);
  } else {
    ops_throw(OPS_RUNTIME_ERROR,
              std::string("Error: random number generation not implemented for data type: ") +
                  dat->type);
  }
  dat->dirty_hd = 1;
  ops_put_data(dat);
}

extern "C" void ops_fill_random_uniform(ops_dat dat) { fill_random(dat, 0); }
extern "C" void ops_fill_random_normal(ops_dat dat) { fill_random(dat, 1); }
extern "C" void ops_randomgen_exit(void) {}
