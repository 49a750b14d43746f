// ORACLE — test infrastructure only. Host-memory implementation of the OPS C API used to
// run the applications on the CPU as the checker for the GPU path, and (with OpenMP) as
// the CPU baseline. Never linked into the product.
//
// Restates the reference's sequential single-node behaviour:
//   dat metadata / base offset   ops/c/src/core/ops_lib_core.cpp:468-560,
//                                ops/c/src/sequential/ops_host_singlenode.cpp:42-103
//                                (no x padding here; dumps carry the layout)
//   arguments / reductions       ops/c/src/core/ops_lib_core.cpp:840-1149
//   halo transfer                ops/c/src/sequential/ops_host_singlenode.cpp:172-446
//   fetch / set / raw pointer    ops/c/src/core/ops_dummy_singlenode.cpp:626-860,
//                                ops/c/src/externlib/ops_util.cpp:353-500
//   tridiagonal solves           ops/c/src/tridiag/ops_tridiag.cpp:78-249 with the
//                                textbook Thomas recurrence
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include <random>
#include <limits>
#include <string>
#include <vector>

#define OPS_2D
#include "ops_seq_v2.h"
#include "ops_macros.h"

// API v1 accessor sizes, filled by the loop driver (ops_seq_v2.h) before each loop
struct ops_v1_shape ops_v1;

// the host backend's state of an OPS session (OPS_instance holds OPS_diags and OPS_soa)
struct ops_cpu_instance : OPS_instance {
  std::vector<ops_block> blocks;
  std::vector<ops_dat> dats;
  std::vector<ops_stencil> stencils;
  std::vector<ops_reduction> reds;
  std::vector<ops_halo> halos;
  std::vector<ops_halo_group> groups;
  std::string dump_dir;
};

static ops_cpu_instance *g = nullptr;
static std::string g_err;
static std::mt19937 g_rng;

static void fail(int code, const std::string &m) {
  g_err = m;
  throw OPSException(code, m.c_str());
}

static char *dup(const char *s) {
  if (!s) return nullptr;
  size_t n = strlen(s) + 1;
  char *d = (char *)malloc(n);
  memcpy(d, s, n);
  return d;
}

extern "C" const char *ops_get_last_error(void) { return g_err.c_str(); }
extern "C" int ops_cpu_soa(void) { return g ? g->OPS_soa : 0; }
// a dat argument that reduces into an edge dat along a collapsed dimension
// (ops_dummy_singlenode.cpp:247-252)
extern "C" int ops_cpu_lowdim_reduction(const ops_arg *a) {
  if (a->argtype != OPS_ARG_DAT || !a->opt || !a->dat || !a->stencil || !a->dat->e_dat) return 0;
  if (a->acc != OPS_INC && a->acc != OPS_MIN && a->acc != OPS_MAX) return 0;
  for (int d = 0; d < a->stencil->dims && d < OPS_MAX_DIM; d++)
    if (a->dat->size[d] == 1) return 1;
  return 0;
}

template <typename T> static void fill_identity(ops_dat d, int acc) {
  T v = acc == OPS_INC ? T(0) : (acc == OPS_MAX ? std::numeric_limits<T>::lowest() : std::numeric_limits<T>::max());
  T *p = (T *)d->data;
  for (size_t i = 0; i < d->mem / sizeof(T); i++) p[i] = v;
}

extern "C" void ops_cpu_before_loop(ops_arg *args, int n) {
  for (int i = 0; i < n; i++)
    if (args[i].argtype == OPS_ARG_DAT && args[i].dat && args[i].dat->locked_hd)
      fail(OPS_RUNTIME_ERROR, "ops_par_loop on a dat locked by ops_dat_get_raw_pointer");
  // the first low-dim reduction of a kind since the dat was last written otherwise starts
  // from the reduction's identity over the whole allocation (ops_dummy_singlenode.cpp:248-345)
  for (int i = 0; i < n; i++) {
    if (!ops_cpu_lowdim_reduction(&args[i])) continue;
    ops_dat d = args[i].dat;
    if (d->edat_acc == args[i].acc) continue;
    d->edat_acc = args[i].acc;
    if (!strcmp(d->type, "double")) fill_identity<double>(d, args[i].acc);
    else if (!strcmp(d->type, "float")) fill_identity<float>(d, args[i].acc);
    else if (!strcmp(d->type, "int")) fill_identity<int>(d, args[i].acc);
    else fail(OPS_NOT_IMPLEMENTED, std::string("low-dim reduction into a dat of type ") + d->type);
  }
}
// Extent of one dat argument (loop range plus stencil reach, per dimension): the message
// for the first dimension that leaves the allocation goes to *why.
static bool arg_in_bounds(int dim, const int *range, const ops_arg &a, std::string *why) {
  if (a.argtype != OPS_ARG_DAT || !a.dat || !a.stencil) return true;
  const ops_stencil s = a.stencil;
  for (int d = 0; d < dim && d < a.dat->block->dims; d++) {
    const int st = d < s->dims ? s->stride[d] : 1;
    const int m = (s->type != 0 && d < s->dims) ? s->mgrid_stride[d] : 1;
    int lo_p = 0, hi_p = 0;
    for (int p = 0; d < s->dims && p < s->points; p++) {
      const int v = s->stencil[p * s->dims + d];
      lo_p = (p == 0 || v < lo_p) ? v : lo_p;
      hi_p = (p == 0 || v > hi_p) ? v : hi_p;
    }
    const long shift = -(long)a.dat->base[d] - a.dat->d_m[d];
    long first = (st ? (long)range[2 * d] * st : 0) + lo_p + shift;
    long last = (st ? (long)(range[2 * d + 1] - 1) * st : 0) + hi_p + shift;
    if (s->type == 2) {  // restrict: index * m
      first = (long)range[2 * d] * st * m + lo_p + shift;
      last = (long)(range[2 * d + 1] - 1) * st * m + hi_p + shift;
    } else if (s->type == 1) {  // prolong: index / m
      first = (long)(range[2 * d] / m) * st + lo_p + shift;
      last = (long)((range[2 * d + 1] - 1) / m) * st + hi_p + shift;
    }
    if (first < 0 || last >= a.dat->size[d]) {
      if (why)
        *why = std::string("ops_par_loop: dat ") + a.dat->name +
               " addressed outside its allocation in dimension " + std::to_string(d) + " (" +
               std::to_string(first) + ".." + std::to_string(last) + " of size " +
               std::to_string(a.dat->size[d]) + ", stencil " + s->name + ")";
      return false;
    }
  }
  return true;
}
extern "C" int ops_cpu_arg_in_bounds(int dim, const int *range, const ops_arg *a) {
  return arg_in_bounds(dim, range, *a, nullptr) ? 1 : 0;
}
// Extent check of every active dat argument: an out-of-allocation access raises
// OPS_INVALID_ARGUMENT, as the device runtime does.
extern "C" void ops_cpu_check_bounds(int dim, const int *range, const ops_arg *args, int n) {
  std::string why;
  for (int i = 0; i < n; i++)
    if (args[i].opt && !arg_in_bounds(dim, range, args[i], &why)) fail(OPS_INVALID_ARGUMENT, why);
}
extern "C" void ops_cpu_after_loop(ops_arg *args, int n) {
  for (int i = 0; i < n; i++)
    if (args[i].argtype == OPS_ARG_DAT && args[i].dat && args[i].acc != OPS_READ) {
      // every dat argument, whatever its opt flag (ops_dummy_singlenode.cpp:140-148)
      args[i].dat->dirty_hd = 1;
      if (args[i].opt && args[i].dat->e_dat) args[i].dat->edat_acc = args[i].acc;  // :204-213
    }
}

static void dump_all(const char *dir);

OPS_instance *OPS_instance::getOPSInstance() {
  if (!g) fail(OPS_RUNTIME_ERROR, "Error: OPS used before ops_init");
  return g;
}

extern "C" void ops_init(const int argc, const char *const argv[], const int diags) {
  if (g) ops_exit();
  g = new ops_cpu_instance();
  g->OPS_diags = diags;
  const char *e = getenv("OPS_DUMP_DIR");
  if (e) g->dump_dir = e;
  for (int i = 1; i < argc; i++) {
    const char *p = argv && argv[i] ? strstr(argv[i], "OPS_DUMP_DIR=") : nullptr;
    if (p) g->dump_dir = p + 13;
  }
}

extern "C" void ops_exit(void) {
  if (!g) return;
  if (!g->dump_dir.empty()) dump_all(g->dump_dir.c_str());
  for (ops_dat d : g->dats) {
    if (!d) continue;
    if (!d->user_managed) free(d->data);
    free((void *)d->name);
    free((void *)d->type);
    delete d;
  }
  for (ops_reduction r : g->reds) {
    free(r->data);
    free(r->name);
    free(r->type);
    free(r);
  }
  for (ops_stencil s : g->stencils) {
    free(s->stencil);
    free(s->stride);
    free(s->mgrid_stride);
    free((void *)s->name);
    free(s);
  }
  for (ops_halo h : g->halos) free(h);
  for (ops_halo_group q : g->groups) {
    free(q->halos);
    free(q);
  }
  for (ops_block b : g->blocks) {
    free((void *)b->name);
    free(b);
  }
  delete g;
  g = nullptr;
}

extern "C" void ops_set_soa(const int v) { g->OPS_soa = v; }
extern "C" void ops_partition(const char *) {}
extern "C" int ops_is_root(void) { return 1; }
extern "C" void ops_execute(OPS_instance *) {}

extern "C" ops_block ops_decl_block(int dims, const char *name) {
  ops_block b = (ops_block)calloc(1, sizeof(ops_block_core));
  b->index = (int)g->blocks.size();
  b->dims = dims;
  b->name = dup(name);
  b->instance = g;
  g->blocks.push_back(b);
  return b;
}

extern "C" ops_dat ops_decl_dat_char(ops_block block, int dim, int *bs, int *base, int *d_m,
                                     int *d_p, int *stride, char *data, int type_size,
                                     char const *type, char const *name) {
  if (dim <= 0) fail(OPS_INVALID_ARGUMENT, "ops_decl_dat: dim must be positive");
  ops_dat d = new ops_dat_core();
  memset((void *)d, 0, sizeof(ops_dat_core));
  d->index = (int)g->dats.size();
  d->block = block;
  d->dim = dim;
  d->type_size = type_size;
  d->elem_size = dim * type_size;
  for (int n = 0; n < OPS_MAX_DIM; n++) {
    if (n < block->dims) {
      if (d_m[n] > 0 || d_p[n] < 0) fail(OPS_INVALID_ARGUMENT, "ops_decl_dat: bad halo depth");
      d->size[n] = bs[n] != 1 ? bs[n] - d_m[n] + d_p[n] : 1;
      if (bs[n] == 1) d->e_dat = 1;
      d->base[n] = d->size[n] == 1 ? 0 : base[n];
      d->d_m[n] = d_m[n];
      d->d_p[n] = d_p[n];
      d->stride[n] = stride ? stride[n] : 1;
    } else {
      d->size[n] = 1;
      d->stride[n] = 1;
    }
  }
  size_t bytes = (size_t)d->elem_size;
  for (int n = 0; n < block->dims; n++) bytes *= (size_t)d->size[n];
  d->mem = bytes;
  d->data = (char *)calloc(1, bytes);
  if (data) memcpy(d->data, data, bytes);
  d->user_managed = 0;
  size_t off = 0, cum = 1;
  for (int n = 0; n < block->dims; n++) {
    off += (size_t)(g->OPS_soa ? d->type_size : d->elem_size) * cum * (size_t)(-d->base[n] - d->d_m[n]);
    cum *= (size_t)d->size[n];
  }
  d->base_offset = off;
  d->name = dup(name);
  d->type = dup(type);
  d->hdf5_file = "none";
  g->dats.push_back(d);
  return d;
}

extern "C" void ops_free_dat(ops_dat d) {
  for (auto &x : g->dats)
    if (x == d) x = nullptr;
  free(d->data);
  free((void *)d->name);
  free((void *)d->type);
  delete d;
}

extern "C" void ops_dat_deep_copy(ops_dat t, ops_dat o) {
  if (t->mem != o->mem) {
    free(t->data);
    t->data = (char *)malloc(o->mem);
  }
  char *keep = t->data;
  const char *nm = t->name, *ty = t->type;
  int idx = t->index;
  *t = *o;
  t->data = keep;
  t->index = idx;
  t->name = nm ? nm : dup(o->name);
  t->type = ty ? ty : dup(o->type);
  t->user_managed = 0;
  memcpy(t->data, o->data, o->mem);
}

extern "C" ops_dat ops_dat_copy(ops_dat o) {
  ops_dat d = new ops_dat_core();
  memset((void *)d, 0, sizeof(ops_dat_core));
  d->index = (int)g->dats.size();
  g->dats.push_back(d);
  ops_dat_deep_copy(d, o);
  return d;
}

static ops_stencil mk_sten(int dims, int points, const int *sten, const int *stride,
                           const char *name, int type) {
  ops_stencil s = (ops_stencil)calloc(1, sizeof(ops_stencil_core));
  s->index = (int)g->stencils.size();
  s->dims = dims;
  s->points = points;
  s->name = dup(name);
  s->stencil = (int *)malloc(sizeof(int) * dims * points + 1);
  memcpy(s->stencil, sten, sizeof(int) * dims * points);
  s->stride = (int *)malloc(sizeof(int) * dims);
  s->mgrid_stride = (int *)malloc(sizeof(int) * dims);
  for (int i = 0; i < dims; i++) {
    s->stride[i] = stride ? stride[i] : 1;
    s->mgrid_stride[i] = 1;
  }
  s->type = type;
  g->stencils.push_back(s);
  return s;
}
extern "C" ops_stencil ops_decl_stencil(int dims, int points, int *sten, char const *name) {
  return mk_sten(dims, points, sten, nullptr, name, 0);
}
extern "C" ops_stencil ops_decl_strided_stencil(int dims, int points, int *sten, int *stride,
                                                char const *name) {
  return mk_sten(dims, points, sten, stride, name, 0);
}
extern "C" ops_stencil ops_decl_restrict_stencil(int dims, int points, int *sten, int *stride,
                                                 char const *name) {
  ops_stencil s = mk_sten(dims, points, sten, nullptr, name, 2);
  for (int i = 0; i < dims; i++) s->mgrid_stride[i] = stride[i];
  return s;
}
extern "C" ops_stencil ops_decl_prolong_stencil(int dims, int points, int *sten, int *stride,
                                                char const *name) {
  ops_stencil s = mk_sten(dims, points, sten, nullptr, name, 1);
  for (int i = 0; i < dims; i++) s->mgrid_stride[i] = stride[i];
  return s;
}

extern "C" ops_arg ops_arg_dat(ops_dat dat, int dim, ops_stencil st, char const *, ops_access acc) {
  ops_arg a;
  memset(&a, 0, sizeof a);
  if (acc == OPS_WRITE && st->points != 1)
    fail(OPS_INVALID_ARGUMENT, "OPS_WRITE with a multi-point stencil");
  a.argtype = OPS_ARG_DAT;
  a.dat = dat;
  a.stencil = st;
  a.data = dat ? dat->data : nullptr;
  a.acc = acc;
  a.opt = 1;
  a.dim = dim;
  return a;
}
extern "C" ops_arg ops_arg_dat_opt(ops_dat dat, int dim, ops_stencil st, char const *t,
                                   ops_access acc, int flag) {
  ops_arg a = ops_arg_dat(dat, dim, st, t, acc);
  a.opt = flag;
  return a;
}
extern "C" ops_arg ops_arg_idx(void) {
  ops_arg a;
  memset(&a, 0, sizeof a);
  a.argtype = OPS_ARG_IDX;
  return a;
}
extern "C" ops_arg ops_arg_gbl_char(char *data, int dim, int size, ops_access acc) {
  ops_arg a;
  memset(&a, 0, sizeof a);
  a.argtype = OPS_ARG_GBL;
  a.data = data;
  a.dim = dim;
  a.elem_size = size;
  a.acc = acc;
  return a;
}

static int red_kind(const char *t) {
  if (!strcmp(t, "double")) return 0;
  if (!strcmp(t, "float")) return 1;
  if (!strcmp(t, "int")) return 2;
  fail(OPS_NOT_IMPLEMENTED, "reduction type not recognised");
  return -1;
}

template <typename T> static void fill(char *p, int bytes, T v) {
  for (int i = 0; i < bytes / (int)sizeof(T); i++) ((T *)p)[i] = v;
}

static void identity(int kind, ops_access acc, char *p, int bytes) {
  if (acc == OPS_INC) {
    memset(p, 0, (size_t)bytes);
    return;
  }
  bool mn = acc == OPS_MIN;
  if (kind == 0) fill<double>(p, bytes, mn ? DBL_MAX : -DBL_MAX);
  if (kind == 1) fill<float>(p, bytes, mn ? FLT_MAX : -FLT_MAX);
  if (kind == 2) fill<int>(p, bytes, mn ? INT_MAX : INT_MIN);
}

extern "C" ops_arg ops_arg_reduce(ops_reduction h, int dim, const char *type, ops_access acc) {
  ops_arg a;
  memset(&a, 0, sizeof a);
  a.argtype = OPS_ARG_GBL;
  a.dim = dim;
  a.data = (char *)h;
  a.acc = acc;
  if (!h->initialized) {
    h->initialized = 1;
    h->acc = acc;
    identity(red_kind(type), acc, h->data, h->size);
  } else if (h->acc != acc) {
    fail(OPS_INVALID_ARGUMENT, "reduction handle used with a different access type");
  }
  return a;
}

namespace ops_cpu {
void red_init(const ops_arg &a, double *buf) {
  ops_reduction h = (ops_reduction)a.data;
  identity(red_kind(h->type), a.acc, (char *)buf, h->size);
}
template <typename T> static void comb(ops_access acc, T *h, const T *b, int n) {
  for (int i = 0; i < n; i++) {
    if (acc == OPS_INC) h[i] = h[i] + b[i];
    else if (acc == OPS_MIN) h[i] = b[i] < h[i] ? b[i] : h[i];
    else h[i] = b[i] > h[i] ? b[i] : h[i];
  }
}
void red_combine(const ops_arg &a, const double *buf) {
  ops_reduction h = (ops_reduction)a.data;
  int k = red_kind(h->type);
  if (k == 0) comb<double>(a.acc, (double *)h->data, (const double *)buf, h->size / 8);
  if (k == 1) comb<float>(a.acc, (float *)h->data, (const float *)buf, h->size / 4);
  if (k == 2) comb<int>(a.acc, (int *)h->data, (const int *)buf, h->size / 4);
}
}  // namespace ops_cpu

extern "C" ops_reduction ops_decl_reduction_handle(int size, const char *type, const char *name) {
  ops_reduction r = (ops_reduction)calloc(1, sizeof(ops_reduction_core));
  r->size = size;
  r->data = (char *)calloc((size_t)size, 1);
  r->name = dup(name);
  r->type = dup(type);
  r->index = (int)g->reds.size();
  r->instance = g;
  g->reds.push_back(r);
  return r;
}

extern "C" void ops_reduction_result_char(ops_reduction h, int, char *ptr) {
  memcpy(ptr, h->data, (size_t)h->size);
  h->initialized = 0;
}

extern "C" void ops_decl_const_char(OPS_instance *, int, char const *, int, char *, char const *) {}

extern "C" ops_halo ops_decl_halo(ops_dat from, ops_dat to, int *iter, int *fb, int *tb,
                                  int *fd, int *td) {
  ops_halo h = (ops_halo)calloc(1, sizeof(ops_halo_core));
  h->index = (int)g->halos.size();
  h->from = from;
  h->to = to;
  for (int i = 0; i < OPS_MAX_DIM; i++) {
    bool in = i < from->block->dims;
    h->iter_size[i] = in ? iter[i] : 1;
    h->from_base[i] = in ? fb[i] : 0;
    h->to_base[i] = in ? tb[i] : 0;
    h->from_dir[i] = in ? fd[i] : i + 1;
    h->to_dir[i] = in ? td[i] : i + 1;
  }
  // window inside each dat's allocation (same check as the device runtime)
  ops_dat dd[2] = {from, to};
  const int *bb[2] = {h->from_base, h->to_base}, *rr[2] = {h->from_dir, h->to_dir};
  for (int s = 0; s < 2; s++)
    for (int i = 0; i < dd[s]->block->dims; i++) {
      const int ax = abs(rr[s][i]) - 1;
      if (ax < 0 || ax >= dd[s]->block->dims) fail(OPS_INVALID_ARGUMENT, "ops_decl_halo: direction");
      const long first = (long)bb[s][i] - dd[s]->d_m[i] - dd[s]->base[i];
      if (first < 0 || first + h->iter_size[ax] > dd[s]->size[i])
        fail(OPS_INVALID_ARGUMENT, "ops_decl_halo: window outside the allocation");
    }
  g->halos.push_back(h);
  return h;
}

extern "C" ops_halo_group ops_decl_halo_group(int n, ops_halo *halos) {
  ops_halo_group q = (ops_halo_group)calloc(1, sizeof(ops_halo_group_core));
  q->nhalos = n;
  q->halos = (ops_halo *)calloc((size_t)(n > 0 ? n : 1), sizeof(ops_halo));
  memcpy(q->halos, halos, sizeof(ops_halo) * (size_t)n);
  q->index = (int)g->groups.size();
  q->instance = g;
  g->groups.push_back(q);
  return q;
}

// Staged through a linear buffer with the reference's ranges/step/buf_strides
// (ops_host_singlenode.cpp:190-446): the buffer is laid out in the halo's own axis order.
static void halo_ranges(ops_dat d, const int *hb, const int *dir, const int *iter, int *r,
                        int *step, int *bs) {
  for (int i = 0; i < OPS_MAX_DIM; i++) {
    if (dir[i] > 0) {
      r[2 * i] = hb[i] - d->d_m[i] - d->base[i];
      r[2 * i + 1] = r[2 * i] + iter[abs(dir[i]) - 1];
      step[i] = 1;
    } else {
      r[2 * i + 1] = hb[i] - 1 - d->d_m[i] - d->base[i];
      r[2 * i] = r[2 * i + 1] + iter[abs(dir[i]) - 1];
      step[i] = -1;
    }
    bs[i] = 1;
    for (int j = 0; j != abs(dir[i]) - 1; j++) bs[i] *= iter[j];
  }
}

static double read_as(ops_dat d, const char *p) {
  if (d->type_size == 8) return *(const double *)p;
  return *(const float *)p;
}

extern "C" void ops_halo_transfer(ops_halo_group grp) {
  for (int hh = 0; hh < grp->nhalos; hh++) {
    ops_halo h = grp->halos[hh];
    ops_dat f = h->from, t = h->to;
    int st = f->type_size < t->type_size ? f->type_size : t->type_size;
    bool mixed = f->type_size != t->type_size;
    size_t pts = 1;
    for (int i = 0; i < OPS_MAX_DIM; i++) pts *= (size_t)h->iter_size[i];
    std::vector<char> buf(pts * (size_t)(f->dim > t->dim ? f->dim : t->dim) * 8 + 8);
    int r[10], step[5], bs[5];
    for (int pass = 0; pass < 2; pass++) {
      ops_dat d = pass == 0 ? f : t;
      halo_ranges(d, pass == 0 ? h->from_base : h->to_base, pass == 0 ? h->from_dir : h->to_dir,
                  h->iter_size, r, step, bs);
      int esz = mixed ? st * d->dim : d->elem_size;
      int lo[5], hi[5];
      for (int i = 0; i < 5; i++) {
        lo[i] = r[2 * i] < r[2 * i + 1] + 1 ? r[2 * i] : r[2 * i + 1] + 1;
        hi[i] = r[2 * i] + 1 > r[2 * i + 1] ? r[2 * i] + 1 : r[2 * i + 1];
      }
      for (int m = lo[4]; m < hi[4]; m++)
        for (int l = lo[3]; l < hi[3]; l++)
          for (int k = lo[2]; k < hi[2]; k++)
            for (int j = lo[1]; j < hi[1]; j++)
              for (int i = lo[0]; i < hi[0]; i++) {
                int c[5] = {i, j, k, l, m};
                size_t bo = 0, doff = 0, cum = 1;
                for (int a = 0; a < 5; a++) {
                  bo += (size_t)((c[a] - r[2 * a]) * step[a] * bs[a]);
                  doff += (size_t)c[a] * cum;
                  cum *= (size_t)d->size[a];
                }
                for (int comp = 0; comp < d->dim; comp++) {
                  // SoA: component comp a whole allocation plane further on
                  // (ops_host_singlenode.cpp:124-160)
                  char *dp = g->OPS_soa ? d->data + ((size_t)comp * cum + doff) * d->type_size
                                    : d->data + doff * d->elem_size + (size_t)comp * d->type_size;
                  char *bp = buf.data() + bo * (size_t)esz + (size_t)comp * (mixed ? st : d->type_size);
                  if (!mixed) {
                    if (pass == 0) memcpy(bp, dp, (size_t)d->type_size);
                    else memcpy(dp, bp, (size_t)d->type_size);
                  } else if (pass == 0) {
                    double v = read_as(d, dp);
                    if (st == 4) { float fv = (float)v; memcpy(bp, &fv, 4); }
                    else memcpy(bp, &v, 8);
                  } else {
                    double v = st == 4 ? (double)*(float *)bp : *(double *)bp;
                    if (d->type_size == 8) *(double *)dp = v;
                    else *(float *)dp = (float)v;
                  }
                }
              }
    }
    t->dirty_hd = 1;
  }
}

extern "C" void ops_printf(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vprintf(fmt, ap);
  va_end(ap);
}
extern "C" void ops_fprintf(FILE *s, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vfprintf(s, fmt, ap);
  va_end(ap);
}
extern "C" void ops_timers_core(double *cpu, double *et) {
  struct timeval t;
  gettimeofday(&t, nullptr);
  *cpu = 0;
  *et = t.tv_sec + t.tv_usec * 1e-6;
}
extern "C" void ops_timers(double *cpu, double *et) { ops_timers_core(cpu, et); }
extern "C" void ops_diagnostic_output(void) {}
extern "C" void ops_timing_output_stdout(void) {}
extern "C" void ops_get_data(ops_dat) {}
extern "C" void ops_print_dat_to_txtfile(ops_dat, const char *) {}

extern "C" int ops_dat_get_local_npartitions(ops_dat) { return 1; }
extern "C" int ops_dat_get_global_npartitions(ops_dat) { return 1; }
extern "C" void ops_dat_get_extents(ops_dat d, int, int *disp, int *size) {
  for (int i = 0; i < d->block->dims; i++) {
    if (disp) disp[i] = 0;
    if (size) size[i] = d->size[i] + d->d_m[i] - d->d_p[i];
  }
}
extern "C" size_t ops_dat_get_slab_extents(ops_dat d, int part, int *disp, int *size, int *slab) {
  int e[OPS_MAX_DIM];
  ops_dat_get_extents(d, part, nullptr, e);
  size_t b = (size_t)d->elem_size;
  for (int i = 0; i < d->block->dims; i++) {
    if (slab[2 * i] < 0 || slab[2 * i + 1] > e[i]) fail(OPS_RUNTIME_ERROR, "slab out of range");
    disp[i] = slab[2 * i];
    size[i] = slab[2 * i + 1] - slab[2 * i];
    b *= (size_t)size[i];
  }
  return b;
}
extern "C" void ops_dat_get_raw_metadata(ops_dat d, int part, int *disp, int *size, int *stride,
                                         int *d_m, int *d_p) {
  ops_dat_get_extents(d, part, disp, size);
  for (int i = 0; i < d->block->dims; i++) {
    if (stride) stride[i] = d->size[i];
    if (d_m) d_m[i] = d->d_m[i];
    if (d_p) d_p[i] = d->d_p[i];
  }
}
extern "C" char *ops_dat_get_raw_pointer(ops_dat d, int, ops_stencil, ops_memspace *ms) {
  *ms = OPS_HOST;
  d->locked_hd = OPS_HOST;
  return d->data + d->base_offset;
}
extern "C" void ops_dat_release_raw_data(ops_dat d, int, ops_access acc) {
  if (!d->locked_hd) fail(OPS_RUNTIME_ERROR, "release without get_raw_pointer");
  if (acc != OPS_READ) d->dirty_hd = 1;
  d->locked_hd = 0;
}
extern "C" void ops_dat_release_raw_data_memspace(ops_dat d, int p, ops_access acc, ops_memspace *) {
  ops_dat_release_raw_data(d, p, acc);
}

// packed slab copy, range in interior coordinates (ops_util.cpp:353-500)
static void slab(ops_dat d, char *buf, const int *range, int dir) {
  int nd = d->block->dims, cnt[OPS_MAX_DIM] = {1, 1, 1, 1, 1};
  for (int i = 0; i < nd; i++) cnt[i] = range[2 * i + 1] - range[2 * i];
  size_t k = 0;
  for (int m = 0; m < cnt[4]; m++)
    for (int l = 0; l < cnt[3]; l++)
      for (int z = 0; z < cnt[2]; z++)
        for (int y = 0; y < cnt[1]; y++)
          for (int x = 0; x < cnt[0]; x++) {
            int c[5] = {x, y, z, l, m};
            size_t off = 0, cum = 1;
            for (int a = 0; a < nd; a++) {
              off += (size_t)(range[2 * a] + c[a] - d->d_m[a]) * cum;
              cum *= (size_t)d->size[a];
            }
            if (g->OPS_soa && d->dim > 1) {
              // component c of the point sits c planes further on; the buffer is packed
              // with components innermost (ops_util.cpp:378-400)
              size_t plane = 1;
              for (int a = 0; a < nd; a++) plane *= (size_t)d->size[a];
              for (int cc = 0; cc < d->dim; cc++) {
                char *dp = d->data + (cc * plane + off) * d->type_size;
                if (dir == 0) memcpy(buf + k + (size_t)cc * d->type_size, dp, (size_t)d->type_size);
                else memcpy(dp, buf + k + (size_t)cc * d->type_size, (size_t)d->type_size);
              }
            } else {
              char *dp = d->data + off * d->elem_size;
              if (dir == 0) memcpy(buf + k, dp, (size_t)d->elem_size);
              else memcpy(dp, buf + k, (size_t)d->elem_size);
            }
            k += (size_t)d->elem_size;
          }
}
extern "C" void ops_dat_fetch_data_slab_memspace(ops_dat d, int, char *data, int *range, ops_memspace) {
  slab(d, data, range, 0);
}
extern "C" void ops_dat_fetch_data_memspace(ops_dat d, int part, char *data, ops_memspace ms) {
  int e[OPS_MAX_DIM], r[2 * OPS_MAX_DIM];
  ops_dat_get_extents(d, part, nullptr, e);
  for (int i = 0; i < d->block->dims; i++) {
    r[2 * i] = 0;
    r[2 * i + 1] = e[i];
  }
  ops_dat_fetch_data_slab_memspace(d, part, data, r, ms);
}
extern "C" void ops_dat_fetch_data(ops_dat d, int p, char *data) {
  ops_dat_fetch_data_memspace(d, p, data, OPS_HOST);
}
extern "C" void ops_dat_set_data_slab_memspace(ops_dat d, int, char *data, int *range, ops_memspace) {
  slab(d, data, range, 1);
  d->dirty_hd = 1;
}
extern "C" void ops_dat_set_data_memspace(ops_dat d, int p, char *data, ops_memspace ms) {
  int r[2 * OPS_MAX_DIM];
  for (int i = 0; i < d->block->dims; i++) {
    r[2 * i] = d->d_m[i];
    r[2 * i + 1] = d->size[i] - d->d_p[i];
  }
  ops_dat_set_data_slab_memspace(d, p, data, r, ms);
}
extern "C" void ops_dat_set_data(ops_dat d, int p, char *data) {
  ops_dat_set_data_memspace(d, p, data, OPS_HOST);
}

extern "C" void ops_randomgen_init(unsigned int seed, int) { g_rng.seed(seed); }
extern "C" void ops_fill_random_uniform(ops_dat d) {
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (size_t i = 0; i < d->mem / 8; i++) ((double *)d->data)[i] = u(g_rng);
}
extern "C" void ops_fill_random_normal(ops_dat d) {
  std::normal_distribution<double> u(0.0, 1.0);
  for (size_t i = 0; i < d->mem / 8; i++) ((double *)d->data)[i] = u(g_rng);
}
extern "C" void ops_randomgen_exit(void) {}

// Thomas sweeps over the solve box (same recurrence as the GPU solver's header comment)
template <bool INC>
static void thomas(int ndim, int sd, const int *range, ops_dat A, ops_dat B, ops_dat C, ops_dat D,
                   ops_dat U) {
  long s[3] = {1, A->size[0], (long)A->size[0] * A->size[1]};
  int n[3] = {1, 1, 1};
  long off = 0;
  for (int i = 0; i < ndim; i++) {
    n[i] = range[2 * i + 1] - range[2 * i];
    if (n[i] <= 0) return;
    off += (long)range[2 * i] * s[i];
  }
  auto at = [&](ops_dat x) { return x ? (double *)(x->data + x->base_offset) + off : nullptr; };
  const double *a = at(A), *b = at(B), *c = at(C);
  double *d = at(D), *u = at(U);
  int o1 = sd == 0 ? 1 : 0, o2 = sd == 2 ? 1 : 2;
  long st = s[sd];
  int len = n[sd];
  // lines are independent: threads split them, each line is swept in the textbook order
#pragma omp parallel for schedule(static)
  for (long line = 0; line < (long)n[o2] * n[o1]; line++) {
      std::vector<double> cp(len), dp(len);
      const int q = (int)(line / n[o1]), p = (int)(line % n[o1]);
      long o = p * s[o1] + q * s[o2];
      cp[0] = c[o] / b[o];
      dp[0] = d[o] / b[o];
      for (int i = 1; i < len; i++) {
        long k = o + i * st;
        double m = 1.0 / (b[k] - a[k] * cp[i - 1]);
        cp[i] = c[k] * m;
        dp[i] = (d[k] - a[k] * dp[i - 1]) * m;
      }
      double x = dp[len - 1];
      long k = o + (long)(len - 1) * st;
      if (INC) u[k] += x; else d[k] = x;
      for (int i = len - 2; i >= 0; i--) {
        x = dp[i] - cp[i] * x;
        k = o + i * st;
        if (INC) u[k] += x; else d[k] = x;
      }
    }
}

extern "C" void ops_tridMultiDimBatch(int ndim, int sd, int *range, ops_dat a, ops_dat b, ops_dat c,
                                      ops_dat d, ops_tridsolver_params *) {
  thomas<false>(ndim, sd, range, a, b, c, d, nullptr);
}
extern "C" void ops_tridMultiDimBatch_Inc(int ndim, int sd, int *range, ops_dat a, ops_dat b,
                                          ops_dat c, ops_dat d, ops_dat u, ops_tridsolver_params *) {
  thomas<true>(ndim, sd, range, a, b, c, d, u);
}
// the allocation-free device form is the same solve on the host
extern "C" void ops_tridMultiDimBatch_noalloc(int ndim, int sd, int *range, ops_dat a, ops_dat b,
                                              ops_dat c, ops_dat d, ops_dat u, ops_tridsolver_params *) {
  if (u) thomas<true>(ndim, sd, range, a, b, c, d, u);
  else thomas<false>(ndim, sd, range, a, b, c, d, nullptr);
}

static void dump_all(const char *dir) {
  std::string idx = std::string(dir) + "/index.txt";
  FILE *fi = fopen(idx.c_str(), "w");
  if (!fi) return;
  for (ops_dat d : g->dats) {
    if (!d) continue;
    fprintf(fi, "%s %d %d %d", d->name, d->dim, d->type_size, d->block->dims);
    for (int i = 0; i < OPS_MAX_DIM; i++) fprintf(fi, " %d", d->size[i]);
    for (int i = 0; i < OPS_MAX_DIM; i++) fprintf(fi, " %d", d->d_m[i]);
    for (int i = 0; i < OPS_MAX_DIM; i++) fprintf(fi, " %d", d->d_p[i]);
    for (int i = 0; i < OPS_MAX_DIM; i++) fprintf(fi, " %d", d->base[i]);
    fprintf(fi, " %d %zu\n", d->x_pad, d->base_offset);
    std::string fn = std::string(dir) + "/" + d->name + ".bin";
    FILE *fb = fopen(fn.c_str(), "wb");
    if (fb) {
      fwrite(d->data, 1, d->mem, fb);
      fclose(fb);
    }
  }
  fclose(fi);
}

// thread count of the OpenMP loops (1 reproduces the sequential reduction order)
extern "C" void ops_cpu_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

// graph capture is a device feature: the host backend reports it unavailable, so callers
// run their launches eagerly
// loop batches only change how launches are grouped: the host backend runs every loop as
// it comes
extern "C" void ops_hip_batch_begin(void) {}
extern "C" void ops_hip_batch_end(void) {}
// loop chains only change the order in which a device runs the points of chained loops: the
// host backend runs every loop whole, in program order
extern "C" void ops_hip_chain_begin(int) {}
extern "C" void ops_hip_chain_end(void) {}
extern "C" void ops_hip_device_sync(void) {}
extern "C" int ops_hip_graph_begin(void) { return 1; }
extern "C" void *ops_hip_graph_end(void) { return nullptr; }
extern "C" int ops_hip_graph_replay(void *, int) { return 1; }
extern "C" void ops_hip_graph_free(void *) {}
// device-side chaining has no host counterpart: the applications' captured-step mode is
// GPU only (their host path reads reductions and uploads constants on the host)
extern "C" void *ops_hip_const_device_ptr(const char *) {
  fail(OPS_NOT_IMPLEMENTED, "ops_hip_const_device_ptr: not available on the host backend");
  return nullptr;
}
extern "C" void *ops_hip_reduction_device_ptr(ops_reduction) {
  fail(OPS_NOT_IMPLEMENTED, "ops_hip_reduction_device_ptr: not available on the host backend");
  return nullptr;
}
extern "C" void ops_hip_reduction_discard(ops_reduction h) {
  h->initialized = 0;
}
