// libops_hip.so device layer: HBM allocation, host<->device movement, the loop-launch
// seam used by generated code, fused halo copies, slab copies, per-kernel event timing.
//
// Behaviour replaced (not ported):
//   device management      ops/c/src/hip/ops_hip_common.cpp:50-180
//   halo pack/unpack       ops/c/src/hip/ops_hip_rt_support_kernels.cpp:55-253 and
//                          ops/c/src/core/ops_device_singlenode_common.cpp:262-346
//   slab copy kernels      ops/c/src/hip/ops_hip_rt_support_kernels.cpp:256-400
//   stub pre/post work     templates/cpp/cuda/loop_host.cpp.j2:299-782
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "ops_hip_kernel.h"
#include "ops_internal.h"

#define HIPCHECK(x) ops_hip_check((int)(x), #x, __FILE__, __LINE__)

namespace {

struct DevState {
  hipStream_t stream = nullptr;
  int device = 0;
  // pinned staging ring for small stream-ordered uploads (consts, gbl arrays)
  char *stage = nullptr;
  size_t stage_cap = 0, stage_off = 0;
  // pinned bounce buffer for reduction results
  char *pinned_small = nullptr;
  // reduction scratch: per-block partials (one region per loop argument) + counters
  char *red_scratch = nullptr;
  size_t red_scratch_cap = 0;
  // outgrown reduction scratch buffers: a captured graph may still write to them, so they
  // are only freed at exit
  std::vector<char *> red_scratch_retired;
  // event timing
  struct Pending {
    int kid;
    hipEvent_t a, b;
    int cand;  // autotuning candidate measured by this launch, -1 for plain timing
    double pts;  // loop points of the launch (candidates are compared per point)
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  bool live = false;
};
DevState g_dev;
int g_timing = 0;
bool g_capturing = false;

hipEvent_t get_event() {
  if (!g_dev.pool.empty()) {
    hipEvent_t e = g_dev.pool.back();
    g_dev.pool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIPCHECK(hipEventCreate(&e));
  return e;
}

}  // namespace

int g_autotune = 0;

extern "C" void ops_hip_check(int status, const char *what, const char *file, int line) {
  if (status != (int)hipSuccess) {
    char buf[1024];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", status,
             hipGetErrorString((hipError_t)status), what, file, line);
    ops_throw(OPS_RUNTIME_ERROR, buf);
  }
}

extern "C" void ops_hip_check_launch(const char *name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    char buf[512];
    snprintf(buf, sizeof buf, "HIP launch of %s failed: %s", name, hipGetErrorString(e));
    ops_throw(OPS_RUNTIME_ERROR, buf);
  }
}

static void loop_batch_reset();

void ops_dev_init(ops_runtime_instance *inst) {
  if (g_dev.live) return;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    ops_throw(OPS_RUNTIME_CONFIGURATION_ERROR,
              "Error: no HIP device available (libops_hip.so needs an MI355X / gfx950 GPU)");
  if (inst->device_id < 0 || inst->device_id >= n)
    ops_throw(OPS_RUNTIME_CONFIGURATION_ERROR, "Error: OPS_DEVICE out of range");
  g_dev.device = inst->device_id;
  HIPCHECK(hipSetDevice(g_dev.device));
  ops_trid_device_init(g_dev.device);
  HIPCHECK(hipStreamCreate(&g_dev.stream));
  g_dev.stage_cap = 8u << 20;
  HIPCHECK(hipHostMalloc((void **)&g_dev.stage, g_dev.stage_cap, hipHostMallocDefault));
  g_dev.stage_off = 0;
  HIPCHECK(hipHostMalloc((void **)&g_dev.pinned_small, 1u << 16, hipHostMallocDefault));
  g_dev.live = true;
  loop_batch_reset();
}

void ops_dev_exit(ops_runtime_instance *inst) {
  (void)inst;
  if (!g_dev.live) return;
  loop_batch_reset();
  ops_trid_flush();
  ops_trid_release();
  ops_dev_timing_flush();
  HIPCHECK(hipStreamSynchronize(g_dev.stream));
  for (hipEvent_t ev : g_dev.pool) (void)hipEventDestroy(ev);
  g_dev.pool.clear();
  if (g_dev.red_scratch) (void)hipFree(g_dev.red_scratch);
  for (char *p : g_dev.red_scratch_retired) (void)hipFree(p);
  g_dev.red_scratch_retired.clear();
  g_dev.red_scratch = nullptr;
  g_dev.red_scratch_cap = 0;
  (void)hipHostFree(g_dev.stage);
  (void)hipHostFree(g_dev.pinned_small);
  (void)hipStreamDestroy(g_dev.stream);
  g_dev = DevState();
}

void *ops_dev_malloc(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) bytes = 8;
  HIPCHECK(hipMalloc(&p, bytes));
  return p;
}
void ops_dev_free(void *p) {
  ops_trid_flush();
  if (p) {
    (void)hipStreamSynchronize(g_dev.stream);
    (void)hipFree(p);
  }
}
void *ops_dev_host_alloc(size_t bytes) {
  void *p = nullptr;
  HIPCHECK(hipHostMalloc(&p, bytes ? bytes : 8, hipHostMallocCoherent | hipHostMallocMapped));
  memset(p, 0, bytes ? bytes : 8);
  return p;
}
void ops_dev_host_free(void *p) {
  if (p) {
    ops_trid_flush();
    (void)hipStreamSynchronize(g_dev.stream);
    (void)hipHostFree(p);
  }
}
void ops_dev_memset(void *p, int v, size_t bytes) {
  ops_trid_flush();
  HIPCHECK(hipMemsetAsync(p, v, bytes, g_dev.stream));
}
void ops_dev_sync() {
  ops_trid_flush();
  if (g_dev.live) HIPCHECK(hipStreamSynchronize(g_dev.stream));
}
void ops_dev_h2d(void *dst, const void *src, size_t bytes) {
  ops_trid_flush();
  HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_dev.stream));
  HIPCHECK(hipStreamSynchronize(g_dev.stream));
}
void ops_dev_d2h(void *dst, const void *src, size_t bytes) {
  ops_trid_flush();
  if (bytes <= (1u << 16)) {
    HIPCHECK(hipMemcpyAsync(g_dev.pinned_small, src, bytes, hipMemcpyDeviceToHost, g_dev.stream));
    HIPCHECK(hipStreamSynchronize(g_dev.stream));
    memcpy(dst, g_dev.pinned_small, bytes);
  } else {
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_dev.stream));
    HIPCHECK(hipStreamSynchronize(g_dev.stream));
  }
}
void ops_dev_d2d(void *dst, const void *src, size_t bytes) {
  ops_trid_flush();
  HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_dev.stream));
}
void ops_dev_stage_h2d(void *dst, const void *src, size_t bytes) {
  ops_trid_flush();
  if (bytes > g_dev.stage_cap) {
    ops_dev_h2d(dst, src, bytes);
    return;
  }
  size_t off = (g_dev.stage_off + 255) & ~(size_t)255;
  if (off + bytes > g_dev.stage_cap) {
    // ring wraps: earlier staged copies must have been consumed
    HIPCHECK(hipStreamSynchronize(g_dev.stream));
    off = 0;
  }
  memcpy(g_dev.stage + off, src, bytes);
  HIPCHECK(hipMemcpyAsync(dst, g_dev.stage + off, bytes, hipMemcpyHostToDevice, g_dev.stream));
  g_dev.stage_off = off + bytes;
}

void *ops_dev_symbol_address(const void *symbol) {
  void *p = nullptr;
  HIPCHECK(hipGetSymbolAddress(&p, symbol));
  return p;
}

extern "C" void *ops_hip_get_stream(void) {
  ops_trid_flush();
  return (void *)g_dev.stream;
}
void *ops_hip_get_stream_raw() { return (void *)g_dev.stream; }
extern "C" void ops_hip_device_sync(void) { ops_dev_sync(); }

// ------------------------------------------------------------------ loop batches

namespace {
bool g_batch_on = false;
// the pending batch's launcher: one per generated stub (a kernel translated for two
// argument signatures has two stubs, and each keeps its own pending launches)
void (*g_batch_fn)(void) = nullptr;
bool g_batch_keep = false;  // loop_begin of the pending stub: do not flush it
}  // namespace

static void chain_flush();
static void chain_off();

void ops_loop_batch_flush() {
  // a recorded loop chain was issued before anything that flushes: it runs first
  chain_flush();
  if (!g_batch_fn || g_batch_keep) return;
  void (*f)(void) = g_batch_fn;
  g_batch_fn = nullptr;
  f();
}

// batching off and nothing pending: at ops_init and ops_exit, so that a batch left open by
// an exception does not outlive its session (ops_exit launches what is pending first)
static void loop_batch_reset() {
  g_batch_keep = false;
  ops_loop_batch_flush();
  g_batch_on = false;
  chain_off();
}

extern "C" void ops_hip_batch_begin(void) {
  ops_trid_flush();
  g_batch_on = true;
}

extern "C" void ops_hip_batch_end(void) {
  g_batch_on = false;
  ops_loop_batch_flush();
}

extern "C" int ops_hip_batch_join(int kernel_id, void (*flush)(void)) {
  if (!g_batch_on || kernel_id < 0 || flush == nullptr) return 0;
  if (g_batch_fn && g_batch_fn != flush) ops_loop_batch_flush();
  g_batch_fn = flush;
  return 1;
}

// ------------------------------------------------------------------ loop chains
// Between ops_hip_chain_begin(tiles) and ops_hip_chain_end(), loops that can be tiled are
// recorded instead of launched. When the chain ends (or anything else needs the device
// first) it runs tile by tile along the loops' last dimension: every loop's range is cut
// into bands and the bands run band-major (band 0 of each loop in program order, then band
// 1, ...), so that what one loop writes is read by the next loops while it is still in the
// caches (a band of a 3840^2 CloverLeaf advection chain fits the 256 MB Infinity Cache).
// Band boundaries are skewed per loop, the dependency rule of the reference's lazy tiling
// (ops/c/src/core/ops_lazy.cpp:380-1186): loop j's boundaries sit s_j rows behind the
// common ones, s_j >= s_i for every earlier loop i sharing a dat with it, plus
//   - the forward reach of j's read of a dat i writes (read after write), and
//   - the backward reach of i's read of a dat j writes (write after read).
// Each point of each loop then runs once, with the arithmetic of the untiled launch, after
// everything it depends on and before anything that overwrites what it reads: the results
// are bitwise those of the loops launched in order. Loops with reductions, multigrid or
// strided stencils along the tiled dimension, low-dim reductions or fewer than two
// dimensions are not recorded (the chain runs first, then they run as usual).
namespace {
struct ChainLoop {
  void (*relaunch)(const char *, ops_block, int, int *, ops_arg *);
  const char *name;
  ops_block block;
  int dim;
  int range[6];
  std::vector<ops_arg> args;
  std::vector<std::vector<char>> gbl;  // the READ gbl values (the caller's may not outlive it)
};
struct Chain {
  bool on = false, running = false;
  int tiles = 1;
  std::vector<ChainLoop> loops;
} g_chain;
}  // namespace

static bool lowdim_reduction(const ops_arg &a);

static bool chain_eligible(const ops_hip_loop *lp, int dim, const ops_arg *args, int nargs) {
  if (lp->relaunch == nullptr || dim < 2 || dim > 3) return false;
  const int ax = dim - 1;
  for (int i = 0; i < nargs; i++) {
    const ops_arg &a = args[i];
    if (a.argtype == OPS_ARG_GBL && a.acc != OPS_READ) return false;
    if (a.argtype != OPS_ARG_DAT || a.dat == nullptr || a.stencil == nullptr) continue;
    if (a.stencil->type != 0 || lowdim_reduction(a)) return false;
    const int st = ax < a.stencil->dims ? a.stencil->stride[ax] : 1;
    if (st != 1 && !(st == 0 && a.acc == OPS_READ)) return false;
    if (a.acc != OPS_READ)  // writes only at the point itself along the tiled dimension
      for (int p = 0; p < a.stencil->points; p++)
        if (ax < a.stencil->dims && a.stencil->stencil[p * a.stencil->dims + ax] != 0) return false;
  }
  return true;
}

// reach of a dat argument's stencil along the tiled dimension
static void chain_reach(const ops_arg &a, int ax, int &lo, int &hi) {
  lo = hi = 0;
  if (ax >= a.stencil->dims) return;
  for (int p = 0; p < a.stencil->points; p++) {
    const int v = a.stencil->stencil[p * a.stencil->dims + ax];
    if (p == 0 || v < lo) lo = v;
    if (p == 0 || v > hi) hi = v;
  }
}

static void chain_record(const ops_hip_loop *lp, const char *name, ops_block block, int dim,
                         const int *range, const ops_arg *args, int nargs) {
  int start[OPS_MAX_DIM] = {0, 0, 0, 0, 0}, end[OPS_MAX_DIM] = {1, 1, 1, 1, 1};
  for (int d = 0; d < dim; d++) {
    start[d] = range[2 * d];
    end[d] = range[2 * d + 1];
  }
  bool empty = false;
  for (int d = 0; d < dim; d++) empty = empty || end[d] <= start[d];
  if (!empty) ops_check_loop_bounds(name, dim, start, end, args, nargs);  // report at the call
  ChainLoop c;
  c.relaunch = lp->relaunch;
  c.name = name;
  c.block = block;
  c.dim = dim;
  for (int d = 0; d < 6; d++) c.range[d] = d < 2 * dim ? range[d] : 0;
  c.args.assign(args, args + nargs);
  c.gbl.resize(nargs);
  for (int i = 0; i < nargs; i++)
    if (args[i].argtype == OPS_ARG_GBL && args[i].data != nullptr) {
      c.gbl[i].assign(args[i].data, args[i].data + (size_t)args[i].dim * args[i].elem_size);
      c.args[i].data = c.gbl[i].data();
    }
  g_chain.loops.push_back(std::move(c));
}

static void chain_flush() {
  if (g_chain.loops.empty() || g_chain.running) return;
  g_chain.running = true;
  std::vector<ChainLoop> loops;
  loops.swap(g_chain.loops);
  try {
    const int n = (int)loops.size();
    const int ax = loops[0].dim - 1;
    bool tile = g_chain.tiles > 1;
    for (const ChainLoop &c : loops) tile = tile && c.dim == loops[0].dim;
    // skews
    std::vector<int> shift(n, 0);
    for (int j = 0; j < n && tile; j++)
      for (int i = 0; i < j && tile; i++)
        for (const ops_arg &aj : loops[j].args) {
          if (aj.argtype != OPS_ARG_DAT || !aj.dat || !aj.stencil) continue;
          for (const ops_arg &ai : loops[i].args) {
            if (ai.argtype != OPS_ARG_DAT || ai.dat != aj.dat || !ai.stencil) continue;
            const bool wi = ai.acc != OPS_READ, wj = aj.acc != OPS_READ;
            const bool ri = ai.acc != OPS_WRITE, rj = aj.acc != OPS_WRITE;
            if (!wi && !wj) continue;
            // a read that does not move along the tiled dimension sees every band
            const int sti = ax < ai.stencil->dims ? ai.stencil->stride[ax] : 1;
            const int stj = ax < aj.stencil->dims ? aj.stencil->stride[ax] : 1;
            if (sti == 0 || stj == 0) {
              tile = false;
              break;
            }
            int ilo, ihi, jlo, jhi;
            chain_reach(ai, ax, ilo, ihi);
            chain_reach(aj, ax, jlo, jhi);
            int need = 0;
            if (wi && rj) need = std::max(need, jhi);   // read after write
            if (wj && ri) need = std::max(need, -ilo);  // write after read
            shift[j] = std::max(shift[j], shift[i] + need);
          }
          if (!tile) break;
        }
    if (!tile) {
      for (ChainLoop &c : loops) c.relaunch(c.name, c.block, c.dim, c.range, c.args.data());
    } else {
      long lo = LONG_MAX, hi = LONG_MIN;
      for (const ChainLoop &c : loops)
        if (c.range[2 * ax + 1] > c.range[2 * ax]) {
          lo = std::min(lo, (long)c.range[2 * ax]);
          hi = std::max(hi, (long)c.range[2 * ax + 1]);
        }
      const int T = g_chain.tiles;
      for (int t = 0; t < T; t++) {
        const long b0 = lo + (hi - lo) * t / T, b1 = lo + (hi - lo) * (t + 1) / T;
        for (int j = 0; j < n; j++) {
          ChainLoop &c = loops[j];
          const long l0 = c.range[2 * ax], l1 = c.range[2 * ax + 1];
          const long s0 = t == 0 ? l0 : std::max(l0, b0 - shift[j]);
          const long s1 = t == T - 1 ? l1 : std::min(l1, b1 - shift[j]);
          if (s1 <= s0) continue;
          int sub[6];
          for (int d = 0; d < 6; d++) sub[d] = c.range[d];
          sub[2 * ax] = (int)s0;
          sub[2 * ax + 1] = (int)s1;
          c.relaunch(c.name, c.block, c.dim, sub, c.args.data());
        }
      }
    }
  } catch (...) {
    g_chain.running = false;
    throw;
  }
  g_chain.running = false;
}

static void chain_off() { g_chain.on = false; }

extern "C" void ops_hip_chain_begin(int tiles) {
  ops_trid_flush();
  g_chain.on = true;
  g_chain.tiles = tiles < 1 ? 1 : tiles;
}

extern "C" void ops_hip_chain_end(void) {
  g_chain.on = false;
  chain_flush();
}

// ------------------------------------------------------------------ timing

extern "C" void ops_hip_set_timing(int on) { g_timing = on; }

int ops_tune_next(int kid, int ncand, int reps) {
  auto &kt = ops_kernel_table();
  if (!g_autotune || kid < 0 || kid >= (int)kt.size()) return -1;
  ops_kernel_entry &e = kt[kid];
  if (e.tune_state < 0) return -1;
  // a decided shape also applies to captured launches; measuring needs events between
  // launches, so a capture never starts or continues it
  if (e.tune_state == 2) return e.tune_choice;
  if (g_capturing) return -1;
  if (ncand > 8) ncand = 8;
  e.tune_state = 1;
  return e.tune_calls % ncand;
  (void)reps;
}

// a measured launch: events recorded on the stream around it (ops_tune_begin/_end), or
// handed to the launch itself (record = false: the generated stubs pass them to
// hipExtLaunchKernelGGL, so that they bracket the kernel and not the queue around it)
static void tune_begin(int kid, int cand, double pts, bool record, hipEvent_t *ev) {
  hipEvent_t a = get_event(), b = get_event();
  if (record) HIPCHECK(hipEventRecord(a, g_dev.stream));
  g_dev.pending.push_back({kid, a, b, cand, pts});
  if (ev) {
    ev[0] = a;
    ev[1] = b;
  }
}

void ops_tune_begin(int kid, int cand, double pts) { tune_begin(kid, cand, pts, true, nullptr); }

static void tune_end(int kid, int ncand, int reps, bool record);

void ops_tune_end(int kid, int ncand, int reps) { tune_end(kid, ncand, reps, true); }

static void tune_end(int kid, int ncand, int reps, bool record) {
  auto &kt = ops_kernel_table();
  ops_kernel_entry &e = kt[kid];
  if (record) HIPCHECK(hipEventRecord(g_dev.pending.back().b, g_dev.stream));
  if (++e.tune_calls < ncand * reps) return;
  ops_dev_timing_flush();
  int best = 0;
  double best_ms = 1e300;
  for (int c = 0; c < ncand && c < 8; c++) {
    // time per loop point: one kernel's launches may differ in size (the blocks of a
    // multi-block mesh, the faces of a halo update), so candidates are compared per point
    const double avg = e.cand_pts[c] > 0 ? e.cand_ms[c] / e.cand_pts[c] : 1e300;
    if (avg < best_ms) {
      best_ms = avg;
      best = c;
    }
  }
  e.tune_choice = best;
  e.tune_state = 2;
  // OPS_TUNE_REPORT=1: one line per decision on stderr (time per point of each candidate)
  static const bool report = getenv("OPS_TUNE_REPORT") && atoi(getenv("OPS_TUNE_REPORT")) > 0;
  if (report) {
    fprintf(stderr, "ops_autotune %s:", e.name.c_str());
    for (int c = 0; c < ncand && c < 8; c++)
      fprintf(stderr, " %d:%.4g", c, e.cand_pts[c] > 0 ? 1e6 * e.cand_ms[c] / e.cand_pts[c] : -1.0);
    fprintf(stderr, " ns/point -> %d\n", best);
  }
}

void ops_dev_timing_flush() {
  if (g_dev.pending.empty()) return;
  HIPCHECK(hipEventSynchronize(g_dev.pending.back().b));
  auto &kt = ops_kernel_table();
  for (auto &p : g_dev.pending) {
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, p.a, p.b));
    kt[p.kid].total_ms += ms;
    // a kernel's first measured launch is not counted: it pays for first use (cold caches
    // and TLBs over newly allocated dats) and would always fall on candidate 0
    if (p.cand >= 0 && p.cand < 8 && !kt[p.kid].tune_warm) {
      kt[p.kid].tune_warm = 1;
    } else if (p.cand >= 0 && p.cand < 8) {
      kt[p.kid].cand_ms[p.cand] += ms;
      kt[p.kid].cand_n[p.cand]++;
      kt[p.kid].cand_pts[p.cand] += p.pts;
    }
    g_dev.pool.push_back(p.a);
    g_dev.pool.push_back(p.b);
  }
  g_dev.pending.clear();
}

extern "C" int ops_hip_register_kernel(const char *name) {
  auto &kt = ops_kernel_table();
  for (size_t i = 0; i < kt.size(); i++)
    if (kt[i].name == name) return (int)i;
  ops_kernel_entry e;
  e.name = name;
  kt.push_back(e);
  return (int)kt.size() - 1;
}

extern "C" void ops_hip_set_block_shape(int kernel_id, int bx, int by, int bz) {
  auto &kt = ops_kernel_table();
  if (kernel_id < 0 || kernel_id >= (int)kt.size()) return;
  if (bx * by * bz > OPS_HIP_BLOCK_THREADS || bx <= 0 || by <= 0 || bz <= 0 ||
      (bx * by * bz) % OPS_WAVE != 0)
    ops_throw(OPS_INVALID_ARGUMENT,
              "ops_hip_set_block_shape: workgroup must be whole 64-lane wavefronts, <= 256 threads");
  kt[kernel_id].bx = bx;
  kt[kernel_id].by = by;
  kt[kernel_id].bz = bz;
}

extern "C" int ops_hip_kernel_count(void) { return (int)ops_kernel_table().size(); }

extern "C" int ops_hip_kernel_stats(int kernel_id, char *name_out, int name_cap, long *calls,
                                    double *total_ms, double *bytes) {
  ops_dev_timing_flush();
  auto &kt = ops_kernel_table();
  if (kernel_id < 0 || kernel_id >= (int)kt.size()) return 0;
  if (name_out && name_cap > 0) {
    strncpy(name_out, kt[kernel_id].name.c_str(), (size_t)name_cap - 1);
    name_out[name_cap - 1] = 0;
  }
  if (calls) *calls = kt[kernel_id].calls;
  if (total_ms) *total_ms = kt[kernel_id].total_ms;
  if (bytes) *bytes = kt[kernel_id].bytes;
  return 1;
}

extern "C" void ops_hip_reset_stats(void) {
  ops_dev_timing_flush();
  for (auto &k : ops_kernel_table()) {
    k.calls = 0;
    k.total_ms = 0;
    k.bytes = 0;
    k.host_s = 0;
  }
}

// ------------------------------------------------------------------ loop seam

static unsigned pow2_at_least(int v) {
  unsigned p = 1;
  while ((int)p < v && p < 256) p <<= 1;
  return p;
}

// bytes per dat argument by ops_compute_transfer's rule (ops_lib_core.cpp:1687-1697)
static double transfer_bytes(int dims, const int *start, const int *end, const ops_arg *a) {
  double size = 1.0;
  for (int i = 0; i < dims; i++)
    if (a->stencil->stride[i] != 0 && (end[i] - start[i]) > 0) size *= end[i] - start[i];
  return size * a->dat->elem_size * ((a->acc == OPS_READ || a->acc == OPS_WRITE) ? 1.0 : 2.0);
}

// ---- low-dimensional reductions (ops_dummy_singlenode.cpp:204-345): a dat argument with
// INC/MIN/MAX access into an edge dat with a collapsed dimension starts, the first time
// that access follows another write, from the reduction's identity over the whole allocation

template <typename T> __global__ __launch_bounds__(256) void lowdim_fill(T *p, size_t n, T v) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T> static void fill_identity(ops_dat d, int acc) {
  const T v = acc == OPS_INC ? T(0)
                             : (acc == OPS_MAX ? std::numeric_limits<T>::lowest() : std::numeric_limits<T>::max());
  const size_t n = d->mem / sizeof(T);
  hipLaunchKernelGGL((lowdim_fill<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g_dev.stream, (T *)d->data_d,
                     n, v);
  ops_hip_check_launch("low-dim reduction initialisation");
}

static bool lowdim_reduction(const ops_arg &a) {
  if (a.argtype != OPS_ARG_DAT || !a.opt || !a.dat || !a.stencil || !a.dat->e_dat) return false;
  if (a.acc != OPS_INC && a.acc != OPS_MIN && a.acc != OPS_MAX) return false;
  for (int d = 0; d < a.stencil->dims && d < OPS_MAX_DIM; d++)
    if (a.dat->size[d] == 1) return true;
  return false;
}

static void lowdim_init(const ops_arg &a) {
  if (!lowdim_reduction(a) || a.dat->edat_acc == a.acc) return;
  ops_dat d = a.dat;
  d->edat_acc = a.acc;
  if (!strcmp(d->type, "double")) fill_identity<double>(d, a.acc);
  else if (!strcmp(d->type, "float")) fill_identity<float>(d, a.acc);
  else if (!strcmp(d->type, "int")) fill_identity<int>(d, a.acc);
  else ops_throw(OPS_NOT_IMPLEMENTED, std::string("low-dim reduction into a dat of type ") + d->type);
}

// a dat argument the loop addresses (active or not: see ops_hip_loop_begin)
static inline bool live(const ops_arg &a) {
  return a.argtype == OPS_ARG_DAT && a.dat != nullptr && a.stencil != nullptr;
}

extern "C" int ops_hip_loop_begin(ops_hip_loop *lp, int kernel_id, const char *name,
                                  ops_block block, int dim, const int *range, ops_arg *args,
                                  int nargs) {
  ops_runtime_instance *in = ops_instance_get();
  lp->deferred = 0;
  if (g_chain.on && !g_chain.running) {
    if (chain_eligible(lp, dim, args, nargs)) {
      chain_record(lp, name, block, dim, range, args, nargs);
      lp->deferred = 1;
      lp->nblocks = 0;
      return 0;
    }
    chain_flush();  // a loop the chain cannot hold runs after it
  }
  // another launch of the stub whose batch is pending may join that batch: everything
  // else pending runs first
  g_batch_keep = g_batch_on && lp->batch != nullptr && lp->batch == g_batch_fn;
  ops_trid_flush();
  g_batch_keep = false;
  if (nargs > OPS_HIP_MAX_ARGS) ops_throw(OPS_NOT_IMPLEMENTED, "too many loop arguments");
  lp->kernel_id = kernel_id;
  lp->ndim = dim;
  lp->stream = (void *)g_dev.stream;
  lp->timed = 0;
  lp->bytes = 0;
  lp->alias = 0;
  // A dat argument is addressed whatever its opt flag, as the reference's sequential path
  // does (ops_seq.h:184-196 computes every dat argument's pointer; ops_lib_core.cpp:
  // 2319-2331 and ops_dummy_singlenode.cpp:140-148 exchange and mark it): a kernel may guard
  // its accesses with another flag than the one its argument was passed with
  // (apps/c/CloverLeaf_3D/update_halo.cpp:187-190 swaps energy0's and energy1's), and
  // then reads and writes the dat there too.
  for (int i = 0; i < nargs; i++)
    for (int j = i + 1; j < nargs; j++)
      if (live(args[i]) && live(args[j]) && args[i].dat == args[j].dat &&
          (args[i].acc != OPS_READ || args[j].acc != OPS_READ))
        lp->alias = 1;
  int start[OPS_MAX_DIM] = {0, 0, 0, 0, 0}, end[OPS_MAX_DIM] = {1, 1, 1, 1, 1};
  long total = 1;
  for (int d = 0; d < 3; d++) {
    lp->start[d] = 0;
    lp->count[d] = 1;
  }
  if (dim > 3) ops_throw(OPS_NOT_IMPLEMENTED, "ops_par_loop over more than 3 dimensions");
  for (int d = 0; d < dim; d++) {
    start[d] = range[2 * d];
    end[d] = range[2 * d + 1];
    int c = end[d] - start[d];
    if (c < 0) c = 0;
    lp->start[d] = start[d];
    lp->count[d] = c;
    total *= c;
  }
  for (int i = 0; i < nargs; i++) {
    ops_arg &a = args[i];
    lp->ptr[i] = nullptr;
    lp->red_part[i] = nullptr;
    lp->red_value[i] = nullptr;
    lp->red_init[i] = 0;
    for (int d = 0; d < 3; d++) {
      lp->step[i][d] = 0;
      lp->mg[i][d] = 1;
    }
    lp->xdim[i] = lp->ydim[i] = lp->zdim[i] = lp->mdim[i] = 0;
    if (live(a)) {
      ops_dat dat = a.dat;
      if (dat->locked_hd != 0)
        ops_throw(OPS_RUNTIME_ERROR,
                  std::string("Error: ops_par_loop called with dat ") + dat->name +
                      " that is locked by ops_dat_get_raw_pointer");
      if (dat->dirty_hd == 1) ops_put_data(dat);
      lowdim_init(a);
      const int *sst = a.stencil->stride;
      // multigrid stencils (templates/cpp/mpi_openmp/loop_host.cpp.j2:230-237): a restrict
      // argument is read at loop index * mgrid stride (affine: the stride scales the step), a
      // prolong argument at loop index / mgrid stride (the kernel divides; the pointer here
      // is the dat's index 0 and the steps are per coarse index)
      const int stype = a.stencil->type;
      if (((lp->prolong_mask >> i) & 1u) != (stype == 1 ? 1u : 0u))
        ops_throw(OPS_INVALID_ARGUMENT,
                  std::string("ops_par_loop ") + (name ? name : "") + ": argument " + std::to_string(i) +
                      (stype == 1 ? " has a prolong stencil but was not translated as one"
                                  : " was translated for a prolong stencil but has another"));
      auto mgs = [&](int d) { return (stype != 0 && d < a.stencil->dims) ? a.stencil->mgrid_stride[d] : 1; };
      size_t esz = in->OPS_soa ? (size_t)dat->type_size : (size_t)dat->elem_size;
      long long off = 0, cum = 1;
      for (int d = 0; d < dim; d++) {
        int s = (d < a.stencil->dims) ? sst[d] : 1;
        if (stype == 2) s *= mgs(d);
        const long long first = stype == 1 ? 0 : (long long)start[d] * s;
        off += (long long)esz * cum * (first - dat->base[d] - dat->d_m[d]);
        cum *= dat->size[d];
      }
      lp->ptr[i] = dat->data_d + off;
      int comp = in->OPS_soa ? 1 : dat->dim;
      long long cumd = 1;
      for (int d = 0; d < 3; d++) {
        int s = (d < dim && d < a.stencil->dims) ? sst[d] : 0;
        if (stype == 2 && d < dim) s *= mgs(d);
        const long long step = (long long)s * cumd * comp;
        if (step < 0 || step > INT_MAX)
          ops_throw(OPS_INVALID_ARGUMENT,
                    std::string("ops_par_loop ") + (name ? name : "") + ": argument " +
                        std::to_string(i) + " steps " + std::to_string(step) +
                        " elements along dimension " + std::to_string(d) +
                        ", beyond the 32-bit step of the generated kernels");
        lp->step[i][d] = (int)step;
        lp->mg[i][d] = (stype == 1 && d < dim) ? mgs(d) : 1;
        cumd *= dat->size[d];
      }
      lp->xdim[i] = dat->size[0];
      lp->ydim[i] = dat->size[1];
      lp->zdim[i] = dat->size[2];
      lp->mdim[i] = dat->dim;
      if (total > 0 && a.opt) lp->bytes += transfer_bytes(dim, start, end, &a);
    }
  }
  if (total == 0) {
    lp->nblocks = 0;
    return 0;
  }
  ops_check_loop_bounds(name, dim, start, end, args, nargs);
  // an inactive argument whose stencil would leave the allocation gets no address (its
  // kernel is expected not to touch it) rather than an error
  for (int i = 0; i < nargs; i++)
    if (lp->ptr[i] != nullptr && !args[i].opt &&
        !ops_loop_arg_in_bounds(name, dim, start, end, args[i], nullptr))
      lp->ptr[i] = nullptr;
  // workgroup shape: tuned, else 64 lanes along x (coalesced, one wavefront per row
  // segment) and the remaining threads stacked along y/z
  auto &kt = ops_kernel_table();
  const int vx = lp->vx >= 1 ? lp->vx : 1;
  // a launch narrower than one wavefront along x (a halo strip a few columns wide) keeps
  // the narrow heuristic shape: tuned shapes belong to a kernel's wide launches, and one
  // kernel may be launched over rows and over columns (a halo kernel's faces)
  const bool narrow = lp->count[0] < 64 * vx;
  unsigned bx, by, bz;
  if (kernel_id >= 0 && kernel_id < (int)kt.size() && kt[kernel_id].bx > 0 && !narrow) {
    bx = kt[kernel_id].bx;
    by = kt[kernel_id].by;
    bz = kt[kernel_id].bz;
  } else {
    // always a full 256-lane workgroup (four whole wavefronts): the wavefront reductions
    // exchange values across all 64 lanes, so a partial wavefront would read lanes that
    // do not exist. Lanes outside the range are masked by the kernel's bounds test.
    bx = pow2_at_least((lp->count[0] + vx - 1) / vx);
    if (bx > 64) bx = 64;
    if (in->block_size_x > 0) bx = in->block_size_x;
    if (dim >= 3) {
      by = pow2_at_least(lp->count[1]);
      if (by > OPS_HIP_BLOCK_THREADS / bx) by = OPS_HIP_BLOCK_THREADS / bx;
      if (in->block_size_y > 0) by = in->block_size_y;
      bz = OPS_HIP_BLOCK_THREADS / (bx * by);
      if (bz < 1) bz = 1;
    } else if (dim == 2) {
      by = OPS_HIP_BLOCK_THREADS / bx;
      if (in->block_size_y > 0) by = in->block_size_y;
      bz = 1;
    } else {
      // 1-D: no second dimension to stack along, so the workgroup is one row of lanes
      bx = OPS_HIP_BLOCK_THREADS;
      by = bz = 1;
      if (in->block_size_x > 0) {
        bx = in->block_size_x;
        by = OPS_HIP_BLOCK_THREADS / bx;  // whole wavefronts, as in 2-D
      }
    }
  }
  // autotuning: workgroup shapes tried on the kernel's first launches. Loops with a sum
  // reduction are left alone, since the shape sets the order of the partial sums.
  lp->tune_cand = -1;
  if (g_autotune && !narrow && kernel_id >= 0 && kernel_id < (int)kt.size() &&
      kt[kernel_id].tune_state >= 0 && kt[kernel_id].bx == 0 && in->block_size_x <= 0 &&
      in->block_size_y <= 0) {
    bool sum_red = false;
    for (int i = 0; i < nargs; i++)
      if (args[i].argtype == OPS_ARG_GBL && args[i].acc == OPS_INC) sum_red = true;
    if (sum_red) {
      kt[kernel_id].tune_state = -1;
    } else {
      // candidate 0 is the heuristic shape above (narrow along x for narrow ranges, such as
      // the halo strips), so tuning never ends worse than the default
      static const unsigned shapes2[][3] = {{0, 0, 0}, {64, 4, 1}, {128, 2, 1}, {256, 1, 1}, {32, 8, 1}};
      static const unsigned shapes3[][3] = {{0, 0, 0}, {64, 4, 1}, {64, 2, 2}, {32, 4, 2}, {64, 1, 4}};
      static const unsigned shapes1[][3] = {{0, 0, 0}, {256, 1, 1}, {128, 1, 1}};
      const int nc = dim >= 3 ? 5 : (dim == 2 ? 5 : 3);
      const int c = ops_tune_next(kernel_id, nc, 2);
      if (c >= 0) {
        const unsigned *sh = dim >= 3 ? shapes3[c] : (dim == 2 ? shapes2[c] : shapes1[c]);
        if (sh[0] != 0) {
          bx = sh[0];
          by = sh[1];
          bz = sh[2];
        }
        if (kt[kernel_id].tune_state == 1) lp->tune_cand = c;
      }
    }
  }
  if ((bx * by * bz) % OPS_WAVE != 0 || bx * by * bz > OPS_HIP_BLOCK_THREADS)
    ops_throw(OPS_INVALID_ARGUMENT,
              "ops_par_loop: workgroup shape must be a whole number of 64-lane wavefronts "
              "and at most 256 threads");
  // The generated kernels address a dat argument from the workgroup's origin (64-bit,
  // wave-uniform) plus the lane's offset inside the workgroup, tx*sx + ty*sy + tz*sz,
  // taken as 24-bit products along x and y (full-rate multiplies): the x step must be below
  // 2^23 elements (signed product: the lane's x offset is negative in aligned tiles), the y
  // step below 2^24, and the lane offset must fit a signed 32-bit sum.
  for (int i = 0; i < nargs; i++) {
    if (lp->ptr[i] == nullptr || args[i].argtype != OPS_ARG_DAT) continue;
    const long long lane_max = (long long)(bx - 1) * lp->step[i][0] +
                               (long long)(by - 1) * lp->step[i][1] +
                               (long long)(bz - 1) * lp->step[i][2];
    if (lp->step[i][0] >= (1 << 23) || lp->step[i][1] >= (1 << 24) || lane_max > INT_MAX)
      ops_throw(OPS_INVALID_ARGUMENT,
                std::string("ops_par_loop ") + (name ? name : "") + ": argument " +
                    std::to_string(i) + " has an x step of 2^23 or a y step of 2^24 elements or more");
  }
  // x tiles on 128-byte lines of the anchor argument's rows: each wavefront's row segment
  // is then whole cache lines (a row of a dat starts wherever its pitch puts it), so stores
  // fill lines rather than splitting them between workgroups. Loops with reductions keep
  // their tiles (the tiling sets the order of the partial sums).
  bool red_loop = false;
  for (int i = 0; i < nargs; i++)
    if (args[i].argtype == OPS_ARG_GBL && args[i].acc != OPS_READ) red_loop = true;
  // loops without reductions alternate their tile order launch by launch (a reduction's
  // tile order sets the order of its partial sums, so those keep theirs)
  static const bool alternate = [] {
    const char *e = getenv("OPS_ALTERNATE");
    return e == nullptr || atoi(e) != 0;
  }();
  static unsigned launches = 0;
  lp->rev = (alternate && !red_loop) ? (int)(launches++ & 1u) : 0;
  lp->align = 0;
  const int aa = lp->align_arg;
  if (aa >= 0 && aa < nargs && !red_loop && !narrow && vx == 1 && args[aa].argtype == OPS_ARG_DAT &&
      lp->ptr[aa] != nullptr && lp->step[aa][0] == 1) {
    const int esz = args[aa].dat->type_size;
    const int per_line = (esz == 4 || esz == 8) ? 128 / esz : 0;
    if (per_line > 0 && bx % (unsigned)per_line == 0) lp->align = per_line;
  }
  lp->block[0] = bx;
  lp->block[1] = by;
  lp->block[2] = bz;
  lp->grid[0] = (lp->count[0] + (lp->align ? lp->align - 1 : 0) + bx * vx - 1) / (bx * vx);
  lp->grid[1] = (lp->count[1] + by - 1) / by;
  lp->grid[2] = (lp->count[2] + bz - 1) / bz;
  lp->nblocks = lp->grid[0] * lp->grid[1] * lp->grid[2];
  lp->tiles = lp->nblocks;
  for (int d = 0; d < 3; d++) lp->tgrid[d] = lp->grid[d];
  // a loop with reductions runs on a 1-D grid of at most OPS_RED_GRID workgroups (a
  // multiple of the 8 XCDs) that walk the tiles grid-stride: its finisher then folds at
  // most that many partials per argument
  bool has_red = false;
  for (int i = 0; i < nargs; i++)
    if (args[i].argtype == OPS_ARG_GBL && args[i].acc != OPS_READ) has_red = true;
  if (has_red) {
    const unsigned g = lp->tiles < OPS_RED_GRID ? lp->tiles : OPS_RED_GRID;
    lp->grid[0] = g;
    lp->grid[1] = lp->grid[2] = 1;
    lp->nblocks = g;
  }

  // reduction scratch
  size_t need = 0;
  for (int i = 0; i < nargs; i++)
    if (args[i].argtype == OPS_ARG_GBL && args[i].acc != OPS_READ) {
      ops_reduction h = (ops_reduction)args[i].data;
      need += ((size_t)lp->nblocks * (size_t)h->size + 255) & ~(size_t)255;
    }
  if (need > g_dev.red_scratch_cap) {
    if (g_capturing)
      ops_throw(OPS_RUNTIME_ERROR, std::string("ops_par_loop ") + name +
                                       ": reduction scratch would grow during graph capture "
                                       "(run the loop once before capturing it)");
    if (g_dev.red_scratch) g_dev.red_scratch_retired.push_back(g_dev.red_scratch);
    size_t cap = need * 2;
    HIPCHECK(hipMalloc((void **)&g_dev.red_scratch, cap));
    g_dev.red_scratch_cap = cap;
  }
  size_t off = 0;
  for (int i = 0; i < nargs; i++)
    if (args[i].argtype == OPS_ARG_GBL && args[i].acc != OPS_READ) {
      ops_reduction h = (ops_reduction)args[i].data;
      lp->red_part[i] = g_dev.red_scratch + off;
      off += ((size_t)lp->nblocks * (size_t)h->size + 255) & ~(size_t)255;
      lp->red_value[i] = h->data_d;
      lp->red_init[i] = h->dirty_d ? 0 : 1;
    }
  // a timed launch runs alone: the pending batch goes first, outside its time. The stub
  // launches it between the two events (hipExtLaunchKernelGGL), so the time is the kernel's
  // own, without the queue around it or the reduction finisher after it
  if (lp->tune_cand >= 0 || g_timing) ops_loop_batch_flush();
  lp->ev_start = lp->ev_stop = nullptr;
  if (lp->tune_cand >= 0 || g_timing) {
    hipEvent_t ev[2];
    tune_begin(kernel_id, lp->tune_cand,
               lp->tune_cand >= 0 ? (double)lp->count[0] * (double)lp->count[1] * (double)lp->count[2] : 0.0,
               false, ev);
    lp->ev_start = (void *)ev[0];
    lp->ev_stop = (void *)ev[1];
    lp->timed = 1;
  }
  return 1;
}

// ------------------------------------------------------------------ reduction finish
// Every workgroup of a loop with reductions writes its partial (at most OPS_RED_GRID per
// argument); one launch then folds the partials of all the loop's reduction arguments, one
// workgroup per argument: each lane combines a lane-strided share in a fixed order, then
// the workgroup tree, and the result is combined into the handle's device value
// (overwritten on the handle's first use). No atomics, no device-scope fences.

struct RedEntry {
  const void *part;
  void *value;
  void *mirror;  // the handle's pinned host copy (written with value), or null
  unsigned *seq_ptr;  // pinned word that receives seq after the copy (release, system scope), or null
  unsigned seq;
  unsigned nb;  // partials (workgroups of the loop)
  int dim;
  int op;       // OPS_INC, OPS_MIN, OPS_MAX
  int type;     // 0 double, 1 float, 2 int
  int init;
};
struct RedTable {
  RedEntry e[OPS_HIP_MAX_ARGS];
};

template <typename OP, typename T>
__device__ __forceinline__ void red_fold(const RedEntry &E, T *lds) {
  const T *part = (const T *)E.part;
  T *value = (T *)E.value;
  const unsigned tid = threadIdx.x;
  for (int d = 0; d < E.dim; d++) {
    T a = OP::identity();
    for (unsigned i = tid; i < E.nb; i += 256) a = OP::op(a, part[(size_t)i * E.dim + d]);
    const T r = ops_block_reduce<OP>(a, lds);
    if (tid == 0) {
      const T v = E.init ? r : OP::op(value[d], r);
      value[d] = v;
      if (E.mirror) ((T *)E.mirror)[d] = v;
    }
  }
}

template <typename T> __device__ __forceinline__ void red_fold_op(const RedEntry &E, T *lds) {
  if (E.op == OPS_INC) red_fold<ops_red_inc<T>, T>(E, lds);
  else if (E.op == OPS_MIN) red_fold<ops_red_min<T>, T>(E, lds);
  else red_fold<ops_red_max<T>, T>(E, lds);
}

__global__ __launch_bounds__(256) void red_finish_kernel(const RedTable t) {
  __shared__ double lds[8];
  const RedEntry &E = t.e[blockIdx.x];
  if (E.type == 0) red_fold_op<double>(E, lds);
  else if (E.type == 1) red_fold_op<float>(E, (float *)lds);
  else red_fold_op<int>(E, (int *)lds);
  // the host may read the pinned copy as soon as it sees this launch's number
  if (threadIdx.x == 0 && E.seq_ptr)
    __hip_atomic_store(E.seq_ptr, E.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// the pinned sequence word of a handle: after the value's bytes, on its own 64-byte line
static unsigned *red_seq_word(ops_reduction h) {
  return (unsigned *)(h->data_m + (((size_t)std::max(h->size, 8) + 63) & ~(size_t)63));
}

void ops_dev_red_wait(ops_reduction h) {
  ops_trid_flush();  // pending batches and chains first: the finisher may be among them
  if (h->seq_valid) {
    // spin up to 200 us on the finisher's number, then block on the stream
    const unsigned *w = red_seq_word(h);
    const auto t0 = std::chrono::steady_clock::now();
    while (__atomic_load_n(w, __ATOMIC_ACQUIRE) != h->seq_expect) {
      if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(200)) break;
    }
    if (__atomic_load_n(w, __ATOMIC_ACQUIRE) == h->seq_expect) return;
  }
  HIPCHECK(hipStreamSynchronize(g_dev.stream));
}

// reduction handles finished inside the current graph capture (a replay gives them a newer
// device value, as the captured loops would)
static std::vector<ops_reduction> g_capture_reds;
static void capture_note_reduction(ops_reduction h) {
  if (std::find(g_capture_reds.begin(), g_capture_reds.end(), h) == g_capture_reds.end())
    g_capture_reds.push_back(h);
}

static void finish_reductions(const ops_hip_loop *lp, const ops_arg *args, int nargs) {
  RedTable t;
  unsigned n = 0;
  for (int i = 0; i < nargs; i++) {
    const ops_arg &a = args[i];
    if (a.argtype != OPS_ARG_GBL || a.acc == OPS_READ) continue;
    if (a.acc != OPS_INC && a.acc != OPS_MIN && a.acc != OPS_MAX)
      ops_throw(OPS_INVALID_ARGUMENT, "reduction with an unsupported access mode");
    const char *ty = ((ops_reduction)a.data)->type;
    int type;
    if (!strcmp(ty, "double")) type = 0;
    else if (!strcmp(ty, "float")) type = 1;
    else if (!strcmp(ty, "int")) type = 2;
    else ops_throw(OPS_NOT_IMPLEMENTED, std::string("reduction of type ") + ty);
    ops_reduction h = (ops_reduction)a.data;
    // a captured launch replays with its arguments fixed: it stores no number, and reads of
    // the handle after it wait for the stream
    unsigned *seq_ptr = nullptr;
    if (h->data_m && !g_capturing) {
      seq_ptr = red_seq_word(h);
      h->seq_valid = 1;
      h->seq_expect++;
    } else {
      h->seq_valid = 0;
      if (g_capturing) capture_note_reduction(h);
    }
    t.e[n++] = {lp->red_part[i], lp->red_value[i], h->data_m, seq_ptr, h->seq_expect, lp->nblocks, a.dim, a.acc,
                type, lp->red_init[i]};
  }
  if (n == 0) return;
  hipLaunchKernelGGL(red_finish_kernel, dim3(n), dim3(256), 0, g_dev.stream, t);
  ops_hip_check_launch("reduction finish");
}

extern "C" void ops_hip_loop_end(ops_hip_loop *lp, ops_arg *args, int nargs) {
  if (lp->deferred) return;  // recorded in a loop chain: its launches do this
  if (lp->nblocks > 0) finish_reductions(lp, args, nargs);
  if (lp->timed && lp->tune_cand >= 0) {
    tune_end(lp->kernel_id, lp->ndim >= 2 ? 5 : 3, 2, false);
  } else if (lp->timed) {
    if (g_dev.pending.size() >= 8192) ops_dev_timing_flush();
  }
  auto &kt = ops_kernel_table();
  if (lp->kernel_id >= 0 && lp->kernel_id < (int)kt.size()) {
    kt[lp->kernel_id].calls++;
    kt[lp->kernel_id].bytes += lp->bytes;
  }
  for (int i = 0; i < nargs; i++) {
    ops_arg &a = args[i];
    if (live(a) && a.acc != OPS_READ) {
      a.dat->dirty_hd = 2;
      if (a.opt && a.dat->e_dat) a.dat->edat_acc = a.acc;
    }
    if (a.argtype == OPS_ARG_GBL && a.acc != OPS_READ && lp->nblocks > 0)
      ((ops_reduction)a.data)->dirty_d = 1;
  }
}

// ------------------------------------------------------------------ affine copies
// One thread per (point, component) of an up-to-5D box. Source and destination
// addresses are affine in the box coordinates (negative strides express flips and
// permutations), so a halo transfer is a single direct copy with no staging buffer.

struct AffineCopy {
  int ndim;
  int count[OPS_MAX_DIM];
  long long sstride[OPS_MAX_DIM];  // in source elements
  long long dstride[OPS_MAX_DIM];  // in destination elements
  long long sbase, dbase;
  int ncomp;
  long long scomp, dcomp;  // component strides
  long long total;         // points * ncomp
};

template <typename S, typename M, typename D>
__global__ __launch_bounds__(256) void affine_copy_kernel(const S *__restrict__ src,
                                                          D *__restrict__ dst, AffineCopy c) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.total) return;
  int comp = (int)(t % c.ncomp);
  long long p = t / c.ncomp;
  long long so = c.sbase + comp * c.scomp, dof = c.dbase + comp * c.dcomp;
  for (int d = 0; d < c.ndim; d++) {
    long long q = p % c.count[d];
    p /= c.count[d];
    so += q * c.sstride[d];
    dof += q * c.dstride[d];
  }
  dst[dof] = (D)(M)src[so];
}

template <typename S, typename M, typename D>
static void launch_affine(const void *src, void *dst, const AffineCopy &c) {
  if (c.total <= 0) return;
  unsigned blocks = (unsigned)((c.total + 255) / 256);
  hipLaunchKernelGGL((affine_copy_kernel<S, M, D>), dim3(blocks), dim3(256), 0, g_dev.stream,
                     (const S *)src, (D *)dst, c);
  ops_hip_check_launch("affine_copy");
}

static int type_code(const char *t) {
  if (!strcmp(t, "double")) return 8;
  if (!strcmp(t, "float")) return 4;
  return 0;
}

static void launch_by_size(int size, const void *src, void *dst, const AffineCopy &c) {
  switch (size) {
    case 1: launch_affine<unsigned char, unsigned char, unsigned char>(src, dst, c); break;
    case 2: launch_affine<unsigned short, unsigned short, unsigned short>(src, dst, c); break;
    case 4: launch_affine<unsigned, unsigned, unsigned>(src, dst, c); break;
    case 8: launch_affine<unsigned long long, unsigned long long, unsigned long long>(src, dst, c); break;
    default: ops_throw(OPS_NOT_IMPLEMENTED, "copy of unsupported element size");
  }
}

// Halo region mapping of ops_halo_transfer (ops_device_singlenode_common.cpp:286-336):
// dat dimension i walks halo axis |dir[i]|-1, forwards from base-d_m-base_dat when
// dir > 0, backwards from the end of that window when dir < 0.
static void halo_side(ops_dat dat, const int *hbase, const int *dir, const int *iter,
                      long long *stride_per_axis, long long *base_elem) {
  long long cum = 1;
  *base_elem = 0;
  for (int a = 0; a < OPS_MAX_DIM; a++) stride_per_axis[a] = 0;
  for (int i = 0; i < OPS_MAX_DIM; i++) {
    int ax = abs(dir[i]) - 1;
    long long first, step;
    if (dir[i] > 0) {
      first = hbase[i] - dat->d_m[i] - dat->base[i];
      step = 1;
    } else {
      first = (long long)(hbase[i] - 1 - dat->d_m[i] - dat->base[i]) + iter[ax];
      step = -1;
    }
    *base_elem += first * cum;
    stride_per_axis[ax] += step * cum;
    cum *= dat->size[i];
  }
}

void ops_dev_halo_copy(ops_halo h) {
  ops_trid_flush();
  ops_dat f = h->from, t = h->to;
  AffineCopy c;
  c.ndim = OPS_MAX_DIM;
  long long pts = 1;
  for (int a = 0; a < OPS_MAX_DIM; a++) {
    c.count[a] = h->iter_size[a];
    pts *= h->iter_size[a];
  }
  long long fs[OPS_MAX_DIM], ts[OPS_MAX_DIM], fb, tb;
  halo_side(f, h->from_base, h->from_dir, h->iter_size, fs, &fb);
  halo_side(t, h->to_base, h->to_dir, h->iter_size, ts, &tb);
  int ncomp = f->dim < t->dim ? f->dim : t->dim;
  // element strides in units of one component: AoS, or SoA with each component one
  // allocation plane further on (ops_host_singlenode.cpp:124-160)
  const int soa = ops_instance_get()->OPS_soa;
  long long fplane = 1, tplane = 1;
  for (int a = 0; a < OPS_MAX_DIM; a++) {
    fplane *= f->size[a];
    tplane *= t->size[a];
  }
  for (int a = 0; a < OPS_MAX_DIM; a++) {
    c.sstride[a] = soa ? fs[a] : fs[a] * f->dim;
    c.dstride[a] = soa ? ts[a] : ts[a] * t->dim;
  }
  c.sbase = soa ? fb : fb * f->dim;
  c.dbase = soa ? tb : tb * t->dim;
  c.ncomp = ncomp;
  c.scomp = soa ? fplane : 1;
  c.dcomp = soa ? tplane : 1;
  c.total = pts * ncomp;
  if (pts <= 0) return;
  int fc = type_code(f->type), tc = type_code(t->type);
  if (f->type_size != t->type_size && fc && tc) {
    // mixed precision: through the narrower storage type, as the staged buffer does
    if (fc == 8 && tc == 4) launch_affine<double, float, float>(f->data_d, t->data_d, c);
    else launch_affine<float, float, double>(f->data_d, t->data_d, c);
  } else {
    if (f->type_size != t->type_size)
      ops_throw(OPS_NOT_IMPLEMENTED, "halo transfer between dats of different types");
    launch_by_size(f->type_size, f->data_d, t->data_d, c);
  }
}

// Slab copy between a dat and a dense, packed buffer (no halo): range is given in
// interior coordinates [begin, end) per dimension (ops_util.cpp:353-500).
void ops_dev_slab_copy(ops_dat dat, char *buf, int buf_on_device, const int *range, int dir) {
  ops_trid_flush();
  ops_runtime_instance *in = ops_instance_get();
  int nd = dat->block->dims;
  AffineCopy c;
  c.ndim = nd;
  long long pts = 1, cum = 1, bcum = 1, base = 0;
  int soa = in->OPS_soa;
  long long plane = 1;
  for (int d = 0; d < nd; d++) plane *= dat->size[d];
  for (int d = 0; d < nd; d++) {
    int n = range[2 * d + 1] - range[2 * d];
    if (n < 0) n = 0;
    c.count[d] = n;
    pts *= n;
    long long datstride = soa ? cum : cum * dat->dim;
    base += (long long)(range[2 * d] - dat->d_m[d]) * datstride;
    long long bufstride = bcum * dat->dim;
    if (dir == 0) {
      c.sstride[d] = datstride;
      c.dstride[d] = bufstride;
    } else {
      c.sstride[d] = bufstride;
      c.dstride[d] = datstride;
    }
    cum *= dat->size[d];
    bcum *= n;
  }
  c.ncomp = dat->dim;
  long long dcompstride = soa ? plane : 1;
  if (dir == 0) {
    c.sbase = base;
    c.dbase = 0;
    c.scomp = dcompstride;
    c.dcomp = 1;
  } else {
    c.sbase = 0;
    c.dbase = base;
    c.scomp = 1;
    c.dcomp = dcompstride;
  }
  c.total = pts * dat->dim;
  size_t bytes = (size_t)pts * (size_t)dat->elem_size;
  if (pts == 0) return;
  char *dbuf = buf;
  if (!buf_on_device) dbuf = (char *)ops_dev_malloc(bytes);
  if (dir == 0) {
    launch_by_size(dat->type_size, dat->data_d, dbuf, c);
    if (!buf_on_device) ops_dev_d2h(buf, dbuf, bytes);
  } else {
    if (!buf_on_device) ops_dev_h2d(dbuf, buf, bytes);
    launch_by_size(dat->type_size, dbuf, dat->data_d, c);
  }
  if (!buf_on_device) {
    ops_dev_sync();
    ops_dev_free(dbuf);
  }
}

// ------------------------------------------------------------------ bandwidth probe

typedef double ops_v2d __attribute__((ext_vector_type(2)));

// one 16-byte element per lane on a full grid, streamed past the caches: the variant of a
// plain copy that reaches the highest HBM bandwidth on gfx950
__global__ __launch_bounds__(256) void copy_probe_kernel(const ops_v2d *__restrict__ a,
                                                         ops_v2d *__restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(&a[i]), &b[i]);
}

extern "C" double ops_hip_copy_probe(size_t bytes, int reps) {
  try {
    ops_trid_flush();
    if (!g_dev.live || reps < 1 || bytes < 16) return -1.0;
    const size_t n = bytes / sizeof(ops_v2d);
    if ((n + 255) / 256 > 0x7fffffffu) return -1.0;
    void *a = ops_dev_malloc(n * sizeof(ops_v2d)), *b = ops_dev_malloc(n * sizeof(ops_v2d));
    ops_dev_memset(a, 0, n * sizeof(ops_v2d));
    hipEvent_t e0 = get_event(), e1 = get_event();
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, g_dev.stream,
                       (const ops_v2d *)a, (ops_v2d *)b, n);  // warm-up
    HIPCHECK(hipEventRecord(e0, g_dev.stream));
    for (int r = 0; r < reps; r++)
      hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, g_dev.stream,
                         (const ops_v2d *)a, (ops_v2d *)b, n);
    HIPCHECK(hipEventRecord(e1, g_dev.stream));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    g_dev.pool.push_back(e0);
    g_dev.pool.push_back(e1);
    ops_dev_free(a);
    ops_dev_free(b);
    return 2.0 * (double)(n * sizeof(ops_v2d)) * reps / (ms * 1e-3) / 1e9;
  } catch (std::exception &) {
    return -1.0;
  }
}

// ------------------------------------------------------------------ graph capture

extern "C" int ops_hip_graph_begin(void) {
  if (!g_dev.live) return 1;
  ops_trid_flush();  // work deferred before the capture runs now, not inside the graph
  if (g_timing) ops_throw(OPS_RUNTIME_ERROR, "ops_hip_graph_begin: per-kernel timing is on");
  if (hipStreamBeginCapture(g_dev.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return 1;
  g_capturing = true;
  g_capture_reds.clear();
  return 0;
}

// the handle ops_hip_graph_end returns: the executable graph and the reduction handles its
// loops finish
struct GraphRec {
  hipGraphExec_t exec;
  std::vector<ops_reduction> reds;
};

extern "C" void *ops_hip_graph_end(void) {
  ops_trid_flush();
  hipGraph_t g = nullptr;
  g_capturing = false;
  if (hipStreamEndCapture(g_dev.stream, &g) != hipSuccess || g == nullptr) return nullptr;
  hipGraphExec_t e = nullptr;
  const hipError_t st = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (st != hipSuccess) return nullptr;
  return new GraphRec{e, g_capture_reds};
}

extern "C" int ops_hip_graph_replay(void *graph, int times) {
  GraphRec *r = (GraphRec *)graph;
  for (int i = 0; i < times; i++)
    if (hipGraphLaunch(r->exec, g_dev.stream) != hipSuccess) return 1;
  // the replayed finishers wrote these handles' device values and pinned copies (with no
  // launch number: reads wait for the stream); the handles count as used in a reduction
  for (ops_reduction h : r->reds) {
    h->dirty_d = 1;
    h->seq_valid = 0;
    h->initialized = 1;
  }
  return 0;
}

extern "C" void ops_hip_graph_free(void *graph) {
  GraphRec *r = (GraphRec *)graph;
  if (!r) return;
  (void)hipGraphExecDestroy(r->exec);
  delete r;
}
