// Internal state of libops_hip.so. Not part of the public ABI.
#ifndef OPS_AMD_INTERNAL_H
#define OPS_AMD_INTERNAL_H

#include <map>
#include <string>
#include <vector>

#include "ops_hip_rt.h"
#include "ops_lib_core.h"

// workgroups of a loop with reductions (a multiple of the 8 XCDs): its tiles are walked
// grid-stride, so the finisher folds at most this many partials per argument
#define OPS_RED_GRID 2048u

struct ops_kernel_entry {
  std::string name;
  long calls = 0;
  double total_ms = 0.0;  // sum of HIP-event durations (timing on)
  double bytes = 0.0;     // ops_compute_transfer bytes summed over calls
  double host_s = 0.0;    // host wall time spent in begin..end (stub overhead)
  int bx = 0, by = 0, bz = 0;  // tuned workgroup shape, 0 = default
  // autotuning (OPS_AUTOTUNE): candidates are tried in turn on real launches, each timed
  // with HIP events, and the fastest average is kept
  int tune_state = 0;  // 0 not started, 1 measuring, 2 decided, -1 not tunable
  int tune_calls = 0;  // measured launches so far
  int tune_choice = -1;
  double cand_ms[8] = {0};
  int cand_n[8] = {0};
  double cand_pts[8] = {0};  // loop points measured per candidate
  int tune_warm = 0;         // first measured launch seen (it is left out of the averages)
};

struct ops_const_entry {
  const void *symbol = nullptr;  // host handle of the __constant__/__device__ symbol
  void *dev = nullptr;           // its device address (resolved on first use)
  long nbytes = 0;      // fixed size, or < 0 for runtime-sized (symbol holds a pointer)
  void *owned = nullptr;  // runtime-sized: current device allocation
  long owned_bytes = 0;
};

// the runtime's state of an OPS session (OPS_instance holds OPS_diags and OPS_soa)
struct ops_runtime_instance : OPS_instance {
  int block_size_x = 0, block_size_y = 0, block_size_z = 0;
  int device_id = 0;
  int timing = 0;  // per-kernel HIP events
  std::vector<ops_block> blocks;
  std::vector<ops_dat> dats;
  std::vector<ops_stencil> stencils;
  std::vector<ops_reduction> reductions;
  std::vector<ops_halo> halos;
  std::vector<ops_halo_group> groups;
  ops_stencil internal0[OPS_MAX_DIM + 1] = {};  // one-point stencils per dimensionality
  std::string dump_dir;                          // OPS_DUMP_DIR= (dats written at ops_exit)
};

ops_runtime_instance *ops_instance_get(bool create = false);
[[noreturn]] void ops_throw(int code, const std::string &msg);
// Whether a dat argument of a loop over [start, end) stays inside its allocation (stencil
// extent included); the message for the first dimension that leaves it goes to *why.
bool ops_loop_arg_in_bounds(const char *name, int dim, const int *start, const int *end,
                            const ops_arg &a, std::string *why);
// Throws OPS_INVALID_ARGUMENT when an active (opt != 0) dat argument of a loop over
// [start, end) would be addressed outside its allocation.
void ops_check_loop_bounds(const char *name, int dim, const int *start, const int *end,
                           const ops_arg *args, int nargs);

// kernel registry (generated stubs register on first call) and the device homes of
// constants, keyed by symbol (each application module has its own), for the whole process
std::vector<ops_kernel_entry> &ops_kernel_table();
std::map<const void *, ops_const_entry> &ops_const_table();

// ---- device layer (ops_device.hip) ----
void ops_dev_init(ops_runtime_instance *inst);
void ops_dev_exit(ops_runtime_instance *inst);
void *ops_dev_malloc(size_t bytes);
void ops_dev_free(void *p);
// pinned, coherent host memory that kernels may write (reduction results read on the host)
void *ops_dev_host_alloc(size_t bytes);
void ops_dev_host_free(void *p);
// wait for a reduction handle's pinned copy (the finisher's sequence number, then the stream)
void ops_dev_red_wait(ops_reduction h);
void ops_dev_memset(void *p, int v, size_t bytes);
void ops_dev_h2d(void *dst, const void *src, size_t bytes);   // synchronous
void ops_dev_d2h(void *dst, const void *src, size_t bytes);   // synchronous (drains stream)
void ops_dev_d2d(void *dst, const void *src, size_t bytes);   // stream-ordered
void ops_dev_stage_h2d(void *dst, const void *src, size_t bytes);  // stream-ordered, src reusable
void ops_dev_sync();
// the runtime stream itself (no flush of deferred work)
void *ops_hip_get_stream_raw();
// launches the deferred tridiagonal batch, if any (every launch, copy and synchronisation
// of the device layer calls it first)
void ops_trid_flush();
// launches the pending loop batch, if any (ops_trid_flush calls it first)
void ops_loop_batch_flush();
// frees the tridiagonal solver's scratch (ops_exit; graphs that used it are gone by then)
void ops_trid_release();
// takes the device address of the allocation-free solve's static scratch on `device`
// (ops_dev_init, with that device current; once per device)
void ops_trid_device_init(int device);
// Autotuning seam. ops_tune_next returns the candidate (0..ncand-1) the next launch of
// kernel kid should use while it is being measured, the chosen one afterwards, or -1 when
// tuning is off (or the kernel is not tunable). A measured launch is bracketed by
// ops_tune_begin/ops_tune_end, which decides once every candidate has `reps` samples;
// candidates are compared by time per point (`pts`: the points the launch covers).
int ops_tune_next(int kid, int ncand, int reps);
void ops_tune_begin(int kid, int cand, double pts);
void ops_tune_end(int kid, int ncand, int reps);
extern int g_autotune;
void ops_dev_halo_copy(ops_halo h);
// packed slab copy between a dat (device) and a dense buffer; dir 0: dat->buf, 1: buf->dat
void ops_dev_slab_copy(ops_dat dat, char *buf, int buf_on_device, const int *range, int dir);
void *ops_dev_symbol_address(const void *symbol);
void ops_dev_timing_flush();

// ---- host-side helpers (ops_core.cpp) ----
void ops_put_data(ops_dat dat);  // host-dirty -> device
void ops_dat_host_alloc(ops_dat dat);
void ops_dump_all_dats(ops_runtime_instance *inst, const char *dir);

#endif
