/*
 * Seam between translator-generated loop code and libops_hip.so (C ABI).
 *
 * It replaces the runtime calls the reference's generated HIP stubs make around each
 * kernel launch (ops/c/include/ops_internal2.h:220-339, ops_device_rt_support.h:49-58:
 * ops_timing_realloc, ops_H_D_exchanges_device, ops_halo_exchanges,
 * reallocReductArrays/mvReductArraysToDevice/mvReductArraysToHost,
 * ops_set_dirtybit_device, ops_set_halo_dirtybit3, ops_compute_transfer) and the
 * const upload of templates/cpp/cuda/master_kernel.cpp.j2:40-67.
 *
 * A generated stub calls ops_hip_loop_begin(), launches its kernel with the geometry and
 * device addresses the descriptor holds, then calls ops_hip_loop_end(). Reductions are
 * finished on the device (per-workgroup partials, folded by a finishing kernel into the
 * handle's device value); only ops_reduction_result moves the value to the host.
 */
#ifndef OPS_AMD_HIP_RT_H
#define OPS_AMD_HIP_RT_H

#include "ops_lib_core.h"

#define OPS_HIP_MAX_ARGS 24
#define OPS_HIP_BLOCK_THREADS 256

#ifdef __cplusplus
extern "C" {
#endif

/* Per-launch descriptor filled by ops_hip_loop_begin. Index i follows the loop's args. */
typedef struct ops_hip_loop {
  int ndim;
  int start[3];       /* first loop index per dimension */
  int count[3];       /* iterations per dimension (1 for unused dimensions) */
  unsigned grid[3];   /* workgroups per dimension */
  unsigned block[3];  /* threads per workgroup (whole 64-lane wavefronts, at most 256) */
  int vx;             /* loop points per lane along x (set by the stub before begin) */
  unsigned nblocks;   /* grid[0]*grid[1]*grid[2] */
  void *stream;       /* hipStream_t every launch goes to */
  /* dat args: device address of the loop's first point, and element steps per loop
     index step in x/y/z (stencil strides and components included). */
  char *ptr[OPS_HIP_MAX_ARGS];
  int step[OPS_HIP_MAX_ARGS][3];
  int xdim[OPS_HIP_MAX_ARGS];  /* size[0] (ACC row pitch, in elements of one component) */
  int ydim[OPS_HIP_MAX_ARGS];  /* size[1] */
  int zdim[OPS_HIP_MAX_ARGS];  /* size[2] */
  int mdim[OPS_HIP_MAX_ARGS];  /* components per point */
  /* reduction args: per-block partials, completion counter, running device value */
  void *red_part[OPS_HIP_MAX_ARGS];  /* per-workgroup partials, [workgroup][dim] */
  void *red_value[OPS_HIP_MAX_ARGS];
  int red_init[OPS_HIP_MAX_ARGS];  /* 1: overwrite the running value (first use) */
  /* bookkeeping */
  int kernel_id;
  int timed;
  int tune_cand; /* autotuning candidate measured by this launch, -1 otherwise */
  unsigned prolong_mask; /* set by the stub: bit i = argument i was translated with
                            prolong addressing (coarse index = loop index / mgrid stride) */
  int mg[OPS_HIP_MAX_ARGS][3]; /* prolong arguments: the mgrid stride per dimension (1 otherwise) */
  int alias;     /* 1 if one dat is passed to two arguments and at least one writes it: the
                    generated stub then launches the kernel variant without no-alias
                    pointer parameters */
  double bytes;  /* ops_compute_transfer bytes of this launch */
  void (*batch)(void); /* set by the stub: the launcher of its pending loop batch (NULL
                          when the loop cannot join one) */
  /* set by the stub: calls the stub again for a sub-range (a loop chain runs a recorded
     loop band by band through it); NULL when the loop cannot be recorded */
  void (*relaunch)(const char *name, ops_block block, int dim, int *range, ops_arg *args);
  int deferred;  /* set by ops_hip_loop_begin: 1 when the loop was recorded in a chain */
  /* timed launches (timed != 0): the stub launches with hipExtLaunchKernelGGL between these
     two events (hipEvent_t), which then bracket the kernel's own execution */
  void *ev_start, *ev_stop;
  /* loops with reductions: the launch is a capped 1-D grid of nblocks workgroups that walk
     the `tiles` tiles of the tile grid tgrid[0] x tgrid[1] x tgrid[2] grid-stride */
  unsigned tiles;
  unsigned tgrid[3];
  /* set by the stub: the dat argument whose rows the x tiles start on 128-byte lines of
     (-1: none). ops_hip_loop_begin sets align to the elements per 128-byte line of that
     argument when the launch uses aligned tiles (a loop without reductions, one point per
     lane, wider than a wavefront, the argument contiguous along x), else to 0; the grid then
     has room for the shift of up to align - 1 points at the start of each row */
  int align_arg;
  int align;
  /* set by ops_hip_loop_begin for loops without reductions: 1 on every other launch, whose
     workgroups then walk each XCD's run of tiles from its end (ops_block_coords_dir) */
  int rev;
} ops_hip_loop;

/* Registers a generated loop name once; returns its id for timing and tuning tables. */
int ops_hip_register_kernel(const char *name);

/*
 * Prepares a launch: argument checks, host->device uploads of host-dirty dats, device
 * addresses, launch geometry (block shape from the tuning table for kernel_id, or a
 * default), reduction scratch. Returns 0 when the range is empty (nothing to launch; the
 * stub must still call ops_hip_loop_end), 1 otherwise.
 */
int ops_hip_loop_begin(ops_hip_loop *lp, int kernel_id, const char *name, ops_block block,
                       int dim, const int *range, ops_arg *args, int nargs);
/* Completes bookkeeping after the launch: dirty bits, reduction state, timing. */
void ops_hip_loop_end(ops_hip_loop *lp, ops_arg *args, int nargs);

/* Constants. A translated application defines ops_decl_const_char and
   ops_hip_const_device_ptr itself, over its own table of device symbols (the reference's
   generated master file defines ops_decl_const_char per application,
   templates/cpp/cuda/master_kernel.cpp.j2:40-67), and calls these with the symbol's host
   handle (HIP_SYMBOL(x)) and its declared size; nbytes < 0: runtime-sized, the symbol
   holds a device pointer the runtime allocates and owns until ops_exit.
   ops_hip_const_set uploads `bytes` bytes of `data` in stream order;
   ops_hip_const_symbol_ptr returns the device address of a fixed-size symbol (throws
   OPS_INVALID_ARGUMENT for a null or runtime-sized one). */
void ops_hip_const_set(const void *symbol, long nbytes, const char *name, const char *data,
                       long bytes);
void *ops_hip_const_symbol_ptr(const void *symbol, long nbytes, const char *name);

/* Checks a HIP status from generated code; throws OPSException(OPS_RUNTIME_ERROR). */
void ops_hip_check(int status, const char *what, const char *file, int line);
/* Checks the launch just made (hipGetLastError). */
void ops_hip_check_launch(const char *name);

/* The runtime's single work stream and a host-side barrier on it. */
void *ops_hip_get_stream(void);
void ops_hip_device_sync(void);

/* Tuning: preferred workgroup shape (bx*by*bz must be 256) for a registered kernel. */
void ops_hip_set_block_shape(int kernel_id, int bx, int by, int bz);
/* Per-kernel HIP-event timing: on/off, and the table the timing output is built from. */
void ops_hip_set_timing(int on);
int ops_hip_kernel_count(void);
int ops_hip_kernel_stats(int kernel_id, char *name_out, int name_cap, long *calls,
                         double *total_ms, double *bytes);
void ops_hip_reset_stats(void);
/* Graph capture of a fixed launch sequence (replayed without host launch cost). Between
   begin and end only launches may be issued on the runtime stream: no host
   synchronisation, no allocation, no ops_reduction_result, no host<->device copies.
   begin returns 0 on success; end returns an executable graph (NULL on failure). */
/* Loop batches (no reference counterpart; the reference launches every ops_par_loop on
   its own). Between ops_hip_batch_begin and ops_hip_batch_end the caller promises that the
   loops it issues are independent: no dat written by one is read or written by another.
   Consecutive launches of one generated kernel (loops without reductions) are then
   collected and launched together, one launch per batch, each workgroup finding its loop
   from the batch's prefix sums of workgroups. Any other runtime operation, another
   kernel's launch and ops_hip_batch_end launch the pending batch first. Results are those
   of the separate launches. */
void ops_hip_batch_begin(void);
void ops_hip_batch_end(void);
/* for generated stubs: 1 if this launch joins a batch (batching on); `flush` launches the
   stub's pending batch (it also identifies the batch: a launch of another stub launches
   the pending one first), and is called by the runtime before any other operation */
int ops_hip_batch_join(int kernel_id, void (*flush)(void));
/* Loop chains, the GPU form of the reference's lazy loop tiling (ops/c/src/core/
   ops_lazy.cpp:186-1401; there the queue and its tiling are driven by OPS_TILING). Between
   ops_hip_chain_begin(tiles) and ops_hip_chain_end the loops that can be tiled (no
   reductions, plain stencils with unit stride along the last loop dimension, two or three
   dimensions) are recorded; the chain then runs band by band along the last dimension,
   `tiles` bands with per-loop skewed boundaries derived from the loops' stencils, so that
   data one loop writes is re-read from the caches by the next. No promise is asked of the
   caller: the dependencies are computed, results are bitwise those of the loops run in
   order, and any other runtime operation runs the recorded chain first. */
void ops_hip_chain_begin(int tiles);
void ops_hip_chain_end(void);
int ops_hip_graph_begin(void);
void *ops_hip_graph_end(void);
int ops_hip_graph_replay(void *graph, int times);
void ops_hip_graph_free(void *graph);
/* Device-side chaining for captured steps (no reference counterpart; the reference reads
   reductions and uploads constants on the host, ops_lib_core.cpp:1143-1149, and
   cuda/master_kernel.cpp.j2:40-67):
   - the device address of a fixed-size ops_decl_const symbol, which a kernel of the
     application may overwrite (later launches read the new value);
   - the device address of a reduction handle's running value (the finished result of the
     last loop that reduced into it);
   - discard: forget the handle's device value without reading it, as ops_reduction_result
     would (the next reducing loop overwrites it instead of combining). */
void *ops_hip_const_device_ptr(const char *name);
void *ops_hip_reduction_device_ptr(ops_reduction h);
void ops_hip_reduction_discard(ops_reduction h);
/* Measured HBM bandwidth: a plain copy of `bytes` between two fresh device buffers,
   `reps` times on the runtime stream; returns (read + write bytes) / time in GB/s, or a
   negative value on error. Allocates and frees its buffers outside the timed region. */
double ops_hip_copy_probe(size_t bytes, int reps);

#ifdef __cplusplus
}
#endif

#endif /* OPS_AMD_HIP_RT_H */
