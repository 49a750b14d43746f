/*
 * Batched scalar tridiagonal solves over OPS dats (replaces the reference's shim to the
 * external Tridsolver library: ops/c/include/ops_tridiag.h:45-82,
 * ops/c/src/tridiag/ops_tridiag.cpp:48-249, ops_tridiag_cuda.cpp:48-251).
 *
 * For every line along `solvedim` inside `range` (loop-index coordinates, half-open
 * {x0,x1,y0,y1,z0,z1}) solve  a_i x_{i-1} + b_i x_i + c_i x_{i+1} = d_i
 * (a of the first row and c of the last row are ignored). ops_tridMultiDimBatch writes
 * x into d; ops_tridMultiDimBatch_Inc adds x to u and leaves d unchanged.
 * a, b, c, d (and u) must have identical sizes and type ("double" or "float").
 * The solves run on the GPU on the dats' HBM copies; nothing moves to the host.
 */
#ifndef OPS_AMD_TRIDIAG_H
#define OPS_AMD_TRIDIAG_H

#include "ops_lib_core.h"

#ifdef __cplusplus
/* Solver context (ops_tridiag.h:45-70). The single-GPU solver has no strategy choices;
   the setters are accepted for source compatibility. */
class ops_tridsolver_params {
public:
  enum SolveStrategy {
    GATHER_SCATTER = 0,
    ALLGATHER,
    JACOBI,
    PCR,
    LATENCY_HIDING_INTERLEAVED,
    LATENCY_HIDING_TWO_STEP
  };
  explicit ops_tridsolver_params(ops_block block) : block_(block), strategy_(PCR) {}
  ops_tridsolver_params(ops_block block, SolveStrategy s) : block_(block), strategy_(s) {}
  void set_jacobi_params(double rtol, double atol, int maxiter) {
    (void)rtol;
    (void)atol;
    (void)maxiter;
  }
  void set_batch_size(int batch_size) { (void)batch_size; }
  void set_cuda_opts(int opt_x, int opt_y, int opt_z) {
    (void)opt_x;
    (void)opt_y;
    (void)opt_z;
  }
  void set_cuda_sync(int sync) { (void)sync; }
  ops_block block_;
  SolveStrategy strategy_;
};
extern "C" {
#else
typedef struct ops_tridsolver_params ops_tridsolver_params;
#endif

void ops_tridMultiDimBatch(int ndim, int solvedim, int *range, ops_dat a, ops_dat b, ops_dat c,
                           ops_dat d, ops_tridsolver_params *tridsolver_ctx);

void ops_tridMultiDimBatch_Inc(int ndim, int solvedim, int *range, ops_dat a, ops_dat b,
                               ops_dat c, ops_dat d, ops_dat u,
                               ops_tridsolver_params *tridsolver_ctx);

/* No counterpart in the reference. For callers that cannot allocate device memory (code
   under graph capture, tight serving loops, harnesses that forbid hipMalloc): the solve of
   ops_tridMultiDimBatch (u == NULL: x overwrites d) or of ops_tridMultiDimBatch_Inc (u
   given: x is added to u), launched at once (not deferred), by the sequential sweeps, whose
   c', d' go to a static 64 MiB device array of the library (one per GPU, held until the
   process ends) instead of a buffer it allocates: results are bitwise those of the
   queueing entry points with OPS_TRID_ALGO=thomas. A solve needing more than 64 MiB of c',
   d' (2 * sizeof(type) per point of the box) throws OPSException(OPS_INVALID_ARGUMENT).
   Calls on one GPU share that array: calls on the runtime's stream are ordered, calls that
   overlap on one GPU from other streams or processes race. Work issued before the call runs
   first (deferred solves of the queueing entry points are launched, with their own
   buffers). */
void ops_tridMultiDimBatch_noalloc(int ndim, int solvedim, int *range, ops_dat a, ops_dat b,
                                   ops_dat c, ops_dat d, ops_dat u,
                                   ops_tridsolver_params *tridsolver_ctx);

#ifdef __cplusplus
}
#endif

#endif /* OPS_AMD_TRIDIAG_H */
