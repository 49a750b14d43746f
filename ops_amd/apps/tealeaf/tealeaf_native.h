// Device-resident CG solvers for the TeaLeaf implicit step (no reference counterpart).
//
// Both solve A u = u0 with A = I + rx*Dx(Kx) + ry*Dy(Ky), the 5-point operator of
// tea_leaf_cg_calc_w_reduce_kernel, on the interior of the mesh, starting from the
// residual r = u0 - A u already computed by the caller. They stop when sqrt(r.r) falls
// below eps * initial_residual, the reference's test (apps/c/TeaLeaf/tea_solve.cpp:349-355).
//   solver 1: FP64 CG; alpha/beta and the stop test stay on the GPU (no per-iteration
//             host round trip), p-update fused into the operator application.
//   solver 2: FP32 storage and arithmetic for the CG iterations inside FP64 refinement
//             by reliable updates (residual replacement): whenever the iterated
//             residual has dropped by a factor delta (OPS_TEA_RELIABLE_DELTA, default
//             0.1) since the last replacement, and when it meets the stop threshold,
//             the FP32 residual is replaced by the FP64 residual u0 - A u, its norm
//             taken in FP64. The search direction is kept, so the Krylov process
//             continues rather than restarting. The FP32 solution increment is folded
//             into the FP64 solution every 32 iterations. The stop follows the
//             reference's rule on the iterated residual; the reported error is the
//             FP64 residual after the final replacement.
#ifndef TEALEAF_NATIVE_H
#define TEALEAF_NATIVE_H

#include "ops_lib_core.h"

typedef struct {
  int solver;
  int nx, ny;
  double rx, ry, eps;
  int max_iters;
  double initial_residual;
  // multigrid PCG only: a correction direction d (device, (nx + 2)^2 with a zero halo, its
  // mean `deflation_mean` removed on use) added to the preconditioner,
  // z = V(r) + d (d^T r) / (d^T A d); nullptr for none
  const double *deflation;
  double deflation_mean;
} tea_native_solve_params;

typedef struct {
  int iterations;     // CG iterations
  int outer;          // FP64 residual replacements (solver 2), 1 otherwise
  double error;       // sqrt(r.r) at exit
} tea_native_result;

void tea_native_cg(ops_dat u, ops_dat u0, ops_dat r, ops_dat p, ops_dat w, ops_dat z, ops_dat Kx,
                   ops_dat Ky, const tea_native_solve_params *P, tea_native_result *R);

// solver 3: FP64 flexible preconditioned CG; the preconditioner is one FP32 multigrid
// V-cycle (2 x 2 aggregation with Galerkin coarse operators, red-black Gauss-Seidel
// smoothing, coarse correction scaled by 1.5). Same stop rule as above. Square meshes.
// R->outer reports the number of multigrid levels.
void tea_native_mgcg(ops_dat u, ops_dat u0, ops_dat r, ops_dat p, ops_dat w, ops_dat Kx, ops_dat Ky,
                     const tea_native_solve_params *P, tea_native_result *R);
// solver 4: the same with the V-cycle in FP64 storage and arithmetic
void tea_native_mgcg64(ops_dat u, ops_dat u0, ops_dat r, ops_dat p, ops_dat w, ops_dat Kx, ops_dat Ky,
                       const tea_native_solve_params *P, tea_native_result *R);

// Learned warm start (the U-Net of tealeaf_unet.h). capture_features stores the network
// input of the solve about to start (density, u0 and its residual r = u0 - A u0) as
// dataset sample `index`; capture_target stores its result u and the operator's face
// coefficients (rx Kx, ry Ky). guess overwrites the interior of u with u0 + alpha p, p the
// network's predicted change (less its mean) evaluated on a net_cells^2 view of the n x n
// mesh and alpha the Galerkin scaling p^T r / p^T A p; out2 (when non-null) receives alpha
// and the drop in the squared energy-norm error against the plain start. All need n x n
// double fields with one halo layer.
void tea_unet_capture_features(int index, ops_dat rho, ops_dat u0, ops_dat r, int n);
void tea_unet_capture_target(int index, ops_dat u0, ops_dat u, ops_dat Kx, ops_dat Ky, double rx, double ry, int n);
void tea_unet_guess_dats(ops_dat rho, ops_dat u0, ops_dat r, ops_dat Kx, ops_dat Ky, double rx, double ry, ops_dat u,
                         int n, int net_cells, double *out2);
// the network's change p of the last tea_unet_guess_dats on the mesh (device, (n + 2)^2,
// zero halo) and its mean
const double *tea_unet_direction(double *mean);

#endif
