// TeaLeaf 2D: implicit heat conduction on one OPS block (the configuration of the
// TeaLeaf time-to-solution benchmark).
//
// The loops, ranges and argument lists of the plain solve follow the reference
// application (apps/c/TeaLeaf: start.cpp, build_field.cpp, initialise_chunk.cpp,
// generate.cpp, set_field.cpp, diffuse.cpp, tea_solve.cpp:30-443, tea_leaf_common.cpp,
// tea_leaf_cg.cpp, tea_leaf_jacobi.cpp, update_halo.cpp, field_summary.cpp), so the
// FP64 CG and Jacobi paths can be compared bit for bit. Chebyshev, PPCG and the block
// preconditioner are not provided. The reflective halo update runs one loop per side.
//
// Solvers (tealeaf_set_solver):
//   0  reference path: CG (or Jacobi) iterations as par_loops with host-side alpha/beta
//   1  FP64 CG with the scalars kept on the device (native kernels, tealeaf_native.hip)
//   2  FP32-storage CG iterations with FP64 reliable updates (residual replacement)
//   3  FP64 flexible PCG with an FP32 multigrid V-cycle preconditioner
//   4  the same with the V-cycle in FP64 (what FP32 storage buys is the difference to 3)
// Solver 1 and 2 stop on the same test as the reference: sqrt(r.r) below eps times the
// residual norm of the step's initial guess.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <initializer_list>
#include <string>
#include <vector>

#define OPS_2D
#include <ops_seq_v2.h>

#include "user_types.h"

field_type field;
grid_type grid;
int number_of_states = 0;
state_type *states = NULL;
int g_rect = 1, g_circ = 2, g_point = 3;

#include "tea_kernels.h"
#include "tealeaf_native.h"

namespace {

struct Tea {
  double dtinit = 0.1, end_time = 10.0, eps = 1e-10;
  int end_step = 640000, max_iters = 1000, coefficient = CONDUCTIVITY, use_cg = 0;
  int reflective = 0, summary_frequency = 10, check_result = 0, test_problem = 0;
  int solver = 0;
  int guess = 0, guess_cells = 0;  // learned warm start mode / network size
  int capture = -1;                // dataset sample the next solve is stored as, -1 none
  std::vector<double> start;       // the next solve's starting field (halo included), if set
  int step = 0;
  double time = 0.0, dt = 0.0;
  // last solve
  int last_iters = 0;
  double last_error = 0, last_initial = 0, last_exact = -1;
  double last_warm = -1;  // residual of the learned warm start, -1 without one
  double last_alpha[2] = {0, 0};  // its Galerkin scaling, energy-norm error drop
  int last_guess_used = 0;
  long total_iters = 0;
  ops_block blk;
  ops_dat density, energy0, energy1, u, u0, r, w, Kx, Ky, p, z, volume, xarea, yarea;
  ops_dat cellx, celly, vertexx, vertexy, celldx, celldy, vertexdx, vertexdy, xx, yy;
  ops_stencil S00, P10, P01, M10_M01, S5, HX, HY, SX, SY, SX_P10, SY_P01, SX_P10_M10, SY_P01_M01;
  ops_reduction r_vol, r_mass, r_ie, r_temp;
} T;

const double g_small = 1.0e-16;

void read_deck(const char *path) {
  T = Tea();
  grid.xmin = 0;
  grid.ymin = 0;
  grid.xmax = 100;
  grid.ymax = 100;
  grid.x_cells = 10;
  grid.y_cells = 10;
  free(states);
  states = NULL;
  number_of_states = 0;
  FILE *f = fopen(path, "r");
  if (!f) throw OPSException(OPS_RUNTIME_CONFIGURATION_ERROR, "tealeaf: cannot open deck");
  char line[1024];
  while (fgets(line, sizeof line, f)) {
    std::vector<std::string> tok;
    for (char *t = strtok(line, " =\t\r\n"); t; t = strtok(NULL, " =\t\r\n")) tok.push_back(t);
    if (tok.empty() || tok[0][0] == '*') continue;
    const std::string &k = tok[0];
    auto val = [&](size_t i) { return i < tok.size() ? atof(tok[i].c_str()) : 0.0; };
    if (k == "initial_timestep") T.dtinit = val(1);
    else if (k == "end_time") T.end_time = val(1);
    else if (k == "end_step") T.end_step = (int)val(1);
    else if (k == "xmin") grid.xmin = val(1);
    else if (k == "xmax") grid.xmax = val(1);
    else if (k == "ymin") grid.ymin = val(1);
    else if (k == "ymax") grid.ymax = val(1);
    else if (k == "x_cells") grid.x_cells = (int)val(1);
    else if (k == "y_cells") grid.y_cells = (int)val(1);
    else if (k == "summary_frequency") T.summary_frequency = (int)val(1);
    else if (k == "tl_check_result") T.check_result = 1;
    else if (k == "tl_use_jacobi") T.use_cg = 0;
    else if (k == "tl_use_cg") T.use_cg = 1;
    else if (k == "tl_use_chebyshev" || k == "tl_use_ppcg")
      throw OPSException(OPS_NOT_IMPLEMENTED, "tealeaf: only the CG and Jacobi solvers are built");
    else if (k == "tl_preconditioner_type" && tok.size() > 1 && tok[1] != "none")
      throw OPSException(OPS_NOT_IMPLEMENTED, "tealeaf: preconditioners are not built");
    else if (k == "reflective_boundary") T.reflective = 1;
    else if (k == "test_problem") T.test_problem = (int)val(1);
    else if (k == "tl_max_iters") T.max_iters = (int)val(1);
    else if (k == "tl_eps") T.eps = val(1);
    else if (k == "tl_coefficient_density") T.coefficient = CONDUCTIVITY;
    else if (k == "tl_coefficient_inverse_density") T.coefficient = RECIP_CONDUCTIVITY;
    else if (k == "state") {
      states = (state_type *)realloc(states, sizeof(state_type) * (number_of_states + 1));
      state_type &s = states[number_of_states];
      memset(&s, 0, sizeof s);
      s.defined = 1;
      for (size_t i = 2; i < tok.size(); i++) {
        const std::string &key = tok[i];
        if (key == "xmin") s.xmin = val(i + 1);
        else if (key == "xmax") s.xmax = val(i + 1);
        else if (key == "ymin") s.ymin = val(i + 1);
        else if (key == "ymax") s.ymax = val(i + 1);
        else if (key == "radius") s.radius = val(i + 1);
        else if (key == "density") s.density = val(i + 1);
        else if (key == "energy") s.energy = val(i + 1);
        else if (key == "geometry" && i + 1 < tok.size()) {
          if (tok[i + 1] == "rectangle") s.geometry = g_rect;
          else if (tok[i + 1] == "circle") s.geometry = g_circ;
          else if (tok[i + 1] == "point") s.geometry = g_point;
        }
      }
      number_of_states++;
    }
  }
  fclose(f);
  if (number_of_states == 0)
    throw OPSException(OPS_RUNTIME_CONFIGURATION_ERROR, "tealeaf: no states defined");
  field.x_min = 1;
  field.y_min = 1;
  field.x_max = grid.x_cells + 1;
  field.y_max = grid.y_cells + 1;
  field.left = 0;
  field.bottom = 0;
  float dx = (grid.xmax - grid.xmin) / (float)(grid.x_cells);
  float dy = (grid.ymax - grid.ymin) / (float)(grid.y_cells);
  for (int i = 0; i < number_of_states; i++) {
    states[i].xmin = states[i].xmin + (dx / 100.00);
    states[i].ymin = states[i].ymin + (dy / 100.00);
    states[i].xmax = states[i].xmax - (dx / 100.00);
    states[i].ymax = states[i].ymax - (dy / 100.00);
  }
}

ops_stencil sten(int n, std::vector<int> pts, const char *name) {
  return ops_decl_stencil(2, n, pts.data(), (char *)name);
}
ops_stencil ssten(int n, std::vector<int> pts, int sx, int sy, const char *name) {
  int st[2] = {sx, sy};
  return ops_decl_strided_stencil(2, n, pts.data(), st, (char *)name);
}

void build_field() {
  const int nx = grid.x_cells, ny = grid.y_cells;
  T.blk = ops_decl_block(2, "tea grid");
  double *none = NULL;
  int *inone = NULL;
  int dm[2] = {-1, -1}, dp[2] = {1, 1}, base[2] = {1, 1}, size[2] = {nx, ny};
  auto dat = [&](const char *name) {
    return ops_decl_dat(T.blk, 1, size, base, dm, dp, none, "double", name);
  };
  T.density = dat("density");
  T.energy0 = dat("energy0");
  T.energy1 = dat("energy1");
  T.u = dat("u");
  T.u0 = dat("u0");
  T.r = dat("vector_r");
  T.w = dat("vector_w");
  T.z = dat("vector_z");
  T.Kx = dat("vector_Kx");
  T.Ky = dat("vector_Ky");
  T.p = dat("vector_p");
  int gb[2] = {-1, -1};
  int s1[2] = {nx + 4, ny + 4}, s2[2] = {nx + 5, ny + 4}, s3[2] = {nx + 4, ny + 5};
  T.volume = ops_decl_dat(T.blk, 1, s1, gb, dm, dp, none, "double", "volume");
  T.xarea = ops_decl_dat(T.blk, 1, s2, gb, dm, dp, none, "double", "xarea");
  T.yarea = ops_decl_dat(T.blk, 1, s3, gb, dm, dp, none, "double", "yarea");
  int xm[2] = {-2, 0}, xp[2] = {2, 0}, ym[2] = {0, -2}, yp[2] = {0, 2};
  int bx[2] = {-1, 0}, by[2] = {0, -1}, b0[2] = {0, 0};
  int cx[2] = {nx + 4, 1}, cy[2] = {1, ny + 4}, vx[2] = {nx + 5, 1}, vy[2] = {1, ny + 5};
  T.cellx = ops_decl_dat(T.blk, 1, cx, bx, xm, xp, none, "double", "cellx");
  T.celldx = ops_decl_dat(T.blk, 1, cx, bx, xm, xp, none, "double", "celldx");
  T.celly = ops_decl_dat(T.blk, 1, cy, by, ym, yp, none, "double", "celly");
  T.celldy = ops_decl_dat(T.blk, 1, cy, by, ym, yp, none, "double", "celldy");
  T.vertexx = ops_decl_dat(T.blk, 1, vx, bx, xm, xp, none, "double", "vertexx");
  T.vertexdx = ops_decl_dat(T.blk, 1, vx, bx, xm, xp, none, "double", "vertexdx");
  T.vertexy = ops_decl_dat(T.blk, 1, vy, by, ym, yp, none, "double", "vertexy");
  T.vertexdy = ops_decl_dat(T.blk, 1, vy, by, ym, yp, none, "double", "vertexdy");
  T.xx = ops_decl_dat(T.blk, 1, vx, b0, xm, xp, inone, "int", "xx");
  T.yy = ops_decl_dat(T.blk, 1, vy, b0, ym, yp, inone, "int", "yy");
  T.S00 = sten(1, {0, 0}, "00");
  T.P10 = sten(2, {0, 0, 1, 0}, "0,0:1,0");
  T.P01 = sten(2, {0, 0, 0, 1}, "0,0:0,1");
  T.M10_M01 = sten(3, {0, 0, -1, 0, 0, -1}, "0,0:-1,0:0,-1");
  T.S5 = sten(5, {0, 0, 0, -1, -1, 0, 1, 0, 0, 1}, "0,0:0,-1:-1,0:1,0:0,1");
  T.HX = sten(3, {0, 0, 1, 0, -1, 0}, "halo mirror x");
  T.HY = sten(3, {0, 0, 0, 1, 0, -1}, "halo mirror y");
  T.SX = ssten(1, {0, 0}, 1, 0, "s2D_00_stride2D_x");
  T.SY = ssten(1, {0, 0}, 0, 1, "s2D_00_stride2D_y");
  T.SX_P10 = ssten(2, {0, 0, 1, 0}, 1, 0, "s2D_00_P10_stride2D_x");
  T.SY_P01 = ssten(2, {0, 0, 0, 1}, 0, 1, "s2D_00_0P1_stride2D_y");
  T.SX_P10_M10 = ssten(3, {0, 0, 1, 0, -1, 0}, 1, 0, "self_stride2D_x");
  T.SY_P01_M01 = ssten(3, {0, 0, 0, 1, 0, -1}, 0, 1, "self_stride2D_y");
  T.r_vol = ops_decl_reduction_handle(sizeof(double), "double", "vol");
  T.r_mass = ops_decl_reduction_handle(sizeof(double), "double", "mass");
  T.r_ie = ops_decl_reduction_handle(sizeof(double), "double", "ie");
  T.r_temp = ops_decl_reduction_handle(sizeof(double), "double", "temp");
}

void initialise_chunk() {
  const int x0 = field.x_min, x1 = field.x_max, y0 = field.y_min, y1 = field.y_max;
  int full[] = {x0 - 2, x1 + 3, y0 - 2, y1 + 3};
  ops_par_loop(tea_xx, "initialise_chunk_kernel_xx", T.blk, 2, full,
               ops_arg_dat(T.xx, 1, T.SX, "int", OPS_WRITE), ops_arg_idx());
  ops_par_loop(tea_yy, "initialise_chunk_kernel_yy", T.blk, 2, full,
               ops_arg_dat(T.yy, 1, T.SY, "int", OPS_WRITE), ops_arg_idx());
  int rx[] = {x0 - 2, x1 + 3, y0 - 2, y1 + 2};
  int ry[] = {x0 - 2, x1 + 2, y0 - 2, y1 + 3};
  ops_par_loop(tea_vertex_x, "initialise_chunk_kernel_x", T.blk, 2, rx,
               ops_arg_dat(T.vertexx, 1, T.SX, "double", OPS_WRITE),
               ops_arg_dat(T.xx, 1, T.SX, "int", OPS_READ),
               ops_arg_dat(T.vertexdx, 1, T.SX, "double", OPS_WRITE));
  ops_par_loop(tea_vertex_y, "initialise_chunk_kernel_y", T.blk, 2, ry,
               ops_arg_dat(T.vertexy, 1, T.SY, "double", OPS_WRITE),
               ops_arg_dat(T.yy, 1, T.SY, "int", OPS_READ),
               ops_arg_dat(T.vertexdy, 1, T.SY, "double", OPS_WRITE));
  int cxr[] = {x0 - 2, x1 + 2, y0 - 2, y1 + 2};
  ops_par_loop(tea_cell_x, "initialise_chunk_kernel_cellx", T.blk, 2, cxr,
               ops_arg_dat(T.vertexx, 1, T.SX_P10, "double", OPS_READ),
               ops_arg_dat(T.cellx, 1, T.SX, "double", OPS_WRITE),
               ops_arg_dat(T.celldx, 1, T.SX, "double", OPS_WRITE));
  int cyr[] = {x0 - 2, x1 + 3, y0 - 2, y1 + 2};
  ops_par_loop(tea_cell_y, "initialise_chunk_kernel_celly", T.blk, 2, cyr,
               ops_arg_dat(T.vertexy, 1, T.SY_P01, "double", OPS_READ),
               ops_arg_dat(T.celly, 1, T.SY, "double", OPS_WRITE),
               ops_arg_dat(T.celldy, 1, T.SY, "double", OPS_WRITE));
  int all[] = {x0 - 2, x1 + 2, y0 - 2, y1 + 2};
  ops_par_loop(tea_volume, "initialise_chunk_kernel_volume", T.blk, 2, all,
               ops_arg_dat(T.volume, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.celldy, 1, T.SY, "double", OPS_READ),
               ops_arg_dat(T.xarea, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.celldx, 1, T.SX, "double", OPS_READ),
               ops_arg_dat(T.yarea, 1, T.S00, "double", OPS_WRITE));
}

void generate() {
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  ops_par_loop(tea_generate, "generate_chunk_kernel", T.blk, 2, in,
               ops_arg_dat(T.vertexx, 1, T.SX_P10_M10, "double", OPS_READ),
               ops_arg_dat(T.vertexy, 1, T.SY_P01_M01, "double", OPS_READ),
               ops_arg_dat(T.energy0, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.density, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.u, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.cellx, 1, T.SX_P10_M10, "double", OPS_READ),
               ops_arg_dat(T.celly, 1, T.SY_P01_M01, "double", OPS_READ));
}

void set_fields(int *f, std::initializer_list<int> on) {
  for (int i = 0; i < NUM_FIELDS; i++) f[i] = 0;
  for (int i : on) f[i] = 1;
}

// reflective boundaries, only with the deck's reflective_boundary flag
// (apps/c/TeaLeaf/update_halo.cpp:242-246)
void update_halo(int *fl, int depth) {
  if (T.reflective != 1) return;
  const int x0 = field.x_min, x1 = field.x_max, y0 = field.y_min, y1 = field.y_max;
  int mb[3] = {1, y0, 1}, mt[3] = {1, y1, 1}, ml[3] = {0, x0, 1}, mr[3] = {0, x1, 1};
  int rb[] = {x0 - depth, x1 + depth, y0 - depth, y0};
  int rt[] = {x0 - depth, x1 + depth, y1, y1 + depth};
  int rl[] = {x0 - depth, x0, y0 - depth, y1 + depth};
  int rr[] = {x1, x1 + depth, y0 - depth, y1 + depth};
  int *ms[4] = {mb, mt, ml, mr};
  int *rs[4] = {rb, rt, rl, rr};
  for (int s = 0; s < 4; s++) {
    ops_stencil st = s < 2 ? T.HY : T.HX;
    ops_par_loop(tea_halo, "update_halo_kernel", T.blk, 2, rs[s],
                 ops_arg_dat_opt(T.density, 1, st, "double", OPS_RW, fl[FIELD_DENSITY]),
                 ops_arg_dat_opt(T.energy0, 1, st, "double", OPS_RW, fl[FIELD_ENERGY0]),
                 ops_arg_dat_opt(T.energy1, 1, st, "double", OPS_RW, fl[FIELD_ENERGY1]),
                 ops_arg_dat_opt(T.u, 1, st, "double", OPS_RW, fl[FIELD_U]),
                 ops_arg_dat_opt(T.p, 1, st, "double", OPS_RW, fl[FIELD_P]),
                 ops_arg_dat_opt(T.z, 1, st, "double", OPS_RW, fl[FIELD_SD]),
                 ops_arg_gbl(fl, NUM_FIELDS, "int", OPS_READ), ops_arg_gbl(ms[s], 3, "int", OPS_READ),
                 ops_arg_idx());
  }
}

void set_field() {
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  ops_par_loop(tea_copy, "set_field_kernel", T.blk, 2, in,
               ops_arg_dat(T.energy0, 1, T.S00, "double", OPS_READ),
               ops_arg_dat(T.energy1, 1, T.S00, "double", OPS_WRITE));
}

double reduce(ops_reduction h) {
  double v;
  ops_reduction_result(h, &v);
  return v;
}

// operator set-up of one solve (tea_leaf_common.cpp:40-144); returns rx, ry
void common_init(double *rx, double *ry) {
  const int hd = 1;
  const int x0 = field.x_min, x1 = field.x_max, y0 = field.y_min, y1 = field.y_max;
  double dx = (grid.xmax - grid.xmin) / (double)grid.x_cells;
  double dy = (grid.xmax - grid.xmin) / (double)grid.x_cells;  // as the reference (x twice)
  *rx = T.dt / (dx * dx);
  *ry = T.dt / (dy * dy);
  int in[] = {x0, x1, y0, y1};
  int ext[] = {x0 - hd + 1, x1 + hd, y0 - hd + 1, y1 + hd};
  int ext2[] = {x0 - hd, x1 + hd, y0 - hd, y1 + hd};
  ops_par_loop(tea_init_u_u0, "tea_leaf_common_init_u_u0_kernel", T.blk, 2, in,
               ops_arg_dat(T.u, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.u0, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.energy1, 1, T.S00, "double", OPS_READ),
               ops_arg_dat(T.density, 1, T.S00, "double", OPS_READ));
  if (T.coefficient == RECIP_CONDUCTIVITY)
    ops_par_loop(tea_recip, "tea_leaf_recip_kernel", T.blk, 2, ext2,
                 ops_arg_dat(T.w, 1, T.S00, "double", OPS_WRITE),
                 ops_arg_dat(T.density, 1, T.S00, "double", OPS_READ));
  else
    ops_par_loop(tea_copy, "tea_leaf_yeqx_kernel", T.blk, 2, ext2,
                 ops_arg_dat(T.density, 1, T.S00, "double", OPS_READ),
                 ops_arg_dat(T.w, 1, T.S00, "double", OPS_WRITE));
  ops_par_loop(tea_init_kx_ky, "tea_leaf_common_init_Kx_Ky_kernel", T.blk, 2, ext,
               ops_arg_dat(T.Kx, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.Ky, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.w, 1, T.M10_M01, "double", OPS_READ));
  if (T.reflective == 0) {
    int left[] = {x0 - hd, x0 + 1, y0 - hd, y1 + hd};
    int right[] = {x1, x1 + hd, y0 - hd, y1 + hd};
    int bottom[] = {x0 - hd, x1 + hd, y0 - hd, y0 + 1};
    int top[] = {x0 - hd, x1 + hd, y1, y1 + hd};
    ops_par_loop(tea_zero, "tea_leaf_init_zero_kernel", T.blk, 2, left,
                 ops_arg_dat(T.Kx, 1, T.S00, "double", OPS_WRITE));
    ops_par_loop(tea_zero, "tea_leaf_init_zero_kernel", T.blk, 2, right,
                 ops_arg_dat(T.Kx, 1, T.S00, "double", OPS_WRITE));
    ops_par_loop(tea_zero, "tea_leaf_init_zero_kernel", T.blk, 2, bottom,
                 ops_arg_dat(T.Ky, 1, T.S00, "double", OPS_WRITE));
    ops_par_loop(tea_zero, "tea_leaf_init_zero_kernel", T.blk, 2, top,
                 ops_arg_dat(T.Ky, 1, T.S00, "double", OPS_WRITE));
  }
  ops_par_loop(tea_init_w_r, "tea_leaf_common_init_kernel", T.blk, 2, in,
               ops_arg_dat(T.w, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.r, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.Kx, 1, T.P10, "double", OPS_READ),
               ops_arg_dat(T.Ky, 1, T.P01, "double", OPS_READ),
               ops_arg_dat(T.u, 1, T.S5, "double", OPS_READ), ops_arg_gbl(rx, 1, "double", OPS_READ),
               ops_arg_gbl(ry, 1, "double", OPS_READ));
}

void calc_residual(double rx, double ry) {
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  ops_par_loop(tea_residual, "tea_leaf_common_residual_kernel", T.blk, 2, in,
               ops_arg_dat(T.r, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.Kx, 1, T.P10, "double", OPS_READ),
               ops_arg_dat(T.Ky, 1, T.P01, "double", OPS_READ),
               ops_arg_dat(T.u, 1, T.S5, "double", OPS_READ),
               ops_arg_dat(T.u0, 1, T.S00, "double", OPS_READ),
               ops_arg_gbl(&rx, 1, "double", OPS_READ), ops_arg_gbl(&ry, 1, "double", OPS_READ));
}

double norm2(ops_dat x) {
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  ops_par_loop(tea_norm2, "tea_leaf_norm2_kernel", T.blk, 2, in,
               ops_arg_dat(x, 1, T.S00, "double", OPS_READ),
               ops_arg_reduce(T.r_temp, 1, "double", OPS_INC));
  return reduce(T.r_temp);
}

// one solve of the implicit step: u such that A u = u0 (tea_solve.cpp:30-443)
void solve() {
  int f[NUM_FIELDS];
  double rx, ry;
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  set_fields(f, {FIELD_ENERGY1, FIELD_DENSITY});
  update_halo(f, 1);
  common_init(&rx, &ry);
  set_fields(f, {FIELD_U});
  update_halo(f, 1);
  calc_residual(rx, ry);
  double initial = norm2(T.r);
  initial = sqrt(fabs(initial));
  T.last_warm = -1;
  T.last_guess_used = 0;
  if (T.capture >= 0) tea_unet_capture_features(T.capture, T.density, T.u0, T.r, grid.x_cells);
  if (!T.start.empty()) {
    // a given starting field (tealeaf_set_start); the stop threshold keeps the residual of
    // the standard start u = u0, as for the learned warm start
    ops_dat d = T.u;
    int range[4] = {d->d_m[0], d->size[0] + d->d_m[0], d->d_m[1], d->size[1] + d->d_m[1]};
    ops_dat_set_data_slab_memspace(d, 0, (char *)T.start.data(), range, OPS_HOST);
    T.start.clear();
    set_fields(f, {FIELD_U});
    update_halo(f, 1);
    calc_residual(rx, ry);
    T.last_warm = sqrt(fabs(norm2(T.r)));
    T.last_guess_used = 1;
  } else if (T.guess) {
    // learned warm start: u = u0 + alpha * the network's predicted change, alpha the Galerkin
    // scaling (energy-norm optimal along that direction). The stop threshold keeps
    // the residual of the standard start u = u0. Mode 1 keeps the warm start only when its
    // own residual is below that one, otherwise u = u0 is restored (the plain start); mode
    // 2 always keeps it.
    tea_unet_guess_dats(T.density, T.u0, T.r, T.Kx, T.Ky, rx, ry, T.u, grid.x_cells, T.guess_cells, T.last_alpha);
    set_fields(f, {FIELD_U});
    update_halo(f, 1);
    calc_residual(rx, ry);
    T.last_warm = sqrt(fabs(norm2(T.r)));
    if (T.guess == 2 || T.guess == 3 || (T.guess == 1 && T.last_warm < initial)) {
      T.last_guess_used = 1;
    } else {
      ops_par_loop(tea_copy, "tea_leaf_warm_start_reset_kernel", T.blk, 2, in,
                   ops_arg_dat(T.u0, 1, T.S00, "double", OPS_READ),
                   ops_arg_dat(T.u, 1, T.S00, "double", OPS_WRITE));
      update_halo(f, 1);
      calc_residual(rx, ry);
    }
  }
  double error = 0.0;
  int n = 1;
  if (T.use_cg && T.solver != 0) {
    // device-resident solvers: start from r = u0 - A u already computed above. They keep
    // the halos of the solve vectors at zero, which is exact only when the boundary faces
    // carry zero conductivity (no reflective boundary).
    if (T.reflective)
      throw OPSException(OPS_NOT_IMPLEMENTED,
                         "tealeaf: the device-resident solvers need reflective_boundary off");
    tea_native_solve_params P;
    P.solver = T.solver;
    P.nx = grid.x_cells;
    P.ny = grid.y_cells;
    P.rx = rx;
    P.ry = ry;
    P.eps = T.eps;
    P.max_iters = T.max_iters;
    P.initial_residual = initial;
    P.deflation = nullptr;
    P.deflation_mean = 0.0;
    // modes 3 and 4: the network's change also corrects the multigrid preconditioner
    if (T.guess >= 3 && (T.solver == 3 || T.solver == 4))
      P.deflation = tea_unet_direction(&P.deflation_mean);
    tea_native_result R;
    if (T.solver == 3)
      tea_native_mgcg(T.u, T.u0, T.r, T.p, T.w, T.Kx, T.Ky, &P, &R);
    else if (T.solver == 4)
      tea_native_mgcg64(T.u, T.u0, T.r, T.p, T.w, T.Kx, T.Ky, &P, &R);
    else
      tea_native_cg(T.u, T.u0, T.r, T.p, T.w, T.z, T.Kx, T.Ky, &P, &R);
    n = R.iterations + 1;
    error = R.error;
  } else {
    double rro = 0.0, pw, rrn, alpha, beta, one = 1.0;
    if (T.use_cg) {
      ops_par_loop(tea_zero, "tea_leaf_init_zero2_kernel", T.blk, 2, in,
                   ops_arg_dat(T.p, 1, T.S00, "double", OPS_WRITE));
      ops_par_loop(tea_zero, "tea_leaf_init_zero2_kernel", T.blk, 2, in,
                   ops_arg_dat(T.z, 1, T.S00, "double", OPS_WRITE));
      ops_par_loop(tea_copy, "tea_leaf_yeqx_kernel", T.blk, 2, in,
                   ops_arg_dat(T.r, 1, T.S00, "double", OPS_READ),
                   ops_arg_dat(T.p, 1, T.S00, "double", OPS_WRITE));
      ops_par_loop(tea_dot, "tea_leaf_dot_kernel", T.blk, 2, in,
                   ops_arg_dat(T.r, 1, T.S00, "double", OPS_READ),
                   ops_arg_dat(T.p, 1, T.S00, "double", OPS_READ),
                   ops_arg_reduce(T.r_temp, 1, "double", OPS_INC));
      rro = reduce(T.r_temp);
      set_fields(f, {FIELD_U, FIELD_P});
      update_halo(f, 1);
      set_fields(f, {FIELD_P});
    } else {
      set_fields(f, {FIELD_U});
    }
    for (n = 1; n <= T.max_iters; n++) {
      if (T.use_cg) {
        ops_par_loop(tea_cg_calc_w, "tea_leaf_cg_calc_w_reduce_kernel", T.blk, 2, in,
                     ops_arg_dat(T.w, 1, T.S00, "double", OPS_WRITE),
                     ops_arg_dat(T.Kx, 1, T.P10, "double", OPS_READ),
                     ops_arg_dat(T.Ky, 1, T.P01, "double", OPS_READ),
                     ops_arg_dat(T.p, 1, T.S5, "double", OPS_READ),
                     ops_arg_gbl(&rx, 1, "double", OPS_READ), ops_arg_gbl(&ry, 1, "double", OPS_READ),
                     ops_arg_reduce(T.r_temp, 1, "double", OPS_INC));
        pw = reduce(T.r_temp);
        alpha = rro / pw;
        ops_par_loop(tea_axpy, "tea_leaf_axpy_kernel", T.blk, 2, in,
                     ops_arg_dat(T.u, 1, T.S00, "double", OPS_INC),
                     ops_arg_dat(T.p, 1, T.S00, "double", OPS_READ),
                     ops_arg_gbl(&alpha, 1, "double", OPS_READ));
        ops_par_loop(tea_cg_calc_r, "tea_leaf_cg_calc_ur_r_reduce_kernel", T.blk, 2, in,
                     ops_arg_dat(T.r, 1, T.S00, "double", OPS_INC),
                     ops_arg_dat(T.w, 1, T.S00, "double", OPS_READ),
                     ops_arg_gbl(&alpha, 1, "double", OPS_READ),
                     ops_arg_reduce(T.r_temp, 1, "double", OPS_INC));
        rrn = reduce(T.r_temp);
        beta = rrn / rro;
        ops_par_loop(tea_axpby, "tea_leaf_axpby_kernel", T.blk, 2, in,
                     ops_arg_dat(T.p, 1, T.S00, "double", OPS_RW),
                     ops_arg_dat(T.r, 1, T.S00, "double", OPS_READ),
                     ops_arg_gbl(&beta, 1, "double", OPS_READ),
                     ops_arg_gbl(&one, 1, "double", OPS_READ));
        error = rrn;
        rro = rrn;
      } else {
        ops_par_loop(tea_copy, "tea_leaf_yeqx_kernel", T.blk, 2, in,
                     ops_arg_dat(T.u, 1, T.S00, "double", OPS_READ),
                     ops_arg_dat(T.r, 1, T.S00, "double", OPS_WRITE));
        ops_par_loop(tea_jacobi, "tea_leaf_jacobi_kernel", T.blk, 2, in,
                     ops_arg_dat(T.u, 1, T.S00, "double", OPS_WRITE),
                     ops_arg_dat(T.Kx, 1, T.P10, "double", OPS_READ),
                     ops_arg_dat(T.Ky, 1, T.P01, "double", OPS_READ),
                     ops_arg_dat(T.r, 1, T.S5, "double", OPS_READ),
                     ops_arg_dat(T.u0, 1, T.S00, "double", OPS_READ),
                     ops_arg_gbl(&rx, 1, "double", OPS_READ), ops_arg_gbl(&ry, 1, "double", OPS_READ),
                     ops_arg_reduce(T.r_temp, 1, "double", OPS_INC));
        error = reduce(T.r_temp);
      }
      update_halo(f, 1);
      error = sqrt(fabs(error));
      if (fabs(error) < T.eps * initial) break;
    }
  }
  if (T.capture >= 0) {
    tea_unet_capture_target(T.capture, T.u0, T.u, T.Kx, T.Ky, rx, ry, grid.x_cells);
    T.capture = -1;
  }
  T.last_iters = n > T.max_iters ? T.max_iters : n;
  T.last_error = error;
  T.last_initial = initial;
  T.total_iters += T.last_iters;
  if (T.check_result) {
    set_fields(f, {FIELD_U});
    update_halo(f, 1);
    calc_residual(rx, ry);
    T.last_exact = sqrt(norm2(T.r));
  }
  ops_par_loop(tea_recip2, "tea_leaf_recip2_kernel", T.blk, 2, in,
               ops_arg_dat(T.energy1, 1, T.S00, "double", OPS_WRITE),
               ops_arg_dat(T.u, 1, T.S00, "double", OPS_READ),
               ops_arg_dat(T.density, 1, T.S00, "double", OPS_READ));
  set_fields(f, {FIELD_ENERGY1});
  update_halo(f, 1);
}

void field_summary(double *out) {
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  ops_par_loop(tea_field_summary, "field_summary_kernel", T.blk, 2, in,
               ops_arg_dat(T.volume, 1, T.S00, "double", OPS_READ),
               ops_arg_dat(T.density, 1, T.S00, "double", OPS_READ),
               ops_arg_dat(T.energy1, 1, T.S00, "double", OPS_READ),
               ops_arg_dat(T.u, 1, T.S00, "double", OPS_READ),
               ops_arg_reduce(T.r_vol, 1, "double", OPS_INC),
               ops_arg_reduce(T.r_mass, 1, "double", OPS_INC),
               ops_arg_reduce(T.r_ie, 1, "double", OPS_INC),
               ops_arg_reduce(T.r_temp, 1, "double", OPS_INC));
  out[0] = reduce(T.r_vol);
  out[1] = reduce(T.r_mass);
  out[2] = reduce(T.r_ie);
  out[3] = reduce(T.r_temp);
}

void start() {
  T.time = 0.0;
  T.step = 0;
  T.dt = T.dtinit;
  build_field();
  ops_decl_const("field", 1, "field_type", &field);
  ops_decl_const("grid", 1, "grid_type", &grid);
  ops_decl_const("number_of_states", 1, "int", &number_of_states);
  ops_decl_const("states", number_of_states, "state_type", states);
  ops_decl_const("g_circ", 1, "int", &g_circ);
  ops_decl_const("g_point", 1, "int", &g_point);
  ops_decl_const("g_rect", 1, "int", &g_rect);
  initialise_chunk();
  generate();
  int f[NUM_FIELDS];
  set_fields(f, {FIELD_DENSITY, FIELD_ENERGY0, FIELD_ENERGY1});
  update_halo(f, 1);
  set_field();
}

void diffuse_step() {
  T.step++;
  T.dt = T.dtinit;
  solve();
  T.time = T.time + T.dt;
}

}  // namespace

extern "C" int tealeaf_init(int argc, const char **argv, const char *deck) {
  try {
    ops_init(argc, argv, 1);
    read_deck(deck);
    start();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "tealeaf_init: %s\n", e.what());
    return 1;
  }
}

extern "C" int tealeaf_set_solver(int solver) {
  if (solver < 0 || solver > 4) return 1;
  T.solver = solver;
  return 0;
}

// learned warm start for the implicit solve: modes 1-4 evaluate the network (tea_unet_*)
// on a net_cells^2 view of the mesh before each solve. 1: its warm start kept only when its
// residual beats the standard start's; 2: always kept; 3: always kept, and the network's
// change also corrects the multigrid preconditioner (solvers 3 and 4: z = V(r) + the
// A-orthogonal projection of r's error onto that direction); 4: the correction alone, from
// the standard start. Mode 0 turns it off.
extern "C" int tealeaf_set_guess(int mode, int net_cells) {
  if (mode < 0 || mode > 4 || (mode && net_cells <= 0)) return 1;
  T.guess = mode;
  T.guess_cells = net_cells;
  return 0;
}

// the next solve starts from `field` (u with its halo, size[0] * size[1] values, the layout
// of tealeaf_fetch) instead of u = u0; its stop threshold stays that of u = u0. Used to
// measure how the iteration count depends on the accuracy of a starting guess.
extern "C" int tealeaf_set_start(const double *field, long count) {
  if (field == nullptr || T.u == nullptr || count != (long)T.u->size[0] * T.u->size[1]) return 1;
  T.start.assign(field, field + count);
  return 0;
}

// store the next solve as training sample `index` of the network's dataset (features of
// density, u0 and the starting residual; target u)
extern "C" int tealeaf_capture_next(int index) {
  if (index < 0) return 1;
  T.capture = index;
  return 0;
}

// last solve: residual of the standard start, residual of the learned warm start (-1 when
// off), 1 if the warm start was kept, its Galerkin scaling, and the drop in the squared
// energy-norm error it makes against the standard start
extern "C" int tealeaf_guess_state(double *out5) {
  out5[0] = T.last_initial;
  out5[1] = T.last_warm;
  out5[2] = T.last_guess_used;
  out5[3] = T.last_alpha[0];
  out5[4] = T.last_alpha[1];
  return 0;
}

extern "C" int tealeaf_step(int n) {
  try {
    for (int i = 0; i < n; i++) diffuse_step();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "tealeaf_step: %s\n", e.what());
    return 1;
  }
}

// steps until step >= end_step (the reference's end test, diffuse.cpp); returns steps
extern "C" int tealeaf_run(int max_steps) {
  try {
    int n = 0;
    while (n < max_steps) {
      diffuse_step();
      n++;
      if (T.step >= T.end_step) break;
    }
    return n;
  } catch (std::exception &e) {
    fprintf(stderr, "tealeaf_run: %s\n", e.what());
    return -1;
  }
}

extern "C" int tealeaf_summary(double *out4) {
  try {
    field_summary(out4);
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "tealeaf_summary: %s\n", e.what());
    return 1;
  }
}

// step, time, last solve's iterations, its error, its initial residual, total iterations,
// exact residual of the last solve (tl_check_result) or -1
extern "C" int tealeaf_state(double *out7) {
  out7[0] = T.step;
  out7[1] = T.time;
  out7[2] = T.last_iters;
  out7[3] = T.last_error;
  out7[4] = T.last_initial;
  out7[5] = (double)T.total_iters;
  out7[6] = T.last_exact;
  return 0;
}

extern "C" long tealeaf_fetch(const char *name, double *out) {
  try {
    ops_dat all[] = {T.density, T.energy0, T.energy1, T.u, T.u0, T.r, T.w, T.Kx, T.Ky, T.p};
    for (ops_dat d : all) {
      if (strcmp(d->name, name)) continue;
      long n = (long)d->size[0] * d->size[1];
      if (out) {
        int range[4] = {d->d_m[0], d->size[0] + d->d_m[0], d->d_m[1], d->size[1] + d->d_m[1]};
        ops_dat_fetch_data_slab_memspace(d, 0, (char *)out, range, OPS_HOST);
      }
      return n;
    }
    return -1;
  } catch (std::exception &e) {
    fprintf(stderr, "tealeaf_fetch: %s\n", e.what());
    return -2;
  }
}

extern "C" int tealeaf_finalize(void) {
  try {
    ops_exit();
    free(states);
    states = NULL;
    number_of_states = 0;
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "tealeaf_finalize: %s\n", e.what());
    return 1;
  }
}
