// CloverLeaf 2D (explicit Lagrangian-Eulerian hydrodynamics) on one OPS block.
//
// The sequence of loops, their ranges and argument lists follow the reference
// application (apps/c/CloverLeaf: clover_leaf.cpp:139-245, start.cpp, build_field.cpp,
// initialise_chunk.cpp, generate.cpp, ideal_gas.cpp, timestep via time_step.cpp and
// calc_dt.cpp, PdV.cpp, revert.cpp, accelerate.cpp, flux_calc.cpp, advection.cpp,
// advec_cell.cpp, advec_mom.cpp, reset_field.cpp, field_summary.cpp) so that results
// can be compared bit for bit. Differences from the reference, none of which changes a
// value: the reflective halo update runs one loop per side and field group covering both
// halo layers (the reference runs one loop per layer), and the diagnostic-only
// position lookup of the controlling cell in calc_dt is not evaluated.
//
// Entry points (C ABI) for drivers and tests:
//   cloverleaf_init(argc, argv, deck)   parse the deck, build and generate the problem
//   cloverleaf_step(n)                  advance n hydro steps
//   cloverleaf_set_graph(on)            replay steps from captured hipGraphs (dt on the GPU)
//   cloverleaf_set_chain(tiles)         run the loops between halo exchanges as loop chains
//   cloverleaf_summary(out[5])          field summary: volume, mass, ie, ke, pressure
//   cloverleaf_state(out[4])            step, time, dt, end_step
//   cloverleaf_finalize()
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

// ---- ops_decl_const globals referenced by the kernels
double g_small = 1.0e-16;
double g_big = 1.0e+21;
double dtc_safe, dtu_safe, dtv_safe, dtdiv_safe;
double dt;
field_type field;
grid_type grid;
int number_of_states = 0;
state_type *states = NULL;
int g_rect = 1, g_circ = 2, g_point = 3;

#include "clover_kernels.h"

// device-side time-step control for captured steps (native/clover_dt.hip)
extern "C" {
void *clover_dt_state_alloc(void);
void clover_dt_state_free(void *st);
void clover_dt_state_set(void *st, double dtold, double time, double dt);
void clover_dt_state_get(void *st, double *out4);
void clover_dt_launch(const void *local_dt, void *dt_const, void *st, double g_big,
                      double dtrise, double dtmax, double dtmin);
int ops_hip_graph_begin(void);
void *ops_hip_graph_end(void);
int ops_hip_graph_replay(void *graph, int times);
void ops_hip_graph_free(void *graph);
void *ops_hip_const_device_ptr(const char *name);
void *ops_hip_reduction_device_ptr(ops_reduction h);
void ops_hip_reduction_discard(ops_reduction h);
void ops_hip_batch_begin(void);
void ops_hip_batch_end(void);
void ops_hip_chain_begin(int tiles);
void ops_hip_chain_end(void);
}

namespace {

struct Sim {
  // deck
  double dtinit = 0.1, dtmax = 1.0, dtmin = 0.0000001, dtrise = 1.5, end_time = 10.0;
  int end_step = 640000, test_problem = 0, summary_frequency = 10;
  // run state
  int step = 0, advect_x = 1;
  double time = 0.0, dtold = 0.0;
  int verbose = 0;
  // captured steps: one graph per sweep order (advect_x), the device dt state
  int graph = 0;
  void *step_graph[2] = {nullptr, nullptr};
  // loop chains: the loops between two halo exchanges run band by band in `chain` bands
  // (0 or 1: launched one by one)
  int chain = 0;
  void *dt_state = nullptr;
  // OPS objects
  ops_block blk;
  ops_dat density0, density1, energy0, energy1, pressure, viscosity, soundspeed, volume;
  ops_dat xvel0, xvel1, yvel0, yvel1, vol_flux_x, vol_flux_y, mass_flux_x, mass_flux_y;
  ops_dat xarea, yarea, work1, work2, work3, work4, work5, work6, work7;
  ops_dat cellx, celly, vertexx, vertexy, celldx, celldy, vertexdx, vertexdy, xx, yy;
  ops_stencil S00, P10, P01, M10, M01, S5, P10_P01_P11, M10_M01_M11, P10_M01_P1M1,
      P01_M10_M1P1, X_M_P, P10_M10_M20, P01_M01_M02, P10_P20_M10, P01_P02_M01;
  // reflective halo sources, one stencil per side: bottom, top, left, right
  ops_stencil H[4];
  ops_stencil SX, SY, SX_P10, SY_P01, SX_P10_M10, SY_P01_M01;
  ops_reduction r_dt, r_vol, r_mass, r_ie, r_ke, r_press, r_xl, r_yl;
  double summary[5] = {0, 0, 0, 0, 0};  // the last field summary of the step loop
} S;

// ---------------------------------------------------------------- input deck
// The deck format of apps/c/CloverLeaf/read_input.cpp: "*clover" ... "*endclover", tokens
// separated by blanks or '=', "state N key=value ..." lines.

void read_deck(const char *path) {
  S.dtinit = 0.1;
  S.dtmax = 1.0;
  S.dtmin = 0.0000001;
  S.dtrise = 1.5;
  S.end_time = 10.0;
  S.end_step = 640000;
  S.test_problem = 0;
  S.summary_frequency = 10;
  dtc_safe = 0.7;
  dtu_safe = 0.5;
  dtv_safe = 0.5;
  dtdiv_safe = 0.7;
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
  if (!f) throw OPSException(OPS_RUNTIME_CONFIGURATION_ERROR, "cloverleaf: cannot open deck");
  char line[1024];
  while (fgets(line, sizeof line, f)) {
    std::vector<std::string> tok;
    for (char *t = strtok(line, " =\t\r\n"); t; t = strtok(NULL, " =\t\r\n")) tok.push_back(t);
    if (tok.empty() || tok[0][0] == '*') continue;
    const std::string &k = tok[0];
    auto val = [&](size_t i) { return i < tok.size() ? atof(tok[i].c_str()) : 0.0; };
    if (k == "initial_timestep") S.dtinit = val(1);
    else if (k == "max_timestep") S.dtmax = val(1);
    else if (k == "timestep_rise") S.dtrise = val(1);
    else if (k == "end_time") S.end_time = val(1);
    else if (k == "end_step") S.end_step = (int)val(1);
    else if (k == "xmin") grid.xmin = val(1);
    else if (k == "xmax") grid.xmax = val(1);
    else if (k == "ymin") grid.ymin = val(1);
    else if (k == "ymax") grid.ymax = val(1);
    else if (k == "x_cells") grid.x_cells = (int)val(1);
    else if (k == "y_cells") grid.y_cells = (int)val(1);
    else if (k == "summary_frequency") S.summary_frequency = atoi(tok.size() > 1 ? tok[1].c_str() : "0");
    else if (k == "test_problem") S.test_problem = atoi(tok.size() > 1 ? tok[1].c_str() : "0");
    else if (k == "state") {
      states = (state_type *)realloc(states, sizeof(state_type) * (number_of_states + 1));
      state_type &s = states[number_of_states];
      memset(&s, 0, sizeof s);
      for (size_t i = 2; i < tok.size(); i++) {
        const std::string &key = tok[i];
        if (key == "xvel") s.xvel = val(i + 1);
        else if (key == "yvel") s.yvel = val(i + 1);
        else if (key == "xmin") s.xmin = val(i + 1);
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
    throw OPSException(OPS_RUNTIME_CONFIGURATION_ERROR, "cloverleaf: no states defined");
  field.x_min = 2;
  field.y_min = 2;
  field.x_max = grid.x_cells + 2;
  field.y_max = grid.y_cells + 2;
  field.left = 0;
  field.bottom = 0;
  // states are shrunk by 1/100 of a cell (float cell size, as read_input.cpp:257-265)
  float dx = (grid.xmax - grid.xmin) / (float)(grid.x_cells);
  float dy = (grid.ymax - grid.ymin) / (float)(grid.y_cells);
  for (int i = 0; i < number_of_states; i++) {
    states[i].xmin = states[i].xmin + (dx / 100.00);
    states[i].ymin = states[i].ymin + (dy / 100.00);
    states[i].xmax = states[i].xmax - (dx / 100.00);
    states[i].ymax = states[i].ymax - (dy / 100.00);
  }
}

// ---------------------------------------------------------------- declarations

ops_stencil sten(int n, std::vector<int> pts, const char *name) {
  return ops_decl_stencil(2, n, pts.data(), (char *)name);
}
ops_stencil ssten(int n, std::vector<int> pts, int sx, int sy, const char *name) {
  int st[2] = {sx, sy};
  return ops_decl_strided_stencil(2, n, pts.data(), st, (char *)name);
}

void build_field() {
  const int nx = grid.x_cells, ny = grid.y_cells;
  S.blk = ops_decl_block(2, "clover grid");
  double *none = NULL;
  int *inone = NULL;
  int base[2] = {0, 0}, dm[2] = {-2, -2}, dp[2] = {2, 2};
  auto dat = [&](int sx, int sy, const char *name) {
    int size[2] = {sx, sy};
    return ops_decl_dat(S.blk, 1, size, base, dm, dp, none, "double", name);
  };
  S.density0 = dat(nx + 5, ny + 5, "density0");
  S.density1 = dat(nx + 5, ny + 5, "density1");
  S.energy0 = dat(nx + 5, ny + 5, "energy0");
  S.energy1 = dat(nx + 5, ny + 5, "energy1");
  S.pressure = dat(nx + 5, ny + 5, "pressure");
  S.viscosity = dat(nx + 5, ny + 5, "viscosity");
  S.soundspeed = dat(nx + 5, ny + 5, "soundspeed");
  S.volume = dat(nx + 5, ny + 5, "volume");
  S.xvel0 = dat(nx + 6, ny + 6, "xvel0");
  S.xvel1 = dat(nx + 6, ny + 6, "xvel1");
  S.yvel0 = dat(nx + 6, ny + 6, "yvel0");
  S.yvel1 = dat(nx + 6, ny + 6, "yvel1");
  S.work1 = dat(nx + 6, ny + 6, "work_array1");
  S.work2 = dat(nx + 6, ny + 6, "work_array2");
  S.work3 = dat(nx + 6, ny + 6, "work_array3");
  S.work4 = dat(nx + 6, ny + 6, "work_array4");
  S.work5 = dat(nx + 6, ny + 6, "work_array5");
  S.work6 = dat(nx + 6, ny + 6, "work_array6");
  S.work7 = dat(nx + 6, ny + 6, "work_array7");
  S.vol_flux_x = dat(nx + 6, ny + 5, "vol_flux_x");
  S.mass_flux_x = dat(nx + 6, ny + 5, "mass_flux_x");
  S.xarea = dat(nx + 6, ny + 5, "xarea");
  S.vol_flux_y = dat(nx + 5, ny + 6, "vol_flux_y");
  S.mass_flux_y = dat(nx + 5, ny + 6, "mass_flux_y");
  S.yarea = dat(nx + 5, ny + 6, "yarea");
  // one-dimensional (edge) dats along x and y
  int xm[2] = {-2, 0}, xp[2] = {2, 0}, ym[2] = {0, -2}, yp[2] = {0, 2};
  auto xdat = [&](int n, const char *name) {
    int size[2] = {n, 1};
    return ops_decl_dat(S.blk, 1, size, base, xm, xp, none, "double", name);
  };
  auto ydat = [&](int n, const char *name) {
    int size[2] = {1, n};
    return ops_decl_dat(S.blk, 1, size, base, ym, yp, none, "double", name);
  };
  S.cellx = xdat(nx + 5, "cellx");
  S.celldx = xdat(nx + 5, "celldx");
  S.celly = ydat(ny + 5, "celly");
  S.celldy = ydat(ny + 5, "celldy");
  S.vertexx = xdat(nx + 6, "vertexx");
  S.vertexdx = xdat(nx + 6, "vertexdx");
  S.vertexy = ydat(ny + 6, "vertexy");
  S.vertexdy = ydat(ny + 6, "vertexdy");
  {
    int size[2] = {nx + 6, 1};
    S.xx = ops_decl_dat(S.blk, 1, size, base, xm, xp, inone, "int", "xx");
  }
  {
    int size[2] = {1, ny + 6};
    S.yy = ops_decl_dat(S.blk, 1, size, base, ym, yp, inone, "int", "yy");
  }
  S.S00 = sten(1, {0, 0}, "00");
  S.P10 = sten(2, {0, 0, 1, 0}, "0,0:1,0");
  S.P01 = sten(2, {0, 0, 0, 1}, "0,0:0,1");
  S.M10 = sten(2, {0, 0, -1, 0}, "0,0:-1,0");
  S.M01 = sten(2, {0, 0, 0, -1}, "0,0:0,-1");
  S.S5 = sten(4, {1, 0, -1, 0, 0, 1, 0, -1}, "1,0:-1,0:0,1:0,-1");
  S.P10_P01_P11 = sten(4, {0, 0, 1, 0, 0, 1, 1, 1}, "0,0:1,0:0,1:1,1");
  S.M10_M01_M11 = sten(4, {0, 0, -1, 0, 0, -1, -1, -1}, "0,0:-1,0:0,-1:-1,-1");
  S.P10_M01_P1M1 = sten(4, {0, 0, 1, 0, 0, -1, 1, -1}, "0,0:1,0:0,-1:1,-1");
  S.P01_M10_M1P1 = sten(4, {0, 0, 0, 1, -1, 0, -1, 1}, "0,0:0,1:-1,0:-1,1");
  S.P10_M10_M20 = sten(4, {0, 0, 1, 0, -1, 0, -2, 0}, "0,0:1,0:-1,0:-2,0");
  S.P01_M01_M02 = sten(4, {0, 0, 0, 1, 0, -1, 0, -2}, "0,0:0,1:0,-1:0,-2");
  S.P10_P20_M10 = sten(4, {0, 0, 1, 0, 2, 0, -1, 0}, "0,0:1,0:2,0:-1,0");
  S.P01_P02_M01 = sten(4, {0, 0, 0, 1, 0, 2, 0, -1}, "0,0:0,1:0,2:0,-1");
  S.H[0] = sten(5, {0, 0, 0, 1, 0, 2, 0, 3, 0, 4}, "halo mirror +y");
  S.H[1] = sten(5, {0, 0, 0, -1, 0, -2, 0, -3, 0, -4}, "halo mirror -y");
  S.H[2] = sten(5, {0, 0, 1, 0, 2, 0, 3, 0, 4, 0}, "halo mirror +x");
  S.H[3] = sten(5, {0, 0, -1, 0, -2, 0, -3, 0, -4, 0}, "halo mirror -x");
  S.SX = ssten(1, {0, 0}, 1, 0, "s2D_00_stride2D_x");
  S.SY = ssten(1, {0, 0}, 0, 1, "s2D_00_stride2D_y");
  S.SX_P10 = ssten(2, {0, 0, 1, 0}, 1, 0, "s2D_00_P10_stride2D_x");
  S.SY_P01 = ssten(2, {0, 0, 0, 1}, 0, 1, "s2D_00_0P1_stride2D_y");
  S.SX_P10_M10 = ssten(3, {0, 0, 1, 0, -1, 0}, 1, 0, "self_stride2D_x");
  S.SY_P01_M01 = ssten(3, {0, 0, 0, 1, 0, -1}, 0, 1, "self_stride2D_y");
  S.r_dt = ops_decl_reduction_handle(sizeof(double), "double", "local_dt");
  S.r_vol = ops_decl_reduction_handle(sizeof(double), "double", "vol");
  S.r_mass = ops_decl_reduction_handle(sizeof(double), "double", "mass");
  S.r_ie = ops_decl_reduction_handle(sizeof(double), "double", "ie");
  S.r_ke = ops_decl_reduction_handle(sizeof(double), "double", "ke");
  S.r_press = ops_decl_reduction_handle(sizeof(double), "double", "press");
  S.r_xl = ops_decl_reduction_handle(sizeof(double), "double", "xl_pos");
  S.r_yl = ops_decl_reduction_handle(sizeof(double), "double", "yl_pos");
}

// ---------------------------------------------------------------- set-up loops

void initialise_chunk() {
  const int x0 = field.x_min, x1 = field.x_max, y0 = field.y_min, y1 = field.y_max;
  int full[] = {-2, grid.x_cells + 8, -2, grid.y_cells + 8};
  ops_par_loop(cl_xx, "initialise_chunk_kernel_xx", S.blk, 2, full,
               ops_arg_dat(S.xx, 1, S.SX, "int", OPS_WRITE), ops_arg_idx());
  ops_par_loop(cl_yy, "initialise_chunk_kernel_yy", S.blk, 2, full,
               ops_arg_dat(S.yy, 1, S.SY, "int", OPS_WRITE), ops_arg_idx());
  int rx[] = {x0 - 2, x1 + 3, y0 - 2, y0 - 1};
  int ry[] = {x0 - 2, x0 - 1, y0 - 2, y1 + 3};
  ops_par_loop(cl_vertex_x, "initialise_chunk_kernel_x", S.blk, 2, rx,
               ops_arg_dat(S.vertexx, 1, S.SX, "double", OPS_WRITE),
               ops_arg_dat(S.xx, 1, S.SX, "int", OPS_READ),
               ops_arg_dat(S.vertexdx, 1, S.SX, "double", OPS_WRITE));
  ops_par_loop(cl_vertex_y, "initialise_chunk_kernel_y", S.blk, 2, ry,
               ops_arg_dat(S.vertexy, 1, S.SY, "double", OPS_WRITE),
               ops_arg_dat(S.yy, 1, S.SY, "int", OPS_READ),
               ops_arg_dat(S.vertexdy, 1, S.SY, "double", OPS_WRITE));
  int cx[] = {x0 - 2, x1 + 2, y0 - 2, y0 - 1};
  int cy[] = {x0 - 2, x0 - 1, y0 - 2, y1 + 2};
  ops_par_loop(cl_cell_x, "initialise_chunk_kernel_cellx", S.blk, 2, cx,
               ops_arg_dat(S.vertexx, 1, S.SX_P10, "double", OPS_READ),
               ops_arg_dat(S.cellx, 1, S.SX, "double", OPS_WRITE),
               ops_arg_dat(S.celldx, 1, S.SX, "double", OPS_WRITE));
  ops_par_loop(cl_cell_y, "initialise_chunk_kernel_celly", S.blk, 2, cy,
               ops_arg_dat(S.vertexy, 1, S.SY_P01, "double", OPS_READ),
               ops_arg_dat(S.celly, 1, S.SY, "double", OPS_WRITE),
               ops_arg_dat(S.celldy, 1, S.SY, "double", OPS_WRITE));
  int all[] = {x0 - 2, x1 + 2, y0 - 2, y1 + 2};
  ops_par_loop(cl_volume, "initialise_chunk_kernel_volume", S.blk, 2, all,
               ops_arg_dat(S.volume, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.celldy, 1, S.SY, "double", OPS_READ),
               ops_arg_dat(S.xarea, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.celldx, 1, S.SX, "double", OPS_READ),
               ops_arg_dat(S.yarea, 1, S.S00, "double", OPS_WRITE));
}

void generate() {
  int all[] = {field.x_min - 2, field.x_max + 2, field.y_min - 2, field.y_max + 2};
  ops_par_loop(cl_generate, "generate_chunk_kernel", S.blk, 2, all,
               ops_arg_dat(S.vertexx, 1, S.SX_P10_M10, "double", OPS_READ),
               ops_arg_dat(S.vertexy, 1, S.SY_P01_M01, "double", OPS_READ),
               ops_arg_dat(S.energy0, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.density0, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.xvel0, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.yvel0, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.cellx, 1, S.SX_P10_M10, "double", OPS_READ),
               ops_arg_dat(S.celly, 1, S.SY_P01_M01, "double", OPS_READ));
}

// ---------------------------------------------------------------- physics loops

void ideal_gas(int predict) {
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  ops_dat r = predict ? S.density1 : S.density0, e = predict ? S.energy1 : S.energy0;
  ops_par_loop(cl_ideal_gas, "ideal_gas_kernel", S.blk, 2, in,
               ops_arg_dat(r, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(e, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.pressure, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.soundspeed, 1, S.S00, "double", OPS_WRITE));
}

bool any(const int *f, std::initializer_list<int> ids) {
  for (int i : ids)
    if (f[i]) return true;
  return false;
}

// reflective boundary update of the flagged fields, `depth` halo layers
void update_halo(int *fl, int depth) {
  const int x0 = field.x_min, x1 = field.x_max, y0 = field.y_min, y1 = field.y_max;
  // the two faces of an axis are independent when the mesh has more than 2*depth+1 cells
  // along it (a face's mirror source then stays clear of the other face's halo layers);
  // on thinner meshes they run one after the other
  const bool batch_axis[2] = {y1 - y0 > 2 * depth + 1, x1 - x0 > 2 * depth + 1};
  if (any(fl, {FIELD_DENSITY0, FIELD_DENSITY1, FIELD_ENERGY0, FIELD_ENERGY1, FIELD_PRESSURE,
               FIELD_VISCOSITY, FIELD_SOUNDSPEED})) {
    // cell centres: layer h mirrors h' = 2*pivot - 1 - h
    int mb[5] = {1, y0, 1, 1, 1}, mt[5] = {1, y1, 1, 1, 1};
    int ml[5] = {0, x0, 1, 1, 1}, mr[5] = {0, x1, 1, 1, 1};
    int rb[] = {x0 - depth, x1 + depth, y0 - depth, y0};
    int rt[] = {x0 - depth, x1 + depth, y1, y1 + depth};
    int rl[] = {x0 - depth, x0, y0 - depth, y1 + depth};
    int rr[] = {x1, x1 + depth, y0 - depth, y1 + depth};
    int *ms[4] = {mb, mt, ml, mr};
    int *rs[4] = {rb, rt, rl, rr};
    for (int s = 0; s < 4; s++) {
      // one launch per axis (faces 0-1: y, 2-3: x), axes in the reference's order
      if (s % 2 == 0 && batch_axis[s / 2]) ops_hip_batch_begin();
      ops_stencil st = S.H[s];
      ops_par_loop(cl_halo_cells, "update_halo_kernel1", S.blk, 2, rs[s],
                   ops_arg_dat_opt(S.density0, 1, st, "double", OPS_RW, fl[FIELD_DENSITY0]),
                   ops_arg_dat_opt(S.density1, 1, st, "double", OPS_RW, fl[FIELD_DENSITY1]),
                   ops_arg_dat_opt(S.energy0, 1, st, "double", OPS_RW, fl[FIELD_ENERGY0]),
                   ops_arg_dat_opt(S.energy1, 1, st, "double", OPS_RW, fl[FIELD_ENERGY1]),
                   ops_arg_dat_opt(S.pressure, 1, st, "double", OPS_RW, fl[FIELD_PRESSURE]),
                   ops_arg_dat_opt(S.viscosity, 1, st, "double", OPS_RW, fl[FIELD_VISCOSITY]),
                   ops_arg_dat_opt(S.soundspeed, 1, st, "double", OPS_RW, fl[FIELD_SOUNDSPEED]),
                   ops_arg_gbl(fl, NUM_FIELDS, "int", OPS_READ),
                   ops_arg_gbl(ms[s], 5, "int", OPS_READ), ops_arg_idx());
      if (s % 2 == 1) ops_hip_batch_end();
    }
  }
  if (any(fl, {FIELD_XVEL0, FIELD_XVEL1, FIELD_YVEL0, FIELD_YVEL1})) {
    // nodes: layer h mirrors 2*pivot - h; the normal velocity changes sign
    int mb[5] = {1, y0, 0, 1, -1}, mt[5] = {1, y1, 0, 1, -1};
    int ml[5] = {0, x0, 0, -1, 1}, mr[5] = {0, x1, 0, -1, 1};
    int rb[] = {x0 - depth, x1 + 1 + depth, y0 - depth, y0};
    int rt[] = {x0 - depth, x1 + 1 + depth, y1 + 1, y1 + 1 + depth};
    int rl[] = {x0 - depth, x0, y0 - depth, y1 + 1 + depth};
    int rr[] = {x1 + 1, x1 + 1 + depth, y0 - depth, y1 + 1 + depth};
    int *ms[4] = {mb, mt, ml, mr};
    int *rs[4] = {rb, rt, rl, rr};
    for (int s = 0; s < 4; s++) {
      // one launch per axis (faces 0-1: y, 2-3: x), axes in the reference's order
      if (s % 2 == 0 && batch_axis[s / 2]) ops_hip_batch_begin();
      ops_stencil st = S.H[s];
      ops_par_loop(cl_halo_vel, "update_halo_kernel2_vel", S.blk, 2, rs[s],
                   ops_arg_dat_opt(S.xvel0, 1, st, "double", OPS_RW, fl[FIELD_XVEL0]),
                   ops_arg_dat_opt(S.xvel1, 1, st, "double", OPS_RW, fl[FIELD_XVEL1]),
                   ops_arg_dat_opt(S.yvel0, 1, st, "double", OPS_RW, fl[FIELD_YVEL0]),
                   ops_arg_dat_opt(S.yvel1, 1, st, "double", OPS_RW, fl[FIELD_YVEL1]),
                   ops_arg_gbl(fl, NUM_FIELDS, "int", OPS_READ),
                   ops_arg_gbl(ms[s], 5, "int", OPS_READ), ops_arg_idx());
      if (s % 2 == 1) ops_hip_batch_end();
    }
  }
  if (any(fl, {FIELD_VOL_FLUX_X, FIELD_MASS_FLUX_X})) {
    int mb[6] = {1, y0, 0, 1, FIELD_VOL_FLUX_X, FIELD_MASS_FLUX_X};
    int mt[6] = {1, y1 - 1, 0, 1, FIELD_VOL_FLUX_X, FIELD_MASS_FLUX_X};
    int ml[6] = {0, x0, 0, -1, FIELD_VOL_FLUX_X, FIELD_MASS_FLUX_X};
    int mr[6] = {0, x1, 0, -1, FIELD_VOL_FLUX_X, FIELD_MASS_FLUX_X};
    int rb[] = {x0 - depth, x1 + 1 + depth, y0 - depth, y0};
    int rt[] = {x0 - depth, x1 + 1 + depth, y1, y1 + depth};
    int rl[] = {x0 - depth, x0, y0 - depth, y1 + depth};
    int rr[] = {x1 + 1, x1 + 1 + depth, y0 - depth, y1 + depth};
    int *ms[4] = {mb, mt, ml, mr};
    int *rs[4] = {rb, rt, rl, rr};
    for (int s = 0; s < 4; s++) {
      // one launch per axis (faces 0-1: y, 2-3: x), axes in the reference's order
      if (s % 2 == 0 && batch_axis[s / 2]) ops_hip_batch_begin();
      ops_stencil st = S.H[s];
      ops_par_loop(cl_halo_flux, "update_halo_kernel3_flux_x", S.blk, 2, rs[s],
                   ops_arg_dat_opt(S.vol_flux_x, 1, st, "double", OPS_RW, fl[FIELD_VOL_FLUX_X]),
                   ops_arg_dat_opt(S.mass_flux_x, 1, st, "double", OPS_RW, fl[FIELD_MASS_FLUX_X]),
                   ops_arg_gbl(fl, NUM_FIELDS, "int", OPS_READ),
                   ops_arg_gbl(ms[s], 6, "int", OPS_READ), ops_arg_idx());
      if (s % 2 == 1) ops_hip_batch_end();
    }
  }
  if (any(fl, {FIELD_VOL_FLUX_Y, FIELD_MASS_FLUX_Y})) {
    int mb[6] = {1, y0, 0, -1, FIELD_VOL_FLUX_Y, FIELD_MASS_FLUX_Y};
    int mt[6] = {1, y1, 0, -1, FIELD_VOL_FLUX_Y, FIELD_MASS_FLUX_Y};
    int ml[6] = {0, x0, 0, 1, FIELD_VOL_FLUX_Y, FIELD_MASS_FLUX_Y};
    int mr[6] = {0, x1 - 1, 0, 1, FIELD_VOL_FLUX_Y, FIELD_MASS_FLUX_Y};
    int rb[] = {x0 - depth, x1 + depth, y0 - depth, y0};
    int rt[] = {x0 - depth, x1 + depth, y1 + 1, y1 + 1 + depth};
    int rl[] = {x0 - depth, x0, y0 - depth, y1 + 1 + depth};
    int rr[] = {x1, x1 + depth, y0 - depth, y1 + 1 + depth};
    int *ms[4] = {mb, mt, ml, mr};
    int *rs[4] = {rb, rt, rl, rr};
    for (int s = 0; s < 4; s++) {
      // one launch per axis (faces 0-1: y, 2-3: x), axes in the reference's order
      if (s % 2 == 0 && batch_axis[s / 2]) ops_hip_batch_begin();
      ops_stencil st = S.H[s];
      ops_par_loop(cl_halo_flux, "update_halo_kernel4_flux_y", S.blk, 2, rs[s],
                   ops_arg_dat_opt(S.vol_flux_y, 1, st, "double", OPS_RW, fl[FIELD_VOL_FLUX_Y]),
                   ops_arg_dat_opt(S.mass_flux_y, 1, st, "double", OPS_RW, fl[FIELD_MASS_FLUX_Y]),
                   ops_arg_gbl(fl, NUM_FIELDS, "int", OPS_READ),
                   ops_arg_gbl(ms[s], 6, "int", OPS_READ), ops_arg_idx());
      if (s % 2 == 1) ops_hip_batch_end();
    }
  }
}

void set_fields(int *f, std::initializer_list<int> on) {
  for (int i = 0; i < NUM_FIELDS; i++) f[i] = 0;
  for (int i : on) f[i] = 1;
}

void viscosity_loop() {
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  ops_par_loop(cl_viscosity, "viscosity_kernel", S.blk, 2, in,
               ops_arg_dat(S.xvel0, 1, S.P10_P01_P11, "double", OPS_READ),
               ops_arg_dat(S.yvel0, 1, S.P10_P01_P11, "double", OPS_READ),
               ops_arg_dat(S.celldx, 1, S.SX_P10, "double", OPS_READ),
               ops_arg_dat(S.celldy, 1, S.SY_P01, "double", OPS_READ),
               ops_arg_dat(S.pressure, 1, S.S5, "double", OPS_READ),
               ops_arg_dat(S.density0, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.viscosity, 1, S.S00, "double", OPS_WRITE));
}

void calc_dt_loops() {
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  ops_par_loop(cl_calc_dt, "calc_dt_kernel", S.blk, 2, in,
               ops_arg_dat(S.celldx, 1, S.SX_P10, "double", OPS_READ),
               ops_arg_dat(S.celldy, 1, S.SY_P01, "double", OPS_READ),
               ops_arg_dat(S.soundspeed, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.viscosity, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.density0, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.xvel0, 1, S.P10_P01_P11, "double", OPS_READ),
               ops_arg_dat(S.xarea, 1, S.P10, "double", OPS_READ),
               ops_arg_dat(S.volume, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.yvel0, 1, S.P10_P01_P11, "double", OPS_READ),
               ops_arg_dat(S.yarea, 1, S.P01, "double", OPS_READ),
               ops_arg_dat(S.work1, 1, S.S00, "double", OPS_WRITE));
  ops_par_loop(cl_dt_min, "calc_dt_kernel_min", S.blk, 2, in,
               ops_arg_dat(S.work1, 1, S.S00, "double", OPS_READ),
               ops_arg_reduce(S.r_dt, 1, "double", OPS_MIN));
  // the controlling cell's position (calc_dt.cpp:72-87): jk_control is the constant 1.1,
  // so it is always the first interior cell
  int pt[] = {field.x_min, field.x_min + 1, field.y_min, field.y_min + 1};
  ops_par_loop(cl_dt_get, "calc_dt_kernel_getx", S.blk, 2, pt,
               ops_arg_dat(S.cellx, 1, S.SX, "double", OPS_READ),
               ops_arg_dat(S.celly, 1, S.SY, "double", OPS_READ),
               ops_arg_reduce(S.r_xl, 1, "double", OPS_INC),
               ops_arg_reduce(S.r_yl, 1, "double", OPS_INC));
}

double calc_dt() {
  calc_dt_loops();
  double local_dt, xl, yl;
  ops_reduction_result(S.r_dt, &local_dt);
  ops_reduction_result(S.r_xl, &xl);
  ops_reduction_result(S.r_yl, &yl);
  local_dt = MIN(local_dt, g_big);
  return local_dt;
}

// device: dt is computed on the GPU from the reduction's device value (captured steps)
void timestep(bool device = false) {
  int f[NUM_FIELDS];
  ideal_gas(0);
  set_fields(f, {FIELD_DENSITY0, FIELD_ENERGY0, FIELD_PRESSURE, FIELD_XVEL0, FIELD_YVEL0});
  update_halo(f, 1);
  viscosity_loop();
  set_fields(f, {FIELD_VISCOSITY});
  update_halo(f, 1);
  if (device) {
    calc_dt_loops();
    clover_dt_launch(ops_hip_reduction_device_ptr(S.r_dt), ops_hip_const_device_ptr("dt"),
                     S.dt_state, g_big, S.dtrise, S.dtmax, S.dtmin);
    return;
  }
  double dtlp = calc_dt();
  if (dtlp < S.dtmin) throw OPSException(OPS_RUNTIME_ERROR, "cloverleaf: timestep below dtmin");
  double dt_temp = g_big;
  if (dtlp <= dt_temp) dt_temp = dtlp;
  dt_temp = MIN(MIN(dt_temp, (S.dtold * S.dtrise)), S.dtmax);
  ops_update_const("dt", 1, "double", &dt_temp);
  dt = dt_temp;
  if (dt < S.dtmin) throw OPSException(OPS_RUNTIME_ERROR, "cloverleaf: small timestep");
  S.dtold = dt;
}

void chain_on() {
  if (S.chain > 1) ops_hip_chain_begin(S.chain);
}
void chain_off() {
  if (S.chain > 1) ops_hip_chain_end();
}

void revert() {
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  ops_par_loop(cl_revert, "revert_kernel", S.blk, 2, in,
               ops_arg_dat(S.density0, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.density1, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.energy0, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.energy1, 1, S.S00, "double", OPS_WRITE));
}

void pdv(int predict) {
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  if (predict) {
    chain_on();
    ops_par_loop(cl_pdv_predict, "PdV_kernel_predict", S.blk, 2, in,
                 ops_arg_dat(S.xarea, 1, S.P10_P01_P11, "double", OPS_READ),
                 ops_arg_dat(S.xvel0, 1, S.P10_P01_P11, "double", OPS_READ),
                 ops_arg_dat(S.yarea, 1, S.P10_P01_P11, "double", OPS_READ),
                 ops_arg_dat(S.yvel0, 1, S.P10_P01_P11, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.pressure, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.density0, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.density1, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.viscosity, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.energy0, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.energy1, 1, S.S00, "double", OPS_WRITE));
  } else {
    ops_par_loop(cl_pdv_full, "PdV_kernel_nopredict", S.blk, 2, in,
                 ops_arg_dat(S.xarea, 1, S.P10_P01_P11, "double", OPS_READ),
                 ops_arg_dat(S.xvel0, 1, S.P10_P01_P11, "double", OPS_READ),
                 ops_arg_dat(S.xvel1, 1, S.P10_P01_P11, "double", OPS_READ),
                 ops_arg_dat(S.yarea, 1, S.P10_P01_P11, "double", OPS_READ),
                 ops_arg_dat(S.yvel0, 1, S.P10_P01_P11, "double", OPS_READ),
                 ops_arg_dat(S.yvel1, 1, S.P10_P01_P11, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.pressure, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.density0, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.density1, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.viscosity, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.energy0, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.energy1, 1, S.S00, "double", OPS_WRITE));
  }
  if (predict) {
    ideal_gas(1);
    chain_off();
    int f[NUM_FIELDS];
    set_fields(f, {FIELD_PRESSURE});
    update_halo(f, 1);
    // revert, accelerate, the full PdV and the fluxes form the next chain (the caller
    // closes it after flux_calc)
    chain_on();
    revert();
  }
}

void accelerate() {
  int in1[] = {field.x_min, field.x_max + 1, field.y_min, field.y_max + 1};
  ops_par_loop(cl_accelerate, "accelerate_kernel", S.blk, 2, in1,
               ops_arg_dat(S.density0, 1, S.M10_M01_M11, "double", OPS_READ),
               ops_arg_dat(S.volume, 1, S.M10_M01_M11, "double", OPS_READ),
               ops_arg_dat(S.work1, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.xvel0, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.xvel1, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.xarea, 1, S.M01, "double", OPS_READ),
               ops_arg_dat(S.pressure, 1, S.M10_M01_M11, "double", OPS_READ),
               ops_arg_dat(S.yvel0, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.yvel1, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.yarea, 1, S.M10, "double", OPS_READ),
               ops_arg_dat(S.viscosity, 1, S.M10_M01_M11, "double", OPS_READ));
}

void flux_calc() {
  int rx[] = {field.x_min, field.x_max + 1, field.y_min, field.y_max};
  ops_par_loop(cl_flux_x, "flux_calc_kernelx", S.blk, 2, rx,
               ops_arg_dat(S.vol_flux_x, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.xarea, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.xvel0, 1, S.P01, "double", OPS_READ),
               ops_arg_dat(S.xvel1, 1, S.P01, "double", OPS_READ));
  int ry[] = {field.x_min, field.x_max, field.y_min, field.y_max + 1};
  ops_par_loop(cl_flux_y, "flux_calc_kernely", S.blk, 2, ry,
               ops_arg_dat(S.vol_flux_y, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.yarea, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.yvel0, 1, S.P10, "double", OPS_READ),
               ops_arg_dat(S.yvel1, 1, S.P10, "double", OPS_READ));
}

const int XDIR = 1, YDIR = 2;

void advec_cell(int sweep, int dir) {
  const int x0 = field.x_min, x1 = field.x_max, y0 = field.y_min, y1 = field.y_max;
  int all[] = {x0 - 2, x1 + 2, y0 - 2, y1 + 2};
  int in[] = {x0, x1, y0, y1};
  if (dir == XDIR) {
    if (sweep == 1)
      ops_par_loop(cl_adv_cell_prevol_1x, "advec_cell_kernel1_xdir", S.blk, 2, all,
                   ops_arg_dat(S.work1, 1, S.S00, "double", OPS_WRITE),
                   ops_arg_dat(S.work2, 1, S.S00, "double", OPS_WRITE),
                   ops_arg_dat(S.volume, 1, S.S00, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_x, 1, S.P10, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_y, 1, S.P01, "double", OPS_READ));
    else
      ops_par_loop(cl_adv_cell_prevol_2x, "advec_cell_kernel2_xdir", S.blk, 2, all,
                   ops_arg_dat(S.work1, 1, S.S00, "double", OPS_WRITE),
                   ops_arg_dat(S.work2, 1, S.S00, "double", OPS_WRITE),
                   ops_arg_dat(S.volume, 1, S.S00, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_x, 1, S.P10, "double", OPS_READ));
    int r2x[] = {x0, x1 + 2, y0, y1};
    ops_par_loop(cl_adv_cell_flux_x, "advec_cell_kernel3_xdir", S.blk, 2, r2x,
                 ops_arg_dat(S.vol_flux_x, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.M10, "double", OPS_READ),
                 ops_arg_dat(S.xx, 1, S.SX_P10, "int", OPS_READ),
                 ops_arg_dat(S.vertexdx, 1, S.SX_P10_M10, "double", OPS_READ),
                 ops_arg_dat(S.density1, 1, S.P10_M10_M20, "double", OPS_READ),
                 ops_arg_dat(S.energy1, 1, S.P10_M10_M20, "double", OPS_READ),
                 ops_arg_dat(S.mass_flux_x, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S00, "double", OPS_WRITE));
    ops_par_loop(cl_adv_cell_update_x, "advec_cell_kernel4_xdir", S.blk, 2, in,
                 ops_arg_dat(S.density1, 1, S.S00, "double", OPS_RW),
                 ops_arg_dat(S.energy1, 1, S.S00, "double", OPS_RW),
                 ops_arg_dat(S.mass_flux_x, 1, S.P10, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_x, 1, S.P10, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.work2, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work4, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work5, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work6, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.P10, "double", OPS_READ));
  } else {
    if (sweep == 1)
      ops_par_loop(cl_adv_cell_prevol_1y, "advec_cell_kernel1_ydir", S.blk, 2, all,
                   ops_arg_dat(S.work1, 1, S.S00, "double", OPS_WRITE),
                   ops_arg_dat(S.work2, 1, S.S00, "double", OPS_WRITE),
                   ops_arg_dat(S.volume, 1, S.S00, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_x, 1, S.P10, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_y, 1, S.P01, "double", OPS_READ));
    else
      ops_par_loop(cl_adv_cell_prevol_2y, "advec_cell_kernel2_ydir", S.blk, 2, all,
                   ops_arg_dat(S.work1, 1, S.S00, "double", OPS_WRITE),
                   ops_arg_dat(S.work2, 1, S.S00, "double", OPS_WRITE),
                   ops_arg_dat(S.volume, 1, S.S00, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_y, 1, S.P01, "double", OPS_READ));
    int r2y[] = {x0, x1, y0, y1 + 2};
    ops_par_loop(cl_adv_cell_flux_y, "advec_cell_kernel3_ydir", S.blk, 2, r2y,
                 ops_arg_dat(S.vol_flux_y, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.M01, "double", OPS_READ),
                 ops_arg_dat(S.yy, 1, S.SY_P01, "int", OPS_READ),
                 ops_arg_dat(S.vertexdy, 1, S.SY_P01_M01, "double", OPS_READ),
                 ops_arg_dat(S.density1, 1, S.P01_M01_M02, "double", OPS_READ),
                 ops_arg_dat(S.energy1, 1, S.P01_M01_M02, "double", OPS_READ),
                 ops_arg_dat(S.mass_flux_y, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S00, "double", OPS_WRITE));
    ops_par_loop(cl_adv_cell_update_y, "advec_cell_kernel4_ydir", S.blk, 2, in,
                 ops_arg_dat(S.density1, 1, S.S00, "double", OPS_RW),
                 ops_arg_dat(S.energy1, 1, S.S00, "double", OPS_RW),
                 ops_arg_dat(S.mass_flux_y, 1, S.P01, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_y, 1, S.P01, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.work2, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work4, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work5, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work6, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.P01, "double", OPS_READ));
  }
}

void advec_mom(int which_vel, int sweep, int dir) {
  const int x0 = field.x_min, x1 = field.x_max, y0 = field.y_min, y1 = field.y_max;
  int all[] = {x0 - 2, x1 + 2, y0 - 2, y1 + 2};
  ops_dat vel = which_vel == 1 ? S.xvel1 : S.yvel1;
  int mom_sweep = dir + 2 * (sweep - 1);
  if (mom_sweep == 1)
    ops_par_loop(cl_mom_vol_x1, "advec_mom_kernel_x1", S.blk, 2, all,
                 ops_arg_dat(S.work6, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_x, 1, S.P10, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_y, 1, S.P01, "double", OPS_READ));
  else if (mom_sweep == 2)
    ops_par_loop(cl_mom_vol_y1, "advec_mom_kernel_y1", S.blk, 2, all,
                 ops_arg_dat(S.work6, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_x, 1, S.P10, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_y, 1, S.P01, "double", OPS_READ));
  else if (mom_sweep == 3)
    ops_par_loop(cl_mom_vol_x2, "advec_mom_kernel_x2", S.blk, 2, all,
                 ops_arg_dat(S.work6, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_y, 1, S.P01, "double", OPS_READ));
  else if (mom_sweep == 4)
    ops_par_loop(cl_mom_vol_y2, "advec_mom_kernel_y2", S.blk, 2, all,
                 ops_arg_dat(S.work6, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_x, 1, S.P10, "double", OPS_READ));
  if (dir == XDIR) {
    int r1[] = {x0 - 2, x1 + 2, y0, y1 + 1};
    ops_par_loop(cl_mom_node_flux_x, "advec_mom_kernel_mass_flux_x", S.blk, 2, r1,
                 ops_arg_dat(S.work1, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.mass_flux_x, 1, S.P10_M01_P1M1, "double", OPS_READ));
    int r2[] = {x0 - 1, x1 + 2, y0, y1 + 1};
    ops_par_loop(cl_mom_node_mass_x, "advec_mom_kernel_post_pre_advec_x", S.blk, 2, r2,
                 ops_arg_dat(S.work2, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.M10_M01_M11, "double", OPS_READ),
                 ops_arg_dat(S.density1, 1, S.M10_M01_M11, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work1, 1, S.M10, "double", OPS_READ));
    int r3[] = {x0 - 1, x1 + 1, y0, y1 + 1};
    ops_par_loop(cl_mom_flux_x, "advec_mom_kernel1_x", S.blk, 2, r3,
                 ops_arg_dat(S.work1, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.P10, "double", OPS_READ),
                 ops_arg_dat(S.work5, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.celldx, 1, S.SX_P10_M10, "double", OPS_READ),
                 ops_arg_dat(vel, 1, S.P10_P20_M10, "double", OPS_READ));
    int r4[] = {x0, x1 + 1, y0, y1 + 1};
    ops_par_loop(cl_mom_update_x, "advec_mom_kernel2_x", S.blk, 2, r4,
                 ops_arg_dat(vel, 1, S.S00, "double", OPS_RW),
                 ops_arg_dat(S.work2, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.work5, 1, S.M10, "double", OPS_READ));
  } else {
    int r1[] = {x0, x1 + 1, y0 - 2, y1 + 2};
    ops_par_loop(cl_mom_node_flux_y, "advec_mom_kernel_mass_flux_y", S.blk, 2, r1,
                 ops_arg_dat(S.work1, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.mass_flux_y, 1, S.P01_M10_M1P1, "double", OPS_READ));
    int r2[] = {x0, x1 + 1, y0 - 1, y1 + 2};
    ops_par_loop(cl_mom_node_mass_y, "advec_mom_kernel_post_pre_advec_y", S.blk, 2, r2,
                 ops_arg_dat(S.work2, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.M10_M01_M11, "double", OPS_READ),
                 ops_arg_dat(S.density1, 1, S.M10_M01_M11, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.work1, 1, S.M01, "double", OPS_READ));
    int r3[] = {x0, x1 + 1, y0 - 1, y1 + 1};
    ops_par_loop(cl_mom_flux_y, "advec_mom_kernel1_y", S.blk, 2, r3,
                 ops_arg_dat(S.work1, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.P01, "double", OPS_READ),
                 ops_arg_dat(S.work5, 1, S.S00, "double", OPS_WRITE),
                 ops_arg_dat(S.celldy, 1, S.SY_P01_M01, "double", OPS_READ),
                 ops_arg_dat(vel, 1, S.P01_P02_M01, "double", OPS_READ));
    int r4[] = {x0, x1 + 1, y0, y1 + 1};
    ops_par_loop(cl_mom_update_y, "advec_mom_kernel2_y", S.blk, 2, r4,
                 ops_arg_dat(vel, 1, S.S00, "double", OPS_RW),
                 ops_arg_dat(S.work2, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.S00, "double", OPS_READ),
                 ops_arg_dat(S.work5, 1, S.M01, "double", OPS_READ));
  }
}

void advection() {
  int f[NUM_FIELDS];
  int dir = S.advect_x ? XDIR : YDIR;
  set_fields(f, {FIELD_DENSITY1, FIELD_ENERGY1, FIELD_VOL_FLUX_X, FIELD_VOL_FLUX_Y});
  update_halo(f, 2);
  chain_on();
  advec_cell(1, dir);
  chain_off();
  if (dir == XDIR)
    set_fields(f, {FIELD_DENSITY1, FIELD_ENERGY1, FIELD_XVEL1, FIELD_YVEL1, FIELD_MASS_FLUX_X});
  else
    set_fields(f, {FIELD_DENSITY1, FIELD_ENERGY1, FIELD_XVEL1, FIELD_YVEL1, FIELD_MASS_FLUX_Y});
  update_halo(f, 2);
  chain_on();
  advec_mom(1, 1, dir);
  advec_mom(2, 1, dir);
  dir = S.advect_x ? YDIR : XDIR;
  advec_cell(2, dir);
  chain_off();
  set_fields(f, {FIELD_DENSITY1, FIELD_ENERGY1, FIELD_XVEL1, FIELD_YVEL1, FIELD_MASS_FLUX_X,
                 FIELD_MASS_FLUX_Y});
  update_halo(f, 2);
  // the last momentum sweeps and reset_field form one chain (the caller closes it)
  chain_on();
  advec_mom(1, 2, dir);
  advec_mom(2, 2, dir);
}

void reset_field() {
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  ops_par_loop(cl_reset_cells, "reset_field_kernel1", S.blk, 2, in,
               ops_arg_dat(S.density0, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.density1, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.energy0, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.energy1, 1, S.S00, "double", OPS_READ));
  int in1[] = {field.x_min, field.x_max + 1, field.y_min, field.y_max + 1};
  ops_par_loop(cl_reset_nodes, "reset_field_kernel2", S.blk, 2, in1,
               ops_arg_dat(S.xvel0, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.xvel1, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.yvel0, 1, S.S00, "double", OPS_WRITE),
               ops_arg_dat(S.yvel1, 1, S.S00, "double", OPS_READ));
}

void field_summary(double *out) {
  ideal_gas(0);
  int in[] = {field.x_min, field.x_max, field.y_min, field.y_max};
  ops_par_loop(cl_field_summary, "field_summary_kernel", S.blk, 2, in,
               ops_arg_dat(S.volume, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.density0, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.energy0, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.pressure, 1, S.S00, "double", OPS_READ),
               ops_arg_dat(S.xvel0, 1, S.P10_P01_P11, "double", OPS_READ),
               ops_arg_dat(S.yvel0, 1, S.P10_P01_P11, "double", OPS_READ),
               ops_arg_reduce(S.r_vol, 1, "double", OPS_INC),
               ops_arg_reduce(S.r_mass, 1, "double", OPS_INC),
               ops_arg_reduce(S.r_ie, 1, "double", OPS_INC),
               ops_arg_reduce(S.r_ke, 1, "double", OPS_INC),
               ops_arg_reduce(S.r_press, 1, "double", OPS_INC));
  ops_reduction_result(S.r_vol, &out[0]);
  ops_reduction_result(S.r_mass, &out[1]);
  ops_reduction_result(S.r_ie, &out[2]);
  ops_reduction_result(S.r_ke, &out[3]);
  ops_reduction_result(S.r_press, &out[4]);
}

void start() {
  S.time = 0.0;
  S.step = 0;
  S.dtold = S.dtinit;
  dt = S.dtinit;
  build_field();
  ops_decl_const("g_small", 1, "double", &g_small);
  ops_decl_const("g_big", 1, "double", &g_big);
  ops_decl_const("dtc_safe", 1, "double", &dtc_safe);
  ops_decl_const("dtu_safe", 1, "double", &dtu_safe);
  ops_decl_const("dtv_safe", 1, "double", &dtv_safe);
  ops_decl_const("dtdiv_safe", 1, "double", &dtdiv_safe);
  ops_decl_const("field", 1, "field_type", &field);
  ops_decl_const("grid", 1, "grid_type", &grid);
  ops_decl_const("number_of_states", 1, "int", &number_of_states);
  ops_decl_const("states", number_of_states, "state_type", states);
  ops_decl_const("g_circ", 1, "int", &g_circ);
  ops_decl_const("g_point", 1, "int", &g_point);
  ops_decl_const("g_rect", 1, "int", &g_rect);
  ops_decl_const("dt", 1, "double", &dt);
  initialise_chunk();
  generate();
  S.advect_x = 1;
  ideal_gas(0);
  int f[NUM_FIELDS];
  set_fields(f, {FIELD_DENSITY0, FIELD_ENERGY0, FIELD_PRESSURE, FIELD_VISCOSITY, FIELD_DENSITY1,
                 FIELD_ENERGY1, FIELD_XVEL0, FIELD_YVEL0, FIELD_XVEL1, FIELD_YVEL1});
  update_halo(f, 2);
  double summary[5];
  field_summary(summary);
}

void hydro_step() {
  S.step++;
  timestep();
  pdv(1);
  accelerate();
  pdv(0);
  flux_calc();
  chain_off();
  advection();
  reset_field();
  chain_off();
  S.advect_x = !S.advect_x;
  S.time = S.time + dt;
  // clover_leaf.cpp:206-208: a field summary every summary_frequency steps
  if (S.summary_frequency != 0 && S.step % S.summary_frequency == 0) field_summary(S.summary);
}

// the launches of one step with dt, dtold and time kept on the device
void device_step() {
  timestep(true);
  pdv(1);
  accelerate();
  pdv(0);
  flux_calc();
  chain_off();
  advection();
  reset_field();
  chain_off();
}

// n steps replayed from two captured graphs (one per sweep order); the results are those
// of hydro_step: the same launches, and dt from the same operations on the same values
void capture_steps() {
  if (!S.dt_state) S.dt_state = clover_dt_state_alloc();
  for (int k = 0; k < 2; k++) {
    const int ax = S.advect_x ^ k;
    if (S.step_graph[ax]) continue;
    const int keep = S.advect_x;
    S.advect_x = ax;
    // the captured reductions start from their own results (nothing reads the cell
    // position of a captured step)
    ops_hip_reduction_discard(S.r_dt);
    ops_hip_reduction_discard(S.r_xl);
    ops_hip_reduction_discard(S.r_yl);
    if (ops_hip_graph_begin())
      throw OPSException(OPS_RUNTIME_ERROR, "cloverleaf: graph capture unavailable");
    try {
      device_step();
    } catch (...) {
      ops_hip_graph_free(ops_hip_graph_end());  // leave capture mode before reporting
      S.advect_x = keep;
      throw;
    }
    S.step_graph[ax] = ops_hip_graph_end();
    ops_hip_reduction_discard(S.r_dt);
    ops_hip_reduction_discard(S.r_xl);
    ops_hip_reduction_discard(S.r_yl);
    S.advect_x = keep;
    if (!S.step_graph[ax]) throw OPSException(OPS_RUNTIME_ERROR, "cloverleaf: graph capture failed");
  }
}

void graph_steps(int n) {
  capture_steps();
  clover_dt_state_set(S.dt_state, S.dtold, S.time, dt);
  for (int i = 0; i < n; i++) {
    if (ops_hip_graph_replay(S.step_graph[S.advect_x], 1))
      throw OPSException(OPS_RUNTIME_ERROR, "cloverleaf: graph replay failed");
    S.advect_x = !S.advect_x;
    S.step++;
    // the field summary of every summary_frequency steps runs eagerly between replays (its
    // sums are read on the host)
    if (S.summary_frequency != 0 && S.step % S.summary_frequency == 0) field_summary(S.summary);
  }
  double st[4];
  clover_dt_state_get(S.dt_state, st);
  S.dtold = st[0];
  S.time = st[1];
  dt = st[3];
  if (st[2] != 0.0) throw OPSException(OPS_RUNTIME_ERROR, "cloverleaf: timestep below dtmin");
}

void free_graphs() {
  for (void *&g : S.step_graph) {
    ops_hip_graph_free(g);
    g = nullptr;
  }
  clover_dt_state_free(S.dt_state);
  S.dt_state = nullptr;
}

}  // namespace

extern "C" int cloverleaf_init(int argc, const char **argv, const char *deck) {
  try {
    ops_init(argc, argv, 1);
    const int graph = S.graph;
    S = Sim();
    S.graph = graph;
    read_deck(deck);
    start();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf_init: %s\n", e.what());
    return 1;
  }
}

// loop chains: the loops between halo exchanges run band by band in `tiles` bands (0 or 1:
// one launch per loop); returns the previous setting. Captured step graphs are rebuilt.
extern "C" int cloverleaf_set_chain(int tiles) {
  const int old = S.chain;
  if (tiles != old) free_graphs();
  S.chain = tiles;
  return old;
}

// captured steps on (1) or off (0) for cloverleaf_step; returns the previous setting
extern "C" int cloverleaf_set_graph(int on) {
  const int old = S.graph;
  S.graph = on;
  return old;
}

// captures the two step graphs now (otherwise the first captured step does), so that a
// timed region replays only; 0 on success
extern "C" int cloverleaf_capture(void) {
  try {
    capture_steps();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf_capture: %s\n", e.what());
    return 1;
  }
}

extern "C" int cloverleaf_step(int n) {
  try {
    if (S.graph && n > 0) {
      graph_steps(n);
      return 0;
    }
    for (int i = 0; i < n; i++) hydro_step();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf_step: %s\n", e.what());
    return 1;
  }
}

// steps until the deck's end condition (clover_leaf.cpp:210) or max_steps; returns the
// number of steps taken, or -1 on error
extern "C" int cloverleaf_run(int max_steps) {
  try {
    int n = 0;
    while (n < max_steps) {
      hydro_step();
      n++;
      if ((S.time + g_small) > S.end_time || (S.step >= S.end_step)) break;
    }
    return n;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf_run: %s\n", e.what());
    return -1;
  }
}

extern "C" int cloverleaf_summary(double *out5) {
  try {
    field_summary(out5);
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf_summary: %s\n", e.what());
    return 1;
  }
}

extern "C" int cloverleaf_state(double *out4) {
  out4[0] = S.step;
  out4[1] = S.time;
  out4[2] = dt;
  out4[3] = S.end_step;
  return 0;
}

// copies one dat, halos included (size[0]*size[1] doubles); returns -1 if unknown
extern "C" long cloverleaf_fetch(const char *name, double *out) {
  try {
    ops_dat all[] = {S.density0, S.density1, S.energy0, S.energy1, S.pressure, S.viscosity,
                     S.soundspeed, S.volume, S.xvel0, S.xvel1, S.yvel0, S.yvel1,
                     S.vol_flux_x, S.vol_flux_y, S.mass_flux_x, S.mass_flux_y, S.xarea,
                     S.yarea, S.cellx, S.celly, S.vertexx, S.vertexy, S.celldx, S.celldy};
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
    fprintf(stderr, "cloverleaf_fetch: %s\n", e.what());
    return -2;
  }
}

extern "C" int cloverleaf_finalize(void) {
  try {
    free_graphs();
    ops_exit();
    free(states);
    states = NULL;
    number_of_states = 0;
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf_finalize: %s\n", e.what());
    return 1;
  }
}

#ifdef OPS_APP_MAIN
int main(int argc, const char **argv) {
  const char *deck = argc > 1 ? argv[1] : "clover.in";
  if (cloverleaf_init(argc, argv, deck)) return 1;
  int steps = cloverleaf_run(1 << 30);
  if (steps < 0) return 1;
  double s[5];
  cloverleaf_summary(s);
  printf("step %d  vol %.15e mass %.15e ie %.15e ke %.15e press %.15e\n", steps, s[0], s[1],
         s[2], s[3], s[4]);
  return cloverleaf_finalize();
}
#endif
