// CloverLeaf 3D (explicit Lagrangian-Eulerian hydrodynamics) on one OPS block, written
// against OPS API v1 (pointer kernels with OPS_ACC<n> accessors).
//
// The sequence of loops, their ranges and argument lists follow the reference
// application (apps/c/CloverLeaf_3D: clover_leaf.cpp:136-216, start.cpp, build_field.cpp,
// initialise_chunk.cpp, generate.cpp, ideal_gas.cpp, time_step.cpp, calc_dt.cpp,
// viscosity.cpp, PdV.cpp, revert.cpp, accelerate.cpp, flux_calc.cpp, advection.cpp,
// advec_cell.cpp, advec_mom.cpp, reset_field.cpp, field_summary.cpp, update_halo.cpp) so
// that results can be compared bit for bit. Differences from the reference, none of which
// changes a value: the reflective halo update runs one loop per face and field group
// covering all halo layers of the face (the reference runs one loop per layer and per
// velocity component), and the diagnostic-only position lookup of the controlling cell
// in calc_dt is not evaluated.
//
// Entry points (C ABI) for drivers and tests:
//   cloverleaf3d_init(argc, argv, deck)   parse the deck, build and generate the problem
//   cloverleaf3d_step(n)                  advance n hydro steps
//   cloverleaf3d_run(max_steps)           advance to the deck's end condition
//   cloverleaf3d_set_graph(on)            replay cloverleaf3d_step's steps from captured
//                                         hipGraphs (dt computed on the GPU)
//   cloverleaf3d_summary(out[5])          field summary: volume, mass, ie, ke, pressure
//   cloverleaf3d_state(out[4])            step, time, dt, end_step
//   cloverleaf3d_fetch(name, out)         one dat, halos included
//   cloverleaf3d_finalize()
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <initializer_list>
#include <string>
#include <vector>

#define OPS_3D
#include <ops_seq.h>

#include "user_types.h"

// ---- ops_decl_const globals referenced by the kernels
double g_small = 1.0e-16;
double g_big = 1.0e+21;
double dtc_safe, dtu_safe, dtv_safe, dtw_safe, dtdiv_safe;
double dt;
field_type field;
grid_type grid;
int number_of_states = 0;
state_type *states = NULL;
int g_cube = 1, g_sphe = 2, g_point = 3;

#include "clover3d_kernels.h"

// device-side time-step control for captured steps (apps/cloverleaf/native/clover_dt.hip)
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
}

namespace {

const int XDIR = 1, YDIR = 2, ZDIR = 3;

struct Sim {
  // deck
  double dtinit = 0.1, dtmax = 1.0, dtmin = 0.0000001, dtrise = 1.05, end_time = 10.0;
  int end_step = 640000, test_problem = 0;
  // run state
  int step = 0, advect_x = 1;
  double time = 0.0, dtold = 0.0;
  // captured steps: one graph per sweep order (advect_x), the device dt state
  int graph = 0;
  void *step_graph[2] = {nullptr, nullptr};
  void *dt_state = nullptr;
  // OPS objects
  ops_block blk;
  ops_dat density0, density1, energy0, energy1, pressure, viscosity, soundspeed, volume;
  ops_dat xvel0, xvel1, yvel0, yvel1, zvel0, zvel1;
  ops_dat vol_flux_x, vol_flux_y, vol_flux_z, mass_flux_x, mass_flux_y, mass_flux_z;
  ops_dat xarea, yarea, zarea, work1, work2, work3, work4, work5, work6, work7;
  ops_dat cellx, celly, cellz, vertexx, vertexy, vertexz, celldx, celldy, celldz;
  ops_dat vertexdx, vertexdy, vertexdz, xx, yy, zz;
  ops_stencil S000, P100, P010, P001, M100, M010, M001, S6;
  ops_stencil F_P1P1P1, F_M1M1M1, F_0M1M1, F_M10M1, F_M1M10, F_0P1P1, F_P10P1, F_P1P10;
  ops_stencil F_P1M1M1, F_M1P1M1, F_M1M1P1;
  ops_stencil X_UP2, Y_UP2, Z_UP2;        // 0, +1, +2, -1 along an axis
  ops_stencil X_DN2, Y_DN2, Z_DN2;        // 0, +1, -1, -2 along an axis
  ops_stencil SX, SY, SZ, SX_P, SY_P, SZ_P, SX_PM, SY_PM, SZ_PM;
  // reflective halo sources, one stencil per face: -y, +y, -x, +x, -z, +z faces
  ops_stencil H[6];
  ops_reduction r_dt, r_vol, r_mass, r_ie, r_ke, r_press;
} S;

// ---------------------------------------------------------------- input deck
// The deck format of apps/c/CloverLeaf_3D/read_input.cpp: "*clover" ... "*endclover",
// tokens separated by blanks or '=', "state N key=value ..." lines.

void read_deck(const char *path) {
  S.dtinit = 0.1;
  S.dtmax = 1.0;
  S.dtmin = 0.0000001;
  S.dtrise = 1.05;
  S.end_time = 10.0;
  S.end_step = 640000;
  S.test_problem = 0;
  dtc_safe = 0.75;
  dtu_safe = 0.5;
  dtv_safe = 0.5;
  dtw_safe = 0.5;
  dtdiv_safe = 0.5;
  grid.xmin = grid.ymin = grid.zmin = 0;
  grid.xmax = grid.ymax = grid.zmax = 100;
  grid.x_cells = grid.y_cells = grid.z_cells = 10;
  free(states);
  states = NULL;
  number_of_states = 0;
  FILE *f = fopen(path, "r");
  if (!f) throw OPSException(OPS_RUNTIME_CONFIGURATION_ERROR, "cloverleaf3d: cannot open deck");
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
    else if (k == "zmin") grid.zmin = val(1);
    else if (k == "zmax") grid.zmax = val(1);
    else if (k == "x_cells") grid.x_cells = (int)val(1);
    else if (k == "y_cells") grid.y_cells = (int)val(1);
    else if (k == "z_cells") grid.z_cells = (int)val(1);
    else if (k == "test_problem") S.test_problem = (int)val(1);
    else if (k == "state") {
      states = (state_type *)realloc(states, sizeof(state_type) * (number_of_states + 1));
      state_type &s = states[number_of_states];
      memset(&s, 0, sizeof s);
      for (size_t i = 2; i < tok.size(); i++) {
        const std::string &key = tok[i];
        if (key == "xvel") s.xvel = val(i + 1);
        else if (key == "yvel") s.yvel = val(i + 1);
        else if (key == "zvel") s.zvel = val(i + 1);
        else if (key == "xmin") s.xmin = val(i + 1);
        else if (key == "xmax") s.xmax = val(i + 1);
        else if (key == "ymin") s.ymin = val(i + 1);
        else if (key == "ymax") s.ymax = val(i + 1);
        else if (key == "zmin") s.zmin = val(i + 1);
        else if (key == "zmax") s.zmax = val(i + 1);
        else if (key == "radius") s.radius = val(i + 1);
        else if (key == "density") s.density = val(i + 1);
        else if (key == "energy") s.energy = val(i + 1);
        else if (key == "geometry" && i + 1 < tok.size()) {
          if (tok[i + 1] == "cuboid") s.geometry = g_cube;
          else if (tok[i + 1] == "sphere") s.geometry = g_sphe;
          else if (tok[i + 1] == "point") s.geometry = g_point;
        }
      }
      number_of_states++;
    }
  }
  fclose(f);
  if (number_of_states == 0)
    throw OPSException(OPS_RUNTIME_CONFIGURATION_ERROR, "cloverleaf3d: no states defined");
  field.x_min = field.y_min = field.z_min = 2;
  field.x_max = grid.x_cells + 2;
  field.y_max = grid.y_cells + 2;
  field.z_max = grid.z_cells + 2;
  field.left = field.bottom = field.back = 0;
  // states are shrunk by 1/100 of a cell (float cell size, as read_input.cpp:304-315)
  float dx = (grid.xmax - grid.xmin) / (float)(grid.x_cells);
  float dy = (grid.ymax - grid.ymin) / (float)(grid.y_cells);
  float dz = (grid.zmax - grid.zmin) / (float)(grid.z_cells);
  for (int i = 0; i < number_of_states; i++) {
    states[i].xmin = states[i].xmin + (dx / 100.00);
    states[i].ymin = states[i].ymin + (dy / 100.00);
    states[i].zmin = states[i].zmin + (dz / 100.00);
    states[i].xmax = states[i].xmax - (dx / 100.00);
    states[i].ymax = states[i].ymax - (dy / 100.00);
    states[i].zmax = states[i].zmax - (dz / 100.00);
  }
}

// ---------------------------------------------------------------- declarations

ops_stencil sten(std::vector<int> pts, const char *name) {
  return ops_decl_stencil(3, (int)pts.size() / 3, pts.data(), (char *)name);
}
ops_stencil ssten(std::vector<int> pts, int sx, int sy, int sz, const char *name) {
  int st[3] = {sx, sy, sz};
  return ops_decl_strided_stencil(3, (int)pts.size() / 3, pts.data(), st, (char *)name);
}

// the 8 points (0|s0) x (0|s1) x (0|s2) of a cell or node corner stencil
ops_stencil corner8(int s0, int s1, int s2, const char *name) {
  std::vector<int> p;
  for (int k = 0; k < 2; k++)
    for (int j = 0; j < 2; j++)
      for (int i = 0; i < 2; i++) p.insert(p.end(), {i * s0, j * s1, k * s2});
  return sten(p, name);
}

void build_field() {
  const int nx = grid.x_cells, ny = grid.y_cells, nz = grid.z_cells;
  S.blk = ops_decl_block(3, "clover grid");
  double *none = NULL;
  int *inone = NULL;
  int base[3] = {0, 0, 0}, dm[3] = {-2, -2, -2}, dp[3] = {2, 2, 2};
  auto dat = [&](int sx, int sy, int sz, const char *name) {
    int size[3] = {sx, sy, sz};
    return ops_decl_dat(S.blk, 1, size, base, dm, dp, none, "double", name);
  };
  S.density0 = dat(nx + 5, ny + 5, nz + 5, "density0");
  S.density1 = dat(nx + 5, ny + 5, nz + 5, "density1");
  S.energy0 = dat(nx + 5, ny + 5, nz + 5, "energy0");
  S.energy1 = dat(nx + 5, ny + 5, nz + 5, "energy1");
  S.pressure = dat(nx + 5, ny + 5, nz + 5, "pressure");
  S.viscosity = dat(nx + 5, ny + 5, nz + 5, "viscosity");
  S.soundspeed = dat(nx + 5, ny + 5, nz + 5, "soundspeed");
  S.volume = dat(nx + 5, ny + 5, nz + 5, "volume");
  S.xvel0 = dat(nx + 6, ny + 6, nz + 6, "xvel0");
  S.xvel1 = dat(nx + 6, ny + 6, nz + 6, "xvel1");
  S.yvel0 = dat(nx + 6, ny + 6, nz + 6, "yvel0");
  S.yvel1 = dat(nx + 6, ny + 6, nz + 6, "yvel1");
  S.zvel0 = dat(nx + 6, ny + 6, nz + 6, "zvel0");
  S.zvel1 = dat(nx + 6, ny + 6, nz + 6, "zvel1");
  S.work1 = dat(nx + 6, ny + 6, nz + 6, "work_array1");
  S.work2 = dat(nx + 6, ny + 6, nz + 6, "work_array2");
  S.work3 = dat(nx + 6, ny + 6, nz + 6, "work_array3");
  S.work4 = dat(nx + 6, ny + 6, nz + 6, "work_array4");
  S.work5 = dat(nx + 6, ny + 6, nz + 6, "work_array5");
  S.work6 = dat(nx + 6, ny + 6, nz + 6, "work_array6");
  S.work7 = dat(nx + 6, ny + 6, nz + 6, "work_array7");
  // face-centred data share the cell size in 3D (build_field.cpp:131-146)
  S.vol_flux_x = dat(nx + 5, ny + 5, nz + 5, "vol_flux_x");
  S.mass_flux_x = dat(nx + 5, ny + 5, nz + 5, "mass_flux_x");
  S.xarea = dat(nx + 5, ny + 5, nz + 5, "xarea");
  S.vol_flux_y = dat(nx + 5, ny + 5, nz + 5, "vol_flux_y");
  S.mass_flux_y = dat(nx + 5, ny + 5, nz + 5, "mass_flux_y");
  S.yarea = dat(nx + 5, ny + 5, nz + 5, "yarea");
  S.vol_flux_z = dat(nx + 5, ny + 5, nz + 5, "vol_flux_z");
  S.mass_flux_z = dat(nx + 5, ny + 5, nz + 5, "mass_flux_z");
  S.zarea = dat(nx + 5, ny + 5, nz + 5, "zarea");
  // one-dimensional (edge) dats along x, y and z
  auto edat = [&](int axis, int n, const char *name, const char *type) {
    int size[3] = {1, 1, 1}, m[3] = {0, 0, 0}, p[3] = {0, 0, 0};
    size[axis] = n;
    m[axis] = -2;
    p[axis] = 2;
    if (!strcmp(type, "int")) return ops_decl_dat(S.blk, 1, size, base, m, p, inone, "int", name);
    return ops_decl_dat(S.blk, 1, size, base, m, p, none, "double", name);
  };
  S.cellx = edat(0, nx + 5, "cellx", "double");
  S.celldx = edat(0, nx + 5, "celldx", "double");
  S.celly = edat(1, ny + 5, "celly", "double");
  S.celldy = edat(1, ny + 5, "celldy", "double");
  S.cellz = edat(2, nz + 5, "cellz", "double");
  S.celldz = edat(2, nz + 5, "celldz", "double");
  S.vertexx = edat(0, nx + 6, "vertexx", "double");
  S.vertexdx = edat(0, nx + 6, "vertexdx", "double");
  S.vertexy = edat(1, ny + 6, "vertexy", "double");
  S.vertexdy = edat(1, ny + 6, "vertexdy", "double");
  S.vertexz = edat(2, nz + 6, "vertexz", "double");
  S.vertexdz = edat(2, nz + 6, "vertexdz", "double");
  S.xx = edat(0, nx + 6, "xx", "int");
  S.yy = edat(1, ny + 6, "yy", "int");
  S.zz = edat(2, nz + 6, "zz", "int");

  S.S000 = sten({0, 0, 0}, "000");
  S.P100 = sten({0, 0, 0, 1, 0, 0}, "000:100");
  S.P010 = sten({0, 0, 0, 0, 1, 0}, "000:010");
  S.P001 = sten({0, 0, 0, 0, 0, 1}, "000:001");
  S.M100 = sten({0, 0, 0, -1, 0, 0}, "000:-100");
  S.M010 = sten({0, 0, 0, 0, -1, 0}, "000:0-10");
  S.M001 = sten({0, 0, 0, 0, 0, -1}, "000:00-1");
  S.S6 = sten({0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1}, "7 point");
  S.F_P1P1P1 = corner8(1, 1, 1, "corners +x+y+z");
  S.F_M1M1M1 = corner8(-1, -1, -1, "corners -x-y-z");
  S.F_0M1M1 = corner8(0, -1, -1, "corners -y-z");
  S.F_M10M1 = corner8(-1, 0, -1, "corners -x-z");
  S.F_M1M10 = corner8(-1, -1, 0, "corners -x-y");
  S.F_0P1P1 = corner8(0, 1, 1, "corners +y+z");
  S.F_P10P1 = corner8(1, 0, 1, "corners +x+z");
  S.F_P1P10 = corner8(1, 1, 0, "corners +x+y");
  S.F_P1M1M1 = corner8(1, -1, -1, "corners +x-y-z");
  S.F_M1P1M1 = corner8(-1, 1, -1, "corners -x+y-z");
  S.F_M1M1P1 = corner8(-1, -1, 1, "corners -x-y+z");
  S.X_UP2 = sten({0, 0, 0, 1, 0, 0, 2, 0, 0, -1, 0, 0}, "x 0:1:2:-1");
  S.Y_UP2 = sten({0, 0, 0, 0, 1, 0, 0, 2, 0, 0, -1, 0}, "y 0:1:2:-1");
  S.Z_UP2 = sten({0, 0, 0, 0, 0, 1, 0, 0, 2, 0, 0, -1}, "z 0:1:2:-1");
  S.X_DN2 = sten({0, 0, 0, 1, 0, 0, -1, 0, 0, -2, 0, 0}, "x 0:1:-1:-2");
  S.Y_DN2 = sten({0, 0, 0, 0, 1, 0, 0, -1, 0, 0, -2, 0}, "y 0:1:-1:-2");
  S.Z_DN2 = sten({0, 0, 0, 0, 0, 1, 0, 0, -1, 0, 0, -2}, "z 0:1:-1:-2");
  S.SX = ssten({0, 0, 0}, 1, 0, 0, "000 stride x");
  S.SY = ssten({0, 0, 0}, 0, 1, 0, "000 stride y");
  S.SZ = ssten({0, 0, 0}, 0, 0, 1, "000 stride z");
  S.SX_P = ssten({0, 0, 0, 1, 0, 0}, 1, 0, 0, "000:100 stride x");
  S.SY_P = ssten({0, 0, 0, 0, 1, 0}, 0, 1, 0, "000:010 stride y");
  S.SZ_P = ssten({0, 0, 0, 0, 0, 1}, 0, 0, 1, "000:001 stride z");
  S.SX_PM = ssten({0, 0, 0, 1, 0, 0, -1, 0, 0}, 1, 0, 0, "000:100:-100 stride x");
  S.SY_PM = ssten({0, 0, 0, 0, 1, 0, 0, -1, 0}, 0, 1, 0, "000:010:0-10 stride y");
  S.SZ_PM = ssten({0, 0, 0, 0, 0, 1, 0, 0, -1}, 0, 0, 1, "000:001:00-1 stride z");
  for (int f = 0; f < 6; f++) {
    const int axis = f == 0 || f == 1 ? 1 : (f == 2 || f == 3 ? 0 : 2);
    const int sgn = (f % 2 == 0) ? 1 : -1;  // a low face mirrors inward (+), a high one (-)
    std::vector<int> p;
    for (int o = 0; o <= 4; o++) {
      int pt[3] = {0, 0, 0};
      pt[axis] = sgn * o;
      p.insert(p.end(), pt, pt + 3);
    }
    S.H[f] = sten(p, "halo mirror");
  }
  S.r_dt = ops_decl_reduction_handle(sizeof(double), "double", "local_dt");
  S.r_vol = ops_decl_reduction_handle(sizeof(double), "double", "vol");
  S.r_mass = ops_decl_reduction_handle(sizeof(double), "double", "mass");
  S.r_ie = ops_decl_reduction_handle(sizeof(double), "double", "ie");
  S.r_ke = ops_decl_reduction_handle(sizeof(double), "double", "ke");
  S.r_press = ops_decl_reduction_handle(sizeof(double), "double", "press");
}

// ---------------------------------------------------------------- set-up loops

void initialise_chunk() {
  const int x0 = field.x_min, x1 = field.x_max, y0 = field.y_min, y1 = field.y_max;
  const int z0 = field.z_min, z1 = field.z_max;
  int full[] = {-2, grid.x_cells + 8, -2, grid.y_cells + 8, -2, grid.z_cells + 8};
  ops_par_loop(c3_xx, "initialise_chunk_kernel_xx", S.blk, 3, full,
               ops_arg_dat(S.xx, 1, S.SX, "int", OPS_WRITE), ops_arg_idx());
  ops_par_loop(c3_yy, "initialise_chunk_kernel_yy", S.blk, 3, full,
               ops_arg_dat(S.yy, 1, S.SY, "int", OPS_WRITE), ops_arg_idx());
  ops_par_loop(c3_zz, "initialise_chunk_kernel_zz", S.blk, 3, full,
               ops_arg_dat(S.zz, 1, S.SZ, "int", OPS_WRITE), ops_arg_idx());
  int vr[] = {x0 - 2, x1 + 3, y0 - 2, y1 + 3, z0 - 2, z1 + 3};
  ops_par_loop(c3_vertex_x, "initialise_chunk_kernel_x", S.blk, 3, vr,
               ops_arg_dat(S.vertexx, 1, S.SX, "double", OPS_WRITE),
               ops_arg_dat(S.xx, 1, S.SX, "int", OPS_READ),
               ops_arg_dat(S.vertexdx, 1, S.SX, "double", OPS_WRITE));
  ops_par_loop(c3_vertex_y, "initialise_chunk_kernel_y", S.blk, 3, vr,
               ops_arg_dat(S.vertexy, 1, S.SY, "double", OPS_WRITE),
               ops_arg_dat(S.yy, 1, S.SY, "int", OPS_READ),
               ops_arg_dat(S.vertexdy, 1, S.SY, "double", OPS_WRITE));
  ops_par_loop(c3_vertex_z, "initialise_chunk_kernel_z", S.blk, 3, vr,
               ops_arg_dat(S.vertexz, 1, S.SZ, "double", OPS_WRITE),
               ops_arg_dat(S.zz, 1, S.SZ, "int", OPS_READ),
               ops_arg_dat(S.vertexdz, 1, S.SZ, "double", OPS_WRITE));
  int cx[] = {x0 - 2, x1 + 2, y0 - 2, y1 + 3, z0 - 2, z1 + 3};
  ops_par_loop(c3_cell_x, "initialise_chunk_kernel_cellx", S.blk, 3, cx,
               ops_arg_dat(S.vertexx, 1, S.SX_P, "double", OPS_READ),
               ops_arg_dat(S.cellx, 1, S.SX, "double", OPS_WRITE),
               ops_arg_dat(S.celldx, 1, S.SX, "double", OPS_WRITE));
  int cy[] = {x0 - 2, x1 + 3, y0 - 2, y1 + 2, z0 - 2, z1 + 3};
  ops_par_loop(c3_cell_y, "initialise_chunk_kernel_celly", S.blk, 3, cy,
               ops_arg_dat(S.vertexy, 1, S.SY_P, "double", OPS_READ),
               ops_arg_dat(S.celly, 1, S.SY, "double", OPS_WRITE),
               ops_arg_dat(S.celldy, 1, S.SY, "double", OPS_WRITE));
  int cz[] = {x0 - 2, x1 + 3, y0 - 2, y1 + 3, z0 - 2, z1 + 2};
  ops_par_loop(c3_cell_z, "initialise_chunk_kernel_cellz", S.blk, 3, cz,
               ops_arg_dat(S.vertexz, 1, S.SZ_P, "double", OPS_READ),
               ops_arg_dat(S.cellz, 1, S.SZ, "double", OPS_WRITE),
               ops_arg_dat(S.celldz, 1, S.SZ, "double", OPS_WRITE));
  int all[] = {x0 - 2, x1 + 2, y0 - 2, y1 + 2, z0 - 2, z1 + 2};
  ops_par_loop(c3_volume, "initialise_chunk_kernel_volume", S.blk, 3, all,
               ops_arg_dat(S.volume, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.celldy, 1, S.SY, "double", OPS_READ),
               ops_arg_dat(S.xarea, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.celldx, 1, S.SX, "double", OPS_READ),
               ops_arg_dat(S.yarea, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.celldz, 1, S.SZ, "double", OPS_READ),
               ops_arg_dat(S.zarea, 1, S.S000, "double", OPS_WRITE));
}

void generate() {
  int all[] = {field.x_min - 2, field.x_max + 2, field.y_min - 2, field.y_max + 2,
               field.z_min - 2, field.z_max + 2};
  ops_par_loop(c3_generate, "generate_chunk_kernel", S.blk, 3, all,
               ops_arg_dat(S.vertexx, 1, S.SX_PM, "double", OPS_READ),
               ops_arg_dat(S.vertexy, 1, S.SY_PM, "double", OPS_READ),
               ops_arg_dat(S.vertexz, 1, S.SZ_PM, "double", OPS_READ),
               ops_arg_dat(S.energy0, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.density0, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.xvel0, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.yvel0, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.zvel0, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.cellx, 1, S.SX_PM, "double", OPS_READ),
               ops_arg_dat(S.celly, 1, S.SY_PM, "double", OPS_READ),
               ops_arg_dat(S.cellz, 1, S.SZ_PM, "double", OPS_READ));
}

// ---------------------------------------------------------------- physics loops

int *inner_range(int *r, int px = 0, int py = 0, int pz = 0) {
  r[0] = field.x_min; r[1] = field.x_max + px;
  r[2] = field.y_min; r[3] = field.y_max + py;
  r[4] = field.z_min; r[5] = field.z_max + pz;
  return r;
}

void ideal_gas(int predict) {
  int in[6];
  ops_dat r = predict ? S.density1 : S.density0, e = predict ? S.energy1 : S.energy0;
  ops_par_loop(c3_ideal_gas, "ideal_gas_kernel", S.blk, 3, inner_range(in),
               ops_arg_dat(r, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(e, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.pressure, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.soundspeed, 1, S.S000, "double", OPS_WRITE));
}

bool any(const int *f, std::initializer_list<int> ids) {
  for (int i : ids)
    if (f[i]) return true;
  return false;
}

void set_fields(int *f, std::initializer_list<int> on) {
  for (int i = 0; i < NUM_FIELDS; i++) f[i] = 0;
  for (int i : on) f[i] = 1;
}

// one face of a halo group: mirror parameters and the loop range
struct Face {
  int m[6];
  int r[6];
};

// reflective boundary update of the flagged fields, `depth` halo layers. Faces run in the
// reference's order (bottom, top, left, right, back, front) within each field group; the
// ranges across the face follow apps/c/CloverLeaf_3D/update_halo.cpp per group.
void update_halo(int *fl, int depth) {
  const int d = depth;
  const int x0 = field.x_min, x1 = field.x_max, y0 = field.y_min, y1 = field.y_max;
  const int z0 = field.z_min, z1 = field.z_max;
  // the two faces of an axis are independent when the mesh has more than 2*d+1 cells
  // along it (a face's mirror source then stays clear of the other face's halo layers);
  // on thinner meshes they run one after the other. Faces 0-1: y, 2-3: x, 4-5: z.
  const bool batch_axis[3] = {y1 - y0 > 2 * d + 1, x1 - x0 > 2 * d + 1, z1 - z0 > 2 * d + 1};
  if (any(fl, {FIELD_DENSITY0, FIELD_DENSITY1, FIELD_ENERGY0, FIELD_ENERGY1, FIELD_PRESSURE,
               FIELD_VISCOSITY, FIELD_SOUNDSPEED})) {
    // cell centres: layer h mirrors h' = 2*pivot - 1 - h
    Face faces[6] = {
        {{1, y0, 1}, {x0 - d, x1 + d, y0 - d, y0, z0 - d, z1 + d}},
        {{1, y1, 1}, {x0 - d, x1 + d, y1, y1 + d, z0 - d, z1 + d}},
        {{0, x0, 1}, {x0 - d, x0, y0 - d, y1 + d, z0 - d, z1 + d}},
        {{0, x1, 1}, {x1, x1 + d, y0 - d, y1 + d, z0 - d, z1 + d}},
        {{2, z0, 1}, {x0 - d, x1 + d, y0 - d, y1 + d, z0 - d, z0}},
        {{2, z1, 1}, {x0 - d, x1 + d, y0 - d, y1 + d, z1, z1 + d}}};
    for (int s = 0; s < 6; s++) {
      // one launch per axis, axes in the reference's order
      if (s % 2 == 0 && batch_axis[s / 2]) ops_hip_batch_begin();
      ops_stencil st = S.H[s];
      ops_par_loop(c3_halo_cells, "update_halo_kernel1", S.blk, 3, faces[s].r,
                   ops_arg_dat_opt(S.density0, 1, st, "double", OPS_RW, fl[FIELD_DENSITY0]),
                   ops_arg_dat_opt(S.density1, 1, st, "double", OPS_RW, fl[FIELD_DENSITY1]),
                   ops_arg_dat_opt(S.energy0, 1, st, "double", OPS_RW, fl[FIELD_ENERGY0]),
                   ops_arg_dat_opt(S.energy1, 1, st, "double", OPS_RW, fl[FIELD_ENERGY1]),
                   ops_arg_dat_opt(S.pressure, 1, st, "double", OPS_RW, fl[FIELD_PRESSURE]),
                   ops_arg_dat_opt(S.viscosity, 1, st, "double", OPS_RW, fl[FIELD_VISCOSITY]),
                   ops_arg_dat_opt(S.soundspeed, 1, st, "double", OPS_RW, fl[FIELD_SOUNDSPEED]),
                   ops_arg_gbl(fl, NUM_FIELDS, "int", OPS_READ),
                   ops_arg_gbl(faces[s].m, 6, "int", OPS_READ), ops_arg_idx());
      if (s % 2 == 1) ops_hip_batch_end();
    }
  }
  if (any(fl, {FIELD_XVEL0, FIELD_XVEL1, FIELD_YVEL0, FIELD_YVEL1, FIELD_ZVEL0, FIELD_ZVEL1})) {
    // nodes: layer h mirrors 2*pivot - h; the component normal to the face changes sign
    Face faces[6] = {
        {{1, y0, 0, 1, -1, 1}, {x0 - d, x1 + 1 + d, y0 - d, y0, z0 - d, z1 + 1 + d}},
        {{1, y1, 0, 1, -1, 1}, {x0 - d, x1 + 1 + d, y1 + 1, y1 + 1 + d, z0 - d, z1 + 1 + d}},
        {{0, x0, 0, -1, 1, 1}, {x0 - d, x0, y0 - d, y1 + 1 + d, z0 - d, z1 + 1 + d}},
        {{0, x1, 0, -1, 1, 1}, {x1 + 1, x1 + 1 + d, y0 - d, y1 + 1 + d, z0 - d, z1 + 1 + d}},
        {{2, z0, 0, 1, 1, -1}, {x0 - d, x1 + 1 + d, y0 - d, y1 + 1 + d, z0 - d, z0}},
        {{2, z1, 0, 1, 1, -1}, {x0 - d, x1 + 1 + d, y0 - d, y1 + 1 + d, z1 + 1, z1 + 1 + d}}};
    for (int s = 0; s < 6; s++) {
      if (s % 2 == 0 && batch_axis[s / 2]) ops_hip_batch_begin();
      ops_stencil st = S.H[s];
      ops_par_loop(c3_halo_vel, "update_halo_kernel2_vel", S.blk, 3, faces[s].r,
                   ops_arg_dat_opt(S.xvel0, 1, st, "double", OPS_RW, fl[FIELD_XVEL0]),
                   ops_arg_dat_opt(S.xvel1, 1, st, "double", OPS_RW, fl[FIELD_XVEL1]),
                   ops_arg_dat_opt(S.yvel0, 1, st, "double", OPS_RW, fl[FIELD_YVEL0]),
                   ops_arg_dat_opt(S.yvel1, 1, st, "double", OPS_RW, fl[FIELD_YVEL1]),
                   ops_arg_dat_opt(S.zvel0, 1, st, "double", OPS_RW, fl[FIELD_ZVEL0]),
                   ops_arg_dat_opt(S.zvel1, 1, st, "double", OPS_RW, fl[FIELD_ZVEL1]),
                   ops_arg_gbl(fl, NUM_FIELDS, "int", OPS_READ),
                   ops_arg_gbl(faces[s].m, 6, "int", OPS_READ), ops_arg_idx());
      if (s % 2 == 1) ops_hip_batch_end();
    }
  }
  // face fluxes: layer h mirrors 2*pivot - h; the flux normal to the face changes sign
  struct Group {
    int vf, mf;
    ops_dat v, m;
    Face faces[6];
  };
  const int VX = FIELD_VOL_FLUX_X, MX = FIELD_MASS_FLUX_X, VY = FIELD_VOL_FLUX_Y;
  const int MY = FIELD_MASS_FLUX_Y, VZ = FIELD_VOL_FLUX_Z, MZ = FIELD_MASS_FLUX_Z;
  Group groups[3] = {
      {VX, MX, S.vol_flux_x, S.mass_flux_x,
       {{{1, y0, 0, 1, VX, MX}, {x0 - d, x1 + 1 + d, y0 - d, y0, z0 - d, z1 + d}},
        {{1, y1 - 1, 0, 1, VX, MX}, {x0 - d, x1 + 1 + d, y1, y1 + d, z0 - d, z1 + d}},
        {{0, x0, 0, -1, VX, MX}, {x0 - d, x0, y0 - d, y1 + d, z0 - d, z1 + d}},
        {{0, x1, 0, -1, VX, MX}, {x1 + 1, x1 + 1 + d, y0 - d, y1 + d, z0 - d, z1 + d}},
        {{2, z0, 0, 1, VX, MX}, {x0 - d, x1 + d, y0 - d, y1 + d, z0 - d, z0}},
        {{2, z1 - 1, 0, 1, VX, MX}, {x0 - d, x1 + d, y0 - d, y1 + d, z1, z1 + d}}}},
      {VY, MY, S.vol_flux_y, S.mass_flux_y,
       {{{1, y0, 0, -1, VY, MY}, {x0 - d, x1 + d, y0 - d, y0, z0 - d, z1 + d}},
        {{1, y1, 0, -1, VY, MY}, {x0 - d, x1 + d, y1 + 1, y1 + 1 + d, z0 - d, z1 + d}},
        {{0, x0, 0, 1, VY, MY}, {x0 - d, x0, y0 - d, y1 + 1 + d, z0 - d, z1 + d}},
        {{0, x1 - 1, 0, 1, VY, MY}, {x1, x1 + d, y0 - d, y1 + 1 + d, z0 - d, z1 + d}},
        {{2, z0, 0, 1, VY, MY}, {x0 - d, x1 + d, y0 - d, y1 + d, z0 - d, z0}},
        {{2, z1 - 1, 0, 1, VY, MY}, {x0 - d, x1 + d, y0 - d, y1 + d, z1, z1 + d}}}},
      {VZ, MZ, S.vol_flux_z, S.mass_flux_z,
       {{{1, y0, 0, 1, VZ, MZ}, {x0 - d, x1 + d, y0 - d, y0, z0 - d, z1 + 1 + d}},
        {{1, y1 - 1, 0, 1, VZ, MZ}, {x0 - d, x1 + d, y1, y1 + d, z0 - d, z1 + 1 + d}},
        {{0, x0, 0, 1, VZ, MZ}, {x0 - d, x0, y0 - d, y1 + d, z0 - d, z1 + 1 + d}},
        {{0, x1 - 1, 0, 1, VZ, MZ}, {x1, x1 + d, y0 - d, y1 + d, z0 - d, z1 + 1 + d}},
        {{2, z0, 0, -1, VZ, MZ}, {x0 - d, x1 + d, y0 - d, y1 + d, z0 - d, z0}},
        {{2, z1, 0, -1, VZ, MZ}, {x0 - d, x1 + d, y0 - d, y1 + d, z1 + 1, z1 + 1 + d}}}}};
  // the groups touch different dats, so they go axis by axis: per axis one launch covers
  // both faces of every group; each group still sees its faces in the reference's order
  for (int s0 = 0; s0 < 6; s0 += 2) {
    if (batch_axis[s0 / 2]) ops_hip_batch_begin();
    for (Group &g : groups) {
      if (!fl[g.vf] && !fl[g.mf]) continue;
      for (int s = s0; s < s0 + 2; s++) {
        ops_stencil st = S.H[s];
        ops_par_loop(c3_halo_flux, "update_halo_kernel_flux", S.blk, 3, g.faces[s].r,
                     ops_arg_dat_opt(g.v, 1, st, "double", OPS_RW, fl[g.vf]),
                     ops_arg_dat_opt(g.m, 1, st, "double", OPS_RW, fl[g.mf]),
                     ops_arg_gbl(fl, NUM_FIELDS, "int", OPS_READ),
                     ops_arg_gbl(g.faces[s].m, 6, "int", OPS_READ), ops_arg_idx());
      }
    }
    ops_hip_batch_end();
  }
}

void viscosity_loop() {
  int in[6];
  ops_par_loop(c3_viscosity, "viscosity_kernel", S.blk, 3, inner_range(in),
               ops_arg_dat(S.xvel0, 1, S.F_P1P1P1, "double", OPS_READ),
               ops_arg_dat(S.yvel0, 1, S.F_P1P1P1, "double", OPS_READ),
               ops_arg_dat(S.celldx, 1, S.SX_P, "double", OPS_READ),
               ops_arg_dat(S.celldy, 1, S.SY_P, "double", OPS_READ),
               ops_arg_dat(S.pressure, 1, S.S6, "double", OPS_READ),
               ops_arg_dat(S.density0, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.viscosity, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.zvel0, 1, S.F_P1P1P1, "double", OPS_READ),
               ops_arg_dat(S.celldz, 1, S.SZ_P, "double", OPS_READ),
               ops_arg_dat(S.xarea, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.yarea, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.zarea, 1, S.S000, "double", OPS_READ));
}

void calc_dt_loops() {
  int in[6];
  ops_par_loop(c3_calc_dt, "calc_dt_kernel", S.blk, 3, inner_range(in),
               ops_arg_dat(S.celldx, 1, S.SX_P, "double", OPS_READ),
               ops_arg_dat(S.celldy, 1, S.SY_P, "double", OPS_READ),
               ops_arg_dat(S.soundspeed, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.viscosity, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.density0, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.xvel0, 1, S.F_P1P1P1, "double", OPS_READ),
               ops_arg_dat(S.xarea, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.yvel0, 1, S.F_P1P1P1, "double", OPS_READ),
               ops_arg_dat(S.yarea, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.celldz, 1, S.SZ_P, "double", OPS_READ),
               ops_arg_dat(S.zvel0, 1, S.F_P1P1P1, "double", OPS_READ),
               ops_arg_dat(S.zarea, 1, S.S000, "double", OPS_READ));
  ops_par_loop(c3_dt_min, "calc_dt_kernel_min", S.blk, 3, in,
               ops_arg_dat(S.work1, 1, S.S000, "double", OPS_READ),
               ops_arg_reduce(S.r_dt, 1, "double", OPS_MIN));
}

double calc_dt() {
  calc_dt_loops();
  double local_dt;
  ops_reduction_result(S.r_dt, &local_dt);
  local_dt = MIN(local_dt, g_big);
  return local_dt;
}

// device: dt is computed on the GPU from the reduction's device value (captured steps)
void timestep(bool device = false) {
  int f[NUM_FIELDS];
  ideal_gas(0);
  // zvel0 is not refreshed here, as in the reference (time_step.cpp:60-80)
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
  if (dtlp < S.dtmin) throw OPSException(OPS_RUNTIME_ERROR, "cloverleaf3d: timestep below dtmin");
  double dt_temp = g_big;
  if (dtlp <= dt_temp) dt_temp = dtlp;
  dt_temp = MIN(MIN(dt_temp, (S.dtold * S.dtrise)), S.dtmax);
  ops_update_const("dt", 1, "double", &dt_temp);
  dt = dt_temp;
  if (dt < S.dtmin) throw OPSException(OPS_RUNTIME_ERROR, "cloverleaf3d: small timestep");
  S.dtold = dt;
}

void revert() {
  int in[6];
  ops_par_loop(c3_revert, "revert_kernel", S.blk, 3, inner_range(in),
               ops_arg_dat(S.density0, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.density1, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.energy0, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.energy1, 1, S.S000, "double", OPS_WRITE));
}

void pdv(int predict) {
  int in[6];
  inner_range(in);
  if (predict) {
    ops_par_loop(c3_pdv_predict, "PdV_kernel_predict", S.blk, 3, in,
                 ops_arg_dat(S.xarea, 1, S.P100, "double", OPS_READ),
                 ops_arg_dat(S.xvel0, 1, S.F_P1P1P1, "double", OPS_READ),
                 ops_arg_dat(S.yarea, 1, S.P010, "double", OPS_READ),
                 ops_arg_dat(S.yvel0, 1, S.F_P1P1P1, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.pressure, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.density0, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.density1, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.viscosity, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.energy0, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.energy1, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.zarea, 1, S.P001, "double", OPS_READ),
                 ops_arg_dat(S.zvel0, 1, S.F_P1P1P1, "double", OPS_READ));
  } else {
    ops_par_loop(c3_pdv_full, "PdV_kernel_nopredict", S.blk, 3, in,
                 ops_arg_dat(S.xarea, 1, S.P100, "double", OPS_READ),
                 ops_arg_dat(S.xvel0, 1, S.F_P1P1P1, "double", OPS_READ),
                 ops_arg_dat(S.xvel1, 1, S.F_P1P1P1, "double", OPS_READ),
                 ops_arg_dat(S.yarea, 1, S.P010, "double", OPS_READ),
                 ops_arg_dat(S.yvel0, 1, S.F_P1P1P1, "double", OPS_READ),
                 ops_arg_dat(S.yvel1, 1, S.F_P1P1P1, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.pressure, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.density0, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.density1, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.viscosity, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.energy0, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.energy1, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.zarea, 1, S.P001, "double", OPS_READ),
                 ops_arg_dat(S.zvel0, 1, S.F_P1P1P1, "double", OPS_READ),
                 ops_arg_dat(S.zvel1, 1, S.F_P1P1P1, "double", OPS_READ));
  }
  if (predict) {
    ideal_gas(1);
    int f[NUM_FIELDS];
    set_fields(f, {FIELD_PRESSURE});
    update_halo(f, 1);
    revert();
  }
}

void accelerate() {
  int in1[6];
  ops_par_loop(c3_accelerate, "accelerate_kernel", S.blk, 3, inner_range(in1, 1, 1, 1),
               ops_arg_dat(S.density0, 1, S.F_M1M1M1, "double", OPS_READ),
               ops_arg_dat(S.volume, 1, S.F_M1M1M1, "double", OPS_READ),
               ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.xvel0, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.xvel1, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.xarea, 1, S.F_0M1M1, "double", OPS_READ),
               ops_arg_dat(S.pressure, 1, S.F_M1M1M1, "double", OPS_READ),
               ops_arg_dat(S.yvel0, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.yvel1, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.yarea, 1, S.F_M10M1, "double", OPS_READ),
               ops_arg_dat(S.viscosity, 1, S.F_M1M1M1, "double", OPS_READ),
               ops_arg_dat(S.zvel0, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.zvel1, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.zarea, 1, S.F_M1M10, "double", OPS_READ));
}

void flux_calc() {
  int rx[6], ry[6], rz[6];
  ops_par_loop(c3_flux_x, "flux_calc_kernelx", S.blk, 3, inner_range(rx, 1, 0, 0),
               ops_arg_dat(S.vol_flux_x, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.xarea, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.xvel0, 1, S.F_0P1P1, "double", OPS_READ),
               ops_arg_dat(S.xvel1, 1, S.F_0P1P1, "double", OPS_READ));
  ops_par_loop(c3_flux_y, "flux_calc_kernely", S.blk, 3, inner_range(ry, 0, 1, 0),
               ops_arg_dat(S.vol_flux_y, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.yarea, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.yvel0, 1, S.F_P10P1, "double", OPS_READ),
               ops_arg_dat(S.yvel1, 1, S.F_P10P1, "double", OPS_READ));
  ops_par_loop(c3_flux_z, "flux_calc_kernelz", S.blk, 3, inner_range(rz, 0, 0, 1),
               ops_arg_dat(S.vol_flux_z, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.zarea, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.zvel0, 1, S.F_P1P10, "double", OPS_READ),
               ops_arg_dat(S.zvel1, 1, S.F_P1P10, "double", OPS_READ));
}

void advec_cell(int sweep, int dir) {
  const int x0 = field.x_min, x1 = field.x_max, y0 = field.y_min, y1 = field.y_max;
  const int z0 = field.z_min, z1 = field.z_max;
  int all[] = {x0 - 2, x1 + 2, y0 - 2, y1 + 2, z0 - 2, z1 + 2};
  int in[] = {x0, x1, y0, y1, z0, z1};
  if (dir == XDIR) {
    if (sweep == 1)
      ops_par_loop(c3_cell_vol_1x, "advec_cell_kernel1_xdir", S.blk, 3, all,
                   ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.work2, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_x, 1, S.P100, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_y, 1, S.P010, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_z, 1, S.P001, "double", OPS_READ));
    else if (sweep == 3)
      ops_par_loop(c3_cell_vol_2x, "advec_cell_kernel2_xdir", S.blk, 3, all,
                   ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.work2, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_x, 1, S.P100, "double", OPS_READ));
    int rf[] = {x0, x1 + 2, y0, y1, z0, z1};
    ops_par_loop(c3_cell_flux_x, "advec_cell_kernel3_xdir", S.blk, 3, rf,
                 ops_arg_dat(S.vol_flux_x, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.M100, "double", OPS_READ),
                 ops_arg_dat(S.xx, 1, S.SX_P, "int", OPS_READ),
                 ops_arg_dat(S.vertexdx, 1, S.SX_PM, "double", OPS_READ),
                 ops_arg_dat(S.density1, 1, S.X_DN2, "double", OPS_READ),
                 ops_arg_dat(S.energy1, 1, S.X_DN2, "double", OPS_READ),
                 ops_arg_dat(S.mass_flux_x, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S000, "double", OPS_WRITE));
    ops_par_loop(c3_cell_update_x, "advec_cell_kernel4_xdir", S.blk, 3, in,
                 ops_arg_dat(S.density1, 1, S.S000, "double", OPS_RW),
                 ops_arg_dat(S.energy1, 1, S.S000, "double", OPS_RW),
                 ops_arg_dat(S.mass_flux_x, 1, S.P100, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_x, 1, S.P100, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work2, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work4, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work5, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work6, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.P100, "double", OPS_READ));
  } else if (dir == YDIR) {
    if (sweep == 2) {
      if (S.advect_x)
        ops_par_loop(c3_cell_vol_1y, "advec_cell_kernel1_ydir", S.blk, 3, all,
                     ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
                     ops_arg_dat(S.work2, 1, S.S000, "double", OPS_WRITE),
                     ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                     ops_arg_dat(S.vol_flux_z, 1, S.P001, "double", OPS_READ),
                     ops_arg_dat(S.vol_flux_y, 1, S.P010, "double", OPS_READ));
      else
        ops_par_loop(c3_cell_vol_2y, "advec_cell_kernel2_ydir", S.blk, 3, all,
                     ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
                     ops_arg_dat(S.work2, 1, S.S000, "double", OPS_WRITE),
                     ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                     ops_arg_dat(S.vol_flux_y, 1, S.P010, "double", OPS_READ),
                     ops_arg_dat(S.vol_flux_x, 1, S.P100, "double", OPS_READ));
    }
    // the reference's range also extends z by 2 (advec_cell.cpp:60)
    int rf[] = {x0, x1, y0, y1 + 2, z0, z1 + 2};
    ops_par_loop(c3_cell_flux_y, "advec_cell_kernel3_ydir", S.blk, 3, rf,
                 ops_arg_dat(S.vol_flux_y, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.M010, "double", OPS_READ),
                 ops_arg_dat(S.yy, 1, S.SY_P, "int", OPS_READ),
                 ops_arg_dat(S.vertexdy, 1, S.SY_PM, "double", OPS_READ),
                 ops_arg_dat(S.density1, 1, S.Y_DN2, "double", OPS_READ),
                 ops_arg_dat(S.energy1, 1, S.Y_DN2, "double", OPS_READ),
                 ops_arg_dat(S.mass_flux_y, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S000, "double", OPS_WRITE));
    ops_par_loop(c3_cell_update_y, "advec_cell_kernel4_ydir", S.blk, 3, in,
                 ops_arg_dat(S.density1, 1, S.S000, "double", OPS_RW),
                 ops_arg_dat(S.energy1, 1, S.S000, "double", OPS_RW),
                 ops_arg_dat(S.mass_flux_y, 1, S.P010, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_y, 1, S.P010, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work2, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work4, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work5, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work6, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.P010, "double", OPS_READ));
  } else {
    if (sweep == 1)
      ops_par_loop(c3_cell_vol_1z, "advec_cell_kernel1_zdir", S.blk, 3, all,
                   ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.work2, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_x, 1, S.P100, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_y, 1, S.P010, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_z, 1, S.P001, "double", OPS_READ));
    else if (sweep == 3)
      ops_par_loop(c3_cell_vol_2z, "advec_cell_kernel2_zdir", S.blk, 3, all,
                   ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.work2, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                   ops_arg_dat(S.vol_flux_z, 1, S.P001, "double", OPS_READ));
    int rf[] = {x0, x1, y0, y1, z0, z1 + 2};
    ops_par_loop(c3_cell_flux_z, "advec_cell_kernel3_zdir", S.blk, 3, rf,
                 ops_arg_dat(S.vol_flux_z, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.M001, "double", OPS_READ),
                 ops_arg_dat(S.zz, 1, S.SZ_P, "int", OPS_READ),
                 ops_arg_dat(S.vertexdz, 1, S.SZ_PM, "double", OPS_READ),
                 ops_arg_dat(S.density1, 1, S.Z_DN2, "double", OPS_READ),
                 ops_arg_dat(S.energy1, 1, S.Z_DN2, "double", OPS_READ),
                 ops_arg_dat(S.mass_flux_z, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S000, "double", OPS_WRITE));
    ops_par_loop(c3_cell_update_z, "advec_cell_kernel4_zdir", S.blk, 3, in,
                 ops_arg_dat(S.density1, 1, S.S000, "double", OPS_RW),
                 ops_arg_dat(S.energy1, 1, S.S000, "double", OPS_RW),
                 ops_arg_dat(S.mass_flux_z, 1, S.P001, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_z, 1, S.P001, "double", OPS_READ),
                 ops_arg_dat(S.work1, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work2, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work4, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work5, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work6, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.P001, "double", OPS_READ));
  }
}

void advec_mom(int which_vel, int sweep, int dir) {
  const int x0 = field.x_min, x1 = field.x_max, y0 = field.y_min, y1 = field.y_max;
  const int z0 = field.z_min, z1 = field.z_max;
  int all[] = {x0 - 2, x1 + 2, y0 - 2, y1 + 2, z0 - 2, z1 + 2};
  ops_dat vel = which_vel == 1 ? S.xvel1 : (which_vel == 2 ? S.yvel1 : S.zvel1);
  if (sweep == 1 && dir == XDIR)
    ops_par_loop(c3_mom_vol_x1, "advec_mom_kernel_x1", S.blk, 3, all,
                 ops_arg_dat(S.work6, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_x, 1, S.P100, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_y, 1, S.P010, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_z, 1, S.P001, "double", OPS_READ));
  else if (sweep == 1 && dir == ZDIR)
    ops_par_loop(c3_mom_vol_z1, "advec_mom_kernel_z1", S.blk, 3, all,
                 ops_arg_dat(S.work6, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_x, 1, S.P100, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_y, 1, S.P010, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_z, 1, S.P001, "double", OPS_READ));
  else if (sweep == 2 && S.advect_x)
    ops_par_loop(c3_mom_vol_x2, "advec_mom_kernel_x2", S.blk, 3, all,
                 ops_arg_dat(S.work6, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_y, 1, S.P010, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_z, 1, S.P001, "double", OPS_READ));
  else if (sweep == 2)
    ops_par_loop(c3_mom_vol_y2, "advec_mom_kernel_y2", S.blk, 3, all,
                 ops_arg_dat(S.work6, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_x, 1, S.P100, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_y, 1, S.P010, "double", OPS_READ));
  else if (sweep == 3 && dir == XDIR)
    ops_par_loop(c3_mom_vol_x3, "advec_mom_kernel_x3", S.blk, 3, all,
                 ops_arg_dat(S.work6, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_x, 1, S.P100, "double", OPS_READ));
  else if (sweep == 3 && dir == ZDIR)
    ops_par_loop(c3_mom_vol_z3, "advec_mom_kernel_z3", S.blk, 3, all,
                 ops_arg_dat(S.work6, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.work7, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.vol_flux_z, 1, S.P001, "double", OPS_READ));
  int r_upd[] = {x0, x1 + 1, y0, y1 + 1, z0, z1 + 1};
  if (dir == XDIR) {
    if (which_vel == 1) {
      int r1[] = {x0 - 2, x1 + 2, y0, y1 + 1, z0, z1 + 1};
      ops_par_loop(c3_node_flux_x, "advec_mom_kernel_mass_flux_x", S.blk, 3, r1,
                   ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.mass_flux_x, 1, S.F_P1M1M1, "double", OPS_READ));
      int r2[] = {x0 - 1, x1 + 2, y0, y1 + 1, z0, z1 + 1};
      ops_par_loop(c3_node_mass_x, "advec_mom_kernel_post_pre_advec_x", S.blk, 3, r2,
                   ops_arg_dat(S.work2, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.work7, 1, S.F_M1M1M1, "double", OPS_READ),
                   ops_arg_dat(S.density1, 1, S.F_M1M1M1, "double", OPS_READ),
                   ops_arg_dat(S.work3, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.work1, 1, S.M100, "double", OPS_READ));
    }
    int r3[] = {x0 - 1, x1 + 1, y0, y1 + 1, z0, z1 + 1};
    ops_par_loop(c3_mom_flux_x, "advec_mom_kernel1_x", S.blk, 3, r3,
                 ops_arg_dat(S.work1, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.P100, "double", OPS_READ),
                 ops_arg_dat(S.work5, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.celldx, 1, S.SX_PM, "double", OPS_READ),
                 ops_arg_dat(vel, 1, S.X_UP2, "double", OPS_READ));
    ops_par_loop(c3_mom_update_x, "advec_mom_kernel2_x", S.blk, 3, r_upd,
                 ops_arg_dat(vel, 1, S.S000, "double", OPS_RW),
                 ops_arg_dat(S.work2, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work5, 1, S.M100, "double", OPS_READ));
  } else if (dir == YDIR) {
    if (which_vel == 1) {
      int r1[] = {x0, x1 + 1, y0 - 2, y1 + 2, z0, z1 + 1};
      ops_par_loop(c3_node_flux_y, "advec_mom_kernel_mass_flux_y", S.blk, 3, r1,
                   ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.mass_flux_y, 1, S.F_M1P1M1, "double", OPS_READ));
      int r2[] = {x0, x1 + 1, y0 - 1, y1 + 2, z0, z1 + 1};
      ops_par_loop(c3_node_mass_y, "advec_mom_kernel_post_pre_advec_y", S.blk, 3, r2,
                   ops_arg_dat(S.work2, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.work7, 1, S.F_M1M1M1, "double", OPS_READ),
                   ops_arg_dat(S.density1, 1, S.F_M1M1M1, "double", OPS_READ),
                   ops_arg_dat(S.work3, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.work1, 1, S.M010, "double", OPS_READ));
    }
    int r3[] = {x0, x1 + 1, y0 - 1, y1 + 1, z0, z1 + 1};
    ops_par_loop(c3_mom_flux_y, "advec_mom_kernel1_y", S.blk, 3, r3,
                 ops_arg_dat(S.work1, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.P010, "double", OPS_READ),
                 ops_arg_dat(S.work5, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.celldy, 1, S.SY_PM, "double", OPS_READ),
                 ops_arg_dat(vel, 1, S.Y_UP2, "double", OPS_READ));
    ops_par_loop(c3_mom_update_y, "advec_mom_kernel2_y", S.blk, 3, r_upd,
                 ops_arg_dat(vel, 1, S.S000, "double", OPS_RW),
                 ops_arg_dat(S.work2, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work5, 1, S.M010, "double", OPS_READ));
  } else {
    if (which_vel == 1) {
      int r1[] = {x0, x1 + 1, y0, y1 + 1, z0 - 2, z1 + 2};
      ops_par_loop(c3_node_flux_z, "advec_mom_kernel_mass_flux_z", S.blk, 3, r1,
                   ops_arg_dat(S.work1, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.mass_flux_z, 1, S.F_M1M1P1, "double", OPS_READ));
      int r2[] = {x0, x1 + 1, y0, y1 + 1, z0 - 1, z1 + 2};
      ops_par_loop(c3_node_mass_z, "advec_mom_kernel_post_pre_advec_z", S.blk, 3, r2,
                   ops_arg_dat(S.work2, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.work7, 1, S.F_M1M1M1, "double", OPS_READ),
                   ops_arg_dat(S.density1, 1, S.F_M1M1M1, "double", OPS_READ),
                   ops_arg_dat(S.work3, 1, S.S000, "double", OPS_WRITE),
                   ops_arg_dat(S.work1, 1, S.M001, "double", OPS_READ));
    }
    int r3[] = {x0, x1 + 1, y0, y1 + 1, z0 - 1, z1 + 1};
    ops_par_loop(c3_mom_flux_z, "advec_mom_kernel1_z", S.blk, 3, r3,
                 ops_arg_dat(S.work1, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.P001, "double", OPS_READ),
                 ops_arg_dat(S.work5, 1, S.S000, "double", OPS_WRITE),
                 ops_arg_dat(S.celldz, 1, S.SZ_PM, "double", OPS_READ),
                 ops_arg_dat(vel, 1, S.Z_UP2, "double", OPS_READ));
    ops_par_loop(c3_mom_update_z, "advec_mom_kernel2_z", S.blk, 3, r_upd,
                 ops_arg_dat(vel, 1, S.S000, "double", OPS_RW),
                 ops_arg_dat(S.work2, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work3, 1, S.S000, "double", OPS_READ),
                 ops_arg_dat(S.work5, 1, S.M001, "double", OPS_READ));
  }
}

// three directional sweeps per step: x, y, z when advect_x, else z, y, x
// (advection.cpp:35-181)
void advection() {
  int f[NUM_FIELDS];
  int dir = S.advect_x ? XDIR : ZDIR;
  auto mflux = [](int d) {
    return d == XDIR ? FIELD_MASS_FLUX_X : (d == YDIR ? FIELD_MASS_FLUX_Y : FIELD_MASS_FLUX_Z);
  };
  set_fields(f, {FIELD_DENSITY1, FIELD_ENERGY1, FIELD_VOL_FLUX_X, FIELD_VOL_FLUX_Y,
                 FIELD_VOL_FLUX_Z});
  update_halo(f, 2);
  advec_cell(1, dir);
  set_fields(f, {FIELD_DENSITY1, FIELD_ENERGY1, FIELD_XVEL1, FIELD_YVEL1, FIELD_ZVEL1, mflux(dir)});
  update_halo(f, 2);
  for (int v = 1; v <= 3; v++) advec_mom(v, 1, dir);
  dir = YDIR;
  set_fields(f, {FIELD_DENSITY1, FIELD_ENERGY1, FIELD_VOL_FLUX_X, FIELD_VOL_FLUX_Y,
                 FIELD_VOL_FLUX_Z});
  update_halo(f, 2);
  advec_cell(2, dir);
  set_fields(f, {FIELD_DENSITY1, FIELD_ENERGY1, FIELD_XVEL1, FIELD_YVEL1, FIELD_ZVEL1, mflux(dir)});
  update_halo(f, 2);
  for (int v = 1; v <= 3; v++) advec_mom(v, 2, dir);
  // the third sweep starts without a halo refresh, as in the reference
  dir = S.advect_x ? ZDIR : XDIR;
  advec_cell(3, dir);
  set_fields(f, {FIELD_DENSITY1, FIELD_ENERGY1, FIELD_XVEL1, FIELD_YVEL1, FIELD_ZVEL1, mflux(dir)});
  update_halo(f, 2);
  for (int v = 1; v <= 3; v++) advec_mom(v, 3, dir);
}

void reset_field() {
  int in[6], in1[6];
  ops_par_loop(c3_reset_cells, "reset_field_kernel1", S.blk, 3, inner_range(in),
               ops_arg_dat(S.density0, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.density1, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.energy0, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.energy1, 1, S.S000, "double", OPS_READ));
  ops_par_loop(c3_reset_nodes, "reset_field_kernel2", S.blk, 3, inner_range(in1, 1, 1, 1),
               ops_arg_dat(S.xvel0, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.xvel1, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.yvel0, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.yvel1, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.zvel0, 1, S.S000, "double", OPS_WRITE),
               ops_arg_dat(S.zvel1, 1, S.S000, "double", OPS_READ));
}

void field_summary(double *out) {
  ideal_gas(0);
  int in[6];
  ops_par_loop(c3_field_summary, "field_summary_kernel", S.blk, 3, inner_range(in),
               ops_arg_dat(S.volume, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.density0, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.energy0, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.pressure, 1, S.S000, "double", OPS_READ),
               ops_arg_dat(S.xvel0, 1, S.F_P1P1P1, "double", OPS_READ),
               ops_arg_dat(S.yvel0, 1, S.F_P1P1P1, "double", OPS_READ),
               ops_arg_dat(S.zvel0, 1, S.F_P1P1P1, "double", OPS_READ),
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
  ops_decl_const("dtw_safe", 1, "double", &dtw_safe);
  ops_decl_const("dtdiv_safe", 1, "double", &dtdiv_safe);
  ops_decl_const("field", 1, "field_type", &field);
  ops_decl_const("grid", 1, "grid_type", &grid);
  ops_decl_const("states", number_of_states, "state_type", states);
  ops_decl_const("number_of_states", 1, "int", &number_of_states);
  ops_decl_const("g_sphe", 1, "int", &g_sphe);
  ops_decl_const("g_point", 1, "int", &g_point);
  ops_decl_const("g_cube", 1, "int", &g_cube);
  ops_decl_const("dt", 1, "double", &dt);
  initialise_chunk();
  generate();
  S.advect_x = 1;
  ideal_gas(0);
  int f[NUM_FIELDS];
  set_fields(f, {FIELD_DENSITY0, FIELD_ENERGY0, FIELD_PRESSURE, FIELD_VISCOSITY, FIELD_DENSITY1,
                 FIELD_ENERGY1, FIELD_XVEL0, FIELD_YVEL0, FIELD_ZVEL0, FIELD_XVEL1, FIELD_YVEL1,
                 FIELD_ZVEL1});
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
  advection();
  reset_field();
  S.advect_x = !S.advect_x;
  S.time = S.time + dt;
}

// the launches of one step with dt, dtold and time kept on the device
void device_step() {
  timestep(true);
  pdv(1);
  accelerate();
  pdv(0);
  flux_calc();
  advection();
  reset_field();
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
    ops_hip_reduction_discard(S.r_dt);  // the captured MIN starts from its own result
    if (ops_hip_graph_begin())
      throw OPSException(OPS_RUNTIME_ERROR, "cloverleaf3d: graph capture unavailable");
    try {
      device_step();
    } catch (...) {
      ops_hip_graph_free(ops_hip_graph_end());  // leave capture mode before reporting
      S.advect_x = keep;
      throw;
    }
    S.step_graph[ax] = ops_hip_graph_end();
    ops_hip_reduction_discard(S.r_dt);
    S.advect_x = keep;
    if (!S.step_graph[ax])
      throw OPSException(OPS_RUNTIME_ERROR, "cloverleaf3d: graph capture failed");
  }
}

void graph_steps(int n) {
  capture_steps();
  clover_dt_state_set(S.dt_state, S.dtold, S.time, dt);
  for (int i = 0; i < n; i++) {
    if (ops_hip_graph_replay(S.step_graph[S.advect_x], 1))
      throw OPSException(OPS_RUNTIME_ERROR, "cloverleaf3d: graph replay failed");
    S.advect_x = !S.advect_x;
    S.step++;
  }
  double st[4];
  clover_dt_state_get(S.dt_state, st);
  S.dtold = st[0];
  S.time = st[1];
  dt = st[3];
  if (st[2] != 0.0) throw OPSException(OPS_RUNTIME_ERROR, "cloverleaf3d: timestep below dtmin");
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

extern "C" int cloverleaf3d_init(int argc, const char **argv, const char *deck) {
  try {
    ops_init(argc, argv, 1);
    const int graph = S.graph;
    S = Sim();
    S.graph = graph;
    read_deck(deck);
    start();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf3d_init: %s\n", e.what());
    return 1;
  }
}

// captured steps on (1) or off (0) for cloverleaf3d_step; returns the previous setting
extern "C" int cloverleaf3d_set_graph(int on) {
  const int old = S.graph;
  S.graph = on;
  return old;
}

// captures the two step graphs now (otherwise the first captured step does), so that a
// timed region replays only; 0 on success
extern "C" int cloverleaf3d_capture(void) {
  try {
    capture_steps();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf3d_capture: %s\n", e.what());
    return 1;
  }
}

extern "C" int cloverleaf3d_step(int n) {
  try {
    if (S.graph && n > 0) {
      graph_steps(n);
      return 0;
    }
    for (int i = 0; i < n; i++) hydro_step();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf3d_step: %s\n", e.what());
    return 1;
  }
}

// steps until the deck's end condition (clover_leaf.cpp:190) or max_steps; returns the
// number of steps taken, or -1 on error
extern "C" int cloverleaf3d_run(int max_steps) {
  try {
    int n = 0;
    while (n < max_steps) {
      hydro_step();
      n++;
      if ((S.time + g_small) > S.end_time || (S.step >= S.end_step)) break;
    }
    return n;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf3d_run: %s\n", e.what());
    return -1;
  }
}

extern "C" int cloverleaf3d_summary(double *out5) {
  try {
    field_summary(out5);
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf3d_summary: %s\n", e.what());
    return 1;
  }
}

extern "C" int cloverleaf3d_state(double *out4) {
  out4[0] = S.step;
  out4[1] = S.time;
  out4[2] = dt;
  out4[3] = S.end_step;
  return 0;
}

// copies one dat, halos included (size[0]*size[1]*size[2] values); returns the count, -1 if
// the name is unknown
extern "C" long cloverleaf3d_fetch(const char *name, double *out) {
  try {
    ops_dat all[] = {S.density0, S.density1, S.energy0, S.energy1, S.pressure, S.viscosity,
                     S.soundspeed, S.volume, S.xvel0, S.xvel1, S.yvel0, S.yvel1, S.zvel0,
                     S.zvel1, S.vol_flux_x, S.vol_flux_y, S.vol_flux_z, S.mass_flux_x,
                     S.mass_flux_y, S.mass_flux_z, S.xarea, S.yarea, S.zarea, S.cellx, S.celly,
                     S.cellz, S.vertexx, S.vertexy, S.vertexz, S.celldx, S.celldy, S.celldz};
    for (ops_dat d : all) {
      if (strcmp(d->name, name)) continue;
      long n = (long)d->size[0] * d->size[1] * d->size[2];
      if (out) {
        int range[6] = {d->d_m[0], d->size[0] + d->d_m[0], d->d_m[1], d->size[1] + d->d_m[1],
                        d->d_m[2], d->size[2] + d->d_m[2]};
        ops_dat_fetch_data_slab_memspace(d, 0, (char *)out, range, OPS_HOST);
      }
      return n;
    }
    return -1;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf3d_fetch: %s\n", e.what());
    return -2;
  }
}

extern "C" int cloverleaf3d_finalize(void) {
  try {
    free_graphs();
    ops_exit();
    free(states);
    states = NULL;
    number_of_states = 0;
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "cloverleaf3d_finalize: %s\n", e.what());
    return 1;
  }
}
