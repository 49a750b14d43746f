// CloverLeaf 3D user kernels, written against OPS API v1: each kernel receives pointers at
// the loop's current point and addresses neighbours with OPS_ACC<n>(x, y, z), n being the
// argument's position in the ops_par_loop call.
//
// The arithmetic of every kernel keeps the operation order of the reference application
// (apps/c/CloverLeaf_3D/*_kernel.h): same operand order, same MIN/MAX/SIGN macros, same
// constants, so the FP64 fields can be compared bit for bit. Where the reference repeats
// one formula per direction, the shared part is an inline helper taking the loaded values.
// Constants (dt, g_small, g_big, dt*_safe, field, grid, states, number_of_states,
// g_cube/g_sphe/g_point) are ops_decl_const globals.
#ifndef CLOVER3D_KERNELS_H
#define CLOVER3D_KERNELS_H

#include "user_types.h"

// ------------------------------------------------------------ mesh set-up
// (initialise_chunk_kernel.h)

void c3_xx(int *xx, int *idx) { xx[OPS_ACC0(0, 0, 0)] = idx[0] - 2; }
void c3_yy(int *yy, int *idx) { yy[OPS_ACC0(0, 0, 0)] = idx[1] - 2; }
void c3_zz(int *zz, int *idx) { zz[OPS_ACC0(0, 0, 0)] = idx[2] - 2; }

void c3_vertex_x(double *vx, const int *xx, double *vdx) {
  int lo = field.x_min - 2;
  double d = (grid.xmax - grid.xmin) / (double)grid.x_cells;
  double origin = grid.xmin + d * field.left;
  vx[OPS_ACC0(0, 0, 0)] = origin + d * (xx[OPS_ACC1(0, 0, 0)] - lo);
  vdx[OPS_ACC2(0, 0, 0)] = (double)d;
}

void c3_vertex_y(double *vy, const int *yy, double *vdy) {
  int lo = field.y_min - 2;
  double d = (grid.ymax - grid.ymin) / (double)grid.y_cells;
  double origin = grid.ymin + d * field.bottom;
  vy[OPS_ACC0(0, 0, 0)] = origin + d * (yy[OPS_ACC1(0, 0, 0)] - lo);
  vdy[OPS_ACC2(0, 0, 0)] = (double)d;
}

void c3_vertex_z(double *vz, const int *zz, double *vdz) {
  int lo = field.z_min - 2;
  double d = (grid.zmax - grid.zmin) / (double)grid.z_cells;
  double origin = grid.zmin + d * field.back;
  vz[OPS_ACC0(0, 0, 0)] = origin + d * (zz[OPS_ACC1(0, 0, 0)] - lo);
  vdz[OPS_ACC2(0, 0, 0)] = (double)d;
}

void c3_cell_x(const double *vx, double *cx, double *cdx) {
  double d = (grid.xmax - grid.xmin) / (double)grid.x_cells;
  cx[OPS_ACC1(0, 0, 0)] = 0.5 * (vx[OPS_ACC0(0, 0, 0)] + vx[OPS_ACC0(1, 0, 0)]);
  cdx[OPS_ACC2(0, 0, 0)] = d;
}

void c3_cell_y(const double *vy, double *cy, double *cdy) {
  double d = (grid.ymax - grid.ymin) / (double)grid.y_cells;
  cy[OPS_ACC1(0, 0, 0)] = 0.5 * (vy[OPS_ACC0(0, 0, 0)] + vy[OPS_ACC0(0, 1, 0)]);
  cdy[OPS_ACC2(0, 0, 0)] = d;
}

void c3_cell_z(const double *vz, double *cz, double *cdz) {
  double d = (grid.zmax - grid.zmin) / (double)grid.z_cells;
  cz[OPS_ACC1(0, 0, 0)] = 0.5 * (vz[OPS_ACC0(0, 0, 0)] + vz[OPS_ACC0(0, 0, 1)]);
  cdz[OPS_ACC2(0, 0, 0)] = d;
}

void c3_volume(double *vol, const double *cdy, double *xa, const double *cdx, double *ya,
               const double *cdz, double *za) {
  double dx = (grid.xmax - grid.xmin) / (double)grid.x_cells;
  double dy = (grid.ymax - grid.ymin) / (double)grid.y_cells;
  double dz = (grid.zmax - grid.zmin) / (double)grid.z_cells;
  vol[OPS_ACC0(0, 0, 0)] = dx * dy * dz;
  xa[OPS_ACC2(0, 0, 0)] = cdy[OPS_ACC1(0, 0, 0)] * cdz[OPS_ACC5(0, 0, 0)];
  ya[OPS_ACC4(0, 0, 0)] = cdx[OPS_ACC3(0, 0, 0)] * cdz[OPS_ACC5(0, 0, 0)];
  za[OPS_ACC6(0, 0, 0)] = cdx[OPS_ACC3(0, 0, 0)] * cdy[OPS_ACC1(0, 0, 0)];
}

// ------------------------------------------------------------ initial state
// (generate_chunk_kernel.h): background state 0, then states 1.. painted in order. The
// node flag is shared by all states (set by one state, it stays set for the later ones),
// as in the reference kernel.

void c3_generate(const double *vx, const double *vy, const double *vz, double *e0,
                 double *r0, double *u0, double *v0, double *w0, const double *cx,
                 const double *cy, const double *cz) {
  e0[OPS_ACC3(0, 0, 0)] = states[0].energy;
  r0[OPS_ACC4(0, 0, 0)] = states[0].density;
  u0[OPS_ACC5(0, 0, 0)] = states[0].xvel;
  v0[OPS_ACC6(0, 0, 0)] = states[0].yvel;
  w0[OPS_ACC7(0, 0, 0)] = states[0].zvel;
  int node_in = 0;
  for (int s = 1; s < number_of_states; s++) {
    const state_type &st = states[s];
    int cell_in = 0;
    if (st.geometry == g_cube) {
      for (int i = -1; i <= 0; i++)
        for (int j = -1; j <= 0; j++)
          for (int k = -1; k <= 0; k++)
            if (vx[OPS_ACC0(1 + i, 0, 0)] >= st.xmin && vx[OPS_ACC0(i, 0, 0)] < st.xmax &&
                vy[OPS_ACC1(0, 1 + j, 0)] >= st.ymin && vy[OPS_ACC1(0, j, 0)] < st.ymax &&
                vz[OPS_ACC2(0, 0, 1 + k)] >= st.zmin && vz[OPS_ACC2(0, 0, k)] < st.zmax)
              node_in = 1;
      cell_in = vx[OPS_ACC0(1, 0, 0)] >= st.xmin && vx[OPS_ACC0(0, 0, 0)] < st.xmax &&
                vy[OPS_ACC1(0, 1, 0)] >= st.ymin && vy[OPS_ACC1(0, 0, 0)] < st.ymax &&
                vz[OPS_ACC2(0, 0, 1)] >= st.zmin && vz[OPS_ACC2(0, 0, 0)] < st.zmax;
    } else if (st.geometry == g_sphe) {
      // the distance of the cell centre (the same for every corner of the reference's
      // corner loop)
      const double dx = cx[OPS_ACC8(0, 0, 0)] - st.xmin;
      const double dy = cy[OPS_ACC9(0, 0, 0)] - st.ymin;
      const double dz = cz[OPS_ACC10(0, 0, 0)] - st.zmin;
      const double radius = sqrt(dx * dx + dy * dy + dz * dz);
      if (radius <= st.radius) node_in = 1;
      cell_in = radius <= st.radius;
    } else if (st.geometry == g_point) {
      for (int i = -1; i <= 0; i++)
        for (int j = -1; j <= 0; j++)
          for (int k = -1; k <= 0; k++)
            if (vx[OPS_ACC0(i, 0, 0)] == st.xmin && vy[OPS_ACC1(0, j, 0)] == st.ymin &&
                vz[OPS_ACC2(0, 0, k)] == st.zmin)
              node_in = 1;
      cell_in = vx[OPS_ACC0(0, 0, 0)] == st.xmin && vy[OPS_ACC1(0, 0, 0)] == st.ymin &&
                vz[OPS_ACC2(0, 0, 0)] == st.zmin;
    } else {
      continue;
    }
    if (cell_in) {
      e0[OPS_ACC3(0, 0, 0)] = st.energy;
      r0[OPS_ACC4(0, 0, 0)] = st.density;
    }
    if (node_in) {
      u0[OPS_ACC5(0, 0, 0)] = st.xvel;
      v0[OPS_ACC6(0, 0, 0)] = st.yvel;
      w0[OPS_ACC7(0, 0, 0)] = st.zvel;
    }
  }
}

// ------------------------------------------------------------ equation of state
// (ideal_gas_kernel.h), gamma = 1.4

void c3_ideal_gas(const double *rho, const double *e, double *p, double *cs) {
  const double r = rho[OPS_ACC0(0, 0, 0)];
  double v = 1.0 / r;
  p[OPS_ACC2(0, 0, 0)] = (1.4 - 1.0) * r * e[OPS_ACC1(0, 0, 0)];
  double dpde = (1.4 - 1.0) * r;
  double dpdv = -1.0 * r * p[OPS_ACC2(0, 0, 0)];
  double c2 = v * v * (p[OPS_ACC2(0, 0, 0)] * dpde - dpdv);
  cs[OPS_ACC3(0, 0, 0)] = sqrt(c2);
}

// ------------------------------------------------------------ reflective halos
// One kernel per field group, one launch per face covering all halo layers of the face.
// The mirror source for halo index h on axis m[0] is h + off with
//   off = 2 * (m[1] - h) - m[2]
// (m[1] the pivot plane, m[2] = 1 for cell-centred and 0 for face/node-centred data); the
// copied value is multiplied by the sign of the group (-1 flips the component normal to
// the face). This reproduces the per-layer strip loops of
// apps/c/CloverLeaf_3D/update_halo.cpp.

#define C3_MIRROR(m, idx)                                                                   \
  const int off = 2 * ((m)[1] - (idx)[(m)[0]]) - (m)[2];                                    \
  const int ox = (m)[0] == 0 ? off : 0, oy = (m)[0] == 1 ? off : 0, oz = (m)[0] == 2 ? off : 0

void c3_halo_cells(double *r0, double *r1, double *e0, double *e1, double *p, double *q,
                   double *cs, const int *fields, const int *m, const int *idx) {
  C3_MIRROR(m, idx);
  if (fields[FIELD_DENSITY0] == 1) r0[OPS_ACC0(0, 0, 0)] = r0[OPS_ACC0(ox, oy, oz)];
  if (fields[FIELD_DENSITY1] == 1) r1[OPS_ACC1(0, 0, 0)] = r1[OPS_ACC1(ox, oy, oz)];
  if (fields[FIELD_ENERGY0] == 1) e0[OPS_ACC2(0, 0, 0)] = e0[OPS_ACC2(ox, oy, oz)];
  if (fields[FIELD_ENERGY1] == 1) e1[OPS_ACC3(0, 0, 0)] = e1[OPS_ACC3(ox, oy, oz)];
  if (fields[FIELD_PRESSURE] == 1) p[OPS_ACC4(0, 0, 0)] = p[OPS_ACC4(ox, oy, oz)];
  if (fields[FIELD_VISCOSITY] == 1) q[OPS_ACC5(0, 0, 0)] = q[OPS_ACC5(ox, oy, oz)];
  if (fields[FIELD_SOUNDSPEED] == 1) cs[OPS_ACC6(0, 0, 0)] = cs[OPS_ACC6(ox, oy, oz)];
}

// nodal velocities: m[3], m[4], m[5] are the factors for xvel, yvel and zvel
void c3_halo_vel(double *u0, double *u1, double *v0, double *v1, double *w0, double *w1,
                 const int *fields, const int *m, const int *idx) {
  C3_MIRROR(m, idx);
  if (fields[FIELD_XVEL0] == 1) u0[OPS_ACC0(0, 0, 0)] = m[3] * u0[OPS_ACC0(ox, oy, oz)];
  if (fields[FIELD_XVEL1] == 1) u1[OPS_ACC1(0, 0, 0)] = m[3] * u1[OPS_ACC1(ox, oy, oz)];
  if (fields[FIELD_YVEL0] == 1) v0[OPS_ACC2(0, 0, 0)] = m[4] * v0[OPS_ACC2(ox, oy, oz)];
  if (fields[FIELD_YVEL1] == 1) v1[OPS_ACC3(0, 0, 0)] = m[4] * v1[OPS_ACC3(ox, oy, oz)];
  if (fields[FIELD_ZVEL0] == 1) w0[OPS_ACC4(0, 0, 0)] = m[5] * w0[OPS_ACC4(ox, oy, oz)];
  if (fields[FIELD_ZVEL1] == 1) w1[OPS_ACC5(0, 0, 0)] = m[5] * w1[OPS_ACC5(ox, oy, oz)];
}

// face fluxes: m[3] the factor, m[4]/m[5] the field ids of the (volume, mass) pair
void c3_halo_flux(double *vf, double *mf, const int *fields, const int *m, const int *idx) {
  C3_MIRROR(m, idx);
  if (fields[m[4]] == 1) vf[OPS_ACC0(0, 0, 0)] = m[3] * vf[OPS_ACC0(ox, oy, oz)];
  if (fields[m[5]] == 1) mf[OPS_ACC1(0, 0, 0)] = m[3] * mf[OPS_ACC1(ox, oy, oz)];
}

// ------------------------------------------------------------ artificial viscosity
// (viscosity_kernel.h)

void c3_viscosity(const double *u, const double *v, const double *cdx, const double *cdy,
                  const double *p, const double *rho, double *q, const double *w,
                  const double *cdz, const double *xa, const double *ya, const double *za) {
  // sums of the four nodal values on each face of the cell, per velocity component
#define C3_FACE(a, n, i0, j0, k0, i1, j1, k1, i2, j2, k2, i3, j3, k3)                       \
  (a[OPS_ACC##n(i0, j0, k0)] + a[OPS_ACC##n(i1, j1, k1)] + a[OPS_ACC##n(i2, j2, k2)] +       \
   a[OPS_ACC##n(i3, j3, k3)])
  double ux1 = C3_FACE(u, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1);
  double ux2 = C3_FACE(u, 0, 1, 0, 0, 1, 1, 0, 1, 0, 1, 1, 1, 1);
  double uy1 = C3_FACE(u, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 1, 0, 1);
  double uy2 = C3_FACE(u, 0, 0, 1, 0, 1, 1, 0, 0, 1, 1, 1, 1, 1);
  double uz1 = C3_FACE(u, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0);
  double uz2 = C3_FACE(u, 0, 0, 0, 1, 1, 0, 1, 0, 1, 1, 1, 1, 1);
  double vx1 = C3_FACE(v, 1, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1);
  double vx2 = C3_FACE(v, 1, 1, 0, 0, 1, 1, 0, 1, 0, 1, 1, 1, 1);
  double vy1 = C3_FACE(v, 1, 0, 0, 0, 1, 0, 0, 0, 0, 1, 1, 0, 1);
  double vy2 = C3_FACE(v, 1, 0, 1, 0, 1, 1, 0, 0, 1, 1, 1, 1, 1);
  double vz1 = C3_FACE(v, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0);
  double vz2 = C3_FACE(v, 1, 0, 0, 1, 1, 0, 1, 0, 1, 1, 1, 1, 1);
  double wx1 = C3_FACE(w, 7, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1);
  double wx2 = C3_FACE(w, 7, 1, 0, 0, 1, 1, 0, 1, 0, 1, 1, 1, 1);
  double wy1 = C3_FACE(w, 7, 0, 0, 0, 1, 0, 0, 0, 0, 1, 1, 0, 1);
  double wy2 = C3_FACE(w, 7, 0, 1, 0, 1, 1, 0, 0, 1, 1, 1, 1, 1);
  double wz1 = C3_FACE(w, 7, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0);
  double wz2 = C3_FACE(w, 7, 0, 0, 1, 1, 0, 1, 0, 1, 1, 1, 1, 1);
#undef C3_FACE
  const double dx = cdx[OPS_ACC2(0, 0, 0)], dy = cdy[OPS_ACC3(0, 0, 0)],
               dz = cdz[OPS_ACC8(0, 0, 0)];
  double div = xa[OPS_ACC9(0, 0, 0)] * (ux2 - ux1) + ya[OPS_ACC10(0, 0, 0)] * (vy2 - vy1) +
               za[OPS_ACC11(0, 0, 0)] * (wz2 - wz1);
  double sxx = 0.25 * (ux2 - ux1) / (dx);
  double syy = 0.25 * (vy2 - vy1) / (dy);
  double szz = 0.25 * (wz2 - wz1) / (dz);
  double sxy = 0.25 * (uy2 - uy1) / (dy) + 0.25 * (vx2 - vx1) / (dx);
  double sxz = 0.25 * (uz2 - uz1) / (dz) + 0.25 * (wx2 - wx1) / (dx);
  double syz = 0.25 * (vz2 - vz1) / (dz) + 0.25 * (wy2 - wy1) / (dy);
  double px = (p[OPS_ACC4(1, 0, 0)] - p[OPS_ACC4(-1, 0, 0)]) / (dx + cdx[OPS_ACC2(1, 0, 0)]);
  double py = (p[OPS_ACC4(0, 1, 0)] - p[OPS_ACC4(0, -1, 0)]) / (dy + cdy[OPS_ACC3(0, 1, 0)]);
  double pz = (p[OPS_ACC4(0, 0, 1)] - p[OPS_ACC4(0, 0, -1)]) / (dz + cdz[OPS_ACC8(0, 0, 1)]);
  double px2 = px * px, py2 = py * py, pz2 = pz * pz;
  double lim = (sxx * px2 + syy * py2 + szz * pz2 + sxy * px * py + sxz * px * pz +
                syz * py * pz) /
               MAX(px2 + py2 + pz2, 1.0e-16);
  if ((lim > 0.0) || (div >= 0.0)) {
    q[OPS_ACC6(0, 0, 0)] = 0.0;
  } else {
    px = SIGN(MAX(1.0e-16, fabs(px)), px);
    py = SIGN(MAX(1.0e-16, fabs(py)), py);
    pz = SIGN(MAX(1.0e-16, fabs(pz)), pz);
    double pgrad = sqrt(px * px + py * py + pz * pz);
    double gx = fabs(dx * pgrad / px);
    double gy = fabs(dy * pgrad / py);
    double gz = fabs(dz * pgrad / pz);
    double g = MIN(gx, MIN(gy, gz));
    double g2 = g * g;
    q[OPS_ACC6(0, 0, 0)] = 2.0 * (rho[OPS_ACC5(0, 0, 0)]) * g2 * lim * lim;
  }
}

// ------------------------------------------------------------ time step control
// (calc_dt_kernel.h)

void c3_calc_dt(const double *cdx, const double *cdy, const double *ss, const double *q,
                const double *rho, const double *u, const double *xa, const double *vol,
                const double *v, const double *ya, double *dtmin, const double *cdz,
                const double *w, const double *za) {
  const double V = vol[OPS_ACC7(0, 0, 0)];
  double ds = MIN(MIN(cdx[OPS_ACC0(0, 0, 0)], cdy[OPS_ACC1(0, 0, 0)]), cdz[OPS_ACC11(0, 0, 0)]);
  ds = 1.0 / (ds * ds);
  double cc = ss[OPS_ACC2(0, 0, 0)] * ss[OPS_ACC2(0, 0, 0)];
  cc = cc + 2.0 * q[OPS_ACC3(0, 0, 0)] / rho[OPS_ACC4(0, 0, 0)];
  double dtct = ds * cc;
  dtct = dtc_safe * 1.0 / MAX(sqrt(dtct), g_small);
  double du1 = (u[OPS_ACC5(0, 0, 0)] + u[OPS_ACC5(0, 1, 0)] + u[OPS_ACC5(0, 0, 1)] +
                u[OPS_ACC5(0, 1, 1)]) * xa[OPS_ACC6(0, 0, 0)];
  double du2 = (u[OPS_ACC5(1, 0, 0)] + u[OPS_ACC5(1, 1, 0)] + u[OPS_ACC5(1, 0, 1)] +
                u[OPS_ACC5(1, 1, 1)]) * xa[OPS_ACC6(0, 0, 0)];
  double dtut = dtu_safe * 4.0 * V / MAX(MAX(fabs(du1), fabs(du2)), 1.0e-5 * V);
  double dv1 = (v[OPS_ACC8(0, 0, 0)] + v[OPS_ACC8(1, 0, 0)] + v[OPS_ACC8(0, 0, 1)] +
                v[OPS_ACC8(1, 0, 1)]) * ya[OPS_ACC9(0, 0, 0)];
  double dv2 = (v[OPS_ACC8(0, 1, 0)] + v[OPS_ACC8(1, 1, 0)] + v[OPS_ACC8(0, 1, 1)] +
                v[OPS_ACC8(1, 1, 1)]) * ya[OPS_ACC9(0, 0, 0)];
  double dtvt = dtv_safe * 4.0 * V / MAX(MAX(fabs(dv1), fabs(dv2)), 1.0e-5 * V);
  double dw1 = (w[OPS_ACC12(0, 0, 0)] + w[OPS_ACC12(0, 1, 0)] + w[OPS_ACC12(1, 0, 0)] +
                w[OPS_ACC12(1, 1, 0)]) * za[OPS_ACC13(0, 0, 0)];
  double dw2 = (w[OPS_ACC12(0, 0, 1)] + w[OPS_ACC12(0, 1, 1)] + w[OPS_ACC12(1, 0, 1)] +
                w[OPS_ACC12(1, 1, 1)]) * za[OPS_ACC13(0, 0, 0)];
  double dtwt = dtw_safe * 4.0 * V / MAX(MAX(fabs(dw1), fabs(dw2)), 1.0e-5 * V);
  double div = du2 - du1 + dv2 - dv1 + dw2 - dw1;
  double dtdivt = dtdiv_safe * 4.0 * (V) / MAX(V * 1.0e-05, fabs(div));
  dtmin[OPS_ACC10(0, 0, 0)] = MIN(MIN(MIN(dtct, dtut), MIN(dtvt, dtdivt)), dtwt);
}

void c3_dt_min(const double *dtmin, double *dt_min_val) {
  *dt_min_val = MIN(*dt_min_val, dtmin[OPS_ACC0(0, 0, 0)]);
}

// ------------------------------------------------------------ Lagrangian step
// (PdV_kernel.h). A face flux is area * (8 nodal velocities summed in order) * 1/8 * dt
// (* 1/2 in the predictor, whose 8 terms are the 4 old velocities twice).

inline double c3_sum8(double a, double b, double c, double d, double e, double f, double g,
                      double h) {
  return a + b + c + d + e + f + g + h;
}

inline void c3_pdv_finish(double total, double V, double p, double r0, double q, double e0,
                          double &vc, double &r1, double &e1) {
  vc = (V) / (V + total);
  double recip = 1.0 / V;
  double de = (p / r0 + q / r0) * total * recip;
  e1 = e0 - de;
  r1 = r0 * vc;
}

void c3_pdv_predict(const double *xa, const double *u, const double *ya, const double *v,
                    double *vc, const double *vol, const double *p, const double *r0,
                    double *r1, const double *q, const double *e0, double *e1,
                    const double *za, const double *w) {
#define U(i, j, k) u[OPS_ACC1(i, j, k)]
#define V_(i, j, k) v[OPS_ACC3(i, j, k)]
#define W(i, j, k) w[OPS_ACC13(i, j, k)]
  double left = (xa[OPS_ACC0(0, 0, 0)] * c3_sum8(U(0, 0, 0), U(0, 1, 0), U(0, 0, 1), U(0, 1, 1),
                                                  U(0, 0, 0), U(0, 1, 0), U(0, 0, 1), U(0, 1, 1))) *
                0.125 * dt * 0.5;
  double right = (xa[OPS_ACC0(1, 0, 0)] * c3_sum8(U(1, 0, 0), U(1, 1, 0), U(1, 0, 1), U(1, 1, 1),
                                                   U(1, 0, 0), U(1, 1, 0), U(1, 0, 1), U(1, 1, 1))) *
                 0.125 * dt * 0.5;
  double bottom = (ya[OPS_ACC2(0, 0, 0)] * c3_sum8(V_(0, 0, 0), V_(1, 0, 0), V_(0, 0, 1),
                                                    V_(1, 0, 1), V_(0, 0, 0), V_(1, 0, 0),
                                                    V_(0, 0, 1), V_(1, 0, 1))) *
                  0.125 * dt * 0.5;
  double top = (ya[OPS_ACC2(0, 1, 0)] * c3_sum8(V_(0, 1, 0), V_(1, 1, 0), V_(0, 1, 1),
                                                 V_(1, 1, 1), V_(0, 1, 0), V_(1, 1, 0),
                                                 V_(0, 1, 1), V_(1, 1, 1))) *
               0.125 * dt * 0.5;
  double back = (za[OPS_ACC12(0, 0, 0)] * c3_sum8(W(0, 0, 0), W(1, 0, 0), W(0, 1, 0), W(1, 1, 0),
                                                   W(0, 0, 0), W(1, 0, 0), W(0, 1, 0), W(1, 1, 0))) *
                0.125 * dt * 0.5;
  double front = (za[OPS_ACC12(0, 0, 1)] * c3_sum8(W(0, 0, 1), W(1, 0, 1), W(0, 1, 1),
                                                    W(1, 1, 1), W(0, 0, 1), W(1, 0, 1),
                                                    W(0, 1, 1), W(1, 1, 1))) *
                 0.125 * dt * 0.5;
#undef U
#undef V_
#undef W
  double total = right - left + top - bottom + front - back;
  c3_pdv_finish(total, vol[OPS_ACC5(0, 0, 0)], p[OPS_ACC6(0, 0, 0)], r0[OPS_ACC7(0, 0, 0)],
                q[OPS_ACC9(0, 0, 0)], e0[OPS_ACC10(0, 0, 0)], vc[OPS_ACC4(0, 0, 0)],
                r1[OPS_ACC8(0, 0, 0)], e1[OPS_ACC11(0, 0, 0)]);
}

void c3_pdv_full(const double *xa, const double *u0, const double *u1, const double *ya,
                 const double *v0, const double *v1, double *vc, const double *vol,
                 const double *p, const double *r0, double *r1, const double *q,
                 const double *e0, double *e1, const double *za, const double *w0,
                 const double *w1) {
#define U0(i, j, k) u0[OPS_ACC1(i, j, k)]
#define U1(i, j, k) u1[OPS_ACC2(i, j, k)]
#define V0(i, j, k) v0[OPS_ACC4(i, j, k)]
#define V1(i, j, k) v1[OPS_ACC5(i, j, k)]
#define W0(i, j, k) w0[OPS_ACC15(i, j, k)]
#define W1(i, j, k) w1[OPS_ACC16(i, j, k)]
  double left = (xa[OPS_ACC0(0, 0, 0)] * c3_sum8(U0(0, 0, 0), U0(0, 1, 0), U0(0, 0, 1),
                                                  U0(0, 1, 1), U1(0, 0, 0), U1(0, 1, 0),
                                                  U1(0, 0, 1), U1(0, 1, 1))) *
                0.125 * dt;
  double right = (xa[OPS_ACC0(1, 0, 0)] * c3_sum8(U0(1, 0, 0), U0(1, 1, 0), U0(1, 0, 1),
                                                   U0(1, 1, 1), U1(1, 0, 0), U1(1, 1, 0),
                                                   U1(1, 0, 1), U1(1, 1, 1))) *
                 0.125 * dt;
  double bottom = (ya[OPS_ACC3(0, 0, 0)] * c3_sum8(V0(0, 0, 0), V0(1, 0, 0), V0(0, 0, 1),
                                                    V0(1, 0, 1), V1(0, 0, 0), V1(1, 0, 0),
                                                    V1(0, 0, 1), V1(1, 0, 1))) *
                  0.125 * dt;
  double top = (ya[OPS_ACC3(0, 1, 0)] * c3_sum8(V0(0, 1, 0), V0(1, 1, 0), V0(0, 1, 1),
                                                 V0(1, 1, 1), V1(0, 1, 0), V1(1, 1, 0),
                                                 V1(0, 1, 1), V1(1, 1, 1))) *
               0.125 * dt;
  double back = (za[OPS_ACC14(0, 0, 0)] * c3_sum8(W0(0, 0, 0), W0(1, 0, 0), W0(0, 1, 0),
                                                   W0(1, 1, 0), W1(0, 0, 0), W1(1, 0, 0),
                                                   W1(0, 1, 0), W1(1, 1, 0))) *
                0.125 * dt;
  double front = (za[OPS_ACC14(0, 0, 1)] * c3_sum8(W0(0, 0, 1), W0(1, 0, 1), W0(0, 1, 1),
                                                    W0(1, 1, 1), W1(0, 0, 1), W1(1, 0, 1),
                                                    W1(0, 1, 1), W1(1, 1, 1))) *
                 0.125 * dt;
#undef U0
#undef U1
#undef V0
#undef V1
#undef W0
#undef W1
  double total = right - left + top - bottom + front - back;
  c3_pdv_finish(total, vol[OPS_ACC7(0, 0, 0)], p[OPS_ACC8(0, 0, 0)], r0[OPS_ACC9(0, 0, 0)],
                q[OPS_ACC11(0, 0, 0)], e0[OPS_ACC12(0, 0, 0)], vc[OPS_ACC6(0, 0, 0)],
                r1[OPS_ACC10(0, 0, 0)], e1[OPS_ACC13(0, 0, 0)]);
}

void c3_revert(const double *r0, double *r1, const double *e0, double *e1) {
  r1[OPS_ACC1(0, 0, 0)] = r0[OPS_ACC0(0, 0, 0)];
  e1[OPS_ACC3(0, 0, 0)] = e0[OPS_ACC2(0, 0, 0)];
}

// ------------------------------------------------------------ nodal acceleration
// (accelerate_kernel.h)

inline double c3_sum4x2(double a0, double b0, double a1, double b1, double a2, double b2,
                        double a3, double b3) {
  return a0 * b0 + a1 * b1 + a2 * b2 + a3 * b3;
}

void c3_accelerate(const double *r, const double *vol, double *sbm, const double *u0,
                   double *u1, const double *xa, const double *p, const double *v0,
                   double *v1, const double *ya, const double *q, const double *w0,
                   double *w1, const double *za) {
#define R(i, j, k) r[OPS_ACC0(i, j, k)]
#define VOL(i, j, k) vol[OPS_ACC1(i, j, k)]
#define P(i, j, k) p[OPS_ACC6(i, j, k)]
#define Q(i, j, k) q[OPS_ACC10(i, j, k)]
#define XA(i, j, k) xa[OPS_ACC5(i, j, k)]
#define YA(i, j, k) ya[OPS_ACC9(i, j, k)]
#define ZA(i, j, k) za[OPS_ACC13(i, j, k)]
  double nodal_mass = (R(-1, -1, 0) * VOL(-1, -1, 0) + R(0, -1, 0) * VOL(0, -1, 0) +
                       R(0, 0, 0) * VOL(0, 0, 0) + R(-1, 0, 0) * VOL(-1, 0, 0) +
                       R(-1, -1, -1) * VOL(-1, -1, -1) + R(0, -1, -1) * VOL(0, -1, -1) +
                       R(0, 0, -1) * VOL(0, 0, -1) + R(-1, 0, -1) * VOL(-1, 0, -1)) *
                      0.125;
  sbm[OPS_ACC2(0, 0, 0)] = 0.25 * dt / nodal_mass;
  const double s = sbm[OPS_ACC2(0, 0, 0)];
  // pressure gradient, then the viscous one (the same stencils on q)
  u1[OPS_ACC4(0, 0, 0)] =
      u0[OPS_ACC3(0, 0, 0)] -
      s * c3_sum4x2(XA(0, 0, 0), (P(0, 0, 0) - P(-1, 0, 0)), XA(0, -1, 0), (P(0, -1, 0) - P(-1, -1, 0)),
                    XA(0, 0, -1), (P(0, 0, -1) - P(-1, 0, -1)), XA(0, -1, -1),
                    (P(0, -1, -1) - P(-1, -1, -1)));
  v1[OPS_ACC8(0, 0, 0)] =
      v0[OPS_ACC7(0, 0, 0)] -
      s * c3_sum4x2(YA(0, 0, 0), (P(0, 0, 0) - P(0, -1, 0)), YA(-1, 0, 0), (P(-1, 0, 0) - P(-1, -1, 0)),
                    YA(0, 0, -1), (P(0, 0, -1) - P(0, -1, -1)), YA(-1, 0, -1),
                    (P(-1, 0, -1) - P(-1, -1, -1)));
  w1[OPS_ACC12(0, 0, 0)] =
      w0[OPS_ACC11(0, 0, 0)] -
      s * c3_sum4x2(ZA(0, 0, 0), (P(0, 0, 0) - P(0, 0, -1)), ZA(0, -1, 0), (P(0, -1, 0) - P(0, -1, -1)),
                    ZA(-1, 0, 0), (P(-1, 0, 0) - P(-1, 0, -1)), ZA(-1, -1, 0),
                    (P(-1, -1, 0) - P(-1, -1, -1)));
  u1[OPS_ACC4(0, 0, 0)] =
      u1[OPS_ACC4(0, 0, 0)] -
      s * c3_sum4x2(XA(0, 0, 0), (Q(0, 0, 0) - Q(-1, 0, 0)), XA(0, -1, 0), (Q(0, -1, 0) - Q(-1, -1, 0)),
                    XA(0, 0, -1), (Q(0, 0, -1) - Q(-1, 0, -1)), XA(0, -1, -1),
                    (Q(0, -1, -1) - Q(-1, -1, -1)));
  v1[OPS_ACC8(0, 0, 0)] =
      v1[OPS_ACC8(0, 0, 0)] -
      s * c3_sum4x2(YA(0, 0, 0), (Q(0, 0, 0) - Q(0, -1, 0)), YA(-1, 0, 0), (Q(-1, 0, 0) - Q(-1, -1, 0)),
                    YA(0, 0, -1), (Q(0, 0, -1) - Q(0, -1, -1)), YA(-1, 0, -1),
                    (Q(-1, 0, -1) - Q(-1, -1, -1)));
  w1[OPS_ACC12(0, 0, 0)] =
      w1[OPS_ACC12(0, 0, 0)] -
      s * c3_sum4x2(ZA(0, 0, 0), (Q(0, 0, 0) - Q(0, 0, -1)), ZA(0, -1, 0), (Q(0, -1, 0) - Q(0, -1, -1)),
                    ZA(-1, 0, 0), (Q(-1, 0, 0) - Q(-1, 0, -1)), ZA(-1, -1, 0),
                    (Q(-1, -1, 0) - Q(-1, -1, -1)));
#undef R
#undef VOL
#undef P
#undef Q
#undef XA
#undef YA
#undef ZA
}

// ------------------------------------------------------------ face volume fluxes
// (flux_calc_kernel.h)

void c3_flux_x(double *fx, const double *xa, const double *u0, const double *u1) {
  fx[OPS_ACC0(0, 0, 0)] =
      0.125 * dt * (xa[OPS_ACC1(0, 0, 0)]) *
      c3_sum8(u0[OPS_ACC2(0, 0, 0)], u0[OPS_ACC2(0, 1, 0)], u0[OPS_ACC2(0, 0, 1)],
              u0[OPS_ACC2(0, 1, 1)], u1[OPS_ACC3(0, 0, 0)], u1[OPS_ACC3(0, 1, 0)],
              u1[OPS_ACC3(0, 0, 1)], u1[OPS_ACC3(0, 1, 1)]);
}

void c3_flux_y(double *fy, const double *ya, const double *v0, const double *v1) {
  fy[OPS_ACC0(0, 0, 0)] =
      0.125 * dt * (ya[OPS_ACC1(0, 0, 0)]) *
      c3_sum8(v0[OPS_ACC2(0, 0, 0)], v0[OPS_ACC2(1, 0, 0)], v0[OPS_ACC2(0, 0, 1)],
              v0[OPS_ACC2(1, 0, 1)], v1[OPS_ACC3(0, 0, 0)], v1[OPS_ACC3(1, 0, 0)],
              v1[OPS_ACC3(0, 0, 1)], v1[OPS_ACC3(1, 0, 1)]);
}

// The old-velocity term set is the reference's: (0,0,0), (1,0,0) twice and (1,1,0)
// (flux_calc_kernel.h, flux_calc_kernelz); kept for bitwise agreement.
void c3_flux_z(double *fz, const double *za, const double *w0, const double *w1) {
  fz[OPS_ACC0(0, 0, 0)] =
      0.125 * dt * (za[OPS_ACC1(0, 0, 0)]) *
      c3_sum8(w0[OPS_ACC2(0, 0, 0)], w0[OPS_ACC2(1, 0, 0)], w0[OPS_ACC2(1, 0, 0)],
              w0[OPS_ACC2(1, 1, 0)], w1[OPS_ACC3(0, 0, 0)], w1[OPS_ACC3(1, 0, 0)],
              w1[OPS_ACC3(0, 1, 0)], w1[OPS_ACC3(1, 1, 0)]);
}

// ------------------------------------------------------------ cell advection
// (advec_cell_kernel.h)

// pre/post sweep volumes
void c3_cell_vol_1x(double *pre, double *post, const double *vol, const double *fx,
                    const double *fy, const double *fz) {
  const double dfx = fx[OPS_ACC3(1, 0, 0)] - fx[OPS_ACC3(0, 0, 0)];
  pre[OPS_ACC0(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)] +
                           (fx[OPS_ACC3(1, 0, 0)] - fx[OPS_ACC3(0, 0, 0)] + fy[OPS_ACC4(0, 1, 0)] -
                            fy[OPS_ACC4(0, 0, 0)] + fz[OPS_ACC5(0, 0, 1)] - fz[OPS_ACC5(0, 0, 0)]);
  post[OPS_ACC1(0, 0, 0)] = pre[OPS_ACC0(0, 0, 0)] - (dfx);
}

void c3_cell_vol_2x(double *pre, double *post, const double *vol, const double *fx) {
  pre[OPS_ACC0(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)] + fx[OPS_ACC3(1, 0, 0)] - fx[OPS_ACC3(0, 0, 0)];
  post[OPS_ACC1(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)];
}

void c3_cell_vol_1y(double *pre, double *post, const double *vol, const double *fz,
                    const double *fy) {
  pre[OPS_ACC0(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)] + fy[OPS_ACC4(0, 1, 0)] - fy[OPS_ACC4(0, 0, 0)] +
                           fz[OPS_ACC3(0, 0, 1)] - fz[OPS_ACC3(0, 0, 0)];
  post[OPS_ACC1(0, 0, 0)] = pre[OPS_ACC0(0, 0, 0)] - (fy[OPS_ACC4(0, 1, 0)] - fy[OPS_ACC4(0, 0, 0)]);
}

void c3_cell_vol_2y(double *pre, double *post, const double *vol, const double *fy,
                    const double *fx) {
  pre[OPS_ACC0(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)] + fy[OPS_ACC3(0, 1, 0)] - fy[OPS_ACC3(0, 0, 0)] +
                           fx[OPS_ACC4(1, 0, 0)] - fx[OPS_ACC4(0, 0, 0)];
  post[OPS_ACC1(0, 0, 0)] = pre[OPS_ACC0(0, 0, 0)] - (fy[OPS_ACC3(0, 1, 0)] - fy[OPS_ACC3(0, 0, 0)]);
}

void c3_cell_vol_1z(double *pre, double *post, const double *vol, const double *fx,
                    const double *fy, const double *fz) {
  pre[OPS_ACC0(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)] +
                           (fx[OPS_ACC3(1, 0, 0)] - fx[OPS_ACC3(0, 0, 0)] + fy[OPS_ACC4(0, 1, 0)] -
                            fy[OPS_ACC4(0, 0, 0)] + fz[OPS_ACC5(0, 0, 1)] - fz[OPS_ACC5(0, 0, 0)]);
  post[OPS_ACC1(0, 0, 0)] = pre[OPS_ACC0(0, 0, 0)] - (fz[OPS_ACC5(0, 0, 1)] - fz[OPS_ACC5(0, 0, 0)]);
}

void c3_cell_vol_2z(double *pre, double *post, const double *vol, const double *fz) {
  pre[OPS_ACC0(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)] + fz[OPS_ACC3(0, 0, 1)] - fz[OPS_ACC3(0, 0, 0)];
  post[OPS_ACC1(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)];
}

// second-order limiter of the donor-cell flux of one quantity q:
// qd/qu/qw are q at the donor, upwind and downwind cells
inline double c3_cell_limiter(double sigma, double sigma3, double sigma4, double qd, double qu,
                              double qw) {
  const double one_by_six = 1.0 / 6.0;
  double diffuw = qd - qu;
  double diffdw = qw - qd;
  if ((diffuw * diffdw) > 0.0)
    return (1.0 - sigma) * SIGN(1.0, diffdw) *
           MIN(MIN(fabs(diffuw), fabs(diffdw)),
               one_by_six * (sigma3 * fabs(diffuw) + sigma4 * fabs(diffdw)));
  return 0.0;
}

// upwind/donor/downwind offsets along the sweep axis and the cell-width ratio index:
// flux > 0 takes the donor below; otherwise above, except on the last face
#define C3_UPWIND(flux, above_last)                                                         \
  int up, don, dn, dif;                                                                     \
  if ((flux) > 0.0) {                                                                       \
    up = -2; don = -1; dn = 0; dif = don;                                                   \
  } else if (above_last) {                                                                  \
    up = 1; don = 0; dn = -1; dif = up;                                                     \
  } else {                                                                                  \
    up = 0; don = 0; dn = -1; dif = up;                                                     \
  }

// mass and energy fluxes through one face; the 3 kernels differ only in the axis
#define C3_CELL_FLUX(AX)                                                                    \
  const double f = vf[OPS_ACC0(0, 0, 0)];                                                   \
  const double pre_d = pre[OPS_ACC1(AX(don))];                                              \
  const double sigmat = fabs(f) / pre_d;                                                    \
  const double sigma3 = (1.0 + sigmat) * (vd[OPS_ACC3(0, 0, 0)] / vd[OPS_ACC3(AX(dif))]);   \
  const double sigma4 = 2.0 - sigmat;                                                       \
  const double rd = rho[OPS_ACC4(AX(don))];                                                 \
  mf[OPS_ACC6(0, 0, 0)] =                                                                   \
      (f) * (rd + c3_cell_limiter(sigmat, sigma3, sigma4, rd, rho[OPS_ACC4(AX(up))],        \
                                  rho[OPS_ACC4(AX(dn))]));                                  \
  const double sigmam = fabs(mf[OPS_ACC6(0, 0, 0)]) / (rd * pre_d);                         \
  const double ed = e[OPS_ACC5(AX(don))];                                                   \
  ef[OPS_ACC7(0, 0, 0)] =                                                                   \
      mf[OPS_ACC6(0, 0, 0)] *                                                               \
      (ed + c3_cell_limiter(sigmam, sigma3, sigma4, ed, e[OPS_ACC5(AX(up))], e[OPS_ACC5(AX(dn))]))

#define C3_AXX(o) o, 0, 0
#define C3_AXY(o) 0, o, 0
#define C3_AXZ(o) 0, 0, o

void c3_cell_flux_x(const double *vf, const double *pre, const int *xx, const double *vd,
                    const double *rho, const double *e, double *mf, double *ef) {
  C3_UPWIND(vf[OPS_ACC0(0, 0, 0)], xx[OPS_ACC2(1, 0, 0)] < field.x_max + 2 - 2)
  C3_CELL_FLUX(C3_AXX);
}

void c3_cell_flux_y(const double *vf, const double *pre, const int *yy, const double *vd,
                    const double *rho, const double *e, double *mf, double *ef) {
  C3_UPWIND(vf[OPS_ACC0(0, 0, 0)], yy[OPS_ACC2(0, 1, 0)] < field.y_max + 2 - 2)
  C3_CELL_FLUX(C3_AXY);
}

void c3_cell_flux_z(const double *vf, const double *pre, const int *zz, const double *vd,
                    const double *rho, const double *e, double *mf, double *ef) {
  C3_UPWIND(vf[OPS_ACC0(0, 0, 0)], zz[OPS_ACC2(0, 0, 1)] < field.z_max + 2 - 2)
  C3_CELL_FLUX(C3_AXZ);
}

// density and energy update from the fluxes through the two faces along the sweep axis
#define C3_CELL_UPDATE(AX)                                                                  \
  pm[OPS_ACC6(0, 0, 0)] = rho[OPS_ACC0(0, 0, 0)] * pre[OPS_ACC4(0, 0, 0)];                  \
  postm[OPS_ACC7(0, 0, 0)] = pm[OPS_ACC6(0, 0, 0)] + mf[OPS_ACC2(0, 0, 0)] - mf[OPS_ACC2(AX(1))]; \
  pe[OPS_ACC9(0, 0, 0)] = (e[OPS_ACC1(0, 0, 0)] * pm[OPS_ACC6(0, 0, 0)] + ef[OPS_ACC10(0, 0, 0)] - \
                           ef[OPS_ACC10(AX(1))]) /                                          \
                          postm[OPS_ACC7(0, 0, 0)];                                         \
  av[OPS_ACC8(0, 0, 0)] = pre[OPS_ACC4(0, 0, 0)] + vf[OPS_ACC3(0, 0, 0)] - vf[OPS_ACC3(AX(1))];   \
  rho[OPS_ACC0(0, 0, 0)] = postm[OPS_ACC7(0, 0, 0)] / av[OPS_ACC8(0, 0, 0)];                \
  e[OPS_ACC1(0, 0, 0)] = pe[OPS_ACC9(0, 0, 0)]

void c3_cell_update_x(double *rho, double *e, const double *mf, const double *vf,
                      const double *pre, const double *post, double *pm, double *postm,
                      double *av, double *pe, const double *ef) {
  C3_CELL_UPDATE(C3_AXX);
}

void c3_cell_update_y(double *rho, double *e, const double *mf, const double *vf,
                      const double *pre, const double *post, double *pm, double *postm,
                      double *av, double *pe, const double *ef) {
  C3_CELL_UPDATE(C3_AXY);
}

void c3_cell_update_z(double *rho, double *e, const double *mf, const double *vf,
                      const double *pre, const double *post, double *pm, double *postm,
                      double *av, double *pe, const double *ef) {
  C3_CELL_UPDATE(C3_AXZ);
}

// ------------------------------------------------------------ momentum advection
// (advec_mom_kernel.h)

// nodal post/pre volumes for the sweep (x1/z1: first sweep along x/z, x2/y2: second sweep
// with/without advect_x, x3/z3: third sweep along x/z)
void c3_mom_vol_x1(double *pre, double *post, const double *vol, const double *fx,
                   const double *fy, const double *fz) {
  post[OPS_ACC1(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)] + fy[OPS_ACC4(0, 1, 0)] - fy[OPS_ACC4(0, 0, 0)] +
                            fz[OPS_ACC5(0, 0, 1)] - fz[OPS_ACC5(0, 0, 0)];
  pre[OPS_ACC0(0, 0, 0)] = post[OPS_ACC1(0, 0, 0)] + fx[OPS_ACC3(1, 0, 0)] - fx[OPS_ACC3(0, 0, 0)];
}

void c3_mom_vol_z1(double *pre, double *post, const double *vol, const double *fx,
                   const double *fy, const double *fz) {
  post[OPS_ACC1(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)] + fx[OPS_ACC3(1, 0, 0)] - fx[OPS_ACC3(0, 0, 0)] +
                            fy[OPS_ACC4(0, 1, 0)] - fy[OPS_ACC4(0, 0, 0)];
  pre[OPS_ACC0(0, 0, 0)] = post[OPS_ACC1(0, 0, 0)] + fz[OPS_ACC5(0, 0, 1)] - fz[OPS_ACC5(0, 0, 0)];
}

void c3_mom_vol_x2(double *pre, double *post, const double *vol, const double *fy,
                   const double *fz) {
  post[OPS_ACC1(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)] + fz[OPS_ACC4(0, 0, 1)] - fz[OPS_ACC4(0, 0, 0)];
  pre[OPS_ACC0(0, 0, 0)] = post[OPS_ACC1(0, 0, 0)] + fy[OPS_ACC3(0, 1, 0)] - fy[OPS_ACC3(0, 0, 0)];
}

void c3_mom_vol_y2(double *pre, double *post, const double *vol, const double *fx,
                   const double *fy) {
  post[OPS_ACC1(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)] + fx[OPS_ACC3(1, 0, 0)] - fx[OPS_ACC3(0, 0, 0)];
  pre[OPS_ACC0(0, 0, 0)] = post[OPS_ACC1(0, 0, 0)] + fy[OPS_ACC4(0, 1, 0)] - fy[OPS_ACC4(0, 0, 0)];
}

void c3_mom_vol_x3(double *pre, double *post, const double *vol, const double *fx) {
  post[OPS_ACC1(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)];
  pre[OPS_ACC0(0, 0, 0)] = post[OPS_ACC1(0, 0, 0)] + fx[OPS_ACC3(1, 0, 0)] - fx[OPS_ACC3(0, 0, 0)];
}

void c3_mom_vol_z3(double *pre, double *post, const double *vol, const double *fz) {
  post[OPS_ACC1(0, 0, 0)] = vol[OPS_ACC2(0, 0, 0)];
  pre[OPS_ACC0(0, 0, 0)] = post[OPS_ACC1(0, 0, 0)] + fz[OPS_ACC3(0, 0, 1)] - fz[OPS_ACC3(0, 0, 0)];
}

// nodal mass flux: the mean of the 8 face fluxes around the node
void c3_node_flux_x(double *nf, const double *m) {
  nf[OPS_ACC0(0, 0, 0)] =
      0.125 * c3_sum8(m[OPS_ACC1(0, -1, 0)], m[OPS_ACC1(0, 0, 0)], m[OPS_ACC1(1, -1, 0)],
                      m[OPS_ACC1(1, 0, 0)], m[OPS_ACC1(0, -1, -1)], m[OPS_ACC1(0, 0, -1)],
                      m[OPS_ACC1(1, -1, -1)], m[OPS_ACC1(1, 0, -1)]);
}

void c3_node_flux_y(double *nf, const double *m) {
  nf[OPS_ACC0(0, 0, 0)] =
      0.125 * c3_sum8(m[OPS_ACC1(-1, 0, 0)], m[OPS_ACC1(0, 0, 0)], m[OPS_ACC1(-1, 1, 0)],
                      m[OPS_ACC1(0, 1, 0)], m[OPS_ACC1(-1, 0, -1)], m[OPS_ACC1(0, 0, -1)],
                      m[OPS_ACC1(-1, 1, -1)], m[OPS_ACC1(0, 1, -1)]);
}

void c3_node_flux_z(double *nf, const double *m) {
  nf[OPS_ACC0(0, 0, 0)] =
      0.125 * c3_sum8(m[OPS_ACC1(-1, 0, 0)], m[OPS_ACC1(0, 0, 0)], m[OPS_ACC1(-1, 0, 1)],
                      m[OPS_ACC1(0, 0, 1)], m[OPS_ACC1(-1, -1, 0)], m[OPS_ACC1(0, -1, 0)],
                      m[OPS_ACC1(-1, -1, 1)], m[OPS_ACC1(0, -1, 1)]);
}

// nodal masses after (post) and before (pre) the sweep; the 3 kernels differ only in the
// neighbour of the node flux
#define C3_NODE_MASS(i, j, k)                                                               \
  nmpost[OPS_ACC0(0, 0, 0)] =                                                               \
      0.125 * (rho[OPS_ACC2(0, -1, 0)] * pv[OPS_ACC1(0, -1, 0)] +                           \
               rho[OPS_ACC2(0, 0, 0)] * pv[OPS_ACC1(0, 0, 0)] +                             \
               rho[OPS_ACC2(-1, -1, 0)] * pv[OPS_ACC1(-1, -1, 0)] +                         \
               rho[OPS_ACC2(-1, 0, 0)] * pv[OPS_ACC1(-1, 0, 0)] +                           \
               rho[OPS_ACC2(0, -1, -1)] * pv[OPS_ACC1(0, -1, -1)] +                         \
               rho[OPS_ACC2(0, 0, -1)] * pv[OPS_ACC1(0, 0, -1)] +                           \
               rho[OPS_ACC2(-1, -1, -1)] * pv[OPS_ACC1(-1, -1, -1)] +                       \
               rho[OPS_ACC2(-1, 0, -1)] * pv[OPS_ACC1(-1, 0, -1)]);                         \
  nmpre[OPS_ACC3(0, 0, 0)] = nmpost[OPS_ACC0(0, 0, 0)] - nf[OPS_ACC4(i, j, k)] + nf[OPS_ACC4(0, 0, 0)]

void c3_node_mass_x(double *nmpost, const double *pv, const double *rho, double *nmpre,
                    const double *nf) {
  C3_NODE_MASS(-1, 0, 0);
}

void c3_node_mass_y(double *nmpost, const double *pv, const double *rho, double *nmpre,
                    const double *nf) {
  C3_NODE_MASS(0, -1, 0);
}

void c3_node_mass_z(double *nmpost, const double *pv, const double *rho, double *nmpre,
                    const double *nf) {
  C3_NODE_MASS(0, 0, -1);
}

// limited momentum flux through a node face: fl the node flux, md the donor's pre-sweep
// nodal mass, w/wd the cell widths at the node and at `dif`, vd/vu/vw the velocity at the
// donor, upwind and downwind nodes
inline double c3_mom_limited_flux(double fl, double md, double w, double wd, double vd,
                                  double vu, double vw) {
  double sigma = fabs(fl) / md;
  double vdiffuw = vd - vu;
  double vdiffdw = vw - vd;
  double limiter = 0.0;
  if (vdiffuw * vdiffdw > 0.0) {
    double auw = fabs(vdiffuw);
    double adw = fabs(vdiffdw);
    double wind = 1.0;
    if (vdiffdw <= 0.0) wind = -1.0;
    limiter = wind * MIN(w * ((2.0 - sigma) * adw / w + (1.0 + sigma) * auw / wd) / 6.0,
                         MIN(auw, adw));
  }
  double advec_vel = vd + (1.0 - sigma) * limiter;
  return advec_vel * fl;
}

#define C3_MOM_FLUX(AX)                                                                     \
  int up, don, dn, dif;                                                                     \
  if ((nf[OPS_ACC0(0, 0, 0)]) < 0.0) {                                                      \
    up = 2; don = 1; dn = 0; dif = don;                                                     \
  } else {                                                                                  \
    up = -1; don = 0; dn = 1; dif = up;                                                     \
  }                                                                                         \
  mom[OPS_ACC2(0, 0, 0)] = c3_mom_limited_flux(                                             \
      nf[OPS_ACC0(0, 0, 0)], nmpre[OPS_ACC1(AX(don))], cd[OPS_ACC3(0, 0, 0)],               \
      cd[OPS_ACC3(AX(dif))], vel[OPS_ACC4(AX(don))], vel[OPS_ACC4(AX(up))],                 \
      vel[OPS_ACC4(AX(dn))])

void c3_mom_flux_x(const double *nf, const double *nmpre, double *mom, const double *cd,
                   const double *vel) {
  C3_MOM_FLUX(C3_AXX);
}

void c3_mom_flux_y(const double *nf, const double *nmpre, double *mom, const double *cd,
                   const double *vel) {
  C3_MOM_FLUX(C3_AXY);
}

void c3_mom_flux_z(const double *nf, const double *nmpre, double *mom, const double *cd,
                   const double *vel) {
  C3_MOM_FLUX(C3_AXZ);
}

#define C3_MOM_UPDATE(AX)                                                                   \
  vel[OPS_ACC0(0, 0, 0)] = (vel[OPS_ACC0(0, 0, 0)] * nmpre[OPS_ACC2(0, 0, 0)] +             \
                            mom[OPS_ACC3(AX(-1))] - mom[OPS_ACC3(0, 0, 0)]) /               \
                           nmpost[OPS_ACC1(0, 0, 0)]

void c3_mom_update_x(double *vel, const double *nmpost, const double *nmpre, const double *mom) {
  C3_MOM_UPDATE(C3_AXX);
}

void c3_mom_update_y(double *vel, const double *nmpost, const double *nmpre, const double *mom) {
  C3_MOM_UPDATE(C3_AXY);
}

void c3_mom_update_z(double *vel, const double *nmpost, const double *nmpre, const double *mom) {
  C3_MOM_UPDATE(C3_AXZ);
}

// ------------------------------------------------------------ end of step
// (reset_field_kernel.h)

void c3_reset_cells(double *r0, const double *r1, double *e0, const double *e1) {
  r0[OPS_ACC0(0, 0, 0)] = r1[OPS_ACC1(0, 0, 0)];
  e0[OPS_ACC2(0, 0, 0)] = e1[OPS_ACC3(0, 0, 0)];
}

void c3_reset_nodes(double *u0, const double *u1, double *v0, const double *v1, double *w0,
                    const double *w1) {
  u0[OPS_ACC0(0, 0, 0)] = u1[OPS_ACC1(0, 0, 0)];
  v0[OPS_ACC2(0, 0, 0)] = v1[OPS_ACC3(0, 0, 0)];
  w0[OPS_ACC4(0, 0, 0)] = w1[OPS_ACC5(0, 0, 0)];
}

// ------------------------------------------------------------ diagnostics
// (field_summary_kernel.h): kinetic energy from the 8 corner velocities, z outermost

void c3_field_summary(const double *vol, const double *rho, const double *e, const double *p,
                      const double *u, const double *v, const double *w, double *svol,
                      double *smass, double *sie, double *ske, double *spress) {
  double vsqrd = 0.0;
  for (int k = 0; k < 2; k++)
    for (int j = 0; j < 2; j++)
      for (int i = 0; i < 2; i++)
        vsqrd += 0.125 * (u[OPS_ACC4(i, j, k)] * u[OPS_ACC4(i, j, k)] +
                          v[OPS_ACC5(i, j, k)] * v[OPS_ACC5(i, j, k)] +
                          w[OPS_ACC6(i, j, k)] * w[OPS_ACC6(i, j, k)]);
  double cell_vol = vol[OPS_ACC0(0, 0, 0)];
  double cell_mass = cell_vol * rho[OPS_ACC1(0, 0, 0)];
  *svol = *svol + cell_vol;
  *smass = *smass + cell_mass;
  *sie = *sie + cell_mass * e[OPS_ACC2(0, 0, 0)];
  *ske = *ske + cell_mass * 0.5 * vsqrd;
  *spress = *spress + cell_vol * p[OPS_ACC3(0, 0, 0)];
}

#endif
