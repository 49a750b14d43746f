// CloverLeaf 2D user kernels (API v2, ACC<T> accessors).
//
// The arithmetic of every kernel is kept operation for operation identical to the
// reference application (apps/c/CloverLeaf/*_kernel.h): same operand order, same
// MIN/MAX/SIGN macros, same constants, so the FP64 fields can be compared bit for bit.
// Constants (dt, g_small, g_big, dt*_safe, field, grid, states, number_of_states,
// g_rect/g_circ/g_point) are ops_decl_const globals.
#ifndef CLOVER_KERNELS_H
#define CLOVER_KERNELS_H

#include "user_types.h"

// ------------------------------------------------------------ mesh set-up
// (initialise_chunk_kernel.h)

void cl_xx(ACC<int> &xx, int *idx) { xx(0, 0) = idx[0] - 2; }
void cl_yy(ACC<int> &yy, int *idx) { yy(0, 0) = idx[1] - 2; }

void cl_vertex_x(ACC<double> &vx, const ACC<int> &xx, ACC<double> &vdx) {
  int lo = field.x_min - 2;
  double d = (grid.xmax - grid.xmin) / (double)grid.x_cells;
  double origin = grid.xmin + d * field.left;
  vx(0, 0) = origin + d * (xx(0, 0) - lo);
  vdx(0, 0) = (double)d;
}

void cl_vertex_y(ACC<double> &vy, const ACC<int> &yy, ACC<double> &vdy) {
  int lo = field.y_min - 2;
  double d = (grid.ymax - grid.ymin) / (double)grid.y_cells;
  double origin = grid.ymin + d * field.bottom;
  vy(0, 0) = origin + d * (yy(0, 0) - lo);
  vdy(0, 0) = (double)d;
}

void cl_cell_x(const ACC<double> &vx, ACC<double> &cx, ACC<double> &cdx) {
  double d = (grid.xmax - grid.xmin) / (double)grid.x_cells;
  cx(0, 0) = 0.5 * (vx(0, 0) + vx(1, 0));
  cdx(0, 0) = d;
}

void cl_cell_y(const ACC<double> &vy, ACC<double> &cy, ACC<double> &cdy) {
  double d = (grid.ymax - grid.ymin) / (double)grid.y_cells;
  cy(0, 0) = 0.5 * (vy(0, 0) + vy(0, 1));
  cdy(0, 0) = d;
}

void cl_volume(ACC<double> &vol, const ACC<double> &cdy, ACC<double> &xa, const ACC<double> &cdx,
               ACC<double> &ya) {
  double dx = (grid.xmax - grid.xmin) / (double)grid.x_cells;
  double dy = (grid.ymax - grid.ymin) / (double)grid.y_cells;
  vol(0, 0) = dx * dy;
  xa(0, 0) = cdy(0, 0);
  ya(0, 0) = cdx(0, 0);
}

// ------------------------------------------------------------ initial state
// (generate_chunk_kernel.h): background state 0, then states 1.. painted in order

void cl_generate(const ACC<double> &vx, const ACC<double> &vy, ACC<double> &e0, ACC<double> &r0,
                 ACC<double> &u0, ACC<double> &v0, const ACC<double> &cx,
                 const ACC<double> &cy) {
  e0(0, 0) = states[0].energy;
  r0(0, 0) = states[0].density;
  u0(0, 0) = states[0].xvel;
  v0(0, 0) = states[0].yvel;
  for (int s = 1; s < number_of_states; s++) {
    const double xc = states[s].xmin, yc = states[s].ymin;
    int node_in = 0, cell_in = 0;
    if (states[s].geometry == g_rect) {
      for (int i = -1; i <= 0; i++)
        for (int j = -1; j <= 0; j++)
          if (vx(1 + i, 0) >= states[s].xmin && vx(0 + i, 0) < states[s].xmax)
            if (vy(0, 1 + j) >= states[s].ymin && vy(0, 0 + j) < states[s].ymax) node_in = 1;
      if (vx(1, 0) >= states[s].xmin && vx(0, 0) < states[s].xmax)
        if (vy(0, 1) >= states[s].ymin && vy(0, 0) < states[s].ymax) cell_in = 1;
    } else if (states[s].geometry == g_circ) {
      double r = 0.0;
      for (int i = -1; i <= 0; i++)
        for (int j = -1; j <= 0; j++) {
          r = sqrt((cx(i, 0) - xc) * (cx(i, 0) - xc) + (cy(0, j) - yc) * (cy(0, j) - yc));
          if (r <= states[s].radius) node_in = 1;
        }
      if (r <= states[s].radius) cell_in = 1;
    } else if (states[s].geometry == g_point) {
      for (int i = -1; i <= 0; i++)
        for (int j = -1; j <= 0; j++)
          if (vx(i, 0) == xc && vy(0, j) == yc) node_in = 1;
      if (vx(0, 0) == xc && vy(0, 0) == yc) cell_in = 1;
    } else {
      continue;
    }
    if (cell_in) {
      e0(0, 0) = states[s].energy;
      r0(0, 0) = states[s].density;
    }
    if (node_in) {
      u0(0, 0) = states[s].xvel;
      v0(0, 0) = states[s].yvel;
    }
  }
}

// ------------------------------------------------------------ equation of state
// (ideal_gas_kernel.h), gamma = 1.4

void cl_ideal_gas(const ACC<double> &rho, const ACC<double> &e, ACC<double> &p, ACC<double> &cs) {
  double v = 1.0 / rho(0, 0);
  p(0, 0) = (1.4 - 1.0) * rho(0, 0) * e(0, 0);
  double dpde = (1.4 - 1.0) * rho(0, 0);
  double dpdv = -1 * rho(0, 0) * p(0, 0);
  double c2 = v * v * (p(0, 0) * dpde - dpdv);
  cs(0, 0) = sqrt(c2);
}

// ------------------------------------------------------------ reflective halos
// One kernel per field group, one launch per side covering both halo layers. The
// mirror source for halo index h on axis m[0] is h + off with
//   off = 2 * (m[1] - h) - m[2]
// and the copied value is multiplied by m[3] (+1 keeps, -1 flips the normal component),
// which reproduces the per-layer strip loops of apps/c/CloverLeaf/update_halo.cpp.

void cl_halo_cells(ACC<double> &r0, ACC<double> &r1, ACC<double> &e0, ACC<double> &e1,
                   ACC<double> &p, ACC<double> &q, ACC<double> &cs, const int *fields,
                   const int *m, const int *idx) {
  const int off = 2 * (m[1] - idx[m[0]]) - m[2];
  const int ox = m[0] == 0 ? off : 0, oy = m[0] == 1 ? off : 0;
  if (fields[FIELD_DENSITY0] == 1) r0(0, 0) = r0(ox, oy);
  if (fields[FIELD_DENSITY1] == 1) r1(0, 0) = r1(ox, oy);
  if (fields[FIELD_ENERGY0] == 1) e0(0, 0) = e0(ox, oy);
  if (fields[FIELD_ENERGY1] == 1) e1(0, 0) = e1(ox, oy);
  if (fields[FIELD_PRESSURE] == 1) p(0, 0) = p(ox, oy);
  if (fields[FIELD_VISCOSITY] == 1) q(0, 0) = q(ox, oy);
  if (fields[FIELD_SOUNDSPEED] == 1) cs(0, 0) = cs(ox, oy);
}

// nodal velocities: m[3] is the factor for xvel, m[4] for yvel
void cl_halo_vel(ACC<double> &u0, ACC<double> &u1, ACC<double> &v0, ACC<double> &v1,
                 const int *fields, const int *m, const int *idx) {
  const int off = 2 * (m[1] - idx[m[0]]) - m[2];
  const int ox = m[0] == 0 ? off : 0, oy = m[0] == 1 ? off : 0;
  if (fields[FIELD_XVEL0] == 1) u0(0, 0) = m[3] * u0(ox, oy);
  if (fields[FIELD_XVEL1] == 1) u1(0, 0) = m[3] * u1(ox, oy);
  if (fields[FIELD_YVEL0] == 1) v0(0, 0) = m[4] * v0(ox, oy);
  if (fields[FIELD_YVEL1] == 1) v1(0, 0) = m[4] * v1(ox, oy);
}

// face fluxes: fv/fm select which of the pair (volume, mass) is refreshed
void cl_halo_flux(ACC<double> &vf, ACC<double> &mf, const int *fields, const int *m,
                  const int *idx) {
  const int off = 2 * (m[1] - idx[m[0]]) - m[2];
  const int ox = m[0] == 0 ? off : 0, oy = m[0] == 1 ? off : 0;
  if (fields[m[4]] == 1) vf(0, 0) = m[3] * vf(ox, oy);
  if (fields[m[5]] == 1) mf(0, 0) = m[3] * mf(ox, oy);
}

// ------------------------------------------------------------ artificial viscosity
// (viscosity_kernel.h)

void cl_viscosity(const ACC<double> &u, const ACC<double> &v, const ACC<double> &cdx,
                  const ACC<double> &cdy, const ACC<double> &p, const ACC<double> &rho,
                  ACC<double> &q) {
  double ugrad = (u(1, 0) + u(1, 1)) - (u(0, 0) + u(0, 1));
  double vgrad = (v(0, 1) + v(1, 1)) - (v(0, 0) + v(1, 0));
  double div = (cdx(0, 0)) * (ugrad) + (cdy(0, 0)) * (vgrad);
  double strain2 = 0.5 * (u(0, 1) + u(1, 1) - u(0, 0) - u(1, 0)) / (cdy(0, 0)) +
                   0.5 * (v(1, 0) + v(1, 1) - v(0, 0) - v(0, 1)) / (cdx(0, 0));
  double px = (p(1, 0) - p(-1, 0)) / (cdx(0, 0) + cdx(1, 0));
  double py = (p(0, 1) - p(0, -1)) / (cdy(0, 0) + cdy(0, 1));
  double px2 = px * px, py2 = py * py;
  double lim = ((0.5 * (ugrad) / cdx(0, 0)) * px2 + (0.5 * (vgrad) / cdy(0, 0)) * py2 +
                strain2 * px * py) /
               MAX(px2 + py2, 1.0e-16);
  if ((lim > 0.0) || (div >= 0.0)) {
    q(0, 0) = 0.0;
  } else {
    px = SIGN(MAX(1.0e-16, fabs(px)), px);
    py = SIGN(MAX(1.0e-16, fabs(py)), py);
    double pg = sqrt(px * px + py * py);
    double xg = fabs(cdx(0, 0) * pg / px);
    double yg = fabs(cdy(0, 0) * pg / py);
    double g = MIN(xg, yg);
    double g2 = g * g;
    q(0, 0) = 2.0 * (rho(0, 0)) * g2 * lim * lim;
  }
}

// ------------------------------------------------------------ time step control
// (calc_dt_kernel.h)

void cl_calc_dt(const ACC<double> &cdx, const ACC<double> &cdy, const ACC<double> &cs,
                const ACC<double> &q, const ACC<double> &rho, const ACC<double> &u,
                const ACC<double> &xa, const ACC<double> &vol, const ACC<double> &v,
                const ACC<double> &ya, ACC<double> &dtc_out) {
  double dsx = cdx(0, 0), dsy = cdy(0, 0);
  double cc = cs(0, 0) * cs(0, 0);
  cc = cc + 2.0 * q(0, 0) / rho(0, 0);
  cc = MAX(sqrt(cc), g_small);
  double dtct = dtc_safe * MIN(dsx, dsy) / cc;
  double div = 0.0;
  double dv1 = (u(0, 0) + u(0, 1)) * xa(0, 0);
  double dv2 = (u(1, 0) + u(1, 1)) * xa(1, 0);
  div = div + dv2 - dv1;
  double dtut = dtu_safe * 2.0 * vol(0, 0) / MAX(MAX(fabs(dv1), fabs(dv2)), g_small * vol(0, 0));
  dv1 = (v(0, 0) + v(1, 0)) * ya(0, 0);
  dv2 = (v(0, 1) + v(1, 1)) * ya(0, 1);
  div = div + dv2 - dv1;
  double dtvt = dtv_safe * 2.0 * vol(0, 0) / MAX(MAX(fabs(dv1), fabs(dv2)), g_small * vol(0, 0));
  div = div / (2.0 * vol(0, 0));
  double dtdivt;
  if (div < -g_small)
    dtdivt = dtdiv_safe * (-1.0 / div);
  else
    dtdivt = g_big;
  dtc_out(0, 0) = MIN(MIN(dtct, dtut), MIN(dtvt, dtdivt));
}

void cl_dt_min(const ACC<double> &dtc, double *dt_min) { *dt_min = MIN(*dt_min, dtc(0, 0)); }

// position of the controlling cell (calc_dt_kernel.h: calc_dt_kernel_get)
void cl_dt_get(const ACC<double> &cellx, const ACC<double> &celly, double *xl_pos, double *yl_pos) {
  *xl_pos = cellx(0, 0);
  *yl_pos = celly(0, 0);
}

// ------------------------------------------------------------ Lagrangian step
// (PdV_kernel.h, revert_kernel.h, accelerate_kernel.h)

void cl_pdv_predict(const ACC<double> &xa, const ACC<double> &u0, const ACC<double> &ya,
                    const ACC<double> &v0, ACC<double> &dvol, const ACC<double> &vol,
                    const ACC<double> &p, const ACC<double> &r0, ACC<double> &r1,
                    const ACC<double> &q, const ACC<double> &e0, ACC<double> &e1) {
  double lf = (xa(0, 0) * (u0(0, 0) + u0(0, 1) + u0(0, 0) + u0(0, 1))) * 0.25 * dt * 0.5;
  double rf = (xa(1, 0) * (u0(1, 0) + u0(1, 1) + u0(1, 0) + u0(1, 1))) * 0.25 * dt * 0.5;
  double bf = (ya(0, 0) * (v0(0, 0) + v0(1, 0) + v0(0, 0) + v0(1, 0))) * 0.25 * dt * 0.5;
  double tf = (ya(0, 1) * (v0(0, 1) + v0(1, 1) + v0(0, 1) + v0(1, 1))) * 0.25 * dt * 0.5;
  double tot = rf - lf + tf - bf;
  dvol(0, 0) = (vol(0, 0)) / (vol(0, 0) + tot);
  double rv = 1.0 / vol(0, 0);
  double de = (p(0, 0) / r0(0, 0) + q(0, 0) / r0(0, 0)) * tot * rv;
  e1(0, 0) = e0(0, 0) - de;
  r1(0, 0) = r0(0, 0) * dvol(0, 0);
}

void cl_pdv_full(const ACC<double> &xa, const ACC<double> &u0, const ACC<double> &u1,
                 const ACC<double> &ya, const ACC<double> &v0, const ACC<double> &v1,
                 ACC<double> &dvol, const ACC<double> &vol, const ACC<double> &p,
                 const ACC<double> &r0, ACC<double> &r1, const ACC<double> &q,
                 const ACC<double> &e0, ACC<double> &e1) {
  double lf = (xa(0, 0) * (u0(0, 0) + u0(0, 1) + u1(0, 0) + u1(0, 1))) * 0.25 * dt;
  double rf = (xa(1, 0) * (u0(1, 0) + u0(1, 1) + u1(1, 0) + u1(1, 1))) * 0.25 * dt;
  double bf = (ya(0, 0) * (v0(0, 0) + v0(1, 0) + v1(0, 0) + v1(1, 0))) * 0.25 * dt;
  double tf = (ya(0, 1) * (v0(0, 1) + v0(1, 1) + v1(0, 1) + v1(1, 1))) * 0.25 * dt;
  double tot = rf - lf + tf - bf;
  dvol(0, 0) = (vol(0, 0)) / (vol(0, 0) + tot);
  double rv = 1.0 / vol(0, 0);
  double de = (p(0, 0) / r0(0, 0) + q(0, 0) / r0(0, 0)) * tot * rv;
  e1(0, 0) = e0(0, 0) - de;
  r1(0, 0) = r0(0, 0) * dvol(0, 0);
}

void cl_revert(const ACC<double> &r0, ACC<double> &r1, const ACC<double> &e0, ACC<double> &e1) {
  r1(0, 0) = r0(0, 0);
  e1(0, 0) = e0(0, 0);
}

void cl_accelerate(const ACC<double> &r0, const ACC<double> &vol, ACC<double> &sbm,
                   const ACC<double> &u0, ACC<double> &u1, const ACC<double> &xa,
                   const ACC<double> &p, const ACC<double> &v0, ACC<double> &v1,
                   const ACC<double> &ya, const ACC<double> &q) {
  double nodal_mass = (r0(-1, -1) * vol(-1, -1) + r0(0, -1) * vol(0, -1) + r0(0, 0) * vol(0, 0) +
                       r0(-1, 0) * vol(-1, 0)) *
                      0.25;
  sbm(0, 0) = 0.5 * dt / nodal_mass;
  u1(0, 0) = u0(0, 0) - sbm(0, 0) * (xa(0, 0) * (p(0, 0) - p(-1, 0)) +
                                     xa(0, -1) * (p(0, -1) - p(-1, -1)));
  v1(0, 0) = v0(0, 0) - sbm(0, 0) * (ya(0, 0) * (p(0, 0) - p(0, -1)) +
                                     ya(-1, 0) * (p(-1, 0) - p(-1, -1)));
  u1(0, 0) = u1(0, 0) - sbm(0, 0) * (xa(0, 0) * (q(0, 0) - q(-1, 0)) +
                                     xa(0, -1) * (q(0, -1) - q(-1, -1)));
  v1(0, 0) = v1(0, 0) - sbm(0, 0) * (ya(0, 0) * (q(0, 0) - q(0, -1)) +
                                     ya(-1, 0) * (q(-1, 0) - q(-1, -1)));
}

// ------------------------------------------------------------ face volume fluxes
// (flux_calc_kernel.h)

void cl_flux_x(ACC<double> &vfx, const ACC<double> &xa, const ACC<double> &u0,
               const ACC<double> &u1) {
  vfx(0, 0) = 0.25 * dt * (xa(0, 0)) * ((u0(0, 0)) + (u0(0, 1)) + (u1(0, 0)) + (u1(0, 1)));
}

void cl_flux_y(ACC<double> &vfy, const ACC<double> &ya, const ACC<double> &v0,
               const ACC<double> &v1) {
  vfy(0, 0) = 0.25 * dt * (ya(0, 0)) * ((v0(0, 0)) + (v0(1, 0)) + (v1(0, 0)) + (v1(1, 0)));
}

// ------------------------------------------------------------ cell advection
// (advec_cell_kernel.h). The x and y sweeps share one body each, parameterised by the
// unit step (sx, sy) of the sweep direction.

void cl_adv_cell_prevol_1x(ACC<double> &pre, ACC<double> &post, const ACC<double> &vol,
                           const ACC<double> &vfx, const ACC<double> &vfy) {
  pre(0, 0) = vol(0, 0) + (vfx(1, 0) - vfx(0, 0) + vfy(0, 1) - vfy(0, 0));
  post(0, 0) = pre(0, 0) - (vfx(1, 0) - vfx(0, 0));
}
void cl_adv_cell_prevol_1y(ACC<double> &pre, ACC<double> &post, const ACC<double> &vol,
                           const ACC<double> &vfx, const ACC<double> &vfy) {
  pre(0, 0) = vol(0, 0) + (vfy(0, 1) - vfy(0, 0) + vfx(1, 0) - vfx(0, 0));
  post(0, 0) = pre(0, 0) - (vfy(0, 1) - vfy(0, 0));
}
void cl_adv_cell_prevol_2x(ACC<double> &pre, ACC<double> &post, const ACC<double> &vol,
                           const ACC<double> &vfx) {
  pre(0, 0) = vol(0, 0) + vfx(1, 0) - vfx(0, 0);
  post(0, 0) = vol(0, 0);
}
void cl_adv_cell_prevol_2y(ACC<double> &pre, ACC<double> &post, const ACC<double> &vol,
                           const ACC<double> &vfy) {
  pre(0, 0) = vol(0, 0) + vfy(0, 1) - vfy(0, 0);
  post(0, 0) = vol(0, 0);
}

// donor-cell fluxes with the van Leer-type limiter; d = +1 along x (sweep_y = 0) or y
inline double cl_adv_cell_limit(double diffuw, double diffdw, double sig, double s3, double s4) {
  const double one_by_six = 1.0 / 6.0;
  if ((diffuw * diffdw) > 0.0)
    return (1.0 - sig) * SIGN(1.0, diffdw) *
           MIN(MIN(fabs(diffuw), fabs(diffdw)), one_by_six * (s3 * fabs(diffuw) + s4 * fabs(diffdw)));
  return 0.0;
}

void cl_adv_cell_flux_x(const ACC<double> &vfx, const ACC<double> &pre, const ACC<int> &xx,
                        const ACC<double> &vdx, const ACC<double> &rho, const ACC<double> &e,
                        ACC<double> &mfx, ACC<double> &ef) {
  int up, don, dn, dif;
  if (vfx(0, 0) > 0.0) {
    up = -2; don = -1; dn = 0; dif = don;
  } else if (xx(1, 0) < field.x_max + 2 - 2) {
    up = 1; don = 0; dn = -1; dif = up;
  } else {
    up = 0; don = 0; dn = -1; dif = up;
  }
  double sigt = fabs(vfx(0, 0)) / pre(don, 0);
  double s3 = (1.0 + sigt) * (vdx(0, 0) / vdx(dif, 0));
  double s4 = 2.0 - sigt;
  double lim = cl_adv_cell_limit(rho(don, 0) - rho(up, 0), rho(dn, 0) - rho(don, 0), sigt, s3, s4);
  mfx(0, 0) = (vfx(0, 0)) * (rho(don, 0) + lim);
  double sigm = fabs(mfx(0, 0)) / (rho(don, 0) * pre(don, 0));
  lim = cl_adv_cell_limit(e(don, 0) - e(up, 0), e(dn, 0) - e(don, 0), sigm, s3, s4);
  ef(0, 0) = mfx(0, 0) * (e(don, 0) + lim);
}

void cl_adv_cell_flux_y(const ACC<double> &vfy, const ACC<double> &pre, const ACC<int> &yy,
                        const ACC<double> &vdy, const ACC<double> &rho, const ACC<double> &e,
                        ACC<double> &mfy, ACC<double> &ef) {
  int up, don, dn, dif;
  if (vfy(0, 0) > 0.0) {
    up = -2; don = -1; dn = 0; dif = don;
  } else if (yy(0, 1) < field.y_max + 2 - 2) {
    up = 1; don = 0; dn = -1; dif = up;
  } else {
    up = 0; don = 0; dn = -1; dif = up;
  }
  double sigt = fabs(vfy(0, 0)) / pre(0, don);
  double s3 = (1.0 + sigt) * (vdy(0, 0) / vdy(0, dif));
  double s4 = 2.0 - sigt;
  double lim = cl_adv_cell_limit(rho(0, don) - rho(0, up), rho(0, dn) - rho(0, don), sigt, s3, s4);
  mfy(0, 0) = (vfy(0, 0)) * (rho(0, don) + lim);
  double sigm = fabs(mfy(0, 0)) / (rho(0, don) * pre(0, don));
  lim = cl_adv_cell_limit(e(0, don) - e(0, up), e(0, dn) - e(0, don), sigm, s3, s4);
  ef(0, 0) = mfy(0, 0) * (e(0, don) + lim);
}

void cl_adv_cell_update_x(ACC<double> &rho, ACC<double> &e, const ACC<double> &mfx,
                          const ACC<double> &vfx, const ACC<double> &pre, const ACC<double> &post,
                          ACC<double> &pre_mass, ACC<double> &post_mass, ACC<double> &adv_vol,
                          ACC<double> &post_e, const ACC<double> &ef) {
  pre_mass(0, 0) = rho(0, 0) * pre(0, 0);
  post_mass(0, 0) = pre_mass(0, 0) + mfx(0, 0) - mfx(1, 0);
  post_e(0, 0) = (e(0, 0) * pre_mass(0, 0) + ef(0, 0) - ef(1, 0)) / post_mass(0, 0);
  adv_vol(0, 0) = pre(0, 0) + vfx(0, 0) - vfx(1, 0);
  rho(0, 0) = post_mass(0, 0) / adv_vol(0, 0);
  e(0, 0) = post_e(0, 0);
}

void cl_adv_cell_update_y(ACC<double> &rho, ACC<double> &e, const ACC<double> &mfy,
                          const ACC<double> &vfy, const ACC<double> &pre, const ACC<double> &post,
                          ACC<double> &pre_mass, ACC<double> &post_mass, ACC<double> &adv_vol,
                          ACC<double> &post_e, const ACC<double> &ef) {
  pre_mass(0, 0) = rho(0, 0) * pre(0, 0);
  post_mass(0, 0) = pre_mass(0, 0) + mfy(0, 0) - mfy(0, 1);
  post_e(0, 0) = (e(0, 0) * pre_mass(0, 0) + ef(0, 0) - ef(0, 1)) / post_mass(0, 0);
  adv_vol(0, 0) = pre(0, 0) + vfy(0, 0) - vfy(0, 1);
  rho(0, 0) = post_mass(0, 0) / adv_vol(0, 0);
  e(0, 0) = post_e(0, 0);
}

// ------------------------------------------------------------ momentum advection
// (advec_mom_kernel.h)

void cl_mom_vol_x1(ACC<double> &pre, ACC<double> &post, const ACC<double> &vol,
                   const ACC<double> &vfx, const ACC<double> &vfy) {
  post(0, 0) = vol(0, 0) + vfy(0, 1) - vfy(0, 0);
  pre(0, 0) = post(0, 0) + vfx(1, 0) - vfx(0, 0);
}
void cl_mom_vol_y1(ACC<double> &pre, ACC<double> &post, const ACC<double> &vol,
                   const ACC<double> &vfx, const ACC<double> &vfy) {
  post(0, 0) = vol(0, 0) + vfx(1, 0) - vfx(0, 0);
  pre(0, 0) = post(0, 0) + vfy(0, 1) - vfy(0, 0);
}
void cl_mom_vol_x2(ACC<double> &pre, ACC<double> &post, const ACC<double> &vol,
                   const ACC<double> &vfy) {
  post(0, 0) = vol(0, 0);
  pre(0, 0) = post(0, 0) + vfy(0, 1) - vfy(0, 0);
}
void cl_mom_vol_y2(ACC<double> &pre, ACC<double> &post, const ACC<double> &vol,
                   const ACC<double> &vfx) {
  post(0, 0) = vol(0, 0);
  pre(0, 0) = post(0, 0) + vfx(1, 0) - vfx(0, 0);
}

void cl_mom_node_flux_x(ACC<double> &nf, const ACC<double> &mfx) {
  nf(0, 0) = 0.25 * (mfx(0, -1) + mfx(0, 0) + mfx(1, -1) + mfx(1, 0));
}
void cl_mom_node_flux_y(ACC<double> &nf, const ACC<double> &mfy) {
  nf(0, 0) = 0.25 * (mfy(-1, 0) + mfy(0, 0) + mfy(-1, 1) + mfy(0, 1));
}

void cl_mom_node_mass_x(ACC<double> &nm_post, const ACC<double> &post, const ACC<double> &rho,
                        ACC<double> &nm_pre, const ACC<double> &nf) {
  nm_post(0, 0) = 0.25 * (rho(0, -1) * post(0, -1) + rho(0, 0) * post(0, 0) +
                          rho(-1, -1) * post(-1, -1) + rho(-1, 0) * post(-1, 0));
  nm_pre(0, 0) = nm_post(0, 0) - nf(-1, 0) + nf(0, 0);
}
void cl_mom_node_mass_y(ACC<double> &nm_post, const ACC<double> &post, const ACC<double> &rho,
                        ACC<double> &nm_pre, const ACC<double> &nf) {
  nm_post(0, 0) = 0.25 * (rho(0, -1) * post(0, -1) + rho(0, 0) * post(0, 0) +
                          rho(-1, -1) * post(-1, -1) + rho(-1, 0) * post(-1, 0));
  nm_pre(0, 0) = nm_post(0, 0) - nf(0, -1) + nf(0, 0);
}

void cl_mom_flux_x(const ACC<double> &nf, const ACC<double> &nm_pre, ACC<double> &mom,
                   const ACC<double> &cdx, const ACC<double> &vel) {
  int up, don, dn, dif;
  if ((nf(0, 0)) < 0.0) {
    up = 2; don = 1; dn = 0; dif = don;
  } else {
    up = -1; don = 0; dn = 1; dif = up;
  }
  double sigma = fabs(nf(0, 0)) / nm_pre(don, 0);
  double width = cdx(0, 0);
  double vdiffuw = vel(don, 0) - vel(up, 0);
  double vdiffdw = vel(dn, 0) - vel(don, 0);
  double lim = 0.0;
  if (vdiffuw * vdiffdw > 0.0) {
    double auw = fabs(vdiffuw), adw = fabs(vdiffdw);
    double wind = 1.0;
    if (vdiffdw <= 0.0) wind = -1.0;
    lim = wind * MIN(width * ((2.0 - sigma) * adw / width + (1.0 + sigma) * auw / cdx(dif, 0)) / 6.0,
                     MIN(auw, adw));
  }
  double adv = vel(don, 0) + (1.0 - sigma) * lim;
  mom(0, 0) = adv * nf(0, 0);
}

void cl_mom_flux_y(const ACC<double> &nf, const ACC<double> &nm_pre, ACC<double> &mom,
                   const ACC<double> &cdy, const ACC<double> &vel) {
  int up, don, dn, dif;
  if ((nf(0, 0)) < 0.0) {
    up = 2; don = 1; dn = 0; dif = don;
  } else {
    up = -1; don = 0; dn = 1; dif = up;
  }
  double sigma = fabs(nf(0, 0)) / nm_pre(0, don);
  double width = cdy(0, 0);
  double vdiffuw = vel(0, don) - vel(0, up);
  double vdiffdw = vel(0, dn) - vel(0, don);
  double lim = 0.0;
  if (vdiffuw * vdiffdw > 0.0) {
    double auw = fabs(vdiffuw), adw = fabs(vdiffdw);
    double wind = 1.0;
    if (vdiffdw <= 0.0) wind = -1.0;
    lim = wind * MIN(width * ((2.0 - sigma) * adw / width + (1.0 + sigma) * auw / cdy(0, dif)) / 6.0,
                     MIN(auw, adw));
  }
  double adv = vel(0, don) + (1.0 - sigma) * lim;
  mom(0, 0) = adv * nf(0, 0);
}

void cl_mom_update_x(ACC<double> &vel, const ACC<double> &nm_post, const ACC<double> &nm_pre,
                     const ACC<double> &mom) {
  vel(0, 0) = (vel(0, 0) * nm_pre(0, 0) + mom(-1, 0) - mom(0, 0)) / nm_post(0, 0);
}
void cl_mom_update_y(ACC<double> &vel, const ACC<double> &nm_post, const ACC<double> &nm_pre,
                     const ACC<double> &mom) {
  vel(0, 0) = (vel(0, 0) * nm_pre(0, 0) + mom(0, -1) - mom(0, 0)) / nm_post(0, 0);
}

// ------------------------------------------------------------ end of step
// (reset_field_kernel.h, field_summary_kernel.h)

void cl_reset_cells(ACC<double> &r0, const ACC<double> &r1, ACC<double> &e0, const ACC<double> &e1) {
  r0(0, 0) = r1(0, 0);
  e0(0, 0) = e1(0, 0);
}
void cl_reset_nodes(ACC<double> &u0, const ACC<double> &u1, ACC<double> &v0, const ACC<double> &v1) {
  u0(0, 0) = u1(0, 0);
  v0(0, 0) = v1(0, 0);
}

void cl_field_summary(const ACC<double> &vol, const ACC<double> &rho, const ACC<double> &e,
                      const ACC<double> &p, const ACC<double> &u, const ACC<double> &v,
                      double *s_vol, double *s_mass, double *s_ie, double *s_ke, double *s_press) {
  double vsq = 0.0;
  vsq = vsq + 0.25 * (u(0, 0) * u(0, 0) + v(0, 0) * v(0, 0));
  vsq = vsq + 0.25 * (u(1, 0) * u(1, 0) + v(1, 0) * v(1, 0));
  vsq = vsq + 0.25 * (u(0, 1) * u(0, 1) + v(0, 1) * v(0, 1));
  vsq = vsq + 0.25 * (u(1, 1) * u(1, 1) + v(1, 1) * v(1, 1));
  double cv = vol(0, 0);
  double cm = cv * rho(0, 0);
  *s_vol = *s_vol + cv;
  *s_mass = *s_mass + cm;
  *s_ie = *s_ie + cm * e(0, 0);
  *s_ke = *s_ke + cm * 0.5 * vsq;
  *s_press = *s_press + cv * p(0, 0);
}

#endif
