// TeaLeaf 2D user kernels (API v2, ACC<T> accessors).
//
// Arithmetic is kept operation for operation identical to the reference application
// (apps/c/TeaLeaf/*_kernels.h: initialise_chunk, generate_chunk, set_field,
// tea_leaf_common, tea_leaf_cg, tea_leaf_jacobi, tea_leaf, field_summary), so FP64
// fields can be compared bit for bit. Consts: field, grid, states, number_of_states,
// g_rect/g_circ/g_point.
#ifndef TEA_KERNELS_H
#define TEA_KERNELS_H

#include "user_types.h"

// ------------------------------------------------------------ mesh
void tea_zero(ACC<double> &v) { v(0, 0) = 0.0; }

void tea_xx(ACC<int> &xx, int *idx) { xx(0, 0) = idx[0] - 2; }
void tea_yy(ACC<int> &yy, int *idx) { yy(0, 0) = idx[1] - 2; }

void tea_vertex_x(ACC<double> &vx, const ACC<int> &xx, ACC<double> &vdx) {
  int lo = field.x_min - 2;
  double d = (grid.xmax - grid.xmin) / (double)grid.x_cells;
  double origin = grid.xmin + d * field.left;
  vx(0, 0) = origin + d * (xx(0, 0) - lo);
  vdx(0, 0) = (double)d;
}
void tea_vertex_y(ACC<double> &vy, const ACC<int> &yy, ACC<double> &vdy) {
  int lo = field.y_min - 2;
  double d = (grid.ymax - grid.ymin) / (double)grid.y_cells;
  double origin = grid.ymin + d * field.bottom;
  vy(0, 0) = origin + d * (yy(0, 0) - lo);
  vdy(0, 0) = (double)d;
}
void tea_cell_x(const ACC<double> &vx, ACC<double> &cx, ACC<double> &cdx) {
  double d = (grid.xmax - grid.xmin) / (double)grid.x_cells;
  cx(0, 0) = 0.5 * (vx(0, 0) + vx(1, 0));
  cdx(0, 0) = d;
}
void tea_cell_y(const ACC<double> &vy, ACC<double> &cy, ACC<double> &cdy) {
  double d = (grid.ymax - grid.ymin) / (double)grid.y_cells;
  cy(0, 0) = 0.5 * (vy(0, 0) + vy(0, 1));
  cdy(0, 0) = d;
}
void tea_volume(ACC<double> &vol, const ACC<double> &cdy, ACC<double> &xa, const ACC<double> &cdx,
                ACC<double> &ya) {
  double dx = (grid.xmax - grid.xmin) / (double)grid.x_cells;
  double dy = (grid.ymax - grid.ymin) / (double)grid.y_cells;
  vol(0, 0) = dx * dy;
  xa(0, 0) = cdy(0, 0);
  ya(0, 0) = cdx(0, 0);
}

// ------------------------------------------------------------ initial state
void tea_generate(const ACC<double> &vx, const ACC<double> &vy, ACC<double> &e0, ACC<double> &rho,
                  ACC<double> &u, const ACC<double> &cx, const ACC<double> &cy) {
  e0(0, 0) = states[0].energy;
  rho(0, 0) = states[0].density;
  for (int s = 1; s < number_of_states; s++) {
    const double xc = states[s].xmin, yc = states[s].ymin;
    int cell_in = 0;
    if (states[s].geometry == g_rect) {
      if (vx(1, 0) >= states[s].xmin && vx(0, 0) < states[s].xmax)
        if (vy(0, 1) >= states[s].ymin && vy(0, 0) < states[s].ymax) cell_in = 1;
    } else if (states[s].geometry == g_circ) {
      double r = 0.0;
      for (int i = -1; i <= 0; i++)
        for (int j = -1; j <= 0; j++)
          r = sqrt((cx(i, 0) - xc) * (cx(i, 0) - xc) + (cy(0, j) - yc) * (cy(0, j) - yc));
      if (r <= states[s].radius) cell_in = 1;
    } else if (states[s].geometry == g_point) {
      if (vx(0, 0) == xc && vy(0, 0) == yc) cell_in = 1;
    }
    if (cell_in) {
      e0(0, 0) = states[s].energy;
      rho(0, 0) = states[s].density;
    }
  }
  u(0, 0) = e0(0, 0) * rho(0, 0);
}

void tea_copy(const ACC<double> &src, ACC<double> &dst) { dst(0, 0) = src(0, 0); }

// ------------------------------------------------------------ reflective halo (one layer)
void tea_halo(ACC<double> &rho, ACC<double> &e0, ACC<double> &e1, ACC<double> &u, ACC<double> &p,
              ACC<double> &sd, const int *fields, const int *m, const int *idx) {
  const int off = 2 * (m[1] - idx[m[0]]) - m[2];
  const int ox = m[0] == 0 ? off : 0, oy = m[0] == 1 ? off : 0;
  if (fields[FIELD_DENSITY] == 1) rho(0, 0) = rho(ox, oy);
  if (fields[FIELD_ENERGY0] == 1) e0(0, 0) = e0(ox, oy);
  if (fields[FIELD_ENERGY1] == 1) e1(0, 0) = e1(ox, oy);
  if (fields[FIELD_U] == 1) u(0, 0) = u(ox, oy);
  if (fields[FIELD_P] == 1) p(0, 0) = p(ox, oy);
  if (fields[FIELD_SD] == 1) sd(0, 0) = sd(ox, oy);
}

// ------------------------------------------------------------ operator set-up
void tea_init_u_u0(ACC<double> &u, ACC<double> &u0, const ACC<double> &e, const ACC<double> &rho) {
  u(0, 0) = e(0, 0) * rho(0, 0);
  u0(0, 0) = e(0, 0) * rho(0, 0);
}
void tea_recip(ACC<double> &w, const ACC<double> &rho) { w(0, 0) = 1.0 / rho(0, 0); }

void tea_init_kx_ky(ACC<double> &Kx, ACC<double> &Ky, const ACC<double> &w) {
  Kx(0, 0) = (w(-1, 0) + w(0, 0)) / (2.0 * w(-1, 0) * w(0, 0));
  Ky(0, 0) = (w(0, -1) + w(0, 0)) / (2.0 * w(0, -1) * w(0, 0));
}

// w = A u and r = u - w, with A = I + rx*Dx + ry*Dy (5-point, face conductivities Kx, Ky)
void tea_init_w_r(ACC<double> &w, ACC<double> &r, const ACC<double> &Kx, const ACC<double> &Ky,
                  const ACC<double> &u, const double *rx, const double *ry) {
  w(0, 0) = (1.0 + (*ry) * (Ky(0, 1) + Ky(0, 0)) + (*rx) * (Kx(1, 0) + Kx(0, 0))) * u(0, 0) -
            (*ry) * (Ky(0, 1) * u(0, 1) + Ky(0, 0) * u(0, -1)) -
            (*rx) * (Kx(1, 0) * u(1, 0) + Kx(0, 0) * u(-1, 0));
  r(0, 0) = u(0, 0) - w(0, 0);
}

void tea_residual(ACC<double> &r, const ACC<double> &Kx, const ACC<double> &Ky,
                  const ACC<double> &u, const ACC<double> &u0, const double *rx, const double *ry) {
  double smvp = 0.0;
  smvp = (1.0 + (*ry) * (Ky(0, 1) + Ky(0, 0)) + (*rx) * (Kx(1, 0) + Kx(0, 0))) * u(0, 0) -
         (*ry) * (Ky(0, 1) * u(0, 1) + Ky(0, 0) * u(0, -1)) -
         (*rx) * (Kx(1, 0) * u(1, 0) + Kx(0, 0) * u(-1, 0));
  r(0, 0) = u0(0, 0) - smvp;
}

void tea_norm2(const ACC<double> &x, double *norm) { *norm = *norm + x(0, 0) * x(0, 0); }
void tea_dot(const ACC<double> &r, const ACC<double> &p, double *rro) {
  *rro = *rro + r(0, 0) * p(0, 0);
}

// ------------------------------------------------------------ CG
void tea_cg_calc_w(ACC<double> &w, const ACC<double> &Kx, const ACC<double> &Ky,
                   const ACC<double> &p, const double *rx, const double *ry, double *pw) {
  w(0, 0) = (1.0 + (*ry) * (Ky(0, 1) + Ky(0, 0)) + (*rx) * (Kx(1, 0) + Kx(0, 0))) * p(0, 0) -
            (*ry) * (Ky(0, 1) * p(0, 1) + Ky(0, 0) * p(0, -1)) -
            (*rx) * (Kx(1, 0) * p(1, 0) + Kx(0, 0) * p(-1, 0));
  *pw = *pw + w(0, 0) * p(0, 0);
}
void tea_axpy(ACC<double> &u, const ACC<double> &p, const double *alpha) {
  u(0, 0) = u(0, 0) + (*alpha) * p(0, 0);
}
void tea_cg_calc_r(ACC<double> &r, const ACC<double> &w, const double *alpha, double *rnn) {
  r(0, 0) = r(0, 0) - (*alpha) * w(0, 0);
  *rnn = *rnn + r(0, 0) * r(0, 0);
}
void tea_axpby(ACC<double> &u, const ACC<double> &p, const double *alpha, const double *beta) {
  u(0, 0) = (*alpha) * u(0, 0) + (*beta) * p(0, 0);
}

// ------------------------------------------------------------ Jacobi
void tea_jacobi(ACC<double> &u1, const ACC<double> &Kx, const ACC<double> &Ky,
                const ACC<double> &un, const ACC<double> &u0, const double *rx, const double *ry,
                double *error) {
  u1(0, 0) = (u0(0, 0) + (*rx) * (Kx(1, 0) * un(1, 0) + Kx(0, 0) * un(-1, 0)) +
              (*ry) * (Ky(0, 1) * un(0, 1) + Ky(0, 0) * un(0, -1))) /
             (1.0 + (*rx) * (Kx(1, 0) + Kx(0, 0)) + (*ry) * (Ky(0, 1) + Ky(0, 0)));
  *error = *error + fabs(u1(0, 0) - un(0, 0));
}

// ------------------------------------------------------------ finish
void tea_recip2(ACC<double> &z, const ACC<double> &x, const ACC<double> &y) {
  z(0, 0) = x(0, 0) / y(0, 0);
}

void tea_field_summary(const ACC<double> &vol, const ACC<double> &rho, const ACC<double> &e,
                       const ACC<double> &u, double *s_vol, double *s_mass, double *s_ie,
                       double *s_temp) {
  double cv = vol(0, 0);
  double cm = cv * rho(0, 0);
  *s_vol = *s_vol + cv;
  *s_mass = *s_mass + cm;
  *s_ie = *s_ie + cm * e(0, 0);
  *s_temp = *s_temp + cm * u(0, 0);
}

#endif
