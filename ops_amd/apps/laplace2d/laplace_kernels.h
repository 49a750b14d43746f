// Laplace2D Jacobi user kernels (API v2). Same arithmetic as the reference tutorial
// (apps/c/laplace2d_tutorial/step7/laplace_kernels.h): a 5-point average scaled by
// 0.25 with a running max of |new - old|.
#ifndef LAPLACE_KERNELS_H
#define LAPLACE_KERNELS_H

void set_zero(ACC<double> &A) { A(0, 0) = 0.0; }

void copy(ACC<double> &A, const ACC<double> &Anew) { A(0, 0) = Anew(0, 0); }

void apply_stencil(const ACC<double> &A, ACC<double> &Anew, double *error) {
  Anew(0, 0) = 0.25f * (A(1, 0) + A(-1, 0) + A(0, -1) + A(0, 1));
  *error = fmax(*error, fabs(Anew(0, 0) - A(0, 0)));
}

// Dirichlet value from the loop index, u = scale * sin(pi * (idx[dir] + 1) / (n + 1));
// exercised by the device-boundary variant of the application.
void sine_bndcon(ACC<double> &A, const int *idx, const double *scale) {
  A(0, 0) = sin(pi * (idx[1] + 1) / (jmax + 1)) * (*scale);
}

#endif
