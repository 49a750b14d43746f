// ADI heat-diffusion user kernels (API v2). Same arithmetic as the reference's
// apps/c/adi/init_kernel.h and preproc_kernel.h; the mesh extents come in as a loop
// argument (n = {nx, ny, nz}) so that blocks of different sizes share one kernel.
#ifndef ADI_KERNELS_H
#define ADI_KERNELS_H

// u = 1 on the boundary layer of the block, 0 inside
void adi_init_kernel(ACC<double> &u, const int *idx, const int *n) {
  if (idx[0] == 0 || idx[0] == n[0] - 1 || idx[1] == 0 || idx[1] == n[1] - 1 || idx[2] == 0 ||
      idx[2] == n[2] - 1)
    u(0, 0, 0) = 1.0;
  else
    u(0, 0, 0) = 0.0;
}

// right-hand side of the implicit step and the three tridiagonal systems' coefficients;
// Dirichlet rows on the boundary layer
void adi_preproc_kernel(const ACC<double> &u, ACC<double> &du, ACC<double> &ax, ACC<double> &bx,
                        ACC<double> &cx, ACC<double> &ay, ACC<double> &by, ACC<double> &cy,
                        ACC<double> &az, ACC<double> &bz, ACC<double> &cz, const int *idx,
                        const int *n, const double *lam) {
  double a, b, c, d;
  const double lambda = *lam;
  if (idx[0] == 0 || idx[0] == n[0] - 1 || idx[1] == 0 || idx[1] == n[1] - 1 || idx[2] == 0 ||
      idx[2] == n[2] - 1) {
    d = 0.0f;
    a = 0.0f;
    b = 1.0f;
    c = 0.0f;
  } else {
    d = lambda * (u(-1, 0, 0) + u(1, 0, 0) + u(0, -1, 0) + u(0, 1, 0) + u(0, 0, -1) + u(0, 0, 1) -
                  6.0f * u(0, 0, 0));
    a = -0.5f * lambda;
    b = 1.0f + lambda;
    c = -0.5f * lambda;
  }
  du(0, 0, 0) = d;
  ax(0, 0, 0) = a;
  bx(0, 0, 0) = b;
  cx(0, 0, 0) = c;
  ay(0, 0, 0) = a;
  by(0, 0, 0) = b;
  cy(0, 0, 0) = c;
  az(0, 0, 0) = a;
  bz(0, 0, 0) = b;
  cz(0, 0, 0) = c;
}

#endif
