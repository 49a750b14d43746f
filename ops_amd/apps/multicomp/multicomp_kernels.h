// Kernels of the multi-component dat test (the reference's apps/c/multiDim: dats of 2 and 3
// components addressed as ACC(component, x, y), AoS by default and SoA under OPS_SOA).
#ifndef MULTICOMP_KERNELS_H
#define MULTICOMP_KERNELS_H

void mc_set(ACC<double> &v, ACC<double> &w, const int *idx) {
  v(0, 0, 0) = (double)idx[0];
  v(1, 0, 0) = (double)(10 * idx[1]);
  for (int c = 0; c < 3; c++) w(c, 0, 0) = (double)(100 * c + idx[0] + 7 * idx[1]);
}

// neighbours across components
void mc_stencil(const ACC<double> &v, const ACC<double> &w, ACC<double> &out) {
  out(0, 0) = v(0, 1, 0) + v(1, 0, -1) + w(2, -1, 0) - w(0, 0, 1);
}

// the same stencil written against API v1: pointers at the current point addressed with
// OPS_ACC_MD<n>(component, x, y) / OPS_ACC<n>(x, y), n the argument position
void mc_stencil_v1(const double *v, const double *w, double *out) {
  out[OPS_ACC2(0, 0)] = v[OPS_ACC_MD0(0, 1, 0)] + v[OPS_ACC_MD0(1, 0, -1)] +
                        w[OPS_ACC_MD1(2, -1, 0)] - w[OPS_ACC_MD1(0, 0, 1)];
}

void mc_reduce(const ACC<double> &v, const ACC<double> &w, double *sum, double *mx) {
  *sum += v(0, 0, 0) + v(1, 0, 0);
  *mx = *mx > w(2, 0, 0) ? *mx : w(2, 0, 0);
}

#endif
