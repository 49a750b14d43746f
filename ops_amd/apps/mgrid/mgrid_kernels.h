// Kernels of the multigrid transfer test (the reference's apps/c/mgrid/mgrid_kernels.h:
// populate, prolong, restrict and their checks), written against ACC<T>. The populate
// kernels take the row length as an argument so that the test runs at any size; the
// checks fold their verdict with a max, as a MAX reduction expects.
#ifndef MGRID_KERNELS_H
#define MGRID_KERNELS_H

void mgrid_populate(ACC<double> &val, const int *idx, const int *nx) {
  val(0, 0) = (double)(idx[0] + nx[0] * idx[1]);
}

// coarse(i) = fine(i * stride): the fine argument carries a restrict stencil
void mgrid_restrict_kernel(const ACC<double> &fine, ACC<double> &coarse) { coarse(0, 0) = fine(0, 0); }

// fine(i) = coarse(i / stride): the coarse argument carries a prolong stencil
void mgrid_prolong_kernel(const ACC<double> &coarse, ACC<double> &fine) { fine(0, 0) = coarse(0, 0); }

// after two prolongations by 2 of a coarse field populated with x + nc * y (nc = n / 4):
// every point and its four periodic neighbours hold x / 4 + (y / 4) * nc
void mgrid_prolong_check(const ACC<double> &val, const int *idx, int *err, const int *sizex,
                         const int *sizey) {
  const int nc = *sizex / 4;
  int lerr = 0;
  lerr |= val(0, 0) != idx[0] / 4 + (idx[1] / 4) * nc;
  const int xm = (idx[0] - 1) < 0 ? *sizex - 1 : idx[0] - 1;
  const int xp = (idx[0] + 1) >= *sizex ? 0 : idx[0] + 1;
  const int ym = (idx[1] - 1) < 0 ? *sizey - 1 : idx[1] - 1;
  const int yp = (idx[1] + 1) >= *sizey ? 0 : idx[1] + 1;
  lerr |= val(1, 0) != xp / 4 + (idx[1] / 4) * nc;
  lerr |= val(-1, 0) != xm / 4 + (idx[1] / 4) * nc;
  lerr |= val(0, 1) != idx[0] / 4 + (yp / 4) * nc;
  lerr |= val(0, -1) != idx[0] / 4 + (ym / 4) * nc;
  *err = *err > lerr ? *err : lerr;
}

// after two restrictions by 2 of a fine field populated with x + n * y: 4 x + 4 y n
void mgrid_restrict_check(const ACC<double> &val, const int *idx, int *err, const int *sizex) {
  const int lerr = val(0, 0) != idx[0] * 4 + idx[1] * 4 * *sizex;
  *err = *err > lerr ? *err : lerr;
}

#endif
