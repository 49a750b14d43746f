// Kernels of the low-dimensional reduction test (the reference's apps/c/lowdim_test:
// reductions of a 3-D dat onto planes and lines through strided stencils with
// ACC::combine_max/min/inc, and reads of the reduced dats broadcast back over 3-D).
#ifndef LOWDIM_KERNELS_H
#define LOWDIM_KERNELS_H

// integer-valued field (sums of it are exact in FP64 in any order), both signs
void lowdim_set(ACC<double> &v, const int *idx) {
  v(0, 0, 0) = (double)((idx[0] * 37 + idx[1] * 101 + idx[2] * 211) % 1009 - 500);
}

void lowdim_max(const ACC<double> &v, ACC<double> &xy, ACC<double> &yz, ACC<double> &xz, ACC<double> &x,
                ACC<double> &y, ACC<double> &z) {
  const double a = v(0, 0, 0);
  xy.combine_max(0, 0, 0, a);
  yz.combine_max(0, 0, 0, a);
  xz.combine_max(0, 0, 0, a);
  x.combine_max(0, 0, 0, a);
  y.combine_max(0, 0, 0, a);
  z.combine_max(0, 0, 0, a);
}

void lowdim_min(const ACC<double> &v, ACC<double> &xy, ACC<double> &yz, ACC<double> &xz, ACC<double> &x,
                ACC<double> &y, ACC<double> &z) {
  const double a = v(0, 0, 0);
  xy.combine_min(0, 0, 0, a);
  yz.combine_min(0, 0, 0, a);
  xz.combine_min(0, 0, 0, a);
  x.combine_min(0, 0, 0, a);
  y.combine_min(0, 0, 0, a);
  z.combine_min(0, 0, 0, a);
}

void lowdim_inc(const ACC<double> &v, ACC<double> &xy, ACC<double> &yz, ACC<double> &xz, ACC<double> &x,
                ACC<double> &y, ACC<double> &z) {
  const double a = v(0, 0, 0);
  xy.combine_inc(0, 0, 0, a);
  yz.combine_inc(0, 0, 0, a);
  xz.combine_inc(0, 0, 0, a);
  x.combine_inc(0, 0, 0, a);
  y.combine_inc(0, 0, 0, a);
  z.combine_inc(0, 0, 0, a);
}

// the reduced plane and line read back over the whole 3-D range
void lowdim_bcast(const ACC<double> &xy, const ACC<double> &z, ACC<double> &out) {
  out(0, 0, 0) = xy(0, 0, 0) + z(0, 0, 0);
}

#endif
