// Kernels of the API test application.
#ifndef APITEST_KERNELS_H
#define APITEST_KERNELS_H

// the reference's mblock test value (apps/c/mblock/mblock_populate_kernel.h)
void api_populate2d(ACC<double> &val, const int *idx) { val(0, 0) = (double)(idx[0] + 20 * idx[1]); }

// reductions over a 2-D dat of doubles and one of ints: sum, min, max, and a 2-component
// sum (dim > 1)
void api_reduce(const ACC<double> &v, const ACC<int> &w, double *sum, double *mn, double *mx,
                int *isum, int *imn, int *imx, double *pair) {
  const double x = v(0, 0);
  *sum = *sum + x;
  *mn = x < *mn ? x : *mn;
  *mx = x > *mx ? x : *mx;
  const int k = w(0, 0);
  *isum = *isum + k;
  *imn = k < *imn ? k : *imn;
  *imx = k > *imx ? k : *imx;
  pair[0] = pair[0] + x * x;
  pair[1] = pair[1] + 1.0;
}

// one dat read and written at the same point through two arguments (the runtime then uses
// the kernel variant without no-alias pointers)
void api_twice_plus_one(const ACC<double> &in, ACC<double> &out) { out(0, 0) = 2.0 * in(0, 0) + 1.0; }

// o = c * i + idx[0] + 1000 * idx[1] (the loop-batch sequence: launched with `o` as WRITE
// and as RW, which the translator turns into two stubs of one kernel)
void api_bset(ACC<double> &o, const ACC<double> &i, const int *idx, const double *c) {
  o(0, 0) = c[0] * i(0, 0) + (double)idx[0] + 1000.0 * (double)idx[1];
}

#endif
