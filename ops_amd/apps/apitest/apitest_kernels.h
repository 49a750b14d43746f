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

// v += 1, then the sum of v (the graph-replayed reduction sequence)
void api_addone(ACC<double> &v) { v(0, 0) = v(0, 0) + 1.0; }
void api_vsum(const ACC<double> &v, double *sum) { *sum = *sum + v(0, 0); }

// one dat read and written at the same point through two arguments (the runtime then uses
// the kernel variant without no-alias pointers)
void api_twice_plus_one(const ACC<double> &in, ACC<double> &out) { out(0, 0) = 2.0 * in(0, 0) + 1.0; }

// o = c * i + idx[0] + 1000 * idx[1] (the loop-batch sequence: launched with `o` as WRITE
// and as RW, which the translator turns into two stubs of one kernel)
void api_bset(ACC<double> &o, const ACC<double> &i, const int *idx, const double *c) {
  o(0, 0) = c[0] * i(0, 0) + (double)idx[0] + 1000.0 * (double)idx[1];
}

// loop chains (api_chain): loops that read, along y (the tiled dimension) and x, what the
// loops before them wrote, and overwrite what earlier loops read
void api_ch_init(ACC<double> &a, ACC<double> &b, ACC<double> &c, ACC<double> &d, const int *idx) {
  const double x = (double)idx[0], y = (double)idx[1];
  a(0, 0) = 1.0 + 0.01 * x - 0.003 * y * y + 0.25 * ((idx[0] + 3 * idx[1]) % 7);
  b(0, 0) = 0.5 - 0.02 * x * y + 0.125 * ((2 * idx[0] + idx[1]) % 5);
  c(0, 0) = 0.0;
  d(0, 0) = 0.0;
}
void api_ch_l1(ACC<double> &c, const ACC<double> &a, const ACC<double> &b, const double *w) {
  c(0, 0) = a(0, -1) + w[0] * a(0, 1) + b(0, 0) + 0.25 * a(1, 0);
}
void api_ch_l2(ACC<double> &a, const ACC<double> &c) { a(0, 0) = 0.5 * c(0, 0) + c(0, 2) - 0.125 * c(-1, 0); }
void api_ch_l3(ACC<double> &d, const ACC<double> &a, const ACC<double> &c) {
  d(0, 0) = a(0, -2) - a(0, 1) + c(0, -1);
}
void api_ch_sum(const ACC<double> &d, double *s) { *s = *s + d(0, 0); }
void api_ch_l4(ACC<double> &b, const ACC<double> &d) { b(0, 0) = b(0, 0) + d(0, 1) - d(0, -1); }
void api_ch_l5(ACC<double> &c, const ACC<double> &d, const ACC<double> &b) {
  c(0, 0) = d(0, 0) + b(0, 2) - b(0, -2);
}

#endif
