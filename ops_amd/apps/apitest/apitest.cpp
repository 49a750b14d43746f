// API test application: exercises the parts of the OPS API the benchmark apps do not
// stress, each through an extern "C" entry point that runs a whole ops_init .. ops_exit
// session and returns plain arrays.
//   api_mblock     the reference's mblock test (apps/c/mblock/mblock.cpp:80-158): two
//                  20 x 20 blocks with 2-deep halos, five halo groups including a flipped
//                  axis and swapped axes
//   api_halo3d     one 3-D halo between two dats, any orientation, any (float/double)
//                  types, values written and read back through slab set/fetch
//   api_tridiag    ops_tridMultiDimBatch[_Inc] on caller-provided coefficients
//   api_reduce     INC/MIN/MAX reductions on double and int, a 2-component reduction,
//                  accumulation across two loops before ops_reduction_result
//   api_reduce_graph  a sum reduction read on the host after every pass of v += 1, sum(v):
//                  eager passes, then passes replayed from a captured graph, then eager again
//   api_random[_normal]  ops_randomgen_init + ops_fill_random_uniform / _normal on a dat, read back with
//                  its halos
//   api_tridiag2   two solves in a row (the second on a smaller box) through the queueing
//                  entry points or the allocation-free one
//   api_chain      loop chains (ops_hip_chain_begin/end): dependent loops run band by band
//   api_batch      loop batches (ops_hip_batch_begin/end): many independent launches of one
//                  kernel of mixed extents, two stubs of that kernel interleaved, a
//                  tridiagonal solve between launches, eagerly and replayed from a graph
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#define OPS_2D
#include <ops_seq_v2.h>
#include <ops_hip_rt.h>
#include <ops_tridiag.h>

#include "apitest_kernels.h"

namespace {
int fail(const char *where, const std::exception &e) {
  fprintf(stderr, "%s: %s\n", where, e.what());
  return 1;
}

// whole allocation (halos included) of a 2-D dat, x fastest
void fetch_all(ops_dat d, double *out) {
  int r[4] = {d->d_m[0], d->size[0] + d->d_m[0], d->d_m[1], d->size[1] + d->d_m[1]};
  ops_dat_fetch_data_slab_memspace(d, 0, (char *)out, r, OPS_HOST);
}
}  // namespace

// group_mask bit g runs halo group g (0..4) in that order; out0/out1 receive the 24 x 24
// allocations of data0/data1
extern "C" int api_mblock(int argc, const char **argv, int group_mask, double *out0, double *out1) {
  try {
    ops_init(argc, argv, 1);
    ops_block grid0 = ops_decl_block(2, "grid0");
    ops_block grid1 = ops_decl_block(2, "grid1");
    int s00[] = {0, 0};
    ops_stencil S00 = ops_decl_stencil(2, 1, s00, "00");
    int d_p[2] = {2, 2}, d_m[2] = {-2, -2}, size[2] = {20, 20}, base[2] = {0, 0};
    double *none = nullptr;
    ops_dat data0 = ops_decl_dat(grid0, 1, size, base, d_m, d_p, none, "double", "data0");
    ops_dat data1 = ops_decl_dat(grid1, 1, size, base, d_m, d_p, none, "double", "data1");
    ops_halo_group groups[5];
    {  // 0: x faces, same orientation
      int it[] = {2, 20}, bf[] = {18, 0}, bt[] = {-2, 0}, dir[] = {1, 2};
      ops_halo h0 = ops_decl_halo(data0, data1, it, bf, bt, dir, dir);
      bf[0] = 0;
      bt[0] = 20;
      ops_halo h1 = ops_decl_halo(data1, data0, it, bf, bt, dir, dir);
      ops_halo g[] = {h0, h1};
      groups[0] = ops_decl_halo_group(2, g);
    }
    {  // 1: y faces, same orientation
      int it[] = {20, 2}, bf[] = {0, 18}, bt[] = {0, -2}, dir[] = {1, 2};
      ops_halo h0 = ops_decl_halo(data0, data1, it, bf, bt, dir, dir);
      bf[1] = 0;
      bt[1] = 20;
      ops_halo h1 = ops_decl_halo(data1, data0, it, bf, bt, dir, dir);
      ops_halo g[] = {h0, h1};
      groups[1] = ops_decl_halo_group(2, g);
    }
    {  // 2: x faces, y flipped
      int it[] = {2, 20}, bf[] = {0, 0}, bt[] = {20, 0}, dir[] = {1, 2}, dto[] = {1, -2};
      ops_halo h0 = ops_decl_halo(data0, data1, it, bf, bt, dir, dto);
      bf[0] = 18;
      bt[0] = -2;
      ops_halo h1 = ops_decl_halo(data1, data0, it, bf, bt, dto, dir);
      ops_halo g[] = {h0, h1};
      groups[2] = ops_decl_halo_group(2, g);
    }
    {  // 3: y faces, x flipped
      int it[] = {20, 2}, bf[] = {0, 0}, bt[] = {0, 20}, dir[] = {1, 2}, dto[] = {-1, 2};
      ops_halo h0 = ops_decl_halo(data0, data1, it, bf, bt, dir, dto);
      bf[1] = 18;
      bt[1] = -2;
      ops_halo h1 = ops_decl_halo(data1, data0, it, bf, bt, dto, dir);
      ops_halo g[] = {h0, h1};
      groups[3] = ops_decl_halo_group(2, g);
    }
    {  // 4: x face of block 0 onto the y face of block 1 (axes swapped)
      int it[] = {2, 20}, bf[] = {18, 0}, bt[] = {0, -2}, dir[] = {1, 2}, dto[] = {2, 1};
      ops_halo h0 = ops_decl_halo(data0, data1, it, bf, bt, dir, dto);
      bf[0] = 0;
      bt[0] = 20;
      bt[1] = 0;
      ops_halo h1 = ops_decl_halo(data1, data0, it, bf, bt, dto, dir);
      ops_halo g[] = {h0, h1};
      groups[4] = ops_decl_halo_group(2, g);
    }
    int range[] = {0, 20, 0, 20};
    ops_par_loop(api_populate2d, "mblock_populate_kernel", grid0, 2, range,
                 ops_arg_dat(data0, 1, S00, "double", OPS_WRITE), ops_arg_idx());
    ops_par_loop(api_populate2d, "mblock_populate_kernel", grid1, 2, range,
                 ops_arg_dat(data1, 1, S00, "double", OPS_WRITE), ops_arg_idx());
    for (int g = 0; g < 5; g++)
      if (group_mask & (1 << g)) ops_halo_transfer(groups[g]);
    fetch_all(data0, out0);
    fetch_all(data1, out1);
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    return fail("api_mblock", e);
  }
}

// One 3-D halo from dat A to dat B. sizes/d_m/d_p give each dat (3 ints each, A then B);
// tA/tB are 0 for double, 1 for float. inA is A's whole allocation (x fastest, as doubles);
// outB receives B's whole allocation after the transfer (B starts at zero).
extern "C" int api_halo3d(int argc, const char **argv, const int *sizeA, const int *dmA,
                          const int *dpA, int tA, const int *sizeB, const int *dmB, const int *dpB,
                          int tB, const int *iter, const int *from_base, const int *to_base,
                          const int *from_dir, const int *to_dir, const double *inA, double *outB) {
  try {
    ops_init(argc, argv, 1);
    ops_block blk = ops_decl_block(3, "b3");
    int base[3] = {0, 0, 0};
    int s[3], m[3], p[3];
    for (int i = 0; i < 3; i++) {
      s[i] = sizeA[i];
      m[i] = dmA[i];
      p[i] = dpA[i];
    }
    ops_dat A = tA ? ops_decl_dat_char(blk, 1, s, base, m, p, nullptr, nullptr, 4, "float", "A")
                   : ops_decl_dat_char(blk, 1, s, base, m, p, nullptr, nullptr, 8, "double", "A");
    for (int i = 0; i < 3; i++) {
      s[i] = sizeB[i];
      m[i] = dmB[i];
      p[i] = dpB[i];
    }
    ops_dat B = tB ? ops_decl_dat_char(blk, 1, s, base, m, p, nullptr, nullptr, 4, "float", "B")
                   : ops_decl_dat_char(blk, 1, s, base, m, p, nullptr, nullptr, 8, "double", "B");
    long nA = (long)A->size[0] * A->size[1] * A->size[2];
    long nB = (long)B->size[0] * B->size[1] * B->size[2];
    int rA[6], rB[6];
    for (int i = 0; i < 3; i++) {
      rA[2 * i] = A->d_m[i];
      rA[2 * i + 1] = A->size[i] + A->d_m[i];
      rB[2 * i] = B->d_m[i];
      rB[2 * i + 1] = B->size[i] + B->d_m[i];
    }
    if (tA) {
      std::vector<float> f(nA);
      for (long i = 0; i < nA; i++) f[i] = (float)inA[i];
      ops_dat_set_data_slab_memspace(A, 0, (char *)f.data(), rA, OPS_HOST);
    } else {
      ops_dat_set_data_slab_memspace(A, 0, (char *)inA, rA, OPS_HOST);
    }
    int it[3], bf[3], bt[3], fd[3], td[3];
    for (int i = 0; i < 3; i++) {
      it[i] = iter[i];
      bf[i] = from_base[i];
      bt[i] = to_base[i];
      fd[i] = from_dir[i];
      td[i] = to_dir[i];
    }
    ops_halo h = ops_decl_halo(A, B, it, bf, bt, fd, td);
    ops_halo_group g = ops_decl_halo_group(1, &h);
    ops_halo_transfer(g);
    if (tB) {
      std::vector<float> f(nB);
      ops_dat_fetch_data_slab_memspace(B, 0, (char *)f.data(), rB, OPS_HOST);
      for (long i = 0; i < nB; i++) outB[i] = f[i];
    } else {
      ops_dat_fetch_data_slab_memspace(B, 0, (char *)outB, rB, OPS_HOST);
    }
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    return fail("api_halo3d", e);
  }
}

// Tridiagonal batch on ndim-D dats of interior size `size` (one halo layer all round);
// a, b, c, d, u are packed interiors (x fastest); the solve box is `range` (interior
// coordinates). out receives d (inc = 0) or u (inc = 1) after the solve, packed.
extern "C" int api_tridiag(int argc, const char **argv, int ndim, int solvedim, const int *size,
                           const int *range, const double *a, const double *b, const double *c,
                           const double *d, const double *u, int inc, double *out) {
  try {
    ops_init(argc, argv, 1);
    ops_block blk = ops_decl_block(ndim, "trid");
    int sz[3], base[3] = {0, 0, 0}, dm[3] = {-1, -1, -1}, dp[3] = {1, 1, 1}, full[6];
    for (int i = 0; i < ndim; i++) {
      sz[i] = size[i];
      full[2 * i] = 0;
      full[2 * i + 1] = size[i];
    }
    double *none = nullptr;
    ops_dat A = ops_decl_dat(blk, 1, sz, base, dm, dp, none, "double", "a");
    ops_dat B = ops_decl_dat(blk, 1, sz, base, dm, dp, none, "double", "b");
    ops_dat C = ops_decl_dat(blk, 1, sz, base, dm, dp, none, "double", "c");
    ops_dat D = ops_decl_dat(blk, 1, sz, base, dm, dp, none, "double", "d");
    ops_dat U = ops_decl_dat(blk, 1, sz, base, dm, dp, none, "double", "u");
    ops_dat_set_data_slab_memspace(A, 0, (char *)a, full, OPS_HOST);
    ops_dat_set_data_slab_memspace(B, 0, (char *)b, full, OPS_HOST);
    ops_dat_set_data_slab_memspace(C, 0, (char *)c, full, OPS_HOST);
    ops_dat_set_data_slab_memspace(D, 0, (char *)d, full, OPS_HOST);
    ops_dat_set_data_slab_memspace(U, 0, (char *)u, full, OPS_HOST);
    int r[6];
    for (int i = 0; i < 2 * ndim; i++) r[i] = range[i];
    ops_tridsolver_params ctx(blk);
    if (inc)
      ops_tridMultiDimBatch_Inc(ndim, solvedim, r, A, B, C, D, U, &ctx);
    else
      ops_tridMultiDimBatch(ndim, solvedim, r, A, B, C, D, &ctx);
    ops_dat_fetch_data_slab_memspace(inc ? U : D, 0, (char *)out, full, OPS_HOST);
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    return fail("api_tridiag", e);
  }
}

// Reductions over the nx x ny interior: vals (double) and ivals (int), packed. The loop
// runs twice over the two halves (rows [0, ny/2) and [ny/2, ny)) before one
// ops_reduction_result per handle. out: sum, min, max, isum, imin, imax (as doubles),
// pair[0], pair[1].
// out[0 .. nx*ny): v after v = 2 v + 1 with v passed as both the read and the written
// argument; out[nx*ny ..): u = 2 v + 1 into a second dat
extern "C" int api_alias(int argc, const char **argv, int nx, int ny, const double *vals, double *out) {
  try {
    ops_init(argc, argv, 1);
    ops_block blk = ops_decl_block(2, "alias");
    int size[2] = {nx, ny}, base[2] = {0, 0}, dm[2] = {0, 0}, dp[2] = {0, 0};
    double *none = nullptr;
    ops_dat V = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "v");
    ops_dat U = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "u");
    int full[4] = {0, nx, 0, ny};
    ops_dat_set_data_slab_memspace(V, 0, (char *)vals, full, OPS_HOST);
    int s00[] = {0, 0};
    ops_stencil S00 = ops_decl_stencil(2, 1, s00, "00");
    ops_par_loop(api_twice_plus_one, "api_twice_plus_one", blk, 2, full,
                 ops_arg_dat(V, 1, S00, "double", OPS_READ), ops_arg_dat(V, 1, S00, "double", OPS_WRITE));
    ops_par_loop(api_twice_plus_one, "api_twice_plus_one", blk, 2, full,
                 ops_arg_dat(V, 1, S00, "double", OPS_READ), ops_arg_dat(U, 1, S00, "double", OPS_WRITE));
    ops_dat_fetch_data_slab_memspace(V, 0, (char *)out, full, OPS_HOST);
    ops_dat_fetch_data_slab_memspace(U, 0, (char *)(out + (long)nx * ny), full, OPS_HOST);
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    return fail("api_alias", e);
  }
}

extern "C" int api_reduce(int argc, const char **argv, int nx, int ny, const double *vals,
                          const int *ivals, double *out) {
  try {
    ops_init(argc, argv, 1);
    ops_block blk = ops_decl_block(2, "red");
    int size[2] = {nx, ny}, base[2] = {0, 0}, dm[2] = {-1, -1}, dp[2] = {1, 1};
    double *none = nullptr;
    int *inone = nullptr;
    ops_dat V = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "v");
    ops_dat W = ops_decl_dat(blk, 1, size, base, dm, dp, inone, "int", "w");
    int full[4] = {0, nx, 0, ny};
    ops_dat_set_data_slab_memspace(V, 0, (char *)vals, full, OPS_HOST);
    ops_dat_set_data_slab_memspace(W, 0, (char *)ivals, full, OPS_HOST);
    int s00[] = {0, 0};
    ops_stencil S00 = ops_decl_stencil(2, 1, s00, "00");
    ops_reduction hs = ops_decl_reduction_handle(sizeof(double), "double", "sum");
    ops_reduction hmn = ops_decl_reduction_handle(sizeof(double), "double", "min");
    ops_reduction hmx = ops_decl_reduction_handle(sizeof(double), "double", "max");
    ops_reduction his = ops_decl_reduction_handle(sizeof(int), "int", "isum");
    ops_reduction himn = ops_decl_reduction_handle(sizeof(int), "int", "imin");
    ops_reduction himx = ops_decl_reduction_handle(sizeof(int), "int", "imax");
    ops_reduction hp = ops_decl_reduction_handle(2 * sizeof(double), "double", "pair");
    for (int half = 0; half < 2; half++) {
      int r[4] = {0, nx, half ? ny / 2 : 0, half ? ny : ny / 2};
      ops_par_loop(api_reduce, "api_reduce", blk, 2, r, ops_arg_dat(V, 1, S00, "double", OPS_READ),
                   ops_arg_dat(W, 1, S00, "int", OPS_READ), ops_arg_reduce(hs, 1, "double", OPS_INC),
                   ops_arg_reduce(hmn, 1, "double", OPS_MIN), ops_arg_reduce(hmx, 1, "double", OPS_MAX),
                   ops_arg_reduce(his, 1, "int", OPS_INC), ops_arg_reduce(himn, 1, "int", OPS_MIN),
                   ops_arg_reduce(himx, 1, "int", OPS_MAX), ops_arg_reduce(hp, 2, "double", OPS_INC));
    }
    double s, mn, mx, pair[2];
    int is, imn, imx;
    ops_reduction_result(hs, &s);
    ops_reduction_result(hmn, &mn);
    ops_reduction_result(hmx, &mx);
    ops_reduction_result(his, &is);
    ops_reduction_result(himn, &imn);
    ops_reduction_result(himx, &imx);
    ops_reduction_result(hp, pair);
    out[0] = s;
    out[1] = mn;
    out[2] = mx;
    out[3] = is;
    out[4] = imn;
    out[5] = imx;
    out[6] = pair[0];
    out[7] = pair[1];
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    return fail("api_reduce", e);
  }
}

// out[k] = the sum read after pass k of {v += 1; sum(v)}: passes 0 and 1 eager, 2 .. passes - 3
// replayed from a graph captured from that pair of loops (eager where the build has no
// graphs), the last two eager again
extern "C" int api_reduce_graph(int argc, const char **argv, int nx, int ny, const double *vals, int passes,
                                double *out) {
  try {
    ops_init(argc, argv, 1);
    ops_block blk = ops_decl_block(2, "redg");
    int size[2] = {nx, ny}, base[2] = {0, 0}, dm[2] = {-1, -1}, dp[2] = {1, 1};
    double *none = nullptr;
    ops_dat V = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "v");
    int full[4] = {0, nx, 0, ny};
    ops_dat_set_data_slab_memspace(V, 0, (char *)vals, full, OPS_HOST);
    int s00[] = {0, 0};
    ops_stencil S00 = ops_decl_stencil(2, 1, s00, "00");
    ops_reduction hs = ops_decl_reduction_handle(sizeof(double), "double", "vsum");
    auto pass = [&]() {
      ops_par_loop(api_addone, "api_addone", blk, 2, full, ops_arg_dat(V, 1, S00, "double", OPS_RW));
      ops_par_loop(api_vsum, "api_vsum", blk, 2, full, ops_arg_dat(V, 1, S00, "double", OPS_READ),
                   ops_arg_reduce(hs, 1, "double", OPS_INC));
    };
    void *g = nullptr;
    for (int k = 0; k < passes; k++) {
      const bool replay = k >= 2 && k < passes - 2;
      if (replay && !g) {
        // the capture records the pair (it does not run it); the first replay is pass k
        if (ops_hip_graph_begin() == 0) {
          pass();
          g = ops_hip_graph_end();
        }
      }
      if (replay && g) {
        if (ops_hip_graph_replay(g, 1) != 0) throw OPSException(OPS_RUNTIME_ERROR, "replay failed");
      } else {
        pass();
      }
      ops_reduction_result(hs, &out[k]);
    }
    if (g) ops_hip_graph_free(g);
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    return fail("api_reduce_graph", e);
  }
}

// ops_fill_random_uniform (normal = 0) or ops_fill_random_normal (normal = 1) on an
// nx x ny double dat with one halo layer; out receives the whole allocation
// ((nx+2)*(ny+2), x fastest)
static int random_fill(int argc, const char **argv, unsigned seed, int nx, int ny, double *out,
                       int normal) {
  try {
    ops_init(argc, argv, 1);
    ops_block blk = ops_decl_block(2, "rng");
    int size[2] = {nx, ny}, base[2] = {0, 0}, dm[2] = {-1, -1}, dp[2] = {1, 1};
    double *none = nullptr;
    ops_dat V = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "v");
    ops_randomgen_init(seed, 0);
    if (normal) ops_fill_random_normal(V);
    else ops_fill_random_uniform(V);
    fetch_all(V, out);
    ops_randomgen_exit();
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    return fail(normal ? "api_random_normal" : "api_random", e);
  }
}

extern "C" int api_random(int argc, const char **argv, unsigned seed, int nx, int ny, double *out) {
  return random_fill(argc, argv, seed, nx, ny, out, 0);
}

extern "C" int api_random_normal(int argc, const char **argv, unsigned seed, int nx, int ny,
                                 double *out) {
  return random_fill(argc, argv, seed, nx, ny, out, 1);
}

// Loop batches. nloops dats O_j (extents x 3 + 37j mod 70, y 2 + 11j mod 9: narrow and wide
// launches) are set to O_j = c_j S_j + x + 1000 y from sources S_j = j + x / 7 - y / 3 by
// api_bset, alternately through its WRITE stub (even j) and its RW stub (odd j), in one
// sequence. mode bit 0: the sequence runs between ops_hip_batch_begin and _end; bit 1: a
// tridiagonal solve along x of O_1 (coefficients a = c = -1, b = 4) is issued after the
// launch of O_1 and before the later launches (the solve reads what the batched launch
// wrote); bit 2: the sequence is captured as a graph (after one eager run) and replayed
// twice, O_j cleared in between, the second replay's result returned. out: every O_j's
// interior, x fastest, j after j.
extern "C" int api_batch(int argc, const char **argv, int nloops, int mode, double *out) {
  try {
    ops_init(argc, argv, 1);
    ops_block blk = ops_decl_block(2, "batch");
    int base[2] = {0, 0}, dm[2] = {-1, -1}, dp[2] = {1, 1};
    double *none = nullptr;
    std::vector<ops_dat> S, O;
    std::vector<std::vector<int>> sizes;
    for (int j = 0; j < nloops; j++) {
      int size[2] = {3 + (37 * j) % 70, 2 + (11 * j) % 9};
      sizes.push_back({size[0], size[1]});
      S.push_back(ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "s"));
      O.push_back(ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "o"));
      std::vector<double> v((size_t)size[0] * size[1]);
      for (int y = 0; y < size[1]; y++)
        for (int x = 0; x < size[0]; x++) v[(size_t)y * size[0] + x] = j + x / 7.0 - y / 3.0;
      int full[4] = {0, size[0], 0, size[1]};
      ops_dat_set_data_slab_memspace(S[j], 0, (char *)v.data(), full, OPS_HOST);
    }
    int s00[] = {0, 0};
    ops_stencil S00 = ops_decl_stencil(2, 1, s00, "00");
    ops_dat TA = nullptr, TB = nullptr, TC = nullptr;
    const bool solve = (mode & 2) && nloops > 1;
    if (solve) {
      int size[2] = {sizes[1][0], sizes[1][1]};
      TA = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "ta");
      TB = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "tb");
      TC = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "tc");
      std::vector<double> m1((size_t)size[0] * size[1], -1.0), m4((size_t)size[0] * size[1], 4.0);
      int full[4] = {0, size[0], 0, size[1]};
      ops_dat_set_data_slab_memspace(TA, 0, (char *)m1.data(), full, OPS_HOST);
      ops_dat_set_data_slab_memspace(TB, 0, (char *)m4.data(), full, OPS_HOST);
      ops_dat_set_data_slab_memspace(TC, 0, (char *)m1.data(), full, OPS_HOST);
    }
    std::vector<double> cs(nloops);
    for (int j = 0; j < nloops; j++) cs[j] = 0.5 + j;
    ops_tridsolver_params ctx(blk);
    auto sequence = [&]() {
      if (mode & 1) ops_hip_batch_begin();
      for (int j = 0; j < nloops; j++) {
        int r[4] = {0, sizes[j][0], 0, sizes[j][1]};
        if (j % 2 == 0)
          ops_par_loop(api_bset, "api_bset", blk, 2, r, ops_arg_dat(O[j], 1, S00, "double", OPS_WRITE),
                       ops_arg_dat(S[j], 1, S00, "double", OPS_READ), ops_arg_idx(),
                       ops_arg_gbl(&cs[j], 1, "double", OPS_READ));
        else
          ops_par_loop(api_bset, "api_bset", blk, 2, r, ops_arg_dat(O[j], 1, S00, "double", OPS_RW),
                       ops_arg_dat(S[j], 1, S00, "double", OPS_READ), ops_arg_idx(),
                       ops_arg_gbl(&cs[j], 1, "double", OPS_READ));
        if (solve && j == 1) ops_tridMultiDimBatch(2, 0, r, TA, TB, TC, O[1], &ctx);
      }
      if (mode & 1) ops_hip_batch_end();
    };
    auto clear = [&]() {
      for (int j = 0; j < nloops; j++) {
        std::vector<double> z((size_t)sizes[j][0] * sizes[j][1], 0.0);
        int full[4] = {0, sizes[j][0], 0, sizes[j][1]};
        ops_dat_set_data_slab_memspace(O[j], 0, (char *)z.data(), full, OPS_HOST);
      }
    };
    sequence();
    if (mode & 4) {
      void *g = nullptr;
      if (ops_hip_graph_begin() == 0) {
        sequence();
        g = ops_hip_graph_end();
      }
      if (g) {
        for (int rep = 0; rep < 2; rep++) {
          clear();
          if (ops_hip_graph_replay(g, 1) != 0) throw OPSException(OPS_RUNTIME_ERROR, "replay failed");
          ops_hip_device_sync();
        }
        ops_hip_graph_free(g);
      } else {
        clear();
        sequence();
      }
    }
    size_t off = 0;
    for (int j = 0; j < nloops; j++) {
      int full[4] = {0, sizes[j][0], 0, sizes[j][1]};
      ops_dat_fetch_data_slab_memspace(O[j], 0, (char *)(out + off), full, OPS_HOST);
      off += (size_t)sizes[j][0] * sizes[j][1];
    }
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    return fail("api_batch", e);
  }
}

// Loop chains. Dats A, B, C, D of nx x ny cells with two halo layers: A and B set from
// their indices, then five loops (api_ch_l1..l5) over different row ranges that read along
// y up to two rows away what earlier loops wrote (read after write) and overwrite what
// earlier loops read two rows away (write after read), once more for C (write after
// write). mode bit 0: the five loops run between ops_hip_chain_begin(tiles) and
// ops_hip_chain_end; bit 1: a sum reduction of D runs between l3 and l4 (a loop a chain
// cannot hold: the chain runs first, recording resumes after it); bit 2: the weight l1
// reads through ops_arg_gbl is changed by the caller after the call (a recorded loop keeps
// the value of its call). out: A, B, C, D interiors (x fastest), then the sum.
extern "C" int api_chain(int argc, const char **argv, int nx, int ny, int tiles, int mode,
                         double *out) {
  try {
    ops_init(argc, argv, 1);
    ops_block blk = ops_decl_block(2, "chain");
    int size[2] = {nx, ny}, base[2] = {0, 0}, dm[2] = {-2, -2}, dp[2] = {2, 2};
    double *none = nullptr;
    ops_dat A = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "A");
    ops_dat B = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "B");
    ops_dat C = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "C");
    ops_dat D = ops_decl_dat(blk, 1, size, base, dm, dp, none, "double", "D");
    int s00[] = {0, 0}, s_y1[] = {0, 0, 0, -1, 0, 1, 1, 0}, s_y2[] = {0, 0, 0, 2, -1, 0};
    int s_a2[] = {0, -2, 0, 1}, s_c1[] = {0, -1}, s_b2[] = {0, 2, 0, -2}, s_d1[] = {0, 1, 0, -1};
    ops_stencil S00 = ops_decl_stencil(2, 1, s00, "00");
    ops_stencil SY1 = ops_decl_stencil(2, 4, s_y1, "a_l1");
    ops_stencil SY2 = ops_decl_stencil(2, 3, s_y2, "c_l2");
    ops_stencil SA2 = ops_decl_stencil(2, 2, s_a2, "a_l3");
    ops_stencil SC1 = ops_decl_stencil(2, 1, s_c1, "c_l3");
    ops_stencil SB2 = ops_decl_stencil(2, 2, s_b2, "b_l5");
    ops_stencil SD1 = ops_decl_stencil(2, 2, s_d1, "d_l4");
    ops_reduction sum = ops_decl_reduction_handle(sizeof(double), "double", "sum");
    int full[4] = {-2, nx + 2, -2, ny + 2};
    ops_par_loop(api_ch_init, "api_ch_init", blk, 2, full, ops_arg_dat(A, 1, S00, "double", OPS_WRITE),
                 ops_arg_dat(B, 1, S00, "double", OPS_WRITE), ops_arg_dat(C, 1, S00, "double", OPS_WRITE),
                 ops_arg_dat(D, 1, S00, "double", OPS_WRITE), ops_arg_idx());
    double w = 2.0, total = 0.0;
    if (mode & 1) ops_hip_chain_begin(tiles);
    int r1[4] = {0, nx, 0, ny}, r2[4] = {0, nx, 1, ny}, r3[4] = {0, nx, 0, ny - 1};
    int r4[4] = {0, nx, 2, ny}, r5[4] = {0, nx, 0, ny};
    ops_par_loop(api_ch_l1, "api_ch_l1", blk, 2, r1, ops_arg_dat(C, 1, S00, "double", OPS_WRITE),
                 ops_arg_dat(A, 1, SY1, "double", OPS_READ), ops_arg_dat(B, 1, S00, "double", OPS_READ),
                 ops_arg_gbl(&w, 1, "double", OPS_READ));
    if (mode & 4) w = -7.0;
    ops_par_loop(api_ch_l2, "api_ch_l2", blk, 2, r2, ops_arg_dat(A, 1, S00, "double", OPS_WRITE),
                 ops_arg_dat(C, 1, SY2, "double", OPS_READ));
    ops_par_loop(api_ch_l3, "api_ch_l3", blk, 2, r3, ops_arg_dat(D, 1, S00, "double", OPS_WRITE),
                 ops_arg_dat(A, 1, SA2, "double", OPS_READ), ops_arg_dat(C, 1, SC1, "double", OPS_READ));
    if (mode & 2)
      ops_par_loop(api_ch_sum, "api_ch_sum", blk, 2, r1, ops_arg_dat(D, 1, S00, "double", OPS_READ),
                   ops_arg_reduce(sum, 1, "double", OPS_INC));
    ops_par_loop(api_ch_l4, "api_ch_l4", blk, 2, r4, ops_arg_dat(B, 1, S00, "double", OPS_RW),
                 ops_arg_dat(D, 1, SD1, "double", OPS_READ));
    ops_par_loop(api_ch_l5, "api_ch_l5", blk, 2, r5, ops_arg_dat(C, 1, S00, "double", OPS_WRITE),
                 ops_arg_dat(D, 1, S00, "double", OPS_READ), ops_arg_dat(B, 1, SB2, "double", OPS_READ));
    if (mode & 1) ops_hip_chain_end();
    if (mode & 2) ops_reduction_result(sum, &total);
    int in[4] = {0, nx, 0, ny};
    const size_t m = (size_t)nx * ny;
    ops_dat all[4] = {A, B, C, D};
    for (int k = 0; k < 4; k++) ops_dat_fetch_data_slab_memspace(all[k], 0, (char *)(out + k * m), in, OPS_HOST);
    out[4 * m] = total;
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    return fail("api_chain", e);
  }
}

// Two solves in a row on the same dats: the box range1, then (coefficients and right-hand
// side reset to the inputs) the box range2; mode 0 through ops_tridMultiDimBatch[_Inc],
// mode 1 through ops_tridMultiDimBatch_noalloc. out1, out2 receive d (or u with inc) after
// each solve, whole interior, x fastest.
extern "C" int api_tridiag2(int argc, const char **argv, int ndim, int solvedim, const int *size,
                            const int *range1, const int *range2, const double *a, const double *b,
                            const double *c, const double *d, const double *u, int inc, int mode,
                            double *out1, double *out2) {
  try {
    ops_init(argc, argv, 1);
    ops_block blk = ops_decl_block(ndim, "trid2");
    int sz[3], base[3] = {0, 0, 0}, dm[3] = {-1, -1, -1}, dp[3] = {1, 1, 1}, full[6];
    for (int i = 0; i < ndim; i++) {
      sz[i] = size[i];
      full[2 * i] = 0;
      full[2 * i + 1] = size[i];
    }
    double *none = nullptr;
    ops_dat A = ops_decl_dat(blk, 1, sz, base, dm, dp, none, "double", "a");
    ops_dat B = ops_decl_dat(blk, 1, sz, base, dm, dp, none, "double", "b");
    ops_dat C = ops_decl_dat(blk, 1, sz, base, dm, dp, none, "double", "c");
    ops_dat D = ops_decl_dat(blk, 1, sz, base, dm, dp, none, "double", "d");
    ops_dat U = ops_decl_dat(blk, 1, sz, base, dm, dp, none, "double", "u");
    ops_tridsolver_params ctx(blk);
    const int *ranges[2] = {range1, range2};
    double *outs[2] = {out1, out2};
    for (int k = 0; k < 2; k++) {
      ops_dat_set_data_slab_memspace(A, 0, (char *)a, full, OPS_HOST);
      ops_dat_set_data_slab_memspace(B, 0, (char *)b, full, OPS_HOST);
      ops_dat_set_data_slab_memspace(C, 0, (char *)c, full, OPS_HOST);
      ops_dat_set_data_slab_memspace(D, 0, (char *)d, full, OPS_HOST);
      ops_dat_set_data_slab_memspace(U, 0, (char *)u, full, OPS_HOST);
      int r[6];
      for (int i = 0; i < 2 * ndim; i++) r[i] = ranges[k][i];
      if (mode == 1)
        ops_tridMultiDimBatch_noalloc(ndim, solvedim, r, A, B, C, D, inc ? U : nullptr, &ctx);
      else if (inc)
        ops_tridMultiDimBatch_Inc(ndim, solvedim, r, A, B, C, D, U, &ctx);
      else
        ops_tridMultiDimBatch(ndim, solvedim, r, A, B, C, D, &ctx);
      ops_dat_fetch_data_slab_memspace(inc ? U : D, 0, (char *)outs[k], full, OPS_HOST);
    }
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    return fail("api_tridiag2", e);
  }
}
