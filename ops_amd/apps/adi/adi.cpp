// 3-D ADI heat diffusion over a batch of independent blocks (the plain-method workload for
// ops_tridMultiDimBatch). One block reproduces the reference app (apps/c/adi/adi.cpp:
// 331-421): per step a 7-point preproc loop writes the right-hand side du and the
// coefficients of the x, y and z systems, then x and y solves overwrite du and the z solve
// adds its result into u (ops_tridMultiDimBatch_Inc). Each block has its own ops_block;
// a step issues each phase for all blocks before the next phase.
//
// Initial u: mode 0 is the reference's init_kernel (1 on the boundary layer, 0 inside);
// mode 1 fills every block, x fastest and blocks in order, with U[0,1) numbers from
// splitmix64 (Steele, Lea & Flood 2014): u = (next() >> 11) * 2^-53.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#define OPS_3D
#include <ops_seq_v2.h>
#include <ops_hip_rt.h>

#include "adi_kernels.h"

namespace {

struct Blk {
  int n[3];
  ops_block blk;
  ops_dat u, du, ax, bx, cx, ay, by, cy, az, bz, cz;
};

struct State {
  std::vector<Blk> blocks;
  ops_stencil S000 = nullptr, S7 = nullptr;
  double lambda = 1.0;
  ops_tridsolver_params *ctx = nullptr;
  void *graph = nullptr;  // one whole step over all blocks, captured once and replayed
} A;

uint64_t splitmix_next(uint64_t &s) {
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void preproc(Blk &b) {
  int range[] = {0, b.n[0], 0, b.n[1], 0, b.n[2]};
  ops_par_loop(adi_preproc_kernel, "preproc_kernel", b.blk, 3, range,
               ops_arg_dat(b.u, 1, A.S7, "double", OPS_READ),
               ops_arg_dat(b.du, 1, A.S000, "double", OPS_WRITE),
               ops_arg_dat(b.ax, 1, A.S000, "double", OPS_WRITE),
               ops_arg_dat(b.bx, 1, A.S000, "double", OPS_WRITE),
               ops_arg_dat(b.cx, 1, A.S000, "double", OPS_WRITE),
               ops_arg_dat(b.ay, 1, A.S000, "double", OPS_WRITE),
               ops_arg_dat(b.by, 1, A.S000, "double", OPS_WRITE),
               ops_arg_dat(b.cy, 1, A.S000, "double", OPS_WRITE),
               ops_arg_dat(b.az, 1, A.S000, "double", OPS_WRITE),
               ops_arg_dat(b.bz, 1, A.S000, "double", OPS_WRITE),
               ops_arg_dat(b.cz, 1, A.S000, "double", OPS_WRITE), ops_arg_idx(),
               ops_arg_gbl(b.n, 3, "int", OPS_READ), ops_arg_gbl(&A.lambda, 1, "double", OPS_READ));
}

// one step of every block: the blocks are independent, so each phase is issued for all
// blocks before the next phase, which lets the runtime solve the lines of all blocks of a
// sweep together
void step_all() {
  // the blocks' preproc loops are independent (each writes only its own block's dats): they
  // go to the runtime as one loop batch (a few launches instead of one per block)
  ops_hip_batch_begin();
  for (Blk &b : A.blocks) preproc(b);
  ops_hip_batch_end();
  for (Blk &b : A.blocks) {
    int range[] = {0, b.n[0], 0, b.n[1], 0, b.n[2]};
    ops_tridMultiDimBatch(3, 0, range, b.ax, b.bx, b.cx, b.du, A.ctx);
  }
  for (Blk &b : A.blocks) {
    int range[] = {0, b.n[0], 0, b.n[1], 0, b.n[2]};
    ops_tridMultiDimBatch(3, 1, range, b.ay, b.by, b.cy, b.du, A.ctx);
  }
  for (Blk &b : A.blocks) {
    int range[] = {0, b.n[0], 0, b.n[1], 0, b.n[2]};
    ops_tridMultiDimBatch_Inc(3, 2, range, b.az, b.bz, b.cz, b.du, b.u, A.ctx);
  }
}

}  // namespace

// nblocks blocks with extents sizes[3*i .. 3*i+2]; mode 0 or 1 (see above)
extern "C" int adi_init(int argc, const char **argv, int nblocks, const int *sizes, int mode,
                        unsigned long long seed, double lambda) {
  try {
    ops_init(argc, argv, 1);
    A.blocks.clear();
    A.lambda = lambda;
    int s000[] = {0, 0, 0};
    A.S000 = ops_decl_stencil(3, 1, s000, "000");
    int s7[] = {0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1, 0, 0, 0, -1, 0, 0, 1};
    A.S7 = ops_decl_stencil(3, 7, s7, "3d7Point");
    uint64_t rng = seed;
    for (int k = 0; k < nblocks; k++) {
      Blk b;
      for (int d = 0; d < 3; d++) b.n[d] = sizes[3 * k + d];
      std::string name = "Heat3D_" + std::to_string(k);
      b.blk = ops_decl_block(3, name.c_str());
      int d_p[3] = {1, 1, 1}, d_m[3] = {-1, -1, -1}, base[3] = {0, 0, 0};
      double *none = nullptr;
      ops_dat *dats[11] = {&b.u, &b.du, &b.ax, &b.bx, &b.cx, &b.ay, &b.by, &b.cy, &b.az, &b.bz, &b.cz};
      const char *names[11] = {"h_u", "h_du", "h_ax", "h_bx", "h_cx", "h_ay",
                               "h_by", "h_cy", "h_az", "h_bz", "h_cz"};
      for (int i = 0; i < 11; i++) {
        std::string nm = std::string(names[i]) + "_" + std::to_string(k);
        *dats[i] = ops_decl_dat(b.blk, 1, b.n, base, d_m, d_p, none, "double", nm.c_str());
      }
      int range[] = {0, b.n[0], 0, b.n[1], 0, b.n[2]};
      if (mode == 0) {
        ops_par_loop(adi_init_kernel, "init_kernel", b.blk, 3, range,
                     ops_arg_dat(b.u, 1, A.S000, "double", OPS_WRITE), ops_arg_idx(),
                     ops_arg_gbl(b.n, 3, "int", OPS_READ));
      } else {
        const size_t cnt = (size_t)b.n[0] * b.n[1] * b.n[2];
        std::vector<double> v(cnt);
        for (size_t i = 0; i < cnt; i++) v[i] = (double)(splitmix_next(rng) >> 11) * (1.0 / 9007199254740992.0);
        ops_dat_set_data_slab_memspace(b.u, 0, (char *)v.data(), range, OPS_HOST);
      }
      A.blocks.push_back(b);
    }
    A.ctx = new ops_tridsolver_params(A.blocks.empty() ? nullptr : A.blocks[0].blk);
    A.graph = nullptr;
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "adi_init: %s\n", e.what());
    return 1;
  }
}

// Steps every block nsteps times. With OPS_ADI_GRAPH unset or non-zero, the first call
// runs one step eagerly (sizing every scratch buffer), then records one step of all blocks
// as a graph that later steps replay; a backend without graph
// capture falls back to eager launches.
extern "C" int adi_step(int nsteps) {
  try {
    const char *env = getenv("OPS_ADI_GRAPH");
    const bool use_graph = env == nullptr || atoi(env) != 0;
    int s = 0;
    if (use_graph && A.graph == nullptr && nsteps > 1) {
      step_all();
      s = 1;
      if (ops_hip_graph_begin() == 0) {
        try {
          step_all();
        } catch (...) {
          ops_hip_graph_free(ops_hip_graph_end());
          throw;
        }
        A.graph = ops_hip_graph_end();
      }
    }
    if (use_graph && A.graph != nullptr) {
      if (ops_hip_graph_replay(A.graph, nsteps - s) != 0)
        throw OPSException(OPS_RUNTIME_ERROR, "adi_step: graph replay failed");
      return 0;
    }
    for (; s < nsteps; s++) step_all();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "adi_step: %s\n", e.what());
    return 1;
  }
}

// packed interior of block k's u (nx*ny*nz doubles, x fastest); returns the count, and
// writes when out is not null
extern "C" long adi_fetch(int k, double *out) {
  try {
    Blk &b = A.blocks.at(k);
    const long cnt = (long)b.n[0] * b.n[1] * b.n[2];
    if (out) {
      int range[] = {0, b.n[0], 0, b.n[1], 0, b.n[2]};
      ops_dat_fetch_data_slab_memspace(b.u, 0, (char *)out, range, OPS_HOST);
    }
    return cnt;
  } catch (std::exception &e) {
    fprintf(stderr, "adi_fetch: %s\n", e.what());
    return -1;
  }
}

// overwrite block k's u interior with a packed array (x fastest)
extern "C" int adi_set(int k, const double *in) {
  try {
    Blk &b = A.blocks.at(k);
    int range[] = {0, b.n[0], 0, b.n[1], 0, b.n[2]};
    ops_dat_set_data_slab_memspace(b.u, 0, (char *)in, range, OPS_HOST);
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "adi_set: %s\n", e.what());
    return 1;
  }
}

extern "C" int adi_finalize(void) {
  try {
    delete A.ctx;
    A.ctx = nullptr;
    ops_hip_graph_free(A.graph);
    A.graph = nullptr;
    A.blocks.clear();
    ops_exit();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "adi_finalize: %s\n", e.what());
    return 1;
  }
}
