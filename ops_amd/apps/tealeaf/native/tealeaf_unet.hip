// Learned warm start for the TeaLeaf implicit solve: a 4-level 2-D U-Net trained on the GPU
// by its own forward and backward kernels (no reference counterpart; interface in
// tealeaf_unet.h).
//
// Network (channels C = {16, 32, 48, 64}, every convolution 3 x 3, zero padding):
//   enc0 = relu(conv(in))                  H   x W,   16 channels (in: 16, 4 used)
//   enc1 = relu(conv(pool(enc0)))          H/2,       32
//   enc2 = relu(conv(pool(enc1)))          H/4,       48
//   enc3 = relu(conv(pool(enc2)))          H/8,       64
//   dec2 = relu(conv([up(enc3), enc2]))    H/4,       48   (112 input channels)
//   dec1 = relu(conv([up(dec2), enc1]))    H/2,       32   (80)
//   dec0 = relu(conv([up(dec1), enc0]))    H,         16   (48)
//   out  = conv(dec0)                      H,         16   (channel 0 is the prediction)
// pool = 2 x 2 average, up = nearest-neighbour copy, [a, b] = channel concatenation (the
// decoder convolutions read both sources directly; nothing is copied).
//
// Storage: activations NHWC in BF16, weights [Cout][3][3][Cin] in BF16 (an FP32 master copy
// is updated by Adam), gradients FP32. Convolutions are implicit GEMMs on gfx950's
// v_mfma_f32_16x16x32_bf16 (FP32 accumulation). Its lane layout (A: lane l holds
// A[l % 16][8 (l / 16) + j], B: B[8 (l / 16) + j][l % 16], j = 0..7; D: D[4 (l / 16) + j]
// [l % 16], j = 0..3) was checked on the device against a host product.
//   forward / input gradient: M = pixels, N = output channels, K = (tap, input channel)
//     flattened; a wavefront covers 64 consecutive pixels (four M tiles) and all output
//     channels (up to 7 N tiles: the input gradient of the first decoder layer has 112), so
//     each weight fragment feeds four products and each input fragment up to seven;
//   weight gradient: M = output channels, N = input channels, K = 32 pixels of a row, from
//     row-chunked copies of dy and of the layer input ([b][y][x / 8][c][8]: a lane's 8
//     pixels of one channel are one 16-byte load); thin layers run all nine taps in one
//     wavefront per band of rows, the others one wavefront per tap with the nine of a band
//     in one workgroup; bands are summed by a second kernel in a fixed order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ops_hip_kernel.h"
#include "ops_hip_rt.h"
#include "tealeaf_native.h"
#include "tealeaf_unet.h"

#define CK(x) ops_hip_check((int)(x), #x, __FILE__, __LINE__)

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));   // 8 BF16 values, one 16-byte load
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));    // the MFMA operand type
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));    // a v_dot2_f32_bf16 operand
// a.lo b.lo + a.hi b.hi + c, the BF16 pairs packed in 32-bit words (gfx950's v_dot2c_f32_bf16)
__device__ __forceinline__ float dot2(unsigned a, unsigned b, float c) {
  return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf2, a), __builtin_bit_cast(bf2, b), c, false);
}
typedef uint16_t bf16;
__device__ __forceinline__ bf8 as_bf8(u4 v) { return __builtin_bit_cast(bf8, v); }

__host__ __device__ inline bf16 f2bf(float f) {
  // round to nearest even
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16 h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

constexpr int NL = 8;  // convolution layers
// layer table: input channels (source 1 = upsampled, source 2 = same resolution), output
// channels, resolution level
struct LayerSpec {
  int c1, c2, cout, level, relu;
};
constexpr LayerSpec SPEC[NL] = {
    {0, 16, 16, 0, 1},   // enc0 (input padded to 16)
    {0, 16, 32, 1, 1},   // enc1 on pool(enc0)
    {0, 32, 48, 2, 1},   // enc2 on pool(enc1)
    {0, 48, 64, 3, 1},   // enc3 on pool(enc2)
    {64, 48, 48, 2, 1},  // dec2 on [up(enc3), enc2]
    {48, 32, 32, 1, 1},  // dec1 on [up(dec2), enc1]
    {32, 16, 16, 0, 1},  // dec0 on [up(dec1), enc0]
    {0, 16, 16, 0, 0},   // out on dec0
};

// ---------------------------------------------------------------- convolution kernels

struct ConvArgs {
  const bf16 *src1;  // coarse source (H/2 x W/2 x c1), read at (y/2, x/2); may be null
  const bf16 *src2;  // same-resolution source (H x W x c2)
  int c1, c2, cout, H, W;
  const bf16 *w;     // [cout][9][c1 + c2]
  const float *bias; // [cout] or null
  bf16 *out;         // [H][W][cout] or null
  float *out32;      // [H][W][cout] (o32c0 = 0) or [H][W], channel 0 only (o32c0 = 1), or null
  int relu;
  int o32c0 = 0;
  // mode 1, the input gradient of a decoder layer routed by its [up(src1), src2] split (out
  // and out32 unused, the values stored are those out32 would hold): channels < sc1 summed
  // over each 2 x 2 block of pixels into g1 ([H/2][W/2][sc1], (a00 + a01) + (a10 + a11)), the
  // others stored into g2 ([H][W][cout - sc1]); a wavefront covers two rows x 32 pixels, so
  // one lane holds all four pixels of a block (k_conv only). (The pooling adjoint of the
  // encoder layers stays a separate pass: its read-modify-writes at the end of the
  // convolution measured slower than the streaming kernel.)
  int mode = 0, sc1 = 0;
  float *g1 = nullptr, *g2 = nullptr;
};

// the two-row pixel blocks of mode 1: wavefront wq of an image covers rows 2 Y2, 2 Y2 + 1,
// pixels xb .. xb + 31 (M tiles 0, 1 on the first row, 2, 3 on the second)
__host__ __device__ inline int split_segs(int W) { return (W + 31) / 32; }

// one output value of a plain convolution, pixel q = y W + x of image b
__device__ __forceinline__ void conv_store(const ConvArgs &a, int b, int q, int co, float v) {
  const int H = a.H, W = a.W, cout = a.cout;
  if (a.relu && v < 0.f) v = 0.f;
  const long p = (long)b * H * W + q, o = p * cout + co;
  if (a.out) a.out[o] = f2bf(v);
  if (a.out32 && !a.o32c0) a.out32[o] = v;
  if (a.out32 && a.o32c0 && co == 0) a.out32[p] = v;
}

// forward convolution (and the input-gradient convolution with transformed weights) as an
// implicit GEMM on v_mfma_f32_16x16x32_bf16: M = pixels, N = output channels, K = (tap,
// input channel) flattened, k = 9-tap index * cin + channel, in chunks of 32 (the tail
// chunk zero-padded). cin is a multiple of 16, so a lane's 8 consecutive k (one 16-byte
// load) are 8 channels of one pixel at one tap. A wavefront computes four 16-pixel M tiles
// (64 consecutive pixels of the image, row-major; any width) by NT 16-channel N tiles, so
// every weight fragment it loads feeds four products and every input fragment NT.
template <int NT, int MODE>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int b = blockIdx.y;
  const int H = a.H, W = a.W, c1 = a.c1, c2 = a.c2, cin = c1 + c2;
  const int K = 9 * cin, nkc = (K + 31) / 32;
  const int npix = H * W;
  const int q0 = (blockIdx.x * 4 + wv) * 64;  // the wavefront's first pixel (mode 0)
  const int wq = blockIdx.x * 4 + wv, nseg = split_segs(W), Y2 = wq / nseg, xb = (wq - Y2 * nseg) * 32;
  int py[4], px[4];
  bool pv[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    if (MODE == 1) {
      py[t] = 2 * Y2 + (t >> 1);
      px[t] = xb + 16 * (t & 1) + m;
      pv[t] = 2 * Y2 < H && px[t] < W;
    } else {
      const int q = q0 + 16 * t + m;
      pv[t] = q < npix;
      py[t] = q / W;
      px[t] = q - py[t] * W;
    }
  }
  f4 acc[4][NT];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int u = 0; u < NT; u++) acc[t][u] = f4{0.f, 0.f, 0.f, 0.f};
  const int Hc = H >> 1, Wc = W >> 1;
  const bf16 *wrow = a.w + (long)m * K;  // weight row of output channel m (+ 16 u)
  for (int kc = 0; kc < nkc; kc++) {
    const int k = kc * 32 + 8 * kq;
    const bool kv = k < K;
    const int tap = kv ? k / cin : 0, ci = k - tap * cin;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    u4 av[4], bv[NT];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int ys = py[t] + dy, xs = px[t] + dx;
      av[t] = u4{0u, 0u, 0u, 0u};
      if (kv && pv[t] && ys >= 0 && ys < H && xs >= 0 && xs < W) {
        const bf16 *src = ci < c1 ? a.src1 + (((long)b * Hc + (ys >> 1)) * Wc + (xs >> 1)) * c1 + ci
                                  : a.src2 + (((long)b * H + ys) * W + xs) * c2 + (ci - c1);
        av[t] = *(const u4 *)src;
      }
    }
#pragma unroll
    for (int u = 0; u < NT; u++) bv[u] = kv ? *(const u4 *)(wrow + (long)16 * u * K + k) : u4{0u, 0u, 0u, 0u};
#pragma unroll
    for (int u = 0; u < NT; u++)
#pragma unroll
      for (int t = 0; t < 4; t++)
        acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(av[t]), as_bf8(bv[u]), acc[t][u], 0, 0, 0);
  }
  // D[pixel 4 kq + j of the tile][channel m of the N tile]
  const int cout = a.cout;
  if (MODE == 1) {
    if (2 * Y2 >= H) return;
    const int sc1 = a.sc1, sc2 = cout - sc1, Hc = H >> 1, Wc = W >> 1;
#pragma unroll
    for (int u = 0; u < NT; u++) {
      const int co = 16 * u + m;  // sc1 is a multiple of 16: a tile is all g1 or all g2
      const float bia = a.bias ? a.bias[co] : 0.f;
      if (16 * u < sc1) {
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
          for (int jj = 0; jj < 2; jj++) {
            const int x = xb + 16 * t + 4 * kq + 2 * jj;  // even; W is even
            if (x >= W) continue;
            const float v = ((acc[t][u][2 * jj] + bia) + (acc[t][u][2 * jj + 1] + bia)) +
                            ((acc[t + 2][u][2 * jj] + bia) + (acc[t + 2][u][2 * jj + 1] + bia));
            a.g1[(((long)b * Hc + Y2) * Wc + (x >> 1)) * sc1 + co] = v;
          }
      } else {
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int x = xb + 16 * (t & 1) + 4 * kq + j;
            if (x >= W) continue;
            a.g2[(((long)b * H + 2 * Y2 + (t >> 1)) * W + x) * sc2 + (co - sc1)] = acc[t][u][j] + bia;
          }
      }
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < NT; u++) {
    const int co = 16 * u + m;
    const float bia = a.bias ? a.bias[co] : 0.f;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int q = q0 + 16 * t + 4 * kq + j;
        if (q >= npix) continue;
        conv_store(a, b, q, co, acc[t][u][j] + bia);
      }
  }
}

// The same convolution for layers with few input channels (cin <= 32) and output tiles
// (NT <= 2): a workgroup computes a tile of CT_ROWS rows x 64 pixels and first copies the
// input rows y0 - 1 .. y0 + CT_ROWS, pixels x0 - 1 .. x0 + 64 (zero outside the image) into
// LDS, one 16-byte load per 8 channels, so the nine taps read each input value from LDS
// instead of nine times through the vector-memory path (the bottleneck of these layers);
// wavefront w computes row y0 + w. With more input channels the tile's LDS leaves room for
// too few workgroups per CU, and the direct kernel above measured faster.
constexpr int CT_ROWS = 4;
constexpr int CT_PX = 66;                         // 64 pixels and the two neighbours
__host__ __device__ inline int ct_pitch(int cin) { return cin + 8; }  // LDS pixel pitch (bf16)
inline size_t ct_lds_bytes(int cin) { return (size_t)(CT_ROWS + 2) * CT_PX * ct_pitch(cin) * 2; }

template <int NT>
__global__ __launch_bounds__(256) void k_conv_lds(ConvArgs a) {
  extern __shared__ __align__(16) bf16 tile[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int b = blockIdx.z, y0 = blockIdx.y * CT_ROWS, x0 = blockIdx.x * 64;
  const int H = a.H, W = a.W, c1 = a.c1, c2 = a.c2, cin = c1 + c2, cp = ct_pitch(cin);
  const int K = 9 * cin, nkc = (K + 31) / 32;
  const int Hc = H >> 1, Wc = W >> 1;
  // the input tile: (row r, pixel px, channels 8 c8 .. 8 c8 + 7) per 16-byte unit
  const int c8n = cin >> 3, units = (CT_ROWS + 2) * CT_PX * c8n;
  for (int e = threadIdx.x; e < units; e += 256) {
    const int c8 = e % c8n, rp = e / c8n, px = rp % CT_PX, r = rp / CT_PX;
    const int ys = y0 - 1 + r, xs = x0 - 1 + px, ci = 8 * c8;
    u4 v = u4{0u, 0u, 0u, 0u};
    if (ys >= 0 && ys < H && xs >= 0 && xs < W)
      v = ci < c1 ? *(const u4 *)(a.src1 + (((long)b * Hc + (ys >> 1)) * Wc + (xs >> 1)) * c1 + ci)
                  : *(const u4 *)(a.src2 + (((long)b * H + ys) * W + xs) * c2 + (ci - c1));
    *(u4 *)(tile + (r * CT_PX + px) * cp + ci) = v;
  }
  __syncthreads();
  f4 acc[4][NT];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int u = 0; u < NT; u++) acc[t][u] = f4{0.f, 0.f, 0.f, 0.f};
  const bf16 *wrow = a.w + (long)m * K;  // weight row of output channel m (+ 16 u)
  for (int kc = 0; kc < nkc; kc++) {
    const int k = kc * 32 + 8 * kq;
    const bool kv = k < K;
    const int tap = kv ? k / cin : 0, ci = k - tap * cin;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    u4 av[4], bv[NT];
    // tile row wv + 1 + dy, pixel 16 t + m + 1 + dx: always inside the tile
    const bf16 *arow = tile + ((wv + 1 + dy) * CT_PX + m + 1 + dx) * cp + ci;
#pragma unroll
    for (int t = 0; t < 4; t++) av[t] = kv ? *(const u4 *)(arow + 16 * t * cp) : u4{0u, 0u, 0u, 0u};
#pragma unroll
    for (int u = 0; u < NT; u++) bv[u] = kv ? *(const u4 *)(wrow + (long)16 * u * K + k) : u4{0u, 0u, 0u, 0u};
#pragma unroll
    for (int u = 0; u < NT; u++)
#pragma unroll
      for (int t = 0; t < 4; t++)
        acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(av[t]), as_bf8(bv[u]), acc[t][u], 0, 0, 0);
  }
  // D[pixel 4 kq + j of the tile][channel m of the N tile]
  const int y = y0 + wv;
  if (y >= H) return;
#pragma unroll
  for (int u = 0; u < NT; u++) {
    const int co = 16 * u + m;
    const float bia = a.bias ? a.bias[co] : 0.f;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int x = x0 + 16 * t + 4 * kq + j;
        if (x >= W) continue;
        conv_store(a, b, y * W + x, co, acc[t][u][j] + bia);
      }
  }
}

// row-chunked copy of a layer input for the weight gradient: dst[b][y][xc][c][8] holds
// pixels 8 xc .. 8 xc + 7 of channel c of row y, so a lane's 8 consecutive pixels of one
// channel are one 16-byte load and the 16 channels of an operand tile are 256 contiguous
// bytes; rows are padded with zeros to whole chunks (Wb = ceil(W / 8)). Input: the
// concatenation [up(src1), src2] (src1 read at (y/2, x/2)). A workgroup moves 64 pixels of
// one row through LDS: 16-byte loads of 8 channels of a pixel into a pixel-major tile, then
// each 16-byte store gathers 8 pixels of one channel from it.
constexpr int TP = 64;
constexpr int RC_MAXC = 112;  // the widest layer input
__global__ __launch_bounds__(256) void k_to_rc(const bf16 *src1, const bf16 *src2, int c1, int c2, int H, int W,
                                               bf16 *dst) {
  __shared__ __align__(16) bf16 t[TP * RC_MAXC];  // t[px * C + c]
  const int C = c1 + c2, G = C >> 3, x0 = blockIdx.x * TP, y = blockIdx.y, b = blockIdx.z;
  const int Hc = H >> 1, Wc = W >> 1, Wb = (W + 7) >> 3;
  for (int e = threadIdx.x; e < TP * G; e += 256) {
    const int px = e / G, c = 8 * (e - px * G), x = x0 + px;
    u4 v = {0u, 0u, 0u, 0u};
    if (x < W)
      v = c < c1 ? *(const u4 *)(src1 + (((long)b * Hc + (y >> 1)) * Wc + (x >> 1)) * c1 + c)
                 : *(const u4 *)(src2 + (((long)b * H + y) * W + x) * c2 + (c - c1));
    *(u4 *)(t + px * C + c) = v;
  }
  __syncthreads();
  // chunks x0 / 8 .. of this row: one contiguous block of nch * C chunks of 8 values
  const int xc0 = x0 >> 3, nch = min(TP / 8, Wb - xc0);
  u4 *out = (u4 *)(dst + (((long)b * H + y) * Wb + xc0) * C * 8);
  for (int e = threadIdx.x; e < nch * C; e += 256) {
    const int xl = e / C, c = e - xl * C;
    const bf16 *col = t + xl * 8 * C + c;
    u4 v;
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = (unsigned)col[2 * k * C] | ((unsigned)col[(2 * k + 1) * C] << 16);
    out[e] = v;
  }
}

struct WgradArgs {
  const bf16 *dyT;   // masked output gradient, row-chunked [B][H][Wb][cout][8]
  const bf16 *xT;    // layer input, row-chunked [B][H][Wb][cin][8]
  int cin, cout, H, W, rows, band;  // rows = B*H, band = rows per wavefront
  float *part;       // [nbands][9][cout][cin]
};

// the 8 values of one channel of a row-chunked row at pixels 8 xc + dx .. 8 xc + dx + 7
// (dx = -1, 0, 1), zero outside the row's Wb chunks (the chunks' own padding is zero):
// `row` points at chunk 0 of the channel, chunks are cs values apart. A shift loads the
// neighbouring chunk's edge pair and moves the values by one BF16 across the 32-bit words
// (v_alignbyte_b32: the low word of {hi, lo} >> 16).
__device__ __forceinline__ u4 rc8(const bf16 *row, int xc, int dx, int Wb, int cs) {
  if (xc >= Wb) return u4{0u, 0u, 0u, 0u};  // past the row: the other operand is zero there
  const u4 h = *(const u4 *)(row + (long)xc * cs);
  if (dx == 0) return h;
  if (dx < 0) {
    const unsigned l = xc > 0 ? *(const unsigned *)(row + (long)(xc - 1) * cs + 6) : 0u;  // pixels -2, -1
    return u4{__builtin_amdgcn_alignbyte(h[0], l, 2), __builtin_amdgcn_alignbyte(h[1], h[0], 2),
              __builtin_amdgcn_alignbyte(h[2], h[1], 2), __builtin_amdgcn_alignbyte(h[3], h[2], 2)};
  }
  const unsigned n = xc + 1 < Wb ? *(const unsigned *)(row + (long)(xc + 1) * cs) : 0u;  // pixels 8, 9
  return u4{__builtin_amdgcn_alignbyte(h[1], h[0], 2), __builtin_amdgcn_alignbyte(h[2], h[1], 2),
            __builtin_amdgcn_alignbyte(h[3], h[2], 2), __builtin_amdgcn_alignbyte(n, h[3], 2)};
}

// partial weight gradient of one band of rows for one tap, on v_mfma_f32_16x16x32_bf16:
// dW[co][tap][ci] += sum over pixels p of dy[p][co] x[p + offset(tap)][ci]
// (M = output channels, NTC tiles; N = input channels, NTN tiles; K = 32 pixels of a row per
// product). A workgroup is nine wavefronts, one per tap, over the same band of rows: they
// read the same rows of dy and the same three rows of x (shifted), so all but the first
// reader find them in the CU's L1. The loads of U consecutive 32-pixel chunks are issued
// before their products, so a wavefront waits for memory once per U chunks.
template <int NTC, int NTN, int U>
__global__ __launch_bounds__(576) void k_wgrad(WgradArgs a) {
  const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
  const int band = blockIdx.x, tap = threadIdx.x >> 6;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const int H = a.H, W = a.W, cin = a.cin;
  f4 acc[NTC][NTN];
#pragma unroll
  for (int i = 0; i < NTC; i++)
#pragma unroll
    for (int j = 0; j < NTN; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const int Wb = (W + 7) >> 3, cso = 8 * a.cout, csi = 8 * cin;  // chunk strides
  const int r0 = band * a.band, r1 = min(a.rows, r0 + a.band);
  for (int r = r0; r < r1; r++) {
    const int b = r / H, y = r % H, ys = y + dy;
    if (ys < 0 || ys >= H) continue;  // the shifted row is zero padding
    const bf16 *dyrow = a.dyT + ((long)b * H + y) * Wb * cso + 8 * m;
    const bf16 *xrow = a.xT + ((long)b * H + ys) * Wb * csi + 8 * m;
    for (int x0 = 0; x0 < 8 * Wb; x0 += 32 * U) {
      // A = dy^T (cout x pixels): lane holds channel 16 tc + m, pixels p .. p + 7;
      // B = shifted input (pixels x cin): lane holds channel 16 tn + m, pixels p + dx ..
      u4 av[U][NTC], bv[U][NTN];
#pragma unroll
      for (int v = 0; v < U; v++) {
        const int xc = (x0 >> 3) + 4 * v + kq;
#pragma unroll
        for (int tc = 0; tc < NTC; tc++) av[v][tc] = rc8(dyrow + 128 * tc, xc, 0, Wb, cso);
#pragma unroll
        for (int tn = 0; tn < NTN; tn++) bv[v][tn] = rc8(xrow + 128 * tn, xc, dx, Wb, csi);
      }
#pragma unroll
      for (int v = 0; v < U; v++)
#pragma unroll
        for (int tn = 0; tn < NTN; tn++)
#pragma unroll
          for (int tc = 0; tc < NTC; tc++)
            acc[tc][tn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(av[v][tc]), as_bf8(bv[v][tn]), acc[tc][tn], 0, 0, 0);
    }
  }
  // D[co = 16 tc + 4 kq + j][ci = 16 tn + m]
  float *out = a.part + ((long)band * 9 + tap) * a.cout * cin;
#pragma unroll
  for (int tc = 0; tc < NTC; tc++)
#pragma unroll
    for (int tn = 0; tn < NTN; tn++)
#pragma unroll
      for (int j = 0; j < 4; j++) out[(long)(16 * tc + 4 * kq + j) * cin + 16 * tn + m] = acc[tc][tn][j];
}

// the same partial weight gradient for layers with few (output, input) channel tiles
// (NTC * NTN <= 3: the full-resolution layers), all nine taps in one wavefront: the dy
// chunk it loads feeds nine taps, and each of the three shifted x rows three, so a wavefront
// issues about one load per product instead of two or more. A workgroup is four wavefronts,
// each with its own band of rows (the partials are per band, as above).
template <int NTC, int NTN, int U>
__global__ __launch_bounds__(256) void k_wgrad9(WgradArgs a, int nbands) {
  const int lane = threadIdx.x & 63, m = lane & 15, kq = lane >> 4;
  const int band = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (band >= nbands) return;
  const int H = a.H, W = a.W, cin = a.cin;
  f4 acc[9][NTC][NTN];
#pragma unroll
  for (int t = 0; t < 9; t++)
#pragma unroll
    for (int i = 0; i < NTC; i++)
#pragma unroll
      for (int j = 0; j < NTN; j++) acc[t][i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const int Wb = (W + 7) >> 3, cso = 8 * a.cout, csi = 8 * cin;
  const int r0 = band * a.band, r1 = min(a.rows, r0 + a.band);
  for (int r = r0; r < r1; r++) {
    const int b = r / H, y = r % H;
    const bf16 *dyrow = a.dyT + ((long)b * H + y) * Wb * cso + 8 * m;
    for (int x0 = 0; x0 < 8 * Wb; x0 += 32 * U) {
      u4 av[U][NTC];
#pragma unroll
      for (int v = 0; v < U; v++)
#pragma unroll
        for (int tc = 0; tc < NTC; tc++) av[v][tc] = rc8(dyrow + 128 * tc, (x0 >> 3) + 4 * v + kq, 0, Wb, cso);
#pragma unroll
      for (int dyi = 0; dyi < 3; dyi++) {
        const int ys = y + dyi - 1;
        if (ys < 0 || ys >= H) continue;  // zero padding: no contribution
        const bf16 *xrow = a.xT + ((long)b * H + ys) * Wb * csi + 8 * m;
        u4 h[U][NTN];
        unsigned lo[U][NTN], hi[U][NTN];
#pragma unroll
        for (int v = 0; v < U; v++)
#pragma unroll
          for (int tn = 0; tn < NTN; tn++) {
            const int xc = (x0 >> 3) + 4 * v + kq;
            const bf16 *c0 = xrow + 128 * tn;
            const bool in = xc < Wb;
            h[v][tn] = in ? *(const u4 *)(c0 + (long)xc * csi) : u4{0u, 0u, 0u, 0u};
            lo[v][tn] = (in && xc > 0) ? *(const unsigned *)(c0 + (long)(xc - 1) * csi + 6) : 0u;
            hi[v][tn] = (xc + 1 < Wb) ? *(const unsigned *)(c0 + (long)(xc + 1) * csi) : 0u;
          }
#pragma unroll
        for (int v = 0; v < U; v++)
#pragma unroll
          for (int tn = 0; tn < NTN; tn++) {
            const u4 q = h[v][tn];
            const u4 bm = u4{__builtin_amdgcn_alignbyte(q[0], lo[v][tn], 2), __builtin_amdgcn_alignbyte(q[1], q[0], 2),
                             __builtin_amdgcn_alignbyte(q[2], q[1], 2), __builtin_amdgcn_alignbyte(q[3], q[2], 2)};
            const u4 bp = u4{__builtin_amdgcn_alignbyte(q[1], q[0], 2), __builtin_amdgcn_alignbyte(q[2], q[1], 2),
                             __builtin_amdgcn_alignbyte(q[3], q[2], 2), __builtin_amdgcn_alignbyte(hi[v][tn], q[3], 2)};
#pragma unroll
            for (int tc = 0; tc < NTC; tc++) {
              acc[3 * dyi][tc][tn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(av[v][tc]), as_bf8(bm), acc[3 * dyi][tc][tn], 0, 0, 0);
              acc[3 * dyi + 1][tc][tn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(av[v][tc]), as_bf8(q), acc[3 * dyi + 1][tc][tn], 0, 0, 0);
              acc[3 * dyi + 2][tc][tn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(av[v][tc]), as_bf8(bp), acc[3 * dyi + 2][tc][tn], 0, 0, 0);
            }
          }
      }
    }
  }
  // D[co = 16 tc + 4 kq + j][ci = 16 tn + m] of every tap
#pragma unroll
  for (int t = 0; t < 9; t++) {
    float *out = a.part + ((long)band * 9 + t) * a.cout * cin;
#pragma unroll
    for (int tc = 0; tc < NTC; tc++)
#pragma unroll
      for (int tn = 0; tn < NTN; tn++)
#pragma unroll
        for (int j = 0; j < 4; j++) out[(long)(16 * tc + 4 * kq + j) * cin + 16 * tn + m] = acc[t][tc][tn][j];
  }
}

// dW[co][tap][ci] = sum over bands, in a fixed order. Pass one (seg > 0): workgroup
// (x, s) sums bands s L .. s L + L - 1 (L = ceil(nbands / nseg)) of 64 consecutive
// (tap, co, ci) entries into segment partial s (four groups of threads take every fourth
// band, then the four group sums are added in order); pass two sums the nseg segment
// partials the same way and writes dw in its [co][tap][ci] layout.
__global__ __launch_bounds__(256) void k_wgrad_sum(const float *part, int nbands, int cout, int cin,
                                                   float *seg, float *dw) {
  __shared__ float lds[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const long i = (long)blockIdx.x * 64 + tx;  // over [9][cout][cin]
  const long n = 9L * cout * cin;
  const int nseg = gridDim.y, L = (nbands + nseg - 1) / nseg;
  const int k0 = blockIdx.y * L, k1 = min(nbands, k0 + L);
  float s = 0.f;
  if (i < n)
    for (int k = k0 + ty; k < k1; k += 4) s += part[(long)k * n + i];
  lds[ty][tx] = s;
  __syncthreads();
  if (ty != 0 || i >= n) return;
  s = ((lds[0][tx] + lds[1][tx]) + lds[2][tx]) + lds[3][tx];
  if (seg) {
    seg[(long)blockIdx.y * n + i] = s;
    return;
  }
  const int tap = (int)(i / ((long)cout * cin)), rem = (int)(i % ((long)cout * cin));
  const int co = rem / cin, ci = rem % cin;
  dw[((long)co * 9 + tap) * cin + ci] = s;
}

// Masked output gradient of a layer (ReLU: zero where the output was not positive), in
// BF16, and the first pass of the bias gradient (per-channel sums of the masked gradient).
// The masked gradient is written twice: NHWC (gm, the input-gradient convolution's operand)
// and row-chunked (gmT, the weight gradient's: [b][y][x / 8][c][8], rows of H x W images
// padded with zeros to whole chunks). A workgroup takes a tile of 64 pixels of a row:
// lanes take 4 channels of a pixel (16-byte loads of g, 8-byte loads of y and stores of gm,
// all contiguous across the wavefront) and put the masked values into a pixel-major LDS
// tile, from which each 16-byte gmT store gathers 8 pixels of one channel. A lane keeps one
// channel quad, so the tile's per-channel sums, bpart[tile][co], are added in a fixed
// order. With g_c0 set, g holds channel 0 alone ([pixels]; the other channels are zero):
// the output layer's gradient.
constexpr int MB_PX = 64;
__global__ __launch_bounds__(256) void k_mask_bias(const float *g, int g_c0, const bf16 *y, int relu, int cout,
                                                   int W, bf16 *gm, bf16 *gmT, float *bpart) {
  __shared__ __align__(16) bf16 t[MB_PX * 64];  // t[px * cout + c], cout <= 64
  __shared__ float lds[256 * 4];
  const int Q = cout >> 2, lanes = 256 - 256 % Q, tid = threadIdx.x;
  const int Wb = (W + 7) >> 3, tpr = (W + MB_PX - 1) / MB_PX;
  const long tile = blockIdx.x, row = tile / tpr;
  const int x0 = (int)(tile - row * tpr) * MB_PX, npx = min(MB_PX, W - x0);
  const int nchk = min(MB_PX / 8, Wb - (x0 >> 3));  // chunks, the row's padding included
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  if (tid < lanes)
    for (int e = tid; e < nchk * 8 * Q; e += lanes) {
      const int px = e / Q, c = 4 * (e - px * Q);
      float v[4] = {0.f, 0.f, 0.f, 0.f};
      if (px < npx) {
        const long p = row * W + x0 + px, i = p * cout + c;
        if (!g_c0) {
          const f4 a = *(const f4 *)(g + i);
          v[0] = a[0];
          v[1] = a[1];
          v[2] = a[2];
          v[3] = a[3];
        } else if (c == 0) {
          v[0] = g[p];
        }
        if (relu) {
          const uint2 yy = *(const uint2 *)(y + i);
#pragma unroll
          for (int k = 0; k < 4; k++)
            if (bf2f((bf16)((k < 2 ? yy.x : yy.y) >> (16 * (k & 1)))) <= 0.f) v[k] = 0.f;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) s[k] += v[k];
      }
      const uint2 o = {(unsigned)f2bf(v[0]) | ((unsigned)f2bf(v[1]) << 16),
                       (unsigned)f2bf(v[2]) | ((unsigned)f2bf(v[3]) << 16)};
      if (px < npx) *(uint2 *)(gm + (row * W + x0 + px) * cout + c) = o;
      *(uint2 *)(t + px * cout + c) = o;
    }
#pragma unroll
  for (int k = 0; k < 4; k++) lds[tid * 4 + k] = s[k];
  __syncthreads();
  u4 *out = (u4 *)(gmT + (row * Wb + (x0 >> 3)) * cout * 8);
  for (int e = tid; e < nchk * cout; e += 256) {
    const int xl = e / cout, c = e - xl * cout;
    const bf16 *col = t + xl * 8 * cout + c;
    u4 v;
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = (unsigned)col[2 * k * cout] | ((unsigned)col[(2 * k + 1) * cout] << 16);
    out[e] = v;
  }
  if (tid < cout) {
    const int q = tid >> 2, k = tid & 3;
    float v = 0.f;
    for (int l = q; l < lanes; l += Q) v += lds[l * 4 + k];
    bpart[tile * cout + tid] = v;
  }
}

// second pass: segment sums of the tiles' partials, seg[s][co] = sum over the tiles of
// segment s in order; a workgroup reads its segment's partials contiguously (lane t takes
// channel t % cout of every (256 / cout)-th tile), then adds its lanes per channel in order
constexpr int BIAS_SEGS = 256;
__global__ __launch_bounds__(256) void k_bias_seg(const float *bpart, long ntiles, int cout, float *seg) {
  __shared__ float lds[256];
  const int t = threadIdx.x, lanes = 256 - 256 % cout, c = t % cout, nr = lanes / cout;
  const long per = (ntiles + BIAS_SEGS - 1) / BIAS_SEGS, k0 = blockIdx.x * per, k1 = min(ntiles, k0 + per);
  float v = 0.f;
  if (t < lanes)
    for (long k = k0 + t / cout; k < k1; k += nr) v += bpart[k * cout + c];
  lds[t] = v;
  __syncthreads();
  if (t < cout) {
    float r = 0.f;
    for (int q = 0; q < nr; q++) r += lds[q * cout + t];
    seg[blockIdx.x * cout + t] = r;
  }
}

// third pass: db[co] = sum over the segments, one workgroup per channel (each thread a
// strided share in order, then the workgroup tree)
__global__ __launch_bounds__(256) void k_bias_fin(const float *bpart, long ntiles, int cout, float *db) {
  __shared__ float lds[8];
  const int c = blockIdx.x;
  float v = 0.f;
  for (long k = threadIdx.x; k < ntiles; k += 256) v += bpart[k * cout + c];
  v = ops_block_reduce<ops_red_inc<float>>(v, lds);
  if (threadIdx.x == 0) db[c] = v;
}

// weights for the input gradient: wt[ci][8 - tap][co] = w[co][tap][ci]
__global__ __launch_bounds__(256) void k_wtrans(const bf16 *w, int cout, int cin, bf16 *wt) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= 9 * cout * cin) return;
  const int co = i / (9 * cin), rem = i % (9 * cin), tap = rem / cin, ci = rem % cin;
  wt[((long)ci * 9 + (8 - tap)) * cout + co] = w[i];
}

// 2 x 2 average pooling, BF16 -> BF16 (FP32 sum), 8 channels per lane (C is a multiple of 8)
__global__ __launch_bounds__(256) void k_pool(const bf16 *in, int B, int H, int W, int C, bf16 *out) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  const int Hc = H / 2, Wc = W / 2, C8 = C >> 3;
  const long n = (long)B * Hc * Wc * C8;
  if (i >= n) return;
  const int c = (int)(i % C8) * 8;
  const long p = i / C8;
  const int X = (int)(p % Wc), Y = (int)((p / Wc) % Hc), b = (int)(p / ((long)Wc * Hc));
  const long base = (((long)b * H + 2 * Y) * W + 2 * X) * C + c, r = (long)W * C;
  const u4 a0 = *(const u4 *)(in + base), a1 = *(const u4 *)(in + base + C);
  const u4 a2 = *(const u4 *)(in + base + r), a3 = *(const u4 *)(in + base + r + C);
  u4 o;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    unsigned h[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int sh = 16 * q;
      const float s = (bf2f((bf16)(a0[k] >> sh)) + bf2f((bf16)(a1[k] >> sh))) +
                      (bf2f((bf16)(a2[k] >> sh)) + bf2f((bf16)(a3[k] >> sh)));
      h[q] = f2bf(0.25f * s);
    }
    o[k] = h[0] | (h[1] << 16);
  }
  *(u4 *)(out + p * C + c) = o;
}

// adjoint of the pooling: gfine += 0.25 * gcoarse(parent), four channels per lane (C is a
// multiple of 4)
__global__ __launch_bounds__(256) void k_unpool_add(const float *gc, int B, int H, int W, int C, float *gf) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  const int C4 = C >> 2;
  const long n = (long)B * H * W * C4;
  if (i >= n) return;
  const int c = (int)(i % C4) * 4;
  const long p = i / C4;
  const int x = (int)(p % W), y = (int)((p / W) % H), b = (int)(p / ((long)W * H));
  f4 *d = (f4 *)(gf + p * C + c);
  *d = *d + 0.25f * *(const f4 *)(gc + (((long)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c);
}

// The output layer (16 input channels, the prediction its channel 0): only channel 0 of its
// output is used, so it runs as a single-output-channel convolution with the same BF16
// operands and FP32 sums as the layers above (channels 1..15 of its output gradient are zero,
// so their weights' and biases' gradients are zero and are never written). w is its
// BF16 weight row of channel 0, [9][16]; x its input (dec0), [B][H][W][16] BF16.
constexpr int OUT_CIN = 16;
static_assert(SPEC[NL - 1].c1 == 0 && SPEC[NL - 1].c2 == OUT_CIN && SPEC[NL - 1].relu == 0 && SPEC[NL - 1].level == 0,
              "the output layer's kernels assume 16 input channels at full resolution and no ReLU");

// prediction: out[p] = bias + sum over taps and channels of w x (zero outside the image)
__global__ __launch_bounds__(256) void k_out_fwd(const bf16 *x, const bf16 *w, const float *bias, int B, int H,
                                                 int W, float *out) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long npix = (long)B * H * W;
  if (p >= npix) return;
  const long img = p / ((long)H * W) * ((long)H * W);
  const int q = (int)(p - img), yq = q / W, xq = q - yq * W;
  const unsigned *w2 = (const unsigned *)w;  // channel pairs (2k, 2k + 1) of each tap
  float acc = 0.f;
#pragma unroll
  for (int tap = 0; tap < 9; tap++) {
    const int ys = yq + tap / 3 - 1, xs = xq + tap % 3 - 1;
    if (ys < 0 || ys >= H || xs < 0 || xs >= W) continue;
    const u4 *src = (const u4 *)(x + (img + (long)ys * W + xs) * OUT_CIN);
    const u4 a0 = src[0], a1 = src[1];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      acc = dot2(a0[k], w2[tap * 8 + k], acc);
      acc = dot2(a1[k], w2[tap * 8 + 4 + k], acc);
    }
  }
  out[p] = acc + bias[0];
}

// input gradient, into dec0's output gradient: gx[q][c] = sum over taps of w[tap][c] G(q - offset(tap)),
// G = the output gradient rounded to BF16 (the operand precision of the other layers)
__global__ __launch_bounds__(256) void k_out_igrad(const float *g, const bf16 *w, int B, int H, int W, float *gx) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long npix = (long)B * H * W;
  if (p >= npix) return;
  const long img = p / ((long)H * W) * ((long)H * W);
  const int q = (int)(p - img), yq = q / W, xq = q - yq * W;
  float acc[OUT_CIN];
#pragma unroll
  for (int c = 0; c < OUT_CIN; c++) acc[c] = 0.f;
#pragma unroll
  for (int tap = 0; tap < 9; tap++) {
    const int yo = yq - (tap / 3 - 1), xo = xq - (tap % 3 - 1);
    if (yo < 0 || yo >= H || xo < 0 || xo >= W) continue;
    const float G = bf2f(f2bf(g[img + (long)yo * W + xo]));
#pragma unroll
    for (int c = 0; c < OUT_CIN; c++) acc[c] = acc[c] + bf2f(w[tap * OUT_CIN + c]) * G;
  }
  f4 *d = (f4 *)(gx + p * OUT_CIN);
#pragma unroll
  for (int k = 0; k < OUT_CIN / 4; k++) d[k] = f4{acc[4 * k], acc[4 * k + 1], acc[4 * k + 2], acc[4 * k + 3]};
}

// weight gradient partials of one band of rows: part[band][tap][c] = sum over the band's
// output pixels o of G(o) x(o + offset(tap))[c]; nine wavefronts, one per tap. A lane takes
// pairs of horizontally adjacent output pixels (W is even, so a pair stays in its row) in a
// pixel-strided order: the pair's two G values and, per channel, its two inputs are packed
// into BF16 pairs, so one dot2 adds both pixels' products; the wavefront's lanes are then
// added by a fixed butterfly.
__global__ __launch_bounds__(576) void k_out_wgrad(const float *g, const bf16 *x, int B, int H, int W, int band,
                                                   float *part) {
  const int tap = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const long rows = (long)B * H, r0 = (long)blockIdx.x * band, r1 = min(rows, r0 + band);
  float acc[OUT_CIN];
#pragma unroll
  for (int c = 0; c < OUT_CIN; c++) acc[c] = 0.f;
  const int Wp = W >> 1;
  const long n = (r1 - r0) * Wp;
  const u4 zero = {0u, 0u, 0u, 0u};
  for (long e = lane; e < n; e += 64) {
    const long row = r0 + e / Wp;  // b * H + y
    const int xo = 2 * (int)(e % Wp), yo = (int)(row % H), ys = yo + dy;
    if (ys < 0 || ys >= H) continue;
    const int xs0 = xo + dx, xs1 = xo + 1 + dx;
    const bool v0 = xs0 >= 0 && xs0 < W, v1 = xs1 >= 0 && xs1 < W;
    // the pair's G (BF16, as the other layers' operands), zero where its input is outside
    const unsigned gp = (v0 ? (unsigned)f2bf(g[row * W + xo]) : 0u) |
                        ((v1 ? (unsigned)f2bf(g[row * W + xo + 1]) : 0u) << 16);
    const bf16 *base = x + (row + dy) * W * OUT_CIN;
    const u4 p0 = v0 ? *(const u4 *)(base + (long)xs0 * OUT_CIN) : zero;
    const u4 p1 = v0 ? *(const u4 *)(base + (long)xs0 * OUT_CIN + 8) : zero;
    const u4 q0 = v1 ? *(const u4 *)(base + (long)xs1 * OUT_CIN) : zero;
    const u4 q1 = v1 ? *(const u4 *)(base + (long)xs1 * OUT_CIN + 8) : zero;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      // channels 2k, 2k + 1 (and 8 + 2k, 9 + 2k) of the two pixels, as pixel pairs
      acc[2 * k] = dot2((p0[k] & 0xffffu) | (q0[k] << 16), gp, acc[2 * k]);
      acc[2 * k + 1] = dot2((p0[k] >> 16) | (q0[k] & 0xffff0000u), gp, acc[2 * k + 1]);
      acc[8 + 2 * k] = dot2((p1[k] & 0xffffu) | (q1[k] << 16), gp, acc[8 + 2 * k]);
      acc[9 + 2 * k] = dot2((p1[k] >> 16) | (q1[k] & 0xffff0000u), gp, acc[9 + 2 * k]);
    }
  }
#pragma unroll
  for (int c = 0; c < OUT_CIN; c++) {
    const float v = ops_wave_reduce<ops_red_inc<float>>(acc[c]);
    if (lane == 0) part[((long)blockIdx.x * 9 + tap) * OUT_CIN + c] = v;
  }
}

// Loss on channel 0 of the output: the error e = prediction - target in the norm
// e^T (I + beta C) e, C the conduction part of the sample's operator (A = I + C), relative
// to the target's, averaged over the batch. beta = 0 gives the relative mean squared
// error, beta = 1 the energy norm of A (the error measure the CG iteration count depends
// on, which weights high-frequency noise by the operator's large eigenvalues).
// coef[pixel] = {rx Kx, ry Ky} of the faces on the low side of the pixel (faces on the
// mesh boundary carry zero); norm[2 b] = t^T t, norm[2 b + 1] = t^T C t. Gradient
// 2 (I + beta C) e / (norm_b * batch) on channel 0 (zero on the others); partial losses per
// workgroup. The prediction and the gradient are the output layer's channel 0 alone
// ([B][H][W]).
__global__ __launch_bounds__(256) void k_loss(const float *pred, const float *target, const float *coef,
                                              const double *norm, float beta, int B, int H, int W,
                                              float *grad, float *part) {
  __shared__ float lds[4];
  const long npix = (long)B * H * W;
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  float l = 0.f;
  if (p < npix) {
    const int x = (int)(p % W), y = (int)((p / W) % H), b = (int)(p / ((long)W * H));
    auto err = [&](long q) { return pred[q] - target[q]; };
    const float e = err(p);
    float ce = 0.f;
    if (beta != 0.f) {
      if (x > 0) ce += coef[2 * p] * (e - err(p - 1));
      if (x + 1 < W) ce += coef[2 * (p + 1)] * (e - err(p + 1));
      if (y > 0) ce += coef[2 * p + 1] * (e - err(p - W));
      if (y + 1 < H) ce += coef[2 * (p + W) + 1] * (e - err(p + W));
    }
    const float me = e + beta * ce;
    const float inv = (float)(1.0 / (norm[2 * b] + (double)beta * norm[2 * b + 1]));
    l = e * me * inv;
    grad[p] = 2.f * me * inv / (float)B;
  }
  l = ops_block_reduce<ops_red_inc<float>>(l, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = l;
}

// Adam on the FP32 master copy, BF16 copy refreshed
__global__ __launch_bounds__(256) void k_adam(float *w, const float *g, float *m1, float *m2, bf16 *wb,
                                              long n, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi, v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = v;
  const float wn = w[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
  w[i] = wn;
  if (wb) wb[i] = f2bf(wn);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *a, bf16 *b, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

unsigned g1d(long n) { return (unsigned)((n + 255) / 256); }

// ---------------------------------------------------------------- host side

struct Net {
  bool live = false;
  // parameters: per layer weights (cout*9*cin) then bias (cout), FP32 master, Adam moments
  long woff[NL], boff[NL], nparam = 0;
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  bf16 *wb = nullptr, *wt = nullptr;  // BF16 weights, transformed weights (largest layer)
  long step = 0;
  // activations for a batch of B images of H x W
  int B = 0, H = 0, W = 0;
  bf16 *act[NL] = {};   // layer outputs (BF16)
  bf16 *pooled[4] = {}; // pooled inputs of enc1..enc3 (index level)
  bf16 *in = nullptr;   // input, 16 channels
  float *out32 = nullptr;  // the output layer's channel 0 (FP32), [B][H][W]
  float *target = nullptr, *coef = nullptr;
  double *norm = nullptr;  // per image of the batch, {t^T t, t^T C t}
  float *grad[NL] = {}; // gradients w.r.t. layer outputs (FP32)
  float *gin = nullptr;  // input gradient of an encoder layer (largest)
  bf16 *gm = nullptr;   // masked output gradient, BF16
  bf16 *gmT = nullptr;  // the same, row-chunked (weight gradient)
  bf16 *xT = nullptr;   // a layer's input, row-chunked (weight gradient)
  float *wpart = nullptr;
  size_t wpart_cap = 0;
  float *wseg = nullptr;  // segment partials of the two-pass band sum (at most 2048 x 64 values)
  float *lpart = nullptr;
  float *bpart = nullptr;  // bias-gradient partials [tile][cout] (tiles of 64 pixels of a row)
  float *bseg = nullptr;   // their segment sums [BIAS_SEGS][cout]
  char *pinned = nullptr;
};
Net N;

hipStream_t stream() { return (hipStream_t)ops_hip_get_stream(); }
void sync() { CK(hipStreamSynchronize(stream())); }

int cin_of(int l) { return SPEC[l].c1 + SPEC[l].c2; }

template <typename T> T *dalloc(size_t n) {
  T *p = nullptr;
  CK(hipMalloc((void **)&p, n * sizeof(T)));
  CK(hipMemset(p, 0, n * sizeof(T)));
  return p;
}
template <typename T> void dfree(T *&p) {
  if (p) CK(hipFree(p));
  p = nullptr;
}

void free_acts() {
  for (int l = 0; l < NL; l++) {
    dfree(N.act[l]);
    dfree(N.grad[l]);
  }
  for (int l = 0; l < 4; l++) dfree(N.pooled[l]);
  dfree(N.in);
  dfree(N.out32);
  dfree(N.target);
  dfree(N.coef);
  dfree(N.norm);
  dfree(N.gin);
  dfree(N.gm);
  dfree(N.gmT);
  dfree(N.xT);
  dfree(N.lpart);
  dfree(N.bpart);
  dfree(N.bseg);
  N.B = N.H = N.W = 0;
}

void ensure_acts(int B, int H, int W) {
  if (!N.live) throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: call tea_unet_init first");
  if (B == N.B && H == N.H && W == N.W) return;
  if (B <= 0 || H <= 0 || W <= 0 || H % 8 || W % 8)
    throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: B > 0, H and W positive multiples of 8");
  sync();
  free_acts();
  N.B = B;
  N.H = H;
  N.W = W;
  auto pix = [&](int level) { return (long)B * (H >> level) * (W >> level); };
  long maxg = 0, maxrc = 0;
  for (int l = 0; l < NL; l++) {
    N.act[l] = dalloc<bf16>(pix(SPEC[l].level) * SPEC[l].cout);
    // the output layer's gradient: channel 0 alone (the prediction's)
    N.grad[l] = dalloc<float>(pix(SPEC[l].level) * (l == NL - 1 ? 1 : SPEC[l].cout));
    maxg = std::max(maxg, pix(SPEC[l].level) * (long)std::max(cin_of(l), SPEC[l].cout));
    // row-chunked copies: rows padded to whole 8-pixel chunks
    const long rows = (long)B * (H >> SPEC[l].level), wb = ((W >> SPEC[l].level) + 7) / 8;
    maxrc = std::max(maxrc, rows * wb * 8 * (long)std::max(cin_of(l), SPEC[l].cout));
  }
  for (int lv = 1; lv < 4; lv++) N.pooled[lv] = dalloc<bf16>(pix(lv) * SPEC[lv].c2);
  N.in = dalloc<bf16>(pix(0) * 16);
  N.out32 = dalloc<float>(pix(0));
  N.target = dalloc<float>(pix(0));
  N.coef = dalloc<float>(pix(0) * 2);
  N.norm = dalloc<double>(2 * B);
  N.gin = dalloc<float>(maxg);
  N.gm = dalloc<bf16>(maxg);
  N.gmT = dalloc<bf16>(maxrc);
  N.xT = dalloc<bf16>(maxrc);
  N.lpart = dalloc<float>(g1d(pix(0)) + 1);
  long maxbp = 0;
  for (int l = 0; l < NL; l++) {
    const long w = W >> SPEC[l].level;
    maxbp = std::max(maxbp, (long)B * (H >> SPEC[l].level) * ((w + MB_PX - 1) / MB_PX) * SPEC[l].cout);
  }
  N.bpart = dalloc<float>(maxbp);
  N.bseg = dalloc<float>((size_t)BIAS_SEGS * 64);
}


// the weight-gradient kernel for the layer's tile counts (one instance per layer shape of
// the network)
void launch_wgrad(const WgradArgs &a, int nbands) {
  const int ntc = a.cout / 16, ntn = a.cin / 16;
  if (a.cout % 16 || a.cin % 16 || ntc < 1 || ntc > 4 || ntn < 1 || ntn > 7)
    throw OPSException(OPS_INTERNAL_ERROR, "tea_unet: weight-gradient shape outside the kernel's limits");
  const dim3 grid((unsigned)nbands), blk(576);
  hipStream_t st = stream();
  const int key = ntc * 10 + ntn;
  const dim3 grid9((unsigned)((nbands + 3) / 4));
  switch (key) {
    case 11: hipLaunchKernelGGL((k_wgrad9<1, 1, 4>), grid9, dim3(256), 0, st, a, nbands); return;
    case 21: hipLaunchKernelGGL((k_wgrad9<2, 1, 2>), grid9, dim3(256), 0, st, a, nbands); return;
    case 13: hipLaunchKernelGGL((k_wgrad9<1, 3, 2>), grid9, dim3(256), 0, st, a, nbands); return;
    default: break;
  }
  switch (key) {
    case 32: hipLaunchKernelGGL((k_wgrad<3, 2, 2>), grid, blk, 0, st, a); break;
    case 43: hipLaunchKernelGGL((k_wgrad<4, 3, 2>), grid, blk, 0, st, a); break;
    case 37: hipLaunchKernelGGL((k_wgrad<3, 7, 1>), grid, blk, 0, st, a); break;
    case 25: hipLaunchKernelGGL((k_wgrad<2, 5, 2>), grid, blk, 0, st, a); break;
    default: throw OPSException(OPS_INTERNAL_ERROR, "tea_unet: no weight-gradient kernel for this layer shape");
  }
}

template <int MODE>
void launch_conv_mode(const ConvArgs &a, int B, int cin) {
  if (MODE != 1 && cin <= 32 && a.cout <= 32) {
    // thin layers: the LDS-tiled kernel (at most 19 KB of LDS per workgroup)
    const dim3 g((unsigned)((a.W + 63) / 64), (unsigned)((a.H + CT_ROWS - 1) / CT_ROWS), (unsigned)B);
    const size_t lds = ct_lds_bytes(cin);
    if (a.cout == 16) hipLaunchKernelGGL(k_conv_lds<1>, g, dim3(256), lds, stream(), a);
    else hipLaunchKernelGGL(k_conv_lds<2>, g, dim3(256), lds, stream(), a);
    return;
  }
  // four wavefronts of 64 pixels per workgroup, pixels of one image row-major (mode 1: the
  // two-row blocks of 2 x 32 pixels)
  const long waves = MODE == 1 ? (long)(a.H / 2) * split_segs(a.W) : ((long)a.H * a.W + 63) / 64;
  const dim3 grid((unsigned)((waves + 3) / 4), (unsigned)B);
  switch (a.cout / 16) {
    case 1: hipLaunchKernelGGL((k_conv<1, MODE>), grid, dim3(256), 0, stream(), a); break;
    case 2: hipLaunchKernelGGL((k_conv<2, MODE>), grid, dim3(256), 0, stream(), a); break;
    case 3: hipLaunchKernelGGL((k_conv<3, MODE>), grid, dim3(256), 0, stream(), a); break;
    case 4: hipLaunchKernelGGL((k_conv<4, MODE>), grid, dim3(256), 0, stream(), a); break;
    case 5: hipLaunchKernelGGL((k_conv<5, MODE>), grid, dim3(256), 0, stream(), a); break;
    case 6: hipLaunchKernelGGL((k_conv<6, MODE>), grid, dim3(256), 0, stream(), a); break;
    default: hipLaunchKernelGGL((k_conv<7, MODE>), grid, dim3(256), 0, stream(), a); break;
  }
}

// k_conv assumes channel counts in multiples of 16, at most 7 output tiles, and src1 present
// exactly when c1 > 0
void launch_conv(const ConvArgs &a, int B) {
  if (a.c1 % 16 || a.c2 % 16 || a.c2 <= 0 || a.cout % 16 || a.cout <= 0 || a.cout > 112 ||
      (a.c1 > 0) != (a.src1 != nullptr) || !a.src2 || !a.w || a.H <= 0 || a.W <= 0 || (a.c1 && (a.H % 2 || a.W % 2)))
    throw OPSException(OPS_INTERNAL_ERROR, "tea_unet: convolution shape outside the kernel's limits");
  if ((a.mode == 1 && (a.H % 2 || a.W % 2 || a.sc1 <= 0 || a.sc1 % 16 || a.sc1 >= a.cout || !a.g1 || !a.g2)) ||
      a.mode < 0 || a.mode > 1)
    throw OPSException(OPS_INTERNAL_ERROR, "tea_unet: input-gradient routing outside the kernel's limits");
  const int cin = a.c1 + a.c2;
  if (a.mode == 0) launch_conv_mode<0>(a, B, cin);
  else launch_conv_mode<1>(a, B, cin);
}

// sources of layer l
void sources(int l, const bf16 *&s1, const bf16 *&s2) {
  s1 = nullptr;
  switch (l) {
    case 0: s2 = N.in; break;
    case 1: s2 = N.pooled[1]; break;
    case 2: s2 = N.pooled[2]; break;
    case 3: s2 = N.pooled[3]; break;
    case 4: s1 = N.act[3]; s2 = N.act[2]; break;
    case 5: s1 = N.act[4]; s2 = N.act[1]; break;
    case 6: s1 = N.act[5]; s2 = N.act[0]; break;
    default: s2 = N.act[6]; break;
  }
}

void forward() {
  hipStream_t st = stream();
  const int B = N.B, H = N.H, W = N.W;
  for (int l = 0; l < NL; l++) {
    const int lv = SPEC[l].level, h = H >> lv, w = W >> lv;
    if (l >= 1 && l <= 3) {  // pool the previous encoder output
      const long n = (long)B * h * w * SPEC[l].c2;
      hipLaunchKernelGGL(k_pool, dim3(g1d(n / 8)), dim3(256), 0, st, N.act[l - 1], B, 2 * h, 2 * w,
                         SPEC[l].c2, N.pooled[lv]);
    }
    if (l == NL - 1) {
      // the output layer: channel 0 alone (the prediction); its weights are read as BF16
      // pairs (32-bit words)
      if (N.woff[l] % 2) throw OPSException(OPS_INTERNAL_ERROR, "tea_unet: output weights not word aligned");
      hipLaunchKernelGGL(k_out_fwd, dim3(g1d((long)B * h * w)), dim3(256), 0, st, N.act[l - 1], N.wb + N.woff[l],
                         N.p + N.boff[l], B, h, w, N.out32);
      continue;
    }
    ConvArgs a;
    sources(l, a.src1, a.src2);
    a.c1 = SPEC[l].c1;
    a.c2 = SPEC[l].c2;
    a.cout = SPEC[l].cout;
    a.H = h;
    a.W = w;
    a.w = N.wb + N.woff[l];
    a.bias = N.p + N.boff[l];
    a.out = N.act[l];
    a.out32 = (l == NL - 1) ? N.out32 : nullptr;
    a.o32c0 = 1;  // the prediction is channel 0 of the output layer
    a.relu = SPEC[l].relu;
    launch_conv(a, B);
  }
  ops_hip_check_launch("tea_unet forward");
}

// gradients of every parameter from N.grad[NL-1] (output gradient). The layers run from
// the output down, and every layer output's gradient is first stored, never cleared and
// added to: the output layer's input gradient is dec0's; a decoder layer's stores its
// coarse source's (that layer's only consumer) and its skip source's (the encoder layer
// whose gradient the pooled path of the next encoder level adds to later in the pass).
void backward() {
  hipStream_t st = stream();
  const int B = N.B, H = N.H, W = N.W;
  for (int l = NL - 1; l >= 0; l--) {
    const int lv = SPEC[l].level, h = H >> lv, w = W >> lv, cout = SPEC[l].cout, cin = cin_of(l);
    const long npix = (long)B * h * w, n = npix * cout;
    if (l == NL - 1) {
      // the output layer, channel 0 alone: bias gradient (the sum of the gradient), weight
      // gradient in bands of 8 rows (partials folded like the other layers'), input gradient
      // straight into dec0's output gradient
      hipLaunchKernelGGL(k_bias_seg, dim3(BIAS_SEGS), dim3(256), 0, st, N.grad[l], npix, 1, N.bseg);
      hipLaunchKernelGGL(k_bias_fin, dim3(1), dim3(256), 0, st, N.bseg, (long)BIAS_SEGS, 1, N.g + N.boff[l]);
      const int band = 8, nbands = (B * h + band - 1) / band;
      const size_t need = (size_t)nbands * 9 * OUT_CIN;
      if (need > N.wpart_cap) {
        sync();
        dfree(N.wpart);
        N.wpart = dalloc<float>(need);
        N.wpart_cap = need;
      }
      hipLaunchKernelGGL(k_out_wgrad, dim3((unsigned)nbands), dim3(576), 0, st, N.grad[l], N.act[l - 1], B, h, w,
                         band, N.wpart);
      const unsigned gx = (unsigned)((9 * OUT_CIN + 63) / 64);
      const int nseg = std::max(1, std::min(nbands / 16, (int)(2048 / gx)));
      if (nseg > 1) {
        hipLaunchKernelGGL(k_wgrad_sum, dim3(gx, (unsigned)nseg), dim3(256), 0, st, N.wpart, nbands, 1, OUT_CIN,
                           N.wseg, nullptr);
        hipLaunchKernelGGL(k_wgrad_sum, dim3(gx, 1), dim3(256), 0, st, N.wseg, nseg, 1, OUT_CIN, nullptr,
                           N.g + N.woff[l]);
      } else {
        hipLaunchKernelGGL(k_wgrad_sum, dim3(gx, 1), dim3(256), 0, st, N.wpart, nbands, 1, OUT_CIN, nullptr,
                           N.g + N.woff[l]);
      }
      hipLaunchKernelGGL(k_out_igrad, dim3(g1d(npix)), dim3(256), 0, st, N.grad[l], N.wb + N.woff[l], B, h, w,
                         N.grad[NL - 2]);
      continue;
    }
    // masked gradient (the kernel's LDS tile holds at most 64 channels)
    if (cout > 64 || cout % 8) throw OPSException(OPS_INTERNAL_ERROR, "tea_unet: layer output too wide");
    const long mtiles = (long)B * h * ((w + MB_PX - 1) / MB_PX);
    hipLaunchKernelGGL(k_mask_bias, dim3((unsigned)mtiles), dim3(256), 0, st, N.grad[l], l == NL - 1 ? 1 : 0,
                       N.act[l], SPEC[l].relu, cout, w, N.gm, N.gmT, N.bpart);
    hipLaunchKernelGGL(k_bias_seg, dim3(BIAS_SEGS), dim3(256), 0, st, N.bpart, mtiles, cout, N.bseg);
    hipLaunchKernelGGL(k_bias_fin, dim3((unsigned)cout), dim3(256), 0, st, N.bseg, (long)BIAS_SEGS, cout,
                       N.g + N.boff[l]);
    // weight gradient from row-chunked copies of the masked gradient and the layer input
    const bf16 *s1, *s2;
    sources(l, s1, s2);
    if (cin > RC_MAXC || cin % 8) throw OPSException(OPS_INTERNAL_ERROR, "tea_unet: layer input too wide");
    const dim3 tgrid((unsigned)((w + TP - 1) / TP), (unsigned)h, (unsigned)B);
    hipLaunchKernelGGL(k_to_rc, tgrid, dim3(256), 0, st, s1, s2, SPEC[l].c1, SPEC[l].c2, h, w, N.xT);
    WgradArgs wa;
    wa.dyT = N.gmT;
    wa.xT = N.xT;
    wa.cin = cin;
    wa.cout = cout;
    wa.H = h;
    wa.W = w;
    wa.rows = B * h;
    // bands of rows: partial sums of ~16 MB (~64 MB for the nine-tap kernel, whose
    // wavefronts each take a band), at least 64 and at most 4096 bands for the nine-tap
    // kernel; the per-tap kernel (one workgroup per band) takes at least 512 bands so that
    // its workgroups fill the GPU (the ~100-200 bands of a 16 MB bound and a 1024-band floor
    // measured slower), at most 1024
    const long per_band = 9L * cout * cin;
    const bool nine = (cout / 16) * (cin / 16) <= 3;
    const long cap = nine ? 4096L : 1024L;
    const int want = (int)std::min(cap, std::max(nine ? 64L : 512L, (nine ? (16L << 20) : (4L << 20)) / per_band));
    wa.band = std::max(1, (B * h + want - 1) / want);
    const int nbands = (B * h + wa.band - 1) / wa.band;
    const size_t need = (size_t)nbands * per_band;
    if (need > N.wpart_cap) {
      sync();
      dfree(N.wpart);
      N.wpart = dalloc<float>(need);
      N.wpart_cap = need;
    }
    wa.part = N.wpart;
    launch_wgrad(wa, nbands);
    // band sums in two passes when the bands are many: enough workgroups to fill the GPU
    const unsigned gx = (unsigned)((per_band + 63) / 64);
    const int nseg = std::max(1, std::min(nbands / 16, (int)(2048 / gx)));
    if (nseg > 1) {
      hipLaunchKernelGGL(k_wgrad_sum, dim3(gx, (unsigned)nseg), dim3(256), 0, st, N.wpart, nbands, cout, cin,
                         N.wseg, nullptr);
      hipLaunchKernelGGL(k_wgrad_sum, dim3(gx, 1), dim3(256), 0, st, N.wseg, nseg, cout, cin, nullptr,
                         N.g + N.woff[l]);
    } else {
      hipLaunchKernelGGL(k_wgrad_sum, dim3(gx, 1), dim3(256), 0, st, N.wpart, nbands, cout, cin, nullptr,
                         N.g + N.woff[l]);
    }
    if (l == 0) break;  // no gradient w.r.t. the network input
    // input gradient: convolution of the masked gradient with the transformed weights
    hipLaunchKernelGGL(k_wtrans, dim3(g1d(9L * cout * cin)), dim3(256), 0, st, N.wb + N.woff[l], cout, cin, N.wt);
    ConvArgs a;
    a.src1 = nullptr;
    a.src2 = N.gm;
    a.c1 = 0;
    a.c2 = cout;
    a.cout = cin;
    a.H = h;
    a.W = w;
    a.w = N.wt;
    a.bias = nullptr;
    a.out = nullptr;
    a.out32 = nullptr;
    a.relu = 0;
    // routed to the layers that produced the input
    if (l >= 1 && l <= 3) {
      a.out32 = N.gin;  // then back through the pooling below
    } else if (l >= 4 && l <= 6) {
      // input = [up(act[s1]), act[s2]]
      const int s1 = (l == 4) ? 3 : l - 1, s2 = (l == 4) ? 2 : (l == 5 ? 1 : 0);
      a.mode = 1;
      a.sc1 = SPEC[l].c1;
      a.g1 = N.grad[s1];
      a.g2 = N.grad[s2];
    } else {
      a.out32 = N.grad[NL - 2];  // the output layer's input is dec0's output
    }
    launch_conv(a, B);
    if (l >= 1 && l <= 3) {
      // input = pool(act[l-1]): the average's adjoint, added to the skip path's part
      const long nf = (long)B * (2 * h) * (2 * w) * SPEC[l].c2 / 4;
      hipLaunchKernelGGL(k_unpool_add, dim3(g1d(nf)), dim3(256), 0, st, N.gin, B, 2 * h, 2 * w, SPEC[l].c2,
                         N.grad[l - 1]);
    }
  }
  ops_hip_check_launch("tea_unet backward");
}

}  // namespace

// ---------------------------------------------------------------- dataset and training

namespace {

struct Data {
  int count = 0, H = 0, W = 0;
  bf16 *in = nullptr;       // [count][H][W][4] features
  float *target = nullptr;  // [count][H][W]
  float *coef = nullptr;    // [count][H][W][2] operator face coefficients {rx Kx, ry Ky}
  double *stats = nullptr;  // [count][3] normalisation of each sample
  double *norm = nullptr;   // [count][2] {t^T t, t^T C t} of each target (A = I + C)
};
Data D;

// batch input: 16 channels (4 features, zeros) from the packed 4-channel samples, one pixel
// per lane (an 8-byte load, two 16-byte stores)
__global__ __launch_bounds__(256) void k_expand(const bf16 *src, long npix, bf16 *dst) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= npix) return;
  const uint2 f = *(const uint2 *)(src + p * 4);
  u4 *d = (u4 *)(dst + p * 16);
  d[0] = u4{f.x, f.y, 0u, 0u};
  d[1] = u4{0u, 0u, 0u, 0u};
}

__global__ __launch_bounds__(256) void k_pack_in(const float *src, long npix, bf16 *dst) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < npix * 4) dst[i] = f2bf(src[i]);
}

uint64_t mix(uint64_t &s) {  // splitmix64
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

float *scratch_f(size_t n) {
  static float *p = nullptr;
  static size_t cap = 0;
  if (n > cap) {
    sync();
    dfree(p);
    p = dalloc<float>(n);
    cap = n;
  }
  return p;
}

// batch loss from the per-workgroup partials of k_loss
double read_loss(long npix, int B) {
  hipStream_t st = stream();
  const unsigned nb = g1d(npix);
  std::vector<float> h(nb);
  CK(hipMemcpyAsync(h.data(), N.lpart, nb * sizeof(float), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  double s = 0;
  for (float v : h) s += v;
  return s / (double)B;
}

void adam_step(float lr) {
  N.step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)N.step), c2 = 1.f - powf(b2, (float)N.step);
  hipLaunchKernelGGL(k_adam, dim3(g1d(N.nparam)), dim3(256), 0, stream(), N.p, N.g, N.m1, N.m2, nullptr,
                     N.nparam, lr, b1, b2, 1e-8f, c1, c2);
  for (int l = 0; l < NL; l++) {
    const long nw = 9L * SPEC[l].cout * cin_of(l);
    hipLaunchKernelGGL(k_to_bf16, dim3(g1d(nw)), dim3(256), 0, stream(), N.p + N.woff[l], N.wb + N.woff[l], nw);
  }
}

void refresh_bf16() {
  for (int l = 0; l < NL; l++) {
    const long nw = 9L * SPEC[l].cout * cin_of(l);
    hipLaunchKernelGGL(k_to_bf16, dim3(g1d(nw)), dim3(256), 0, stream(), N.p + N.woff[l], N.wb + N.woff[l], nw);
  }
}

template <typename F> int guarded(const char *what, F f) {
  try {
    f();
    return 0;
  } catch (std::exception &e) {
    fprintf(stderr, "%s: %s\n", what, e.what());
    return 1;
  }
}

}  // namespace

// ---------------------------------------------------------------- TeaLeaf states as samples
//
// A TeaLeaf solve finds u with A u = u0, A = I + rx Dx(Kx) + ry Dy(Ky) (Kx, Ky from the
// density), starting from u = u0 with residual r0 = u0 - A u0. The change it makes is
// delta = u - u0 = A^-1 r0. The network sees, per pixel,
//   channel 0: (log10(density) - 0.5) / 1.5
//   channel 1: energy / max(energy)          (energy = u0 / density)
//   channel 2: u0 / max(u0)
//   channel 3: r0 / rms(r0)
// and predicts delta / rms(r0): the problem is linear in (u0, r0), so rms(r0) carries the
// magnitude and the network only the shape. A mesh of n x n cells is seen through f x f
// averages (n = f x network size; the statistics are taken on the averages); the
// prediction returns to the mesh by bilinear interpolation between coarse cell centres.

namespace {

struct Mesh {
  const double *rho, *u0, *r;  // dat base pointers, element (x, y) at (x + 1) + (y + 1) * stride
  int n, stride, f;            // n x n cells, viewed as (n / f)^2 network pixels
};

__device__ __forceinline__ long cell_at(const Mesh &m, int x, int y) { return (x + 1) + (long)(y + 1) * m.stride; }

// f x f averages of the four raw features: raw[pixel][4]
__global__ __launch_bounds__(256) void k_raw(Mesh m, float *raw) {
  const int nc = m.n / m.f;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)nc * nc) return;
  const int X = (int)(i % nc), Y = (int)(i / nc);
  double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0;
  for (int dy = 0; dy < m.f; dy++)
    for (int dx = 0; dx < m.f; dx++) {
      const long c = cell_at(m, X * m.f + dx, Y * m.f + dy);
      const double rho = m.rho[c], u = m.u0[c];
      f0 += (log10(rho) - 0.5) / 1.5;
      f1 += u / rho;
      f2 += u;
      f3 += m.r[c];
    }
  const double w = 1.0 / ((double)m.f * m.f);
  raw[4 * i] = (float)(f0 * w);
  raw[4 * i + 1] = (float)(f1 * w);
  raw[4 * i + 2] = (float)(f2 * w);
  raw[4 * i + 3] = (float)(f3 * w);
}

constexpr int STAT_BLOCKS = 512;

// per-workgroup max |energy|, max |u0|, sum r0^2 over the raw features
__global__ __launch_bounds__(256) void k_stats_part(const float *raw, long npix, double *part) {
  __shared__ double lds[4];
  double a = 0.0, b = 0.0, q = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < npix; i += (long)gridDim.x * 256) {
    a = fmax(a, fabs((double)raw[4 * i + 1]));
    b = fmax(b, fabs((double)raw[4 * i + 2]));
    q += (double)raw[4 * i + 3] * raw[4 * i + 3];
  }
  a = ops_block_reduce<ops_red_max<double>>(a, lds);
  b = ops_block_reduce<ops_red_max<double>>(b, lds);
  q = ops_block_reduce<ops_red_inc<double>>(q, lds);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = a;
    part[3 * blockIdx.x + 1] = b;
    part[3 * blockIdx.x + 2] = q;
  }
}

// stats = {max energy, max u0, rms r0}, each 1 where it would be zero
__global__ __launch_bounds__(256) void k_stats_fin(const double *part, int nb, long npix, double *stats) {
  __shared__ double lds[4];
  double a = 0.0, b = 0.0, q = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) {
    a = fmax(a, part[3 * i]);
    b = fmax(b, part[3 * i + 1]);
    q += part[3 * i + 2];
  }
  a = ops_block_reduce<ops_red_max<double>>(a, lds);
  b = ops_block_reduce<ops_red_max<double>>(b, lds);
  q = ops_block_reduce<ops_red_inc<double>>(q, lds);
  if (threadIdx.x == 0) {
    const double rms = sqrt(q / (double)npix);
    stats[0] = a > 0.0 ? a : 1.0;
    stats[1] = b > 0.0 ? b : 1.0;
    stats[2] = rms > 0.0 ? rms : 1.0;
  }
}

// normalised network input: 4 features, zeros up to `cs` channels
__global__ __launch_bounds__(256) void k_pack(const float *raw, long npix, const double *stats, int cs, bf16 *dst) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= npix) return;
  bf16 *o = dst + i * cs;
  o[0] = f2bf(raw[4 * i]);
  o[1] = f2bf((float)(raw[4 * i + 1] / stats[0]));
  o[2] = f2bf((float)(raw[4 * i + 2] / stats[1]));
  o[3] = f2bf((float)(raw[4 * i + 3] / stats[2]));
  for (int c = 4; c < cs; c++) o[c] = 0;
}

// training target u - u0 (brought to unit rms afterwards: the warm start fits the scale
// itself) and the operator's face coefficients {rx Kx, ry Ky} (the low-side faces of each
// cell), on a mesh seen at its own size (f = 1)
__global__ __launch_bounds__(256) void k_target(Mesh m, const double *u, const double *kx, const double *ky,
                                                double rx, double ry, float *dst, float *coef) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)m.n * m.n) return;
  const long c = cell_at(m, (int)(i % m.n), (int)(i / m.n));
  dst[i] = (float)(u[c] - m.u0[c]);
  coef[2 * i] = (float)(rx * kx[c]);
  coef[2 * i + 1] = (float)(ry * ky[c]);
}

// per-workgroup {t^T t, t^T C t} of one sample (C as in k_loss)
__global__ __launch_bounds__(256) void k_tnorm_part(const float *t, const float *coef, int H, int W, double *part) {
  __shared__ double lds[4];
  const long npix = (long)H * W;
  double v0 = 0.0, v1 = 0.0;
  for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < npix; p += (long)gridDim.x * 256) {
    const int x = (int)(p % W), y = (int)(p / W);
    const double e = t[p];
    double ce = 0.0;
    if (x > 0) ce += coef[2 * p] * (e - t[p - 1]);
    if (x + 1 < W) ce += coef[2 * (p + 1)] * (e - t[p + 1]);
    if (y > 0) ce += coef[2 * p + 1] * (e - t[p - W]);
    if (y + 1 < H) ce += coef[2 * (p + W) + 1] * (e - t[p + W]);
    v0 += e * e;
    v1 += e * ce;
  }
  v0 = ops_block_reduce<ops_red_inc<double>>(v0, lds);
  v1 = ops_block_reduce<ops_red_inc<double>>(v1, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = v0;
    part[2 * blockIdx.x + 1] = v1;
  }
}

// norm = the summed partials (t^T t = 1 for a zero target)
__global__ __launch_bounds__(256) void k_tnorm_fin(const double *part, int nb, double *norm) {
  __shared__ double lds[4];
  double v0 = 0.0, v1 = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) {
    v0 += part[2 * i];
    v1 += part[2 * i + 1];
  }
  v0 = ops_block_reduce<ops_red_inc<double>>(v0, lds);
  v1 = ops_block_reduce<ops_red_inc<double>>(v1, lds);
  if (threadIdx.x == 0) {
    norm[0] = v0 > 0.0 ? v0 : 1.0;
    norm[1] = v1;
  }
}

// t /= rms(t), rms from norm[0] = t^T t
__global__ __launch_bounds__(256) void k_unit_rms(float *t, long npix, const double *norm) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < npix) t[i] = (float)(t[i] / sqrt(norm[0] / (double)npix));
}

// u = u0 + s * bilinear(prediction) written to pf (the (n + 2)^2 layout of the mesh, zero
// halo): the network's change on every cell
__global__ __launch_bounds__(256) void k_apply(Mesh m, const float *pred, int ps, const double *stats, double *pf) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)m.n * m.n) return;
  const int x = (int)(i % m.n), y = (int)(i / m.n), nc = m.n / m.f;
  const double cx = (x + 0.5) / m.f - 0.5, cy = (y + 0.5) / m.f - 0.5;
  const int x0 = (int)floor(cx), y0 = (int)floor(cy);
  const double wx = cx - x0, wy = cy - y0;
  const int xa = max(x0, 0), xb = min(x0 + 1, nc - 1), ya = max(y0, 0), yb = min(y0 + 1, nc - 1);
  const double v = (1.0 - wy) * ((1.0 - wx) * pred[((long)ya * nc + xa) * ps] + wx * pred[((long)ya * nc + xb) * ps]) +
                   wy * ((1.0 - wx) * pred[((long)yb * nc + xa) * ps] + wx * pred[((long)yb * nc + xb) * ps]);
  pf[cell_at(m, x, y)] = stats[2] * v;
}

// Galerkin correction from the network's change p: u = u0 + alpha (p - m), m the mean of p
// and alpha = (p - m)^T r0 / (p - m)^T A (p - m), the energy-norm optimum over the span of
// p and the constant field (the constant's own coefficient is zero: A 1 = 1 with
// zero-conductance boundary faces, and r0 sums to zero). The plain start is alpha = 0, so
// the energy-norm error never grows. A is the solve's operator, the 5-point form of
// tea_leaf_cg_calc_w_reduce_kernel. Partials {sum p, sum r0, p^T r0, p^T A p}.
__global__ __launch_bounds__(256) void k_galerkin_part(Mesh m, const double *pf, const double *kx, const double *ky,
                                                       double rx, double ry, double *part) {
  __shared__ double lds[4];
  double sp = 0.0, sr = 0.0, pr = 0.0, pap = 0.0;
  const long nn = (long)m.n * m.n;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < nn; i += (long)gridDim.x * 256) {
    const long c = cell_at(m, (int)(i % m.n), (int)(i / m.n)), s = m.stride;
    const double p = pf[c], r = m.r[c];
    const double ap = (1.0 + ry * (ky[c + s] + ky[c]) + rx * (kx[c + 1] + kx[c])) * p -
                      ry * (ky[c + s] * pf[c + s] + ky[c] * pf[c - s]) - rx * (kx[c + 1] * pf[c + 1] + kx[c] * pf[c - 1]);
    sp += p;
    sr += r;
    pr += p * r;
    pap += p * ap;
  }
  sp = ops_block_reduce<ops_red_inc<double>>(sp, lds);
  sr = ops_block_reduce<ops_red_inc<double>>(sr, lds);
  pr = ops_block_reduce<ops_red_inc<double>>(pr, lds);
  pap = ops_block_reduce<ops_red_inc<double>>(pap, lds);
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x] = sp;
    part[4 * blockIdx.x + 1] = sr;
    part[4 * blockIdx.x + 2] = pr;
    part[4 * blockIdx.x + 3] = pap;
  }
}

// stats[3] = alpha, stats[4] = m, stats[5] = the drop in the squared energy-norm error
// against the plain start, ((p - m)^T r0)^2 / (p - m)^T A (p - m)
__global__ __launch_bounds__(256) void k_galerkin_fin(const double *part, int nb, long ncell, double *stats) {
  __shared__ double lds[4];
  double sp = 0.0, sr = 0.0, pr = 0.0, pap = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) {
    sp += part[4 * i];
    sr += part[4 * i + 1];
    pr += part[4 * i + 2];
    pap += part[4 * i + 3];
  }
  sp = ops_block_reduce<ops_red_inc<double>>(sp, lds);
  sr = ops_block_reduce<ops_red_inc<double>>(sr, lds);
  pr = ops_block_reduce<ops_red_inc<double>>(pr, lds);
  pap = ops_block_reduce<ops_red_inc<double>>(pap, lds);
  if (threadIdx.x == 0) {
    const double m = sp / (double)ncell;
    const double num = pr - m * sr, den = pap - m * m * (double)ncell;
    stats[3] = den > 0.0 ? num / den : 0.0;
    stats[4] = m;
    stats[5] = den > 0.0 ? num * num / den : 0.0;
  }
}

// u = u0 + alpha (p - m) on every cell
__global__ __launch_bounds__(256) void k_update(Mesh m, const double *pf, const double *stats, double *u) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)m.n * m.n) return;
  const long c = cell_at(m, (int)(i % m.n), (int)(i / m.n));
  u[c] = m.u0[c] + stats[3] * (pf[c] - stats[4]);
}

struct SampleScratch {
  double *part = nullptr;
  double *stats = nullptr;  // guess statistics {max energy, max u0, rms r0, alpha, mean of p, energy drop}
  float *raw = nullptr;
  long raw_cap = 0;
  double *pf = nullptr;     // the network's change on the mesh, (n + 2)^2, zero halo
  long pf_cap = 0;
};
SampleScratch SS;

void scratch_init() {
  if (!SS.part) {
    SS.part = dalloc<double>(4 * STAT_BLOCKS);
    SS.stats = dalloc<double>(6);
  }
}

// norm[index] = {t^T t, t^T C t} of dataset sample `index`
void sample_norm(int index) {
  const long npix = (long)D.H * D.W;
  scratch_init();
  hipLaunchKernelGGL(k_tnorm_part, dim3(STAT_BLOCKS), dim3(256), 0, stream(), D.target + index * npix,
                     D.coef + index * npix * 2, D.H, D.W, SS.part);
  hipLaunchKernelGGL(k_tnorm_fin, dim3(1), dim3(256), 0, stream(), SS.part, STAT_BLOCKS, D.norm + 2L * index);
}

const double *dev_of(ops_dat d) {
  ops_memspace ms = OPS_DEVICE;
  ops_dat_get_raw_pointer(d, 0, nullptr, &ms);  // uploads a host-newer copy
  ops_dat_release_raw_data(d, 0, OPS_READ);
  return (const double *)d->data_d;
}

void check_field(ops_dat d, int n) {
  if (d->size[0] != n + 2 || d->size[1] != n + 2 || d->elem_size != 8 || d->x_pad)
    throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: expected n x n double fields with one halo layer");
}

// network input of a mesh (stats written to `stats`, 3 doubles)
Mesh features(ops_dat rho, ops_dat u0, ops_dat r, int n, int net, double *stats, int cs, bf16 *dst) {
  if (net <= 0 || net % 8 || n % net)
    throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: mesh size must be a multiple of the network size (a multiple of 8)");
  for (ops_dat d : {rho, u0, r}) check_field(d, n);
  const long npix = (long)net * net;
  scratch_init();
  if (npix * 4 > SS.raw_cap) {
    sync();
    dfree(SS.raw);
    SS.raw = dalloc<float>(npix * 4);
    SS.raw_cap = npix * 4;
  }
  Mesh m{dev_of(rho), dev_of(u0), dev_of(r), n, n + 2, n / net};
  hipStream_t st = stream();
  hipLaunchKernelGGL(k_raw, dim3(g1d(npix)), dim3(256), 0, st, m, SS.raw);
  hipLaunchKernelGGL(k_stats_part, dim3(STAT_BLOCKS), dim3(256), 0, st, SS.raw, npix, SS.part);
  hipLaunchKernelGGL(k_stats_fin, dim3(1), dim3(256), 0, st, SS.part, STAT_BLOCKS, npix, stats);
  hipLaunchKernelGGL(k_pack, dim3(g1d(npix)), dim3(256), 0, st, SS.raw, npix, stats, cs, dst);
  return m;
}

}  // namespace

void tea_unet_capture_features(int index, ops_dat rho, ops_dat u0, ops_dat r, int n) {
  if (index < 0 || index >= D.count) throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: dataset index");
  const long npix = (long)D.H * D.W;
  features(rho, u0, r, n, D.H, D.stats + 3L * index, 4, D.in + index * npix * 4);
  ops_hip_check_launch("tea_unet capture (features)");
}

void tea_unet_capture_target(int index, ops_dat u0, ops_dat u, ops_dat Kx, ops_dat Ky, double rx, double ry, int n) {
  if (index < 0 || index >= D.count) throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: dataset index");
  if (n != D.H || n != D.W) throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: training samples are captured at the dataset size");
  for (ops_dat d : {u0, u, Kx, Ky}) check_field(d, n);
  const Mesh m{nullptr, dev_of(u0), nullptr, n, n + 2, 1};
  const long npix = (long)D.H * D.W;
  hipLaunchKernelGGL(k_target, dim3(g1d(npix)), dim3(256), 0, stream(), m, dev_of(u), dev_of(Kx), dev_of(Ky), rx, ry,
                     D.target + index * npix, D.coef + index * npix * 2);
  sample_norm(index);
  hipLaunchKernelGGL(k_unit_rms, dim3(g1d(npix)), dim3(256), 0, stream(), D.target + index * npix, npix,
                     D.norm + 2L * index);
  sample_norm(index);
  ops_hip_check_launch("tea_unet capture (target)");
}

void tea_unet_guess_dats(ops_dat rho, ops_dat u0, ops_dat r, ops_dat Kx, ops_dat Ky, double rx, double ry, ops_dat u,
                         int n, int net_cells, double *out2) {
  if (!N.live) throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: no network (tea_unet_init)");
  if (net_cells <= 0 || net_cells % 8 || n % net_cells)
    throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: mesh size must be a multiple of the network size (a multiple of 8)");
  for (ops_dat d : {u, Kx, Ky}) check_field(d, n);
  ensure_acts(1, net_cells, net_cells);
  scratch_init();
  const long cells = (long)(n + 2) * (n + 2);
  if (cells > SS.pf_cap) {
    sync();
    dfree(SS.pf);
    SS.pf = dalloc<double>(cells);  // zeroed: the halo stays zero
    SS.pf_cap = cells;
  }
  const Mesh m = features(rho, u0, r, n, net_cells, SS.stats, 16, N.in);
  forward();
  hipStream_t st = stream();
  const double *kx = dev_of(Kx), *ky = dev_of(Ky);
  hipLaunchKernelGGL(k_apply, dim3(g1d((long)n * n)), dim3(256), 0, st, m, N.out32, 1, SS.stats, SS.pf);
  hipLaunchKernelGGL(k_galerkin_part, dim3(STAT_BLOCKS), dim3(256), 0, st, m, SS.pf, kx, ky, rx, ry, SS.part);
  hipLaunchKernelGGL(k_galerkin_fin, dim3(1), dim3(256), 0, st, SS.part, STAT_BLOCKS, (long)n * n, SS.stats);
  ops_memspace ms = OPS_DEVICE;
  ops_dat_get_raw_pointer(u, 0, nullptr, &ms);
  hipLaunchKernelGGL(k_update, dim3(g1d((long)n * n)), dim3(256), 0, st, m, SS.pf, SS.stats, (double *)u->data_d);
  ops_dat_release_raw_data(u, 0, OPS_WRITE);
  ops_hip_check_launch("tea_unet warm start");
  if (out2) {
    double h[6];
    CK(hipMemcpy(h, SS.stats, sizeof h, hipMemcpyDeviceToHost));
    out2[0] = h[3];
    out2[1] = h[5];
  }
}

const double *tea_unet_direction(double *mean) {
  if (SS.pf == nullptr || SS.stats == nullptr)
    throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet_direction: no warm start evaluated");
  if (mean) {
    double h[6];
    CK(hipMemcpy(h, SS.stats, sizeof h, hipMemcpyDeviceToHost));
    *mean = h[4];
  }
  return SS.pf;
}

extern "C" int tea_unet_init(unsigned long long seed) {
  return guarded("tea_unet_init", [&] {
    if (!N.live) {
      long off = 0;
      long maxw = 0;
      for (int l = 0; l < NL; l++) {
        N.woff[l] = off;
        off += 9L * SPEC[l].cout * cin_of(l);
        N.boff[l] = off;
        off += SPEC[l].cout;
        maxw = std::max(maxw, 9L * SPEC[l].cout * cin_of(l));
      }
      N.nparam = off;
      N.p = dalloc<float>(off);
      N.g = dalloc<float>(off);
      N.m1 = dalloc<float>(off);
      N.m2 = dalloc<float>(off);
      N.wb = dalloc<bf16>(off);
      N.wt = dalloc<bf16>(maxw);
      N.wseg = dalloc<float>(2048L * 64);
      N.live = true;
    }
    // He initialisation from splitmix64 (uniform with the normal's variance), zero biases
    std::vector<float> h(N.nparam, 0.f);
    uint64_t s = seed;
    for (int l = 0; l < NL; l++) {
      const long nw = 9L * SPEC[l].cout * cin_of(l);
      const int fan_in = 9 * (l == 0 ? 4 : cin_of(l));
      const float a = sqrtf(6.f / (float)fan_in) * (l == NL - 1 ? 0.1f : 1.f);
      for (long i = 0; i < nw; i++) {
        const double u = (double)(mix(s) >> 11) * (1.0 / 9007199254740992.0);
        float v = (float)((2.0 * u - 1.0) * a);
        if (l == 0 && (i % cin_of(l)) >= 4) v = 0.f;  // padded input channels
        h[N.woff[l] + i] = v;
      }
    }
    CK(hipMemcpy(N.p, h.data(), N.nparam * sizeof(float), hipMemcpyHostToDevice));
    CK(hipMemset(N.m1, 0, N.nparam * sizeof(float)));
    CK(hipMemset(N.m2, 0, N.nparam * sizeof(float)));
    N.step = 0;
    refresh_bf16();
    sync();
  });
}

extern "C" long tea_unet_num_params(void) { return N.nparam; }

extern "C" int tea_unet_get_params(float *host) {
  return guarded("tea_unet_get_params", [&] {
    if (!N.live) throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: call tea_unet_init first");
    sync();
    CK(hipMemcpy(host, N.p, N.nparam * sizeof(float), hipMemcpyDeviceToHost));
  });
}

extern "C" int tea_unet_set_params(const float *host) {
  return guarded("tea_unet_set_params", [&] {
    if (!N.live) throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: call tea_unet_init first");
    sync();
    CK(hipMemcpy(N.p, host, N.nparam * sizeof(float), hipMemcpyHostToDevice));
    refresh_bf16();
    sync();
  });
}

// in: [B][H][W][16] FP32 (rounded to BF16), out: [B][H][W] prediction (channel 0)
extern "C" int tea_unet_forward_host(int B, int H, int W, const float *in, float *out) {
  return guarded("tea_unet_forward_host", [&] {
    ensure_acts(B, H, W);
    const long npix = (long)B * H * W;
    float *tmp = scratch_f(npix * 16);
    CK(hipMemcpy(tmp, in, npix * 16 * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf16, dim3(g1d(npix * 16)), dim3(256), 0, stream(), tmp, N.in, npix * 16);
    forward();
    sync();
    CK(hipMemcpy(out, N.out32, npix * sizeof(float), hipMemcpyDeviceToHost));
  });
}

// gradients of every parameter (FP32, the get_params order) for d loss / d prediction =
// grad_out ([B][H][W]); runs the forward pass on `in` first
extern "C" int tea_unet_backward_host(int B, int H, int W, const float *in, const float *grad_out,
                                      float *gparams) {
  return guarded("tea_unet_backward_host", [&] {
    ensure_acts(B, H, W);
    const long npix = (long)B * H * W;
    float *tmp = scratch_f(npix * 16);
    CK(hipMemcpy(tmp, in, npix * 16 * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf16, dim3(g1d(npix * 16)), dim3(256), 0, stream(), tmp, N.in, npix * 16);
    forward();
    sync();
    CK(hipMemcpy(N.grad[NL - 1], grad_out, npix * sizeof(float), hipMemcpyHostToDevice));
    backward();
    sync();
    CK(hipMemcpy(gparams, N.g, N.nparam * sizeof(float), hipMemcpyDeviceToHost));
  });
}

extern "C" int tea_unet_dataset_reserve(int count, int H, int W) {
  return guarded("tea_unet_dataset_reserve", [&] {
    if (count <= 0 || H <= 0 || W <= 0 || H % 8 || W % 8)
      throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: dataset count > 0, H and W positive multiples of 8");
    sync();
    dfree(D.in);
    dfree(D.target);
    D.count = count;
    D.H = H;
    D.W = W;
    dfree(D.stats);
    dfree(D.coef);
    dfree(D.norm);
    D.in = dalloc<bf16>((size_t)count * H * W * 4);
    D.target = dalloc<float>((size_t)count * H * W);
    D.coef = dalloc<float>((size_t)count * H * W * 2);
    D.stats = dalloc<double>((size_t)count * 3);
    D.norm = dalloc<double>((size_t)count * 2);
  });
}

// in4: [H][W][4] features, target: [H][W] (host FP32)
extern "C" int tea_unet_dataset_set(int index, const float *in4, const float *target) {
  return guarded("tea_unet_dataset_set", [&] {
    if (index < 0 || index >= D.count) throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: dataset index");
    const long npix = (long)D.H * D.W;
    float *tmp = scratch_f(npix * 4);
    CK(hipMemcpy(tmp, in4, npix * 4 * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_pack_in, dim3(g1d(npix * 4)), dim3(256), 0, stream(), tmp, npix, D.in + index * npix * 4);
    CK(hipMemcpy(D.target + index * npix, target, npix * sizeof(float), hipMemcpyHostToDevice));
    CK(hipMemsetAsync(D.coef + index * npix * 2, 0, npix * 2 * sizeof(float), stream()));
    sample_norm(index);
    sync();
  });
}

// Adam over the dataset in batches, sample order shuffled by splitmix64(seed) each epoch;
// loss_per_epoch[e] = mean over the epoch's batches of the batch loss (k_loss)
extern "C" int tea_unet_train(int epochs, int batch, float lr, float beta, unsigned long long seed,
                              float *loss_per_epoch) {
  return guarded("tea_unet_train", [&] {
    if (D.count <= 0) throw OPSException(OPS_INVALID_ARGUMENT, "tea_unet: empty dataset");
    if (batch < 1 || batch > D.count) batch = D.count < 4 ? D.count : 4;
    ensure_acts(batch, D.H, D.W);
    const long spix = (long)D.H * D.W, npix = spix * batch;
    std::vector<int> order(D.count);
    uint64_t s = seed;
    hipStream_t st = stream();
    for (int e = 0; e < epochs; e++) {
      for (int i = 0; i < D.count; i++) order[i] = i;
      for (int i = D.count - 1; i > 0; i--) std::swap(order[i], order[(int)(mix(s) % (uint64_t)(i + 1))]);
      double lsum = 0;
      int nb = 0;
      for (int i0 = 0; i0 + batch <= D.count; i0 += batch) {
        for (int k = 0; k < batch; k++) {
          const int idx = order[i0 + k];
          hipLaunchKernelGGL(k_expand, dim3(g1d(spix)), dim3(256), 0, st, D.in + (long)idx * spix * 4, spix,
                             N.in + k * spix * 16);
          CK(hipMemcpyAsync(N.target + k * spix, D.target + (long)idx * spix, spix * sizeof(float),
                            hipMemcpyDeviceToDevice, st));
          CK(hipMemcpyAsync(N.coef + k * spix * 2, D.coef + (long)idx * spix * 2, spix * 2 * sizeof(float),
                            hipMemcpyDeviceToDevice, st));
          CK(hipMemcpyAsync(N.norm + 2 * k, D.norm + 2L * idx, 2 * sizeof(double), hipMemcpyDeviceToDevice, st));
        }
        forward();
        hipLaunchKernelGGL(k_loss, dim3(g1d(npix)), dim3(256), 0, st, N.out32, N.target, N.coef, N.norm, beta, batch, D.H,
                           D.W, N.grad[NL - 1], N.lpart);
        backward();
        adam_step(lr);
        lsum += read_loss(npix, batch);
        nb++;
      }
      if (loss_per_epoch) loss_per_epoch[e] = (float)(lsum / (nb > 0 ? nb : 1));
    }
    sync();
  });
}

extern "C" int tea_unet_free(void) {
  return guarded("tea_unet_free", [&] {
    sync();
    free_acts();
    dfree(D.in);
    dfree(D.target);
    dfree(D.stats);
    dfree(D.coef);
    dfree(D.norm);
    D.count = 0;
    dfree(N.p);
    dfree(N.g);
    dfree(N.m1);
    dfree(N.m2);
    dfree(N.wb);
    dfree(N.wt);
    dfree(N.wpart);
    N.wpart_cap = 0;
    dfree(N.wseg);
    N.live = false;
  });
}
