// Learned preconditioner on gfx950: a 3x3 convolutional network with BF16
// weights/activations and FP32 accumulation on the matrix cores, its
// backward pass, the Adam trainer, the training-set generator (batched
// Jacobi-PCG) and the ARKLsPrec callbacks.
//
// Convolution as implicit GEMM.  For a block of 16 consecutive pixels of one
// row, D[co][px] = sum_k A[co][k] B[k][px] with k running over (tap, input
// channel).
//  - 16 input channels: v_mfma_f32_16x16x32_bf16, one MFMA per pair of taps
//    (K = 2 taps x 16 channels; 5 MFMAs, the tenth tap zero).  Operand
//    layout (lane l, g = l/16, n = l%16): A holds A[n][8g..8g+7], B holds
//    B[8g..8g+7][n], D register i holds D[4g+i][n].  Lane group g reads tap
//    2p + g/2, channels 8(g%2)..+7 of pixel n: one 16-byte load.
//  - 1 input channel: v_mfma_f32_16x16x16bf16_1k with the 9 taps on K (A
//    holds A[n][4g..4g+3], B holds B[4g..4g+3][n]).
// The weights are the A operand (packed once per layer, loaded once per
// wavefront), and each lane's D registers are 4 consecutive channels of one
// pixel: one 8-byte store.  Activations are 16 BF16 channels per pixel in a
// blocked layout: rows padded to whole blocks of 16 pixels, and within a
// block channels 0-7 of the 16 pixels, then channels 8-15, so a tap-pair
// operand is one contiguous 16-byte load per lane.  The network input and
// output are FP64 images.
//
// Backward.  The input gradient of a layer is the same convolution with the
// taps flipped and the channel roles exchanged (a separately packed A
// operand); the ReLU derivative of the layer below is applied on store (the
// stored activation is positive exactly where the pre-activation was).  The
// weight gradient dW[co][ci][t] = sum_p dY[p][co] x[p+off_t][ci] is a GEMM
// with K = pixels: each wavefront accumulates 9 tap tiles over its pixel
// blocks, and the per-wavefront partials are summed in a fixed order
// (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "diffusion_internal.hpp"
#include "learned_prec.h"

using ceda::kBlock;
using ceda::kWave;

namespace {

constexpr int kC = 16;        // hidden channels
constexpr int kMaxL = 8;      // layers
constexpr int kWavesPB = kBlock / kWave;
constexpr int kConvU = 2;     // 16-pixel blocks per wavefront step of the convolution
constexpr int kPW = 9 * 256 + 16;  // weight-gradient partial of one wavefront
constexpr int kPack = 5 * 512;      // packed A operands of one layer (BF16 elements)
constexpr int kWGradWaves = 1024;

typedef short bfx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float fx4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 rounded to nearest even, in one instruction
// (v_cvt_pk_bf16_f32); a NaN stays a NaN
__device__ inline unsigned short f2bf(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }
__device__ inline float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// Blocked activation layout: image rows (row = b*ny + y) of bpr blocks of 16
// pixels; in a block, channels 0-7 of the 16 pixels, then channels 8-15
// (16 bytes per pixel and half).  chunk_off: offset (in BF16 elements) of
// channels 8h..8h+7 of pixel x; act_off: of channels 4g..4g+3.
__device__ inline int64_t chunk_off(int64_t row, int bpr, int x, int h)
{
  return ((row * bpr + (x >> 4)) * 2 + h) * 128 + (x & 15) * 8;
}
__device__ inline int64_t act_off(int64_t row, int bpr, int x, int g)
{
  return chunk_off(row, bpr, x, g >> 1) + 4 * (g & 1);
}
__device__ inline unsigned short d2bf(double x) { return f2bf((float)x); }


// ------------------------------------------------------------ forward conv
struct ConvArgs {
  int nb, nx, ny, dil;
  const void* in;             // NHWC16 BF16, or an FP64 image (CIN = 1)
  void* out;                  // NHWC16 BF16, or an FP64 image (COUT = 1)
  const unsigned short* apack;
  const float* bias;          // nullptr: none
  const double* inv_scale;    // CIN = 1: per-image input multiplier (nullptr: 1)
  const double* scale;        // COUT = 1: per-image output multiplier (nullptr: 1)
  double out_mul;             // COUT = 1: constant output multiplier
  const unsigned short* mask; // NHWC16: outputs zeroed where the mask is not > 0
  int addto;                  // COUT = 1: out += result
  int rows;                   // row-chain kernels: output rows per workgroup (set by conv)
};

template <int CIN, int COUT, bool RELU>
__global__ __launch_bounds__(kBlock) void k_conv(const ConvArgs a)
{
  const int lane = threadIdx.x & (kWave - 1), g = lane >> 4, n = lane & 15;
  // CIN = 16: v_mfma_f32_16x16x32_bf16 over tap pairs, K = (tap 2p + g/2,
  // channels 8(g%2)..+7) for lane group g; the 10th tap is zero.  CIN = 1:
  // v_mfma_f32_16x16x16bf16_1k with the 9 taps on K.
  bf16x8 A8[5];
  bfx4 A1 = {0, 0, 0, 0};
  if (CIN == kC) {
#pragma unroll
    for (int p = 0; p < 5; p++) A8[p] = __builtin_bit_cast(bf16x8, reinterpret_cast<const uint4*>(a.apack)[p * kWave + lane]);
  } else {
    A1 = reinterpret_cast<const bfx4*>(a.apack)[lane];
  }
  float bv[4] = {0.f, 0.f, 0.f, 0.f};
  if (a.bias) {
    if (COUT == kC) {
#pragma unroll
      for (int i = 0; i < 4; i++) bv[i] = a.bias[4 * g + i];
    } else bv[0] = a.bias[0];
  }
  const int nx = a.nx, ny = a.ny, d = a.dil;
  // XCD-aware order: workgroups go round-robin to the 8 XCDs, so XCD k
  // (blockIdx % 8) takes the k-th of 8 contiguous bands of image rows and
  // its workgroups sweep that band row by row, one segment of
  // kWavesPB * kConvU 16-pixel blocks at a time (kConvU consecutive blocks
  // per wavefront).  The rows y - d .. y + d a tile reads are then read by
  // the same XCD at about the same time and stay in its L2.
  const int bpr = (nx + 15) >> 4;
  const int spr = (bpr + kWavesPB * kConvU - 1) / (kWavesPB * kConvU);
  const int rows = a.nb * ny;
  const int band = (rows + 7) >> 3;
  const int r0 = (int)(blockIdx.x & 7) * band;
  const int nrow = (r0 < rows) ? min(rows - r0, band) : 0;
  const int nseg = nrow * spr;
  const int wave = threadIdx.x >> 6;
  for (int c = (int)(blockIdx.x >> 3); c < nseg; c += (int)(gridDim.x >> 3)) {
    const int q = c / spr;
    const int bx0 = (c - q * spr) * (kWavesPB * kConvU) + wave * kConvU;
    if (bx0 >= bpr) continue;  // wavefront-uniform: past the row end
    const int rowi = r0 + q;
    const int bimg = rowi / ny, y = rowi - bimg * ny;
    const int64_t img = (int64_t)bimg * ny;  // first row of the image
    fx4 acc[kConvU];
    // interior segment (wavefront-uniform): every tap of every pixel is
    // inside the image, so the taps are constant offsets from one address
    const bool interior = bx0 * 16 - d >= 0 && (bx0 + kConvU) * 16 + d <= nx && y - d >= 0 && y + d < ny;
    // Elsewhere every load is issued unconditionally from a clamped
    // (in-bounds) address and the zero padding applied by a select
    // afterwards, so the loads of all kConvU blocks are in flight together
    // instead of one branch and one wait per tap.
    if (CIN == kC) {
      // lane group g reads tap 2p + g/2 (channels 8(g%2)..+7, one 16-byte
      // load per pixel); interior segments need no clamp
      const unsigned short* in = static_cast<const unsigned short*>(a.in);
      const int h = g & 1;
      uint4 Bt[kConvU][5];
#pragma unroll
      for (int u = 0; u < kConvU; u++) {
        const int x = (bx0 + u) * 16 + n;
#pragma unroll
        for (int p = 0; p < 5; p++) {
          const int t = min(2 * p + (g >> 1), 8);
          const int xs = x + (t % 3 - 1) * d, ys = y + (t / 3 - 1) * d;
          const int xc = interior ? xs : min(max(xs, 0), nx - 1);
          const int yc = interior ? ys : min(max(ys, 0), ny - 1);
          Bt[u][p] = *reinterpret_cast<const uint4*>(in + chunk_off(img + yc, bpr, xc, h));
        }
      }
#pragma unroll
      for (int u = 0; u < kConvU; u++) {
        const int x = (bx0 + u) * 16 + n;
        acc[u] = fx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int p = 0; p < 5; p++) {
          const int t = 2 * p + (g >> 1);
          const int xs = x + (t % 3 - 1) * d, ys = y + (t / 3 - 1) * d;
          const bool ok = t < 9 && x < nx && xs >= 0 && xs < nx && ys >= 0 && ys < ny;
          const uint4 zero = {0u, 0u, 0u, 0u};
          acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A8[p], __builtin_bit_cast(bf16x8, ok ? Bt[u][p] : zero),
                                                           acc[u], 0, 0, 0);
        }
      }
    } else {
      const double* in = static_cast<const double*>(a.in);
      const double is = a.inv_scale ? a.inv_scale[bimg] : 1.0;
      double v[kConvU][4];
#pragma unroll
      for (int u = 0; u < kConvU; u++) {
        const int x = (bx0 + u) * 16 + n;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int t = min(4 * g + j, 8);
          const int xs = x + (t % 3 - 1) * d, ys = y + (t / 3 - 1) * d;
          const int xc = min(max(xs, 0), nx - 1), yc = min(max(ys, 0), ny - 1);
          v[u][j] = in[(img + yc) * nx + xc];
        }
      }
#pragma unroll
      for (int u = 0; u < kConvU; u++) {
        const int x = (bx0 + u) * 16 + n;
        bfx4 B;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int t = 4 * g + j;
          const int xs = x + (t % 3 - 1) * d, ys = y + (t / 3 - 1) * d;
          const bool ok = t < 9 && x < nx && xs >= 0 && xs < nx && ys >= 0 && ys < ny;
          B[j] = (short)(ok ? d2bf(v[u][j] * is) : (unsigned short)0);
        }
        acc[u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A1, B, fx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      }
    }
#pragma unroll
    for (int u = 0; u < kConvU; u++) {
      const int x = (bx0 + u) * 16 + n;
      if (x >= nx) continue;
      const int64_t pix = (img + y) * nx + x;
      if (COUT == kC) {
        bfx4 o;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          float v = acc[u][i] + bv[i];
          if (RELU) v = fmaxf(v, 0.0f);
          if (a.mask && !(bf2f(a.mask[act_off(img + y, bpr, x, g) + i]) > 0.0f)) v = 0.0f;
          o[i] = (short)f2bf(v);
        }
        *reinterpret_cast<bfx4*>(static_cast<unsigned short*>(a.out) + act_off(img + y, bpr, x, g)) = o;
      } else if (g == 0) {
        const double v = (double)(acc[u][0] + bv[0]) * a.out_mul * (a.scale ? a.scale[bimg] : 1.0);
        double* out = static_cast<double*>(a.out);
        out[pix] = a.addto ? out[pix] + v : v;
      }
    }
  }
}

// ------------------------------------- 16-channel layers, LDS row chains
// A layer of dilation d couples row y only with rows y - d and y + d, so the
// rows of one residue class mod d form a chain that can be swept on its own
// (for d = 1 the chain is the whole column).  A workgroup takes R (kChainR,
// fewer on small images) consecutive output rows of one chain in one column strip of kChainW
// pixels and keeps a ring of four input rows of the strip (plus d halo
// pixels on each side) in LDS: three in use (y - d, y, y + d) and the next
// one, loaded into registers while the current row is computed and stored
// into the ring afterwards.  Each input row is read from memory about once
// per strip (a (kChainW + 2d) / kChainW horizontal halo and two halo rows
// per R rows), against three times plus the taps' reuse through the
// caches for k_conv.  Ring layout: [slot][half h][pixel] of 16-byte chunks
// (channels 8h..8h+7 of one pixel), so the 16 pixels of a tap-pair operand
// are 256 contiguous bytes (no bank conflicts).  The operands, the MFMA
// sequence and the stores are those of k_conv: the same results bit for bit.
constexpr int kChainW = kWavesPB * kConvU * 16;  // pixels per column strip
constexpr int kChainR = 16;                      // output rows per workgroup (at most)
constexpr int kChainMaxD = 64;                   // ring rows of at most kChainW + 128 pixels
constexpr int kChainLoads = (2 * (kChainW + 2 * kChainMaxD) + kBlock - 1) / kBlock;  // chunks per thread

// workgroups of one chain launch: per image, strips x min(d, ny) chains x
// segments of R rows (the longest chain has ceil(ny / d) rows)
inline int64_t chain_tasks(int nb, int nx, int ny, int d, int R)
{
  const int64_t strips = (nx + kChainW - 1) / kChainW;
  const int64_t chains = std::min(d, ny);
  const int64_t segs = ((ny + d - 1) / d + R - 1) / R;
  return (int64_t)nb * strips * chains * segs;
}
// rows per workgroup: kChainR, halved (down to 2) while the launch has fewer
// than g_chain_tasks workgroups, so that small images still spread over the
// chip (each halving adds two halo rows per R rows of the strip's reads)
int g_chain_tasks = 1024;
inline int chain_rows(int nb, int nx, int ny, int d)
{
  int R = kChainR;
  while (R > 2 && chain_tasks(nb, nx, ny, d, R) < g_chain_tasks) R /= 2;
  return R;
}

template <int COUT, bool RELU>
__global__ __launch_bounds__(kBlock) void k_conv_chain(const ConvArgs a)
{
  extern __shared__ uint4 ring[];
  const int lane = threadIdx.x & (kWave - 1), g = lane >> 4, n = lane & 15, wave = threadIdx.x >> 6;
  const int nx = a.nx, ny = a.ny, d = a.dil;
  const int span = kChainW + 2 * d;  // pixels of a ring row
  const int bpr = (nx + 15) >> 4;
  // task: strip fastest, then chain, then segment, then image
  const int strips = (nx + kChainW - 1) / kChainW;
  const int chains = min(d, ny);
  const int R = a.rows;
  const int segs = ((ny + d - 1) / d + R - 1) / R;
  int64_t t = blockIdx.x;
  const int strip = (int)(t % strips);
  t /= strips;
  const int c = (int)(t % chains);
  t /= chains;
  const int seg = (int)(t % segs);
  const int bimg = (int)(t / segs);
  const int64_t img = (int64_t)bimg * ny;
  const int x0 = strip * kChainW;
  const int y0 = c + seg * R * d;
  if (y0 >= ny) return;  // workgroup-uniform: an empty tail segment of a short chain

  bf16x8 A8[5];
#pragma unroll
  for (int p = 0; p < 5; p++) A8[p] = __builtin_bit_cast(bf16x8, reinterpret_cast<const uint4*>(a.apack)[p * kWave + lane]);
  float bv[4] = {0.f, 0.f, 0.f, 0.f};
  if (a.bias) {
    if (COUT == kC) {
#pragma unroll
      for (int i = 0; i < 4; i++) bv[i] = a.bias[4 * g + i];
    } else bv[0] = a.bias[0];
  }
  const unsigned short* in = static_cast<const unsigned short*>(a.in);
  // input row yr (zero outside the image) of the strip into registers
  auto fetch = [&](int yr, uint4 (&v)[kChainLoads]) {
#pragma unroll
    for (int k = 0; k < kChainLoads; k++) {
      const int ci = threadIdx.x + k * kBlock;
      const int h = ci >= span ? 1 : 0, px = ci - h * span;
      const int x = x0 - d + px;
      v[k] = uint4{0u, 0u, 0u, 0u};
      if (ci < 2 * span && yr >= 0 && yr < ny && x >= 0 && x < nx)
        v[k] = *reinterpret_cast<const uint4*>(in + chunk_off(img + yr, bpr, x, h));
    }
  };
  auto store = [&](int slot, const uint4 (&v)[kChainLoads]) {
#pragma unroll
    for (int k = 0; k < kChainLoads; k++) {
      const int ci = threadIdx.x + k * kBlock;
      if (ci < 2 * span) ring[slot * 2 * span + ci] = v[k];
    }
  };
  uint4 v[kChainLoads];
  // chain index m of the segment is row y0 + m d, kept in ring slot m & 3
  fetch(y0 - d, v);
  store(3, v);
  fetch(y0, v);
  store(0, v);
  fetch(y0 + d, v);
  store(1, v);
  __syncthreads();
  const int h = g & 1;
  // Per-lane tap constants: operand p of lane group g is tap
  // tp = min(2p + g/2, 8), in ring row m + tp/3 - 1 at the element offset
  // col (its half and column).  Pixels past the image's last column only feed
  // their own (unstored) outputs, and out-of-image taps are zeros in the
  // ring, so the only operand replaced by zero is the missing 10th tap.
  int rowt[5], colt[5];
#pragma unroll
  for (int p = 0; p < 5; p++) {
    const int tp = min(2 * p + (g >> 1), 8);
    rowt[p] = tp / 3;
    colt[p] = h * span + n + d + (tp % 3 - 1) * d;
  }
  const bool tap9 = (g >> 1) == 1;
  const uint4 zero = {0u, 0u, 0u, 0u};
  // the lane's output element offsets within an image row (kConvU pixels)
  int oofs[kConvU];
#pragma unroll
  for (int u = 0; u < kConvU; u++) {
    const int x = x0 + (wave * kConvU + u) * 16 + n;
    oofs[u] = COUT == kC ? (int)act_off(0, bpr, x, g) : x;
  }
  for (int m = 0; m < R; m++) {
    const int y = y0 + m * d;
    if (y >= ny) break;  // workgroup-uniform
    const bool more = m + 1 < R && y + d < ny;
    if (more) fetch(y + 2 * d, v);
    const int rb[3] = {((m + 3) & 3) * 2 * span, (m & 3) * 2 * span, ((m + 1) & 3) * 2 * span};
    int off[5];
#pragma unroll
    for (int p = 0; p < 5; p++) off[p] = (rowt[p] == 0 ? rb[0] : (rowt[p] == 1 ? rb[1] : rb[2])) + colt[p];
    const int64_t row = img + y;
#pragma unroll
    for (int u = 0; u < kConvU; u++) {
      const int bx = wave * kConvU + u;
      const int x = x0 + bx * 16 + n;
      fx4 acc = fx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int p = 0; p < 5; p++) {
        uint4 bt = ring[off[p] + bx * 16];
        if (p == 4 && tap9) bt = zero;
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A8[p], __builtin_bit_cast(bf16x8, bt), acc, 0, 0, 0);
      }
      if (x >= nx) continue;
      if (COUT == kC) {
        float r[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          r[i] = acc[i] + bv[i];
          if (RELU) r[i] = fmaxf(r[i], 0.0f);
        }
        const int64_t o = row * bpr * 256 + oofs[u];
        if (a.mask) {  // workgroup-uniform
          const bfx4 mk = *reinterpret_cast<const bfx4*>(a.mask + o);
#pragma unroll
          for (int i = 0; i < 4; i++)
            if (!(bf2f((unsigned short)mk[i]) > 0.0f)) r[i] = 0.0f;
        }
        bfx4 ob;
#pragma unroll
        for (int i = 0; i < 4; i++) ob[i] = (short)f2bf(r[i]);
        *reinterpret_cast<bfx4*>(static_cast<unsigned short*>(a.out) + o) = ob;
      } else if (g == 0) {
        const int64_t pix = row * nx + oofs[u];
        const double val = (double)(acc[0] + bv[0]) * a.out_mul * (a.scale ? a.scale[bimg] : 1.0);
        double* out = static_cast<double*>(a.out);
        out[pix] = a.addto ? out[pix] + val : val;
      }
    }
    if (more) store((m + 2) & 3, v);
    __syncthreads();
  }
}

// The first layer (one FP64 input channel, CIN = 1) on the same row chains:
// the ring holds FP64 input rows of the strip; each lane converts its four
// taps (4g..4g+3 of the 9) to BF16 after the image's input scale and runs
// v_mfma_f32_16x16x16bf16_1k, exactly as k_conv<1, 16> does (the same
// results bit for bit).
template <bool RELU>
__global__ __launch_bounds__(kBlock) void k_conv1_chain(const ConvArgs a)
{
  extern __shared__ double ring1[];
  const int lane = threadIdx.x & (kWave - 1), g = lane >> 4, n = lane & 15, wave = threadIdx.x >> 6;
  const int nx = a.nx, ny = a.ny, d = a.dil;
  const int span = kChainW + 2 * d;
  const int bpr = (nx + 15) >> 4;
  const int strips = (nx + kChainW - 1) / kChainW;
  const int chains = min(d, ny);
  const int R = a.rows;
  const int segs = ((ny + d - 1) / d + R - 1) / R;
  int64_t t = blockIdx.x;
  const int strip = (int)(t % strips);
  t /= strips;
  const int c = (int)(t % chains);
  t /= chains;
  const int seg = (int)(t % segs);
  const int bimg = (int)(t / segs);
  const int64_t img = (int64_t)bimg * ny;
  const int x0 = strip * kChainW;
  const int y0 = c + seg * R * d;
  if (y0 >= ny) return;  // workgroup-uniform

  const bfx4 A1 = reinterpret_cast<const bfx4*>(a.apack)[lane];
  float bv[4] = {0.f, 0.f, 0.f, 0.f};
  if (a.bias) {
#pragma unroll
    for (int i = 0; i < 4; i++) bv[i] = a.bias[4 * g + i];
  }
  const double is = a.inv_scale ? a.inv_scale[bimg] : 1.0;
  const double* in = static_cast<const double*>(a.in);
  constexpr int kLoads1 = (kChainW + 2 * kChainMaxD + kBlock - 1) / kBlock;  // pixels per thread
  auto fetch = [&](int yr, double (&v)[kLoads1]) {
#pragma unroll
    for (int k = 0; k < kLoads1; k++) {
      const int px = threadIdx.x + k * kBlock;
      const int x = x0 - d + px;
      v[k] = 0.0;
      if (px < span && yr >= 0 && yr < ny && x >= 0 && x < nx) v[k] = in[(img + yr) * nx + x];
    }
  };
  auto store = [&](int slot, const double (&v)[kLoads1]) {
#pragma unroll
    for (int k = 0; k < kLoads1; k++) {
      const int px = threadIdx.x + k * kBlock;
      if (px < span) ring1[slot * span + px] = v[k];
    }
  };
  // the ring of k_conv_chain: chain index m in slot m & 3, the next row
  // fetched while the current one is computed
  double v[kLoads1];
  fetch(y0 - d, v);
  store(3, v);
  fetch(y0, v);
  store(0, v);
  fetch(y0 + d, v);
  store(1, v);
  __syncthreads();
  // per-lane tap constants (taps 4g .. 4g + 3; out-of-image taps are zeros
  // in the ring, and the K entries past the 9th tap are zero)
  int rowt[4], colt[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int tc = min(4 * g + j, 8);
    rowt[j] = tc / 3;
    colt[j] = n + d + (tc % 3 - 1) * d;
  }
  int oofs[kConvU];
#pragma unroll
  for (int u = 0; u < kConvU; u++) oofs[u] = (int)act_off(0, bpr, x0 + (wave * kConvU + u) * 16 + n, g);
  for (int m = 0; m < R; m++) {
    const int y = y0 + m * d;
    if (y >= ny) break;  // workgroup-uniform
    const bool more = m + 1 < R && y + d < ny;
    if (more) fetch(y + 2 * d, v);
    const int rb[3] = {((m + 3) & 3) * span, (m & 3) * span, ((m + 1) & 3) * span};
    int off[4];
#pragma unroll
    for (int j = 0; j < 4; j++) off[j] = (rowt[j] == 0 ? rb[0] : (rowt[j] == 1 ? rb[1] : rb[2])) + colt[j];
    const int64_t rowo = (img + y) * bpr * 256;
#pragma unroll
    for (int u = 0; u < kConvU; u++) {
      const int bx = wave * kConvU + u;
      const int x = x0 + bx * 16 + n;
      bfx4 B;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const double val = ring1[off[j] + bx * 16];
        B[j] = (short)(4 * g + j < 9 ? d2bf(val * is) : (unsigned short)0);
      }
      const fx4 acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A1, B, fx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      if (x >= nx) continue;
      float r[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        r[i] = acc[i] + bv[i];
        if (RELU) r[i] = fmaxf(r[i], 0.0f);
      }
      const int64_t o = rowo + oofs[u];
      if (a.mask) {  // workgroup-uniform
        const bfx4 mk = *reinterpret_cast<const bfx4*>(a.mask + o);
#pragma unroll
        for (int i = 0; i < 4; i++)
          if (!(bf2f((unsigned short)mk[i]) > 0.0f)) r[i] = 0.0f;
      }
      bfx4 ob;
#pragma unroll
      for (int i = 0; i < 4; i++) ob[i] = (short)f2bf(r[i]);
      *reinterpret_cast<bfx4*>(static_cast<unsigned short*>(a.out) + o) = ob;
    }
    if (more) store((m + 2) & 3, v);
    __syncthreads();
  }
}

// ------------------------------- two layers per sweep (network application)
// Two consecutive convolutions A (dilation da) and B (dilation db) of the
// symmetric network's application (no bias, no activation function) in one
// kernel, the intermediate rows kept in LDS instead of HBM: 64 bytes per
// pixel less traffic per fused pair.  When one dilation divides the other,
// both layers act along the rows of one residue class mod g = min(da, db):
// A couples chain steps k, k +- sa (sa = da / g), B couples k, k +- sb.  A
// workgroup takes kPairW output columns and `rows` chain steps of B; it
// sweeps the chain with two rings in LDS,
//   input rows of the strip (the B columns plus db + da halo pixels on each
//   side), 2 sa + 2 slots, and
//   A's output rows (the B columns plus db halo pixels on each side), BF16
//   exactly as k_conv(_chain) stores them to HBM, 2 sb + 2 slots.
// Step k computes A's row k and B's row k - sb - 1 (whose A rows k - 2sb - 1
// .. k - 1 were written in earlier steps) and stores the input row k + sa + 1
// fetched into registers at the start of the step; the slot counts keep the
// rows written in a step apart from the rows read in it, so that one barrier
// per step suffices.  B takes A's output as zero outside the image, as the
// separate layers' zero padding does.  The operands, the MFMA sequences and
// the roundings are those of the separate layers: the same results bit for
// bit.  Workgroups are laid out XCD-major (workgroup b runs on XCD b % 8 and
// takes the (b / 8)-th task of that XCD's contiguous share), so that the
// strips of one row segment, which share halo columns, meet in one L2.
constexpr int kPairThreads = 512;                // 8 wavefronts: one 16-pixel block of B each
constexpr int kPairWaves = kPairThreads / kWave;
constexpr int kPairW = kPairWaves * 16;          // output pixels per strip (B)
constexpr int kPairMaxD = 32;                    // larger dilations use the separate layers
constexpr int kPairLoads16 =
    (2 * (16 * ((kPairW + 2 * kPairMaxD + 15) / 16) + 2 * kPairMaxD) + kPairThreads - 1) / kPairThreads;
constexpr int kPairLoads1 = (16 * ((kPairW + 2 * kPairMaxD + 15) / 16) + 2 * kPairMaxD + kPairThreads - 1) / kPairThreads;

struct PairArgs {
  int nb, nx, ny, da, db, rows;
  int64_t tasks;                 // workgroups with work (the grid is rounded up to 8)
  const void* in;                // FP64 image (CIN_A = 1) or NHWC16 BF16
  const double* inv_scale;       // CIN_A = 1: per-image input multiplier
  const unsigned short* apack_a; // packed A operands of layer A
  const unsigned short* apack_b; // and of layer B
  void* out;                     // NHWC16 BF16, or an FP64 image (COUT_B = 1)
  const double* scale;           // COUT_B = 1: per-image output multiplier (nullptr: 1)
  double out_mul;                // COUT_B = 1: constant output multiplier
  int addto;                     // COUT_B = 1: out += result
};

struct PairGeom {
  int g, sa, sb;  // chain modulus, A and B offsets in chain steps
  int nba, wa;    // layer A's row: 16-pixel blocks, pixels (kPairW + 2 db rounded up)
  int wi;         // input row pixels (wa + 2 da)
  int mi, ma;     // ring slots
};
__host__ __device__ inline PairGeom pair_geom(int da, int db)
{
  PairGeom q;
  q.g = da < db ? da : db;
  q.sa = da / q.g;
  q.sb = db / q.g;
  q.nba = (kPairW + 2 * db + 15) / 16;
  q.wa = 16 * q.nba;
  q.wi = q.wa + 2 * da;
  q.mi = 2 * q.sa + 2;
  q.ma = 2 * q.sb + 2;
  return q;
}
inline size_t pair_lds(int da, int db, int cin_a)
{
  const PairGeom q = pair_geom(da, db);
  return sizeof(uint4) * 2 * (size_t)q.ma * q.wa + (cin_a == 1 ? sizeof(double) : 2 * sizeof(uint4)) * (size_t)q.mi * q.wi;
}
// chain steps of a segment: every residue class has ceil(ny / g) rows at most
inline int64_t pair_tasks(int nb, int nx, int ny, int da, int db, int rows)
{
  const PairGeom q = pair_geom(da, db);
  const int64_t strips = (nx + kPairW - 1) / kPairW;
  const int64_t chains = std::min(q.g, ny);
  const int64_t segs = ((ny + q.g - 1) / q.g + rows - 1) / rows;
  return (int64_t)nb * strips * chains * segs;
}
__device__ inline int ring_slot(int k, int m) { return ((k % m) + m) % m; }

template <int CIN_A, int COUT_B, int AHEAD>
__global__ __launch_bounds__(kPairThreads) void k_conv_pair(const PairArgs a)
{
  extern __shared__ uint4 lds[];
  const int lane = threadIdx.x & (kWave - 1), g = lane >> 4, n = lane & 15, wave = threadIdx.x >> 6;
  const int h = g & 1;
  const int nx = a.nx, ny = a.ny, da = a.da, db = a.db, R = a.rows;
  const PairGeom q = pair_geom(da, db);
  const int bpr = (nx + 15) >> 4;
  const int64_t per_xcd = (a.tasks + 7) >> 3;
  int64_t t = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if ((blockIdx.x >> 3) >= per_xcd || t >= a.tasks) return;  // workgroup-uniform
  // task: strip fastest, then chain, then segment, then image
  const int strips = (nx + kPairW - 1) / kPairW;
  const int chains = min(q.g, ny);
  const int segs = ((ny + q.g - 1) / q.g + R - 1) / R;
  const int strip = (int)(t % strips);
  t /= strips;
  const int c = (int)(t % chains);
  t /= chains;
  const int seg = (int)(t % segs);
  const int bimg = (int)(t / segs);
  const int64_t img = (int64_t)bimg * ny;
  const int x0 = strip * kPairW;
  const int y0 = c + seg * R * q.g;  // row of chain step 0 of the segment
  if (y0 >= ny) return;              // workgroup-uniform: an empty tail segment of a short chain
  const int xi0 = x0 - db - da;      // first input column of the strip

  uint4* ringA = lds;                                                      // [ma][half][wa]
  uint4* ringI16 = lds + 2 * q.ma * q.wa;                                  // CIN_A = 16: [mi][half][wi]
  double* ringI1 = reinterpret_cast<double*>(lds + 2 * q.ma * q.wa);       // CIN_A = 1: [mi][wi]

  bf16x8 A8a[5], A8b[5];
  bfx4 A1a = {0, 0, 0, 0};
  if (CIN_A == kC) {
#pragma unroll
    for (int p = 0; p < 5; p++) A8a[p] = __builtin_bit_cast(bf16x8, reinterpret_cast<const uint4*>(a.apack_a)[p * kWave + lane]);
  } else {
    A1a = reinterpret_cast<const bfx4*>(a.apack_a)[lane];
  }
#pragma unroll
  for (int p = 0; p < 5; p++) A8b[p] = __builtin_bit_cast(bf16x8, reinterpret_cast<const uint4*>(a.apack_b)[p * kWave + lane]);
  const double is = (CIN_A == 1 && a.inv_scale) ? a.inv_scale[bimg] : 1.0;

  // input row of chain step k (zeros outside the image) into one of AHEAD
  // register sets: rows are fetched AHEAD steps before they are stored
  constexpr int kLoads = CIN_A == kC ? kPairLoads16 : kPairLoads1;
  uint4 v16s[AHEAD][kLoads];
  double v1s[AHEAD][kLoads];
  auto fetch = [&](int k, uint4 (&v16)[kLoads], double (&v1)[kLoads]) {
    const int yr = y0 + k * q.g;
    const bool row_in = yr >= 0 && yr < ny;
#pragma unroll
    for (int i = 0; i < kLoads; i++) {
      const int ci = threadIdx.x + i * kPairThreads;
      if (CIN_A == kC) {
        const int hh = ci >= q.wi ? 1 : 0, px = ci - hh * q.wi;
        const int x = xi0 + px;
        v16[i] = uint4{0u, 0u, 0u, 0u};
        if (ci < 2 * q.wi && row_in && x >= 0 && x < nx)
          v16[i] = *reinterpret_cast<const uint4*>(static_cast<const unsigned short*>(a.in) + chunk_off(img + yr, bpr, x, hh));
      } else {
        const int x = xi0 + ci;
        v1[i] = 0.0;
        if (ci < q.wi && row_in && x >= 0 && x < nx) v1[i] = static_cast<const double*>(a.in)[(img + yr) * nx + x];
      }
    }
  };
  auto store = [&](int slot, const uint4 (&v16)[kLoads], const double (&v1)[kLoads]) {
#pragma unroll
    for (int i = 0; i < kLoads; i++) {
      const int ci = threadIdx.x + i * kPairThreads;
      if (CIN_A == kC) {
        if (ci < 2 * q.wi) ringI16[slot * 2 * q.wi + ci] = v16[i];
      } else {
        if (ci < q.wi) ringI1[slot * q.wi + ci] = v1[i];
      }
    }
  };

  // A's output row at chain step k (row yA inside the image), columns
  // x0 - db .. x0 - db + wa - 1, into ring slot k mod ma
  // Per-lane tap constants, the same in every step: the ring row (0, 1, 2
  // for k - s, k, k + s) and the element offset within the two halves of a
  // slot of each operand a lane reads.  Out-of-image input columns and rows
  // are zeros in the input ring, and A writes zeros for its pixels outside
  // the image, so the only operand ever replaced by zero is the missing 10th
  // tap (p = 4, lane groups 2 and 3).
  int rowB[5], colB[5], rowA[5], colA[5];
#pragma unroll
  for (int p = 0; p < 5; p++) {
    const int tp = min(2 * p + (g >> 1), 8);
    rowB[p] = tp / 3;
    colB[p] = h * q.wa + n + db + (tp % 3 - 1) * db;
    rowA[p] = tp / 3;
    colA[p] = h * q.wi + n + da + (tp % 3 - 1) * da;
  }
  int rowA1[4], colA1[4];  // CIN_A = 1: taps 4g .. 4g + 3
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int tc = min(4 * g + j, 8);
    rowA1[j] = tc / 3;
    colA1[j] = n + da + (tc % 3 - 1) * da;
  }
  const bool tap9 = (g >> 1) == 1;  // lane groups whose operand p = 4 is the missing tap
  const uint4 zero = {0u, 0u, 0u, 0u};
  auto pick = [](int r, const int (&b)[3]) { return r == 0 ? b[0] : (r == 1 ? b[1] : b[2]); };

  // A's output row at chain step k, columns x0 - db .. x0 - db + wa - 1, into
  // ring slot sA; sI: the ring slots of input rows k - sa, k, k + sa.  Rows
  // outside the image are written as zeros.
  auto layer_a = [&](int yA, const int (&sI)[3], int sA) {
    unsigned short* dst = reinterpret_cast<unsigned short*>(ringA + (size_t)sA * 2 * q.wa) + (g >> 1) * q.wa * 8 +
                          n * 8 + 4 * (g & 1);
    if (yA < 0 || yA >= ny) {  // workgroup-uniform
      for (int bA = wave; bA < q.nba; bA += kPairWaves) *reinterpret_cast<bfx4*>(dst + bA * 128) = bfx4{0, 0, 0, 0};
      return;
    }
    const int bI[3] = {sI[0] * 2 * q.wi, sI[1] * 2 * q.wi, sI[2] * 2 * q.wi};
    const int bI1[3] = {sI[0] * q.wi, sI[1] * q.wi, sI[2] * q.wi};
    int off[5], off1[4];
#pragma unroll
    for (int p = 0; p < 5; p++) off[p] = pick(rowA[p], bI) + colA[p];
#pragma unroll
    for (int jj = 0; jj < 4; jj++) off1[jj] = pick(rowA1[jj], bI1) + colA1[jj];
    for (int bA = wave; bA < q.nba; bA += kPairWaves) {  // wavefront-uniform
      const int xa = x0 - db + 16 * bA + n;
      fx4 acc = fx4{0.f, 0.f, 0.f, 0.f};
      if (CIN_A == kC) {
#pragma unroll
        for (int p = 0; p < 5; p++) {
          uint4 bt = ringI16[off[p] + 16 * bA];
          if (p == 4 && tap9) bt = zero;
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A8a[p], __builtin_bit_cast(bf16x8, bt), acc, 0, 0, 0);
        }
      } else {
        bfx4 B;
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
          const double val = ringI1[off1[jj] + 16 * bA];
          B[jj] = (short)(4 * g + jj < 9 ? d2bf(val * is) : (unsigned short)0);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A1a, B, acc, 0, 0, 0);
      }
      // "+ 0.0f": the separate layers add their (zero) bias, which turns a
      // -0 into +0
      bfx4 o;
#pragma unroll
      for (int i = 0; i < 4; i++) o[i] = (short)f2bf(acc[i] + 0.0f);
      if (xa < 0 || xa >= nx) o = bfx4{0, 0, 0, 0};
      *reinterpret_cast<bfx4*>(dst + bA * 128) = o;
    }
  };

  // B's output row at chain step j (row yB inside the image); sR: the ring
  // slots of A's rows j - sb, j, j + sb
  auto layer_b = [&](int yB, const int (&sR)[3]) {
    const int bR[3] = {sR[0] * 2 * q.wa, sR[1] * 2 * q.wa, sR[2] * 2 * q.wa};
    int off[5];
#pragma unroll
    for (int p = 0; p < 5; p++) off[p] = pick(rowB[p], bR) + colB[p];
    {
      const int bx = wave;
      const int x = x0 + bx * 16 + n;
      fx4 acc = fx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int p = 0; p < 5; p++) {
        uint4 bt = ringA[off[p] + 16 * bx];
        if (p == 4 && tap9) bt = zero;
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A8b[p], __builtin_bit_cast(bf16x8, bt), acc, 0, 0, 0);
      }
      if (x < nx && COUT_B == kC) {
        bfx4 o;
#pragma unroll
        for (int i = 0; i < 4; i++) o[i] = (short)f2bf(acc[i] + 0.0f);
        *reinterpret_cast<bfx4*>(static_cast<unsigned short*>(a.out) + act_off(img + yB, bpr, x, g)) = o;
      } else if (x < nx && g == 0) {
        const int64_t pix = (img + yB) * nx + x;
        const double val = (double)(acc[0] + 0.0f) * a.out_mul * (a.scale ? a.scale[bimg] : 1.0);
        double* out = static_cast<double*>(a.out);
        out[pix] = a.addto ? out[pix] + val : val;
      }
    }
  };

  // B's rows of the segment: chain steps 0 .. nB - 1 (the chain may end
  // first); A's rows -sb .. nB - 1 + sb; input rows up to kF1
  const int nB = min(R, (ny - y0 + q.g - 1) / q.g);
  const int kA0 = -q.sb, kA1 = nB - 1 + q.sb, kLast = nB + q.sb, kF1 = kA1 + q.sa;
  for (int k = kA0 - q.sa; k <= kA0 + q.sa; k++) {
    fetch(k, v16s[0], v1s[0]);
    store(ring_slot(k, q.mi), v16s[0], v1s[0]);
  }
#pragma unroll
  for (int j = 1; j < AHEAD; j++)
    if (kA0 + q.sa + j <= kF1) fetch(kA0 + q.sa + j, v16s[j], v1s[j]);
  __syncthreads();
  // ring slots, kept as k mod mi and k mod ma and moved along with k (no
  // division in the loop); the offsets below are taken mod the ring sizes
  // (mi = 2 sa + 2, ma = 2 sb + 2)
  auto addm = [](int x, int c, int m) { return x + c >= m ? x + c - m : x + c; };
  int ki = ring_slot(kA0, q.mi), ka = ring_slot(kA0, q.ma);
  // step s = k - kA0 fetches row k + sa + AHEAD into set s % AHEAD and stores
  // row k + sa + 1 (fetched AHEAD - 1 steps earlier) from set (s + 1) % AHEAD
  for (int k0 = kA0; k0 <= kLast; k0 += AHEAD) {
#pragma unroll
    for (int i = 0; i < AHEAD; i++) {
      const int k = k0 + i;
      if (k <= kLast) {  // workgroup-uniform
        const int kf = k + q.sa + AHEAD;
        if (kf <= kF1) fetch(kf, v16s[i], v1s[i]);
        if (k <= kA1) {
          const int sI[3] = {addm(ki, q.mi - q.sa, q.mi), ki, addm(ki, q.sa, q.mi)};
          layer_a(y0 + k * q.g, sI, ka);
        }
        const int jB = k - q.sb - 1;  // A's rows jB - sb, jB, jB + sb are k - 2sb - 1, k - sb - 1, k - 1
        if (jB >= 0) {
          const int sR[3] = {addm(ka, 1, q.ma), addm(ka, q.sb + 1, q.ma), addm(ka, q.ma - 1, q.ma)};
          layer_b(y0 + jB * q.g, sR);
        }
        if (k + q.sa + 1 <= kF1) store(addm(ki, q.sa + 1, q.mi), v16s[(i + 1) % AHEAD], v1s[(i + 1) % AHEAD]);
        __syncthreads();
        ki = addm(ki, 1, q.mi);
        ka = addm(ka, 1, q.ma);
      }
    }
  }
}

// ---------------------------------------------------------- weight gradient
struct WGradArgs {
  int nb, nx, ny, dil;
  const void* x;              // layer input: NHWC16 BF16, or FP64 image (CIN = 1)
  const double* inv_scale;    // CIN = 1
  const void* dy;             // gradient of the pre-activation: NHWC16, or FP64 (COUT = 1)
  float* partial;             // [waves][kPW]
};

template <int CIN, int COUT>
__global__ __launch_bounds__(kBlock) void k_wgrad(const WGradArgs a)
{
  const int lane = threadIdx.x & (kWave - 1), g = lane >> 4, n = lane & 15;
  constexpr int NT = (CIN == kC) ? 9 : 1;
  fx4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = fx4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.0f;
  const int nx = a.nx, ny = a.ny, d = a.dil;
  const int bpr = (nx + 15) >> 4;
  const int64_t nblk = (int64_t)a.nb * ny * bpr;
  const int wave = blockIdx.x * kWavesPB + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * kWavesPB;
  for (int64_t blk = wave; blk < nblk; blk += nwaves) {
    const int bx = (int)(blk % bpr);
    const int64_t rowi = blk / bpr;
    const int y = (int)(rowi % ny);
    const int64_t b = rowi / ny;
    // A[m = co = n][k = pixel 4g+j]
    bfx4 Aop;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int px = bx * 16 + 4 * g + j;
      unsigned short v = 0;
      if (px < nx) {
        const int64_t pix = (b * ny + y) * nx + px;
        if (COUT == kC) v = static_cast<const unsigned short*>(a.dy)[act_off(b * ny + y, bpr, px, n >> 2) + (n & 3)];
        else if (n == 0) v = d2bf(static_cast<const double*>(a.dy)[pix]);
      }
      Aop[j] = (short)v;
      bsum += bf2f(v);
    }
    if (CIN == kC) {
      const unsigned short* xin = static_cast<const unsigned short*>(a.x);
#pragma unroll
      for (int t = 0; t < 9; t++) {
        bfx4 Bop;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int px = bx * 16 + 4 * g + j;
          const int xs = px + (t % 3 - 1) * d, ys = y + (t / 3 - 1) * d;
          unsigned short v = 0;
          if (px < nx && xs >= 0 && xs < nx && ys >= 0 && ys < ny)
            v = xin[act_off(b * ny + ys, bpr, xs, n >> 2) + (n & 3)];
          Bop[j] = (short)v;
        }
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Aop, Bop, acc[t], 0, 0, 0);
      }
    } else {
      // B[k = pixel][n = tap]
      const double* xin = static_cast<const double*>(a.x);
      const double is = a.inv_scale ? a.inv_scale[b] : 1.0;
      bfx4 Bop;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int px = bx * 16 + 4 * g + j;
        unsigned short v = 0;
        if (n < 9 && px < nx) {
          const int xs = px + (n % 3 - 1) * d, ys = y + (n / 3 - 1) * d;
          if (xs >= 0 && xs < nx && ys >= 0 && ys < ny) v = d2bf(xin[(b * ny + ys) * nx + xs] * is);
        }
        Bop[j] = (short)v;
      }
      acc[0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Aop, Bop, acc[0], 0, 0, 0);
    }
  }
  float* P = a.partial + (int64_t)wave * kPW;
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int i = 0; i < 4; i++) P[t * 256 + (4 * g + i) * 16 + n] = acc[t][i];
  bsum += __shfl_xor(bsum, 16, kWave);
  bsum += __shfl_xor(bsum, 32, kWave);
  if (g == 0) P[9 * 256 + n] = bsum;
}

// sum of the per-wavefront partials into the parameter layout W[cout][cin][3][3],
// b[cout]: 64 consecutive partial offsets per block (coalesced reads), each
// summed over the wavefronts in 4 interleaved groups, the groups combined in
// a fixed order (deterministic), then scattered to the parameter it belongs to
// (added to it when accum is set; no bias gradient when gb is null)
__global__ __launch_bounds__(kBlock) void k_wreduce(const float* __restrict__ partial, int nwaves,
                                                    int cin, int cout, float* __restrict__ gw,
                                                    float* __restrict__ gb, int accum)
{
  const int q = threadIdx.x / 64, off = blockIdx.x * 64 + (threadIdx.x & 63);
  float s = 0.0f;
  if (off < kPW)
    for (int w = q; w < nwaves; w += 4) s += partial[(int64_t)w * kPW + off];
  __shared__ float sh[4][64];
  sh[q][threadIdx.x & 63] = s;
  __syncthreads();
  if (q != 0 || off >= kPW) return;
  const float v = ((sh[0][threadIdx.x] + sh[1][threadIdx.x]) + sh[2][threadIdx.x]) + sh[3][threadIdx.x];
  if (off >= 9 * 256) {
    const int b = off - 9 * 256;
    if (gb && b < cout) gb[b] = accum ? gb[b] + v : v;
    return;
  }
  const int t = off / 256, m = (off % 256) / 16, n = off % 16;
  int k = -1;
  if (cin == kC) {
    if (m < cout) k = (m * cin + n) * 9 + t;  // tile t: m = co, n = ci
  } else if (t == 0 && n < 9 && m < cout) {
    k = m * 9 + n;  // single tile: m = co, n = tap
  }
  if (k >= 0) gw[k] = accum ? gw[k] + v : v;
}

// ------------------------------------------------- packing of the A operands
// The A operand of a layer (cin -> cout) and of its adjoint (cout -> cin,
// taps flipped, channels exchanged), in the layouts k_conv reads, kPack
// elements per layer:
//  16 input channels (v_mfma_f32_16x16x32_bf16, taps in pairs): element
//    p*512 + lane*8 + j holds A[m = lane%16][tap 2p + g/2, channel 8(g%2)+j],
//    g = lane/16 (tap 9: zero);
//  1 input channel (v_mfma_f32_16x16x16bf16_1k): element lane*4 + j holds
//    A[m = lane%16][tap 4g+j] (taps 9..15: zero).
__global__ void k_pack(const float* __restrict__ w, int cin, int cout, unsigned short* __restrict__ afw,
                       unsigned short* __restrict__ abw)
{
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= kPack) return;
  auto W = [&](int co, int ci, int tap) -> float {
    return (co < cout && ci < cin && tap >= 0 && tap < 9) ? w[(co * cin + ci) * 9 + tap] : 0.0f;
  };
  // K = 32 layout indices
  const int p = idx / 512, lane = (idx / 8) % 64, j = idx % 8;
  const int g = lane / 16, m = lane % 16, tap = 2 * p + g / 2, ch = 8 * (g % 2) + j;
  // K = 16 layout indices (first 256 elements)
  const int lane1 = idx / 4, k1 = 4 * (lane1 / 16) + idx % 4, m1 = lane1 % 16;
  // forward: A[m = co][k = (tap, ci)]
  if (cin == kC) afw[idx] = f2bf(tap < 9 ? W(m, ch, tap) : 0.0f);
  else afw[idx] = (idx < 256) ? f2bf(W(m1, 0, k1)) : (unsigned short)0;
  // adjoint: A[m = ci][k = (tap, co)], weights of the flipped tap
  if (abw) {
    if (cout == kC) abw[idx] = f2bf(tap < 9 ? W(ch, m, 8 - tap) : 0.0f);
    else abw[idx] = (idx < 256) ? f2bf(W(0, m1, 8 - k1)) : (unsigned short)0;
  }
}


// ------------------------------------------------------------- small kernels
// per-image sums of squares (partials: P blocks per image)
__global__ __launch_bounds__(kBlock) void k_sumsq(int64_t npix, int P, const double* __restrict__ x,
                                                  double* __restrict__ part)
{
  const int b = blockIdx.y, p = blockIdx.x;
  const double* xb = x + (int64_t)b * npix;
  double s = 0.0;
  for (int64_t i = (int64_t)p * kBlock + threadIdx.x; i < npix; i += (int64_t)P * kBlock) s += xb[i] * xb[i];
  __shared__ double sh[kWavesPB];
  for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[(int64_t)b * P + p] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}
// s = RMS, inv = 1/s (an all-zero image keeps s = 1); one workgroup per
// image sums its P partials in a fixed order
__global__ __launch_bounds__(kBlock) void k_rms_final(int nb, int P, int64_t npix, const double* part,
                                                      double* s, double* inv)
{
  const int b = blockIdx.x;
  double t = 0.0;
  for (int p = threadIdx.x; p < P; p += kBlock) t += part[(int64_t)b * P + p];
  __shared__ double sh[kWavesPB];
  for (int off = kWave / 2; off > 0; off >>= 1) t += __shfl_xor(t, off, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = t;
  __syncthreads();
  if (threadIdx.x != 0) return;
  t = ((sh[0] + sh[1]) + sh[2]) + sh[3];
  double v = sqrt(t / (double)npix);
  if (!(v > 0.0)) v = 1.0;
  s[b] = v;
  inv[b] = 1.0 / v;
}

// (I - rx Dxx - ry Dyy) at one cell of a unit-spacing grid from its five
// values; rx == ry takes the isotropic association c - r((((w + e) + s) + n) - 4c)
__device__ inline double op5(double c, double w, double e, double s, double no, double rx, double ry)
{
  if (rx == ry) return c - rx * ((((w + e) + s) + no) - 4.0 * c);
  return c - (rx * ((w + e) - 2.0 * c) + ry * ((s + no) - 2.0 * c));
}
// y = (I - rx Dxx - ry Dyy) x on nb stacked nx*ny Dirichlet grids (unit spacing)
__global__ __launch_bounds__(kBlock) void k_applyA(int nb, int nx, int ny, double rx, double ry,
                                                   const double* __restrict__ x, double* __restrict__ y)
{
  const int64_t n = (int64_t)nb * nx * ny;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock) {
    const int i = (int)(k % nx);
    const int j = (int)((k / nx) % ny);
    const double c = x[k];
    const double w = i > 0 ? x[k - 1] : 0.0, e = i < nx - 1 ? x[k + 1] : 0.0;
    const double s = j > 0 ? x[k - nx] : 0.0, no = j < ny - 1 ? x[k + nx] : 0.0;
    y[k] = op5(c, w, e, s, no, rx, ry);
  }
}

// d = u - e*inv_s (u: network output in normalised units)
__global__ __launch_bounds__(kBlock) void k_diff(int64_t npix, int nb, const double* __restrict__ u,
                                                 const double* __restrict__ e, const double* __restrict__ inv,
                                                 double* __restrict__ dd)
{
  const int64_t n = npix * nb;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock)
    dd[k] = u[k] - e[k] * inv[k / npix];
}

// per-image partial dot products d.Ad and e.r
__global__ __launch_bounds__(kBlock) void k_dot2(int64_t npix, int P, const double* __restrict__ a1,
                                                 const double* __restrict__ b1, const double* __restrict__ a2,
                                                 const double* __restrict__ b2, double* __restrict__ part)
{
  const int b = blockIdx.y, p = blockIdx.x;
  const int64_t o = (int64_t)b * npix;
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = (int64_t)p * kBlock + threadIdx.x; i < npix; i += (int64_t)P * kBlock) {
    s1 += a1[o + i] * b1[o + i];
    s2 += a2[o + i] * b2[o + i];
  }
  __shared__ double sh[2][kWavesPB];
  for (int off = kWave / 2; off > 0; off >>= 1) {
    s1 += __shfl_xor(s1, off, kWave);
    s2 += __shfl_xor(s2, off, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    sh[0][threadIdx.x / kWave] = s1;
    sh[1][threadIdx.x / kWave] = s2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[((int64_t)b * P + p) * 2] = ((sh[0][0] + sh[0][1]) + sh[0][2]) + sh[0][3];
    part[((int64_t)b * P + p) * 2 + 1] = ((sh[1][0] + sh[1][1]) + sh[1][2]) + sh[1][3];
  }
}
// loss_b = (d.Ad) / (e.r * inv_s^2); coef_b = 2*out_mul / (nb * e.r * inv_s^2)
__global__ void k_loss_final(int nb, int P, const double* part, const double* inv, double out_mul,
                             double* loss_b, double* coef)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double s1 = 0.0, s2 = 0.0;
  for (int p = 0; p < P; p++) {
    s1 += part[((int64_t)b * P + p) * 2];
    s2 += part[((int64_t)b * P + p) * 2 + 1];
  }
  const double er = s2 * inv[b] * inv[b];
  loss_b[b] = s1 / er;
  coef[b] = 2.0 * out_mul / ((double)nb * er);
}
__global__ __launch_bounds__(kBlock) void k_scale_img(int64_t npix, int nb, const double* __restrict__ c,
                                                      const double* __restrict__ x, double* __restrict__ y)
{
  const int64_t n = npix * nb;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock)
    y[k] = c[k / npix] * x[k];
}
__global__ void k_mean(int nb, const double* v, double* out)
{
  if (blockIdx.x || threadIdx.x) return;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += v[b];
  *out = s / (double)nb;
}

// Adam on the FP32 parameters
__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float g = gr[k];
  m[k] = b1 * m[k] + (1.0f - b1) * g;
  v[k] = b2 * v[k] + (1.0f - b2) * g * g;
  const float mh = m[k] / c1, vh = v[k] / c2;
  p[k] = p[k] - lr * mh / (sqrtf(vh) + eps);
}

// gather training pairs idx[0..nb) into contiguous batch buffers
__global__ __launch_bounds__(kBlock) void k_gather(int64_t npix, int nb, const int* __restrict__ idx,
                                                   const double* __restrict__ r, const double* __restrict__ e,
                                                   double* __restrict__ rb, double* __restrict__ eb)
{
  const int64_t n = npix * nb;
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock) {
    const int64_t src = (int64_t)idx[k / npix] * npix + k % npix;
    rb[k] = r[src];
    eb[k] = e[src];
  }
}

// ---------------------------------------------------------- training data
__device__ inline uint64_t smix(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double u01(uint64_t h) { return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
// sine-mode coefficients c[i][k][l] = N(0,1)/|k| up to the sample's cut-off
// (Box-Muller on counter-based hashes of (seed + i, k, l))
__global__ __launch_bounds__(kBlock) void k_coeffs(int nsamp, int n, uint64_t seed, double* __restrict__ c)
{
  const int64_t tot = (int64_t)nsamp * n * n;
  for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < tot; q += (int64_t)gridDim.x * kBlock) {
    const int64_t i = q / ((int64_t)n * n);
    const int k = (int)((q / n) % n) + 1, l = (int)(q % n) + 1;
    const uint64_t si = smix(seed + (uint64_t)i);
    const double cut = 4.0 + (n - 4.0) * u01(smix(si ^ 0x5851F42D4C957F2Dull));
    const double kk = sqrt((double)(k * k + l * l));
    double v = 0.0;
    if (kk <= cut) {
      const uint64_t h1 = smix(si + 2ull * ((uint64_t)k * 4096ull + (uint64_t)l));
      const uint64_t h2 = smix(h1);
      v = sqrt(-2.0 * log(u01(h1))) * cos(6.283185307179586 * u01(h2)) / kk;
    }
    c[q] = v;
  }
}
// the two passes of the separable sine synthesis: pass 0
// t[i][k][x] = sum_l c[i][k][l] S[x][l], pass 1 r[i][y][x] = sum_k S[y][k] t[i][k][x].
// S is symmetric (its entries are sin(pi (k+1)(x+1)/(n+1)), the same double
// either way round), so pass 0 reads S[l][x], consecutive across the threads
__global__ __launch_bounds__(kBlock) void k_synth(int nsamp, int n, int pass, const double* __restrict__ S,
                                                  const double* __restrict__ in, double* __restrict__ out)
{
  const int64_t tot = (int64_t)nsamp * n * n;
  for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < tot; q += (int64_t)gridDim.x * kBlock) {
    const int64_t i = q / ((int64_t)n * n);
    const int a = (int)((q / n) % n), bcol = (int)(q % n);
    const double* base = in + i * n * n;
    double s = 0.0;
    if (pass == 0)
      for (int l = 0; l < n; l++) s += base[(int64_t)a * n + l] * S[(int64_t)l * n + bcol];
    else
      for (int k = 0; k < n; k++) s += S[(int64_t)a * n + k] * base[(int64_t)k * n + bcol];
    out[q] = s;
  }
}

// r[i] = a[i] for the odd samples i (the operator-applied fields), in place
__global__ __launch_bounds__(kBlock) void k_take_odd(int64_t npix, int nsamp, const double* __restrict__ a,
                                                     double* __restrict__ r)
{
  const int64_t tot = npix * nsamp;
  for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < tot; q += (int64_t)gridDim.x * kBlock)
    if ((q / npix) & 1) r[q] = a[q];
}

// batched Jacobi-PCG, one workgroup per pair (constant diagonal dg)
__device__ inline double block_sum(double v, double* sh)
{
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = v;
  __syncthreads();
  return ((sh[0] + sh[1]) + sh[2]) + sh[3];
}
// (a, b) summed over the workgroup, each exactly as block_sum sums it
__device__ inline void block_sum2(double& a, double& b, double* sh)
{
  for (int off = kWave / 2; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, kWave);
    b += __shfl_xor(b, off, kWave);
  }
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) {
    sh[threadIdx.x / kWave] = a;
    sh[kWavesPB + threadIdx.x / kWave] = b;
  }
  __syncthreads();
  a = ((sh[0] + sh[1]) + sh[2]) + sh[3];
  b = ((sh[kWavesPB] + sh[kWavesPB + 1]) + sh[kWavesPB + 2]) + sh[kWavesPB + 3];
}
struct CGState {
  double rz, bb, beta;  // r.z, b.b, the next direction's beta
  int done, iters;
};
// x = 0, r = b, z = b * idg (idg = 1 / dg, the constant inverse Jacobi
// diagonal), and the direction before the first one zero (so that the first
// direction r * idg + beta * 0 is z itself)
__global__ __launch_bounds__(kBlock) void k_cg_init(int n, double idg, const double* __restrict__ bvec,
                                                    double* __restrict__ x, double* __restrict__ r,
                                                    double* __restrict__ p, CGState* __restrict__ st)
{
  const int64_t o = (int64_t)blockIdx.x * n * n;
  __shared__ double sh[2 * kWavesPB];
  double rz = 0.0, bb = 0.0;
  for (int k = threadIdx.x; k < n * n; k += kBlock) {
    const double b = bvec[o + k];
    x[o + k] = 0.0;
    r[o + k] = b;
    p[o + k] = 0.0;
    rz += b * (b * idg);
    bb += b * b;
  }
  block_sum2(rz, bb, sh);
  if (threadIdx.x == 0) st[blockIdx.x] = CGState{rz, bb, 0.0, bb == 0.0 ? 1 : 0, 0};
}
// One Jacobi-PCG iteration on (I - rx Dxx - ry Dyy) of one pair per
// workgroup, in two sweeps over the grid (constant diagonal, z = r * idg):
//   sweep 1  p = z + beta p_old written to the other direction buffer, and
//            p.Ap with Ap formed from the new directions of the 5 points
//            (recomputed from r and p_old, so Ap is never stored);
//   sweep 2  Ap again from the stored p, x += alpha p, r -= alpha Ap, and
//            r.r and r.z.
// Eight vector sweeps of HBM per iteration instead of the textbook loop's
// eleven (ap = A p; alpha; x, r; beta; p = z + beta p).  PAIRS (even n): a
// thread takes two cells of a row with 16-byte accesses, and the cells west
// and east of the pair with 8-byte ones.
__device__ inline double2 cg_ld2(const double* __restrict__ a, int q) { return *reinterpret_cast<const double2*>(a + q); }
__device__ inline void cg_st2(double* __restrict__ a, int q, double u, double v)
{
  *reinterpret_cast<double2*>(a + q) = double2{u, v};
}
template <bool PAIRS>
__global__ __launch_bounds__(kBlock) void k_cg_iter(int n, double rx, double ry, double idg, double tol2,
                                                    double* __restrict__ x, double* __restrict__ r,
                                                    double* __restrict__ pa, double* __restrict__ pb,
                                                    CGState* __restrict__ st)
{
  const CGState S = st[blockIdx.x];
  if (S.done) return;
  const int64_t o = (int64_t)blockIdx.x * n * n;
  double* __restrict__ xs = x + o;
  double* __restrict__ rs = r + o;
  const double* __restrict__ pin = ((S.iters & 1) ? pb : pa) + o;
  double* __restrict__ pout = ((S.iters & 1) ? pa : pb) + o;
  const double beta = S.beta;
  __shared__ double sh[2 * kWavesPB];
  auto dir = [&](int q) { return rs[q] * idg + beta * pin[q]; };
  double pap = 0.0, rr = 0.0, rz = 0.0, alpha;
  if (PAIRS) {
    // cells (2m, 2m + 1); m walks the n/2 pairs of each row
    const int h = n / 2, npair = h * n, di = kBlock % h, dj = kBlock / h, i0 = threadIdx.x % h,
              j0 = threadIdx.x / h;
    for (int m = threadIdx.x, i = i0, j = j0; m < npair; m += kBlock) {
      const int q = 2 * m, c = 2 * i;
      const double2 rc = cg_ld2(rs, q), pc = cg_ld2(pin, q);
      const double c0 = rc.x * idg + beta * pc.x, c1 = rc.y * idg + beta * pc.y;
      const double w = c > 0 ? dir(q - 1) : 0.0, e = c + 2 < n ? dir(q + 2) : 0.0;
      double s0 = 0.0, s1 = 0.0, n0 = 0.0, n1 = 0.0;
      if (j > 0) {
        const double2 ra = cg_ld2(rs, q - n), pv = cg_ld2(pin, q - n);
        s0 = ra.x * idg + beta * pv.x;
        s1 = ra.y * idg + beta * pv.y;
      }
      if (j < n - 1) {
        const double2 ra = cg_ld2(rs, q + n), pv = cg_ld2(pin, q + n);
        n0 = ra.x * idg + beta * pv.x;
        n1 = ra.y * idg + beta * pv.y;
      }
      cg_st2(pout, q, c0, c1);
      pap += c0 * op5(c0, w, c1, s0, n0, rx, ry);
      pap += c1 * op5(c1, c0, e, s1, n1, rx, ry);
      i += di;
      j += dj;
      if (i >= h) i -= h, j++;
    }
    // the barriers of the sum also make every sweep-1 store of pout visible
    // to the workgroup before sweep 2 reads its neighbours
    pap = block_sum(pap, sh);
    alpha = S.rz / pap;
    for (int m = threadIdx.x, i = i0, j = j0; m < npair; m += kBlock) {
      const int q = 2 * m, c = 2 * i;
      const double2 pc = cg_ld2(pout, q);
      const double w = c > 0 ? pout[q - 1] : 0.0, e = c + 2 < n ? pout[q + 2] : 0.0;
      const double2 ps = j > 0 ? cg_ld2(pout, q - n) : double2{0.0, 0.0};
      const double2 pn = j < n - 1 ? cg_ld2(pout, q + n) : double2{0.0, 0.0};
      const double v0 = op5(pc.x, w, pc.y, ps.x, pn.x, rx, ry);
      const double v1 = op5(pc.y, pc.x, e, ps.y, pn.y, rx, ry);
      const double2 xc = cg_ld2(xs, q), rc = cg_ld2(rs, q);
      cg_st2(xs, q, xc.x + alpha * pc.x, xc.y + alpha * pc.y);
      const double r0 = rc.x - alpha * v0, r1 = rc.y - alpha * v1;
      cg_st2(rs, q, r0, r1);
      rr += r0 * r0;
      rr += r1 * r1;
      rz += r0 * (r0 * idg);
      rz += r1 * (r1 * idg);
      i += di;
      j += dj;
      if (i >= h) i -= h, j++;
    }
  } else {
    const int nn = n * n, di = kBlock % n, dj = kBlock / n, i0 = threadIdx.x % n, j0 = threadIdx.x / n;
    for (int k = threadIdx.x, i = i0, j = j0; k < nn; k += kBlock) {
      const double c = dir(k);
      const double w = i > 0 ? dir(k - 1) : 0.0, e = i < n - 1 ? dir(k + 1) : 0.0;
      const double s = j > 0 ? dir(k - n) : 0.0, no = j < n - 1 ? dir(k + n) : 0.0;
      pout[k] = c;
      pap += c * op5(c, w, e, s, no, rx, ry);
      i += di;
      j += dj;
      if (i >= n) i -= n, j++;
    }
    pap = block_sum(pap, sh);  // its barriers order the pout stores before sweep 2
    alpha = S.rz / pap;
    for (int k = threadIdx.x, i = i0, j = j0; k < nn; k += kBlock) {
      const double c = pout[k];
      const double w = i > 0 ? pout[k - 1] : 0.0, e = i < n - 1 ? pout[k + 1] : 0.0;
      const double s = j > 0 ? pout[k - n] : 0.0, no = j < n - 1 ? pout[k + n] : 0.0;
      const double v = op5(c, w, e, s, no, rx, ry);
      xs[k] += alpha * c;
      const double rk = rs[k] - alpha * v;
      rs[k] = rk;
      rr += rk * rk;
      rz += rk * (rk * idg);
      i += di;
      j += dj;
      if (i >= n) i -= n, j++;
    }
  }
  block_sum2(rr, rz, sh);
  if (threadIdx.x == 0) st[blockIdx.x] = CGState{rz, S.bb, rz / S.rz, rr <= tol2 * S.bb ? 1 : 0, S.iters + 1};
}

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, 8192); }

}  // namespace

// ============================================ fixed workspace (no allocation)
// LearnedPrec_ApplyNoAlloc keeps the network application's workspace here
// instead of in allocations: two activation buffers (ping-pong) of up to
// 2^20 padded pixels (image rows padded to whole 16-pixel blocks, summed over
// the images) with 16 BF16 channels each, and for up to 64 images the RMS
// scale, its inverse and the per-image partial sums.  65 MiB per GPU the
// library runs on, held until the process ends.
namespace {
constexpr int kNAMaxImages = 64;
constexpr int64_t kNAMaxPixels = (int64_t)1 << 20;
constexpr int64_t kNAActElems = kNAMaxPixels * 16;  // BF16 elements per activation buffer
constexpr int kNAMaxParts = 512;                    // RMS partials per image (rms_parts)
constexpr int64_t kNAScratchBytes = 2 * kNAActElems * (int64_t)sizeof(unsigned short) +
                                    (int64_t)kNAMaxImages * (2 + kNAMaxParts) * (int64_t)sizeof(double);
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNAScratchBytes];

namespace {
constexpr int kNAMaxDevices = 64;
std::mutex g_na_mutex;
void* g_na_base[kNAMaxDevices] = {};

// device address of the fixed workspace on the current device, resolved by
// hipGetSymbolAddress once per device (host code never uses the array's
// name, which is only a host-side stand-in for it)
int na_scratch(void** base)
{
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNAMaxDevices) return -1;
  std::lock_guard<std::mutex> g(g_na_mutex);
  if (!g_na_base[dev]) {
    void* p = nullptr;
    hipError_t e;
    {
      // may load the code object: kept out of other threads' graph captures
      std::lock_guard<std::recursive_mutex> d(ceda::device_sync_mutex());
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    }
    if (e != hipSuccess || !p) {
      fprintf(stderr, "learned preconditioner: hipGetSymbolAddress of the fixed workspace failed on device %d: %s\n",
              dev, hipGetErrorName(e));
      return -1;
    }
    g_na_base[dev] = p;
  }
  *base = g_na_base[dev];
  return 0;
}
}  // namespace

// ============================================================== host side
struct LearnedPrec {
  int nl = 0;
  bool sym = false;  // M = K^T K with K the layers (linear, no biases)
  int dil[kMaxL], cin[kMaxL], cout[kMaxL], woff[kMaxL], boff[kMaxL];
  int np = 0;
  double out_scale = 1.0;
  ceda::Exec* ex = nullptr;
  float* d_p = nullptr;      // FP32 parameters
  float* d_g = nullptr;      // gradient
  float *d_m = nullptr, *d_v = nullptr;
  unsigned short *d_afw = nullptr, *d_abw = nullptr;  // [layer][9*256]
  int adam_t = 0;
  // workspace
  int64_t cap_pix = 0, cap_act = 0, cap_g = 0;
  int cap_nb = 0;
  unsigned short* act[kMaxL] = {};  // act[l] = input of layer l (l >= 1), NHWC16; sym: act[nl] = K x
  unsigned short* uact[kMaxL] = {}; // sym: uact[l] = input of the adjoint of layer l - 1
  unsigned short *gA = nullptr, *gB = nullptr;  // gradient ping-pong, NHWC16
  double *u = nullptr, *dd = nullptr, *ad = nullptr, *gy = nullptr;
  double *sc = nullptr, *isc = nullptr, *part = nullptr, *lossb = nullptr, *coef = nullptr;
  float* wpart = nullptr;
  int part_cap = 0;
};

namespace {

void lp_pack(LearnedPrec* L)
{
  hipStream_t s = L->ex->stream;
  for (int l = 0; l < L->nl; l++)
    hipLaunchKernelGGL(k_pack, dim3((kPack + 255) / 256), dim3(256), 0, s, L->d_p + L->woff[l], L->cin[l], L->cout[l],
                       L->d_afw + (int64_t)l * kPack,
                       (l > 0 || L->sym) ? L->d_abw + (int64_t)l * kPack : nullptr);
  CEDA_HIP_CHECK(hipGetLastError());
}

constexpr int kMaxParts = 512;  // workgroups per image of the per-image reductions
static_assert(kNAMaxParts >= kMaxParts, "the fixed workspace holds every image's RMS partials");
// one workgroup per 2048 pixels (8 per thread), at most kMaxParts: enough
// workgroups in flight that the sum of a small image is not latency-bound
int rms_parts(int64_t npix) { return (int)std::min<int64_t>(kMaxParts, (npix + 2047) / 2048); }

// activation buffers hold nb*ny rows of whole 16-pixel blocks
int64_t act_elems(int nb, int nx, int ny) { return (int64_t)nb * ny * ((nx + 15) / 16) * 16 * kC; }

// device workspace for nb images of nx*ny pixels, grown on demand: the two
// activation ping-pong buffers and the per-image scale and partial sums that
// an application needs, and with train the buffers only training needs (the
// activations kept for the backward pass, the FP64 residual/error images)
void lp_workspace(LearnedPrec* L, int nb, int nx, int ny, bool train)
{
  const int64_t npix = (int64_t)nx * ny;
  const int64_t tot = npix * nb;
  const int64_t ae = act_elems(nb, nx, ny);
  const int P = rms_parts(npix);
  if (ae > L->cap_g) {
    CEDA_HIP_CHECK(hipStreamSynchronize(L->ex->stream));
    void* ps[] = {L->gA, L->gB};
    for (void* p : ps)
      if (p) CEDA_HIP_CHECK(ceda::dev_free(p));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L->gA, sizeof(unsigned short) * ae));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L->gB, sizeof(unsigned short) * ae));
    L->cap_g = ae;
  }
  if (train && (tot > L->cap_pix || ae > L->cap_act)) {
    CEDA_HIP_CHECK(hipStreamSynchronize(L->ex->stream));
    for (int l = 0; l < kMaxL; l++) {
      if (L->act[l]) CEDA_HIP_CHECK(ceda::dev_free(L->act[l]));
      if (L->uact[l]) CEDA_HIP_CHECK(ceda::dev_free(L->uact[l]));
      L->act[l] = L->uact[l] = nullptr;
    }
    void* ps[] = {L->u, L->dd, L->ad, L->gy};
    for (void* p : ps)
      if (p) CEDA_HIP_CHECK(ceda::dev_free(p));
    for (int l = 1; l < L->nl + (L->sym ? 1 : 0); l++)
      CEDA_HIP_CHECK(ceda::dev_malloc(&L->act[l], sizeof(unsigned short) * ae));
    if (L->sym)
      for (int l = 1; l < L->nl; l++) CEDA_HIP_CHECK(ceda::dev_malloc(&L->uact[l], sizeof(unsigned short) * ae));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L->u, sizeof(double) * tot));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L->dd, sizeof(double) * tot));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L->ad, sizeof(double) * tot));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L->gy, sizeof(double) * tot));
    L->cap_pix = tot;
    L->cap_act = ae;
  }
  if (nb > L->cap_nb || nb * P * 2 > L->part_cap) {
    CEDA_HIP_CHECK(hipStreamSynchronize(L->ex->stream));
    void* ps[] = {L->sc, L->isc, L->part, L->lossb, L->coef};
    for (void* p : ps)
      if (p) CEDA_HIP_CHECK(ceda::dev_free(p));
    const int cnb = std::max(nb, L->cap_nb);
    // room for every image's largest partial count: images that grow later
    // need no new allocation
    const int pc = std::max(cnb * kMaxParts * 2, L->part_cap);
    CEDA_HIP_CHECK(ceda::dev_malloc(&L->sc, sizeof(double) * cnb));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L->isc, sizeof(double) * cnb));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L->part, sizeof(double) * pc));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L->lossb, sizeof(double) * cnb));
    CEDA_HIP_CHECK(ceda::dev_malloc(&L->coef, sizeof(double) * cnb));
    L->cap_nb = cnb;
    L->part_cap = pc;
  }
  if (train && !L->wpart) CEDA_HIP_CHECK(ceda::dev_malloc(&L->wpart, sizeof(float) * kPW * kWGradWaves));
}


// per-image scale s = RMS(in) and 1/s
void lp_scale(LearnedPrec* L, int nb, int64_t npix, const double* in)
{
  hipStream_t s = L->ex->stream;
  const int P = rms_parts(npix);
  hipLaunchKernelGGL(k_sumsq, dim3(P, nb), dim3(kBlock), 0, s, npix, P, in, L->part);
  hipLaunchKernelGGL(k_rms_final, dim3(nb), dim3(kBlock), 0, s, nb, P, npix, L->part, L->sc, L->isc);
}

// 1: the 16-input-channel layers and the first layer use the LDS row-chain
// kernels (dilations up to kChainMaxD); 2: as 1, and the symmetric network's
// application runs two layers per kernel (k_conv_pair) where their dilations
// allow; 0: every layer uses k_conv.  1 is the fastest: the row-chain kernels
// stream near the HBM limit, and k_conv_pair, which saves 64 bytes of traffic
// per pixel and pair, is held back by the occupancy its two LDS rings allow.
int g_conv_chain = 1;
int g_pair_rows = 16;  // chain steps per workgroup of k_conv_pair
int g_pair_ahead = 2;  // input rows in flight per workgroup of k_conv_pair (1..3)

template <int CIN, int COUT, bool RELU>
void conv(LearnedPrec* L, const ConvArgs& a)
{
  if (CIN == kC && g_conv_chain && a.dil >= 1 && a.dil <= kChainMaxD) {
    ConvArgs b = a;
    b.rows = chain_rows(a.nb, a.nx, a.ny, a.dil);
    const int64_t tasks = chain_tasks(a.nb, a.nx, a.ny, a.dil, b.rows);
    const size_t lds = sizeof(uint4) * 4 * 2 * (size_t)(kChainW + 2 * a.dil);
    hipLaunchKernelGGL((k_conv_chain<COUT, RELU>), dim3((unsigned)tasks), dim3(kBlock), lds, L->ex->stream, b);
    return;
  }
  if (CIN == 1 && COUT == kC && g_conv_chain && a.dil >= 1 && a.dil <= kChainMaxD) {
    ConvArgs b = a;
    b.rows = chain_rows(a.nb, a.nx, a.ny, a.dil);
    const int64_t tasks = chain_tasks(a.nb, a.nx, a.ny, a.dil, b.rows);
    const size_t lds = sizeof(double) * 4 * (size_t)(kChainW + 2 * a.dil);
    hipLaunchKernelGGL((k_conv1_chain<RELU>), dim3((unsigned)tasks), dim3(kBlock), lds, L->ex->stream, b);
    return;
  }
  // one row segment (kWavesPB * kConvU blocks of 16 pixels) per workgroup
  // and step; a multiple of 8 workgroups (the kernel's XCD bands)
  const int64_t segs = (int64_t)a.nb * a.ny * (((a.nx + 15) / 16 + kWavesPB * kConvU - 1) / (kWavesPB * kConvU));
  const unsigned grid = (unsigned)std::min<int64_t>((segs + 7) / 8 * 8, 4096);
  hipLaunchKernelGGL((k_conv<CIN, COUT, RELU>), dim3(grid), dim3(kBlock), 0, L->ex->stream, a);
}

// LDS a workgroup may use on the current device
size_t lds_per_block()
{
  static int v = -1;
  if (v < 0) {
    int dev = 0, x = 0;
    CEDA_HIP_CHECK(hipGetDevice(&dev));
    CEDA_HIP_CHECK(hipDeviceGetAttribute(&x, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    v = x;
  }
  return (size_t)v;
}

// whether k_conv_pair<CIN_A, COUT_B> can run layers of dilations da, db:
// one divides the other, both at most kPairMaxD, the rings fit in LDS (above
// 64 KiB once the kernel has been allowed that much dynamic LDS)
template <int CIN_A, int COUT_B>
bool pair_ok(int da, int db)
{
  if (da < 1 || db < 1 || da > kPairMaxD || db > kPairMaxD) return false;
  if ((da < db ? db % da : da % db) != 0) return false;
  const size_t lds = pair_lds(da, db, CIN_A);
  if (lds > lds_per_block()) return false;
  if (lds <= 65536) return true;
  static std::atomic<int> state{0};  // 0 not tried, 1 allowed, -1 refused
  if (state.load() == 0) {
    const void* fs[] = {(const void*)k_conv_pair<CIN_A, COUT_B, 1>, (const void*)k_conv_pair<CIN_A, COUT_B, 2>,
                        (const void*)k_conv_pair<CIN_A, COUT_B, 3>};
    bool allowed = true;
    for (const void* f : fs) {
      const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_per_block());
      if (e != hipSuccess) {
        (void)hipGetLastError();
        allowed = false;
      }
    }
    state.store(allowed ? 1 : -1);
  }
  return state.load() == 1;
}

template <int CIN_A, int COUT_B>
void conv_pair(LearnedPrec* L, PairArgs a)
{
  a.rows = g_pair_rows;
  a.tasks = pair_tasks(a.nb, a.nx, a.ny, a.da, a.db, a.rows);
  const int64_t grid = (a.tasks + 7) / 8 * 8;
  const size_t lds = pair_lds(a.da, a.db, CIN_A);
  if (g_pair_ahead == 1)
    hipLaunchKernelGGL((k_conv_pair<CIN_A, COUT_B, 1>), dim3((unsigned)grid), dim3(kPairThreads), lds, L->ex->stream, a);
  else if (g_pair_ahead == 2)
    hipLaunchKernelGGL((k_conv_pair<CIN_A, COUT_B, 2>), dim3((unsigned)grid), dim3(kPairThreads), lds, L->ex->stream, a);
  else
    hipLaunchKernelGGL((k_conv_pair<CIN_A, COUT_B, 3>), dim3((unsigned)grid), dim3(kPairThreads), lds, L->ex->stream, a);
}

// symmetric network: h_0 = in/s, h_{l+1} = W_l * h_l (l < nl), then the
// adjoint layers u_nl = h_nl, u_l = W_l^T * u_{l+1} (the input-gradient
// convolution: flipped taps, exchanged channels), out = [out +] scale *
// out_scale * u_0.  No biases and no activation function, so the map is
// linear, symmetric and positive semi-definite up to the BF16 roundings.
void lp_forward_sym(LearnedPrec* L, int nb, int nx, int ny, const double* in, double* out, int addto,
                    bool normalised)
{
  ConvArgs a{};
  a.nb = nb;
  a.nx = nx;
  a.ny = ny;
  const int nl = L->nl;
  for (int l = 0; l < nl; l++) {
    a.dil = L->dil[l];
    a.apack = L->d_afw + (int64_t)l * kPack;
    a.in = (l == 0) ? (const void*)in : (const void*)L->act[l];
    a.inv_scale = (l == 0) ? L->isc : nullptr;
    a.out = L->act[l + 1];
    if (l == 0) conv<1, kC, false>(L, a);
    else conv<kC, kC, false>(L, a);
  }
  a.inv_scale = nullptr;
  for (int l = nl - 1; l >= 0; l--) {
    a.dil = L->dil[l];
    a.apack = L->d_abw + (int64_t)l * kPack;
    a.in = (l == nl - 1) ? (const void*)L->act[nl] : (const void*)L->uact[l + 1];
    if (l == 0) {
      a.out = out;
      a.scale = normalised ? nullptr : L->sc;
      a.out_mul = L->out_scale;
      a.addto = addto;
      conv<kC, 1, false>(L, a);
    } else {
      a.out = L->uact[l];
      conv<kC, kC, false>(L, a);
    }
  }
  CEDA_HIP_CHECK(hipGetLastError());
}

// forward pass; the output (FP64) goes to out with out = [out +] scale*out_scale*N
// (scale = per-image RMS, or 1 in normalised units when normalised = true)
void lp_forward(LearnedPrec* L, int nb, int nx, int ny, const double* in, double* out, int addto,
                bool normalised)
{
  if (L->sym) {
    lp_forward_sym(L, nb, nx, ny, in, out, addto, normalised);
    return;
  }
  ConvArgs a{};
  a.nb = nb;
  a.nx = nx;
  a.ny = ny;
  for (int l = 0; l < L->nl; l++) {
    a.dil = L->dil[l];
    a.apack = L->d_afw + (int64_t)l * kPack;
    a.bias = L->d_p + L->boff[l];
    a.mask = nullptr;
    const bool first = (l == 0), last = (l == L->nl - 1);
    a.in = first ? (const void*)in : (const void*)L->act[l];
    a.inv_scale = first ? L->isc : nullptr;
    if (last) {
      a.out = out;
      a.scale = normalised ? nullptr : L->sc;
      a.out_mul = L->out_scale;
      a.addto = addto;
      if (first) conv<1, 1, false>(L, a);
      else conv<kC, 1, false>(L, a);
    } else {
      a.out = L->act[l + 1];
      if (first) conv<1, kC, true>(L, a);
      else conv<kC, kC, true>(L, a);
    }
  }
  CEDA_HIP_CHECK(hipGetLastError());
}

// the network applied with its workspace in given device memory: two
// activation buffers used in turn (each layer reads one and writes the
// other), the per-image scale and inverse, the RMS partials.  The same
// kernels and operands as lp_scale + lp_forward, so the same result bit for
// bit.
void lp_apply_in(LearnedPrec* L, int nb, int nx, int ny, const double* in, double* out, int addto,
                 unsigned short* buf0, unsigned short* buf1, double* sc, double* isc, double* part)
{
  hipStream_t s = L->ex->stream;
  const int64_t npix = (int64_t)nx * ny;
  const int P = rms_parts(npix);
  hipLaunchKernelGGL(k_sumsq, dim3(P, nb), dim3(kBlock), 0, s, npix, P, in, part);
  hipLaunchKernelGGL(k_rms_final, dim3(nb), dim3(kBlock), 0, s, nb, P, npix, part, sc, isc);
  ConvArgs a{};
  a.nb = nb;
  a.nx = nx;
  a.ny = ny;
  const int nl = L->nl;
  const void* src = in;
  unsigned short* dst = buf0;
  auto advance = [&]() {
    src = dst;
    dst = (dst == buf0) ? buf1 : buf0;
  };
  if (L->sym) {
    // the 2 nl layers of the application: layer i < nl is the forward layer
    // i, layer i >= nl the adjoint of layer 2 nl - 1 - i
    const int nops = 2 * nl;
    auto op_dil = [&](int i) { return L->dil[i < nl ? i : nops - 1 - i]; };
    auto op_pack = [&](int i) {
      return i < nl ? L->d_afw + (int64_t)i * kPack : L->d_abw + (int64_t)(nops - 1 - i) * kPack;
    };
    for (int i = 0; i < nops;) {
      const bool first = (i == 0);
      if (g_conv_chain == 2 && i + 1 < nops) {
        const bool last = (i + 2 == nops);
        const int da = op_dil(i), db = op_dil(i + 1);
        const bool ok = first ? (last ? pair_ok<1, 1>(da, db) : pair_ok<1, kC>(da, db))
                              : (last ? pair_ok<kC, 1>(da, db) : pair_ok<kC, kC>(da, db));
        if (ok) {
          PairArgs pa{};
          pa.nb = nb;
          pa.nx = nx;
          pa.ny = ny;
          pa.da = da;
          pa.db = db;
          pa.in = src;
          pa.inv_scale = first ? isc : nullptr;
          pa.apack_a = op_pack(i);
          pa.apack_b = op_pack(i + 1);
          if (last) {
            pa.out = out;
            pa.scale = sc;
            pa.out_mul = L->out_scale;
            pa.addto = addto;
          } else {
            pa.out = dst;
          }
          if (first) {
            if (last) conv_pair<1, 1>(L, pa);
            else conv_pair<1, kC>(L, pa);
          } else {
            if (last) conv_pair<kC, 1>(L, pa);
            else conv_pair<kC, kC>(L, pa);
          }
          if (!last) advance();
          i += 2;
          continue;
        }
      }
      ConvArgs c{};
      c.nb = nb;
      c.nx = nx;
      c.ny = ny;
      c.dil = op_dil(i);
      c.apack = op_pack(i);
      c.in = src;
      c.inv_scale = first ? isc : nullptr;
      if (i == nops - 1) {
        c.out = out;
        c.scale = sc;
        c.out_mul = L->out_scale;
        c.addto = addto;
        conv<kC, 1, false>(L, c);
      } else {
        c.out = dst;
        if (first) conv<1, kC, false>(L, c);
        else conv<kC, kC, false>(L, c);
        advance();
      }
      i += 1;
    }
  } else {
    for (int l = 0; l < nl; l++) {
      a.dil = L->dil[l];
      a.apack = L->d_afw + (int64_t)l * kPack;
      a.bias = L->d_p + L->boff[l];
      a.mask = nullptr;
      const bool first = (l == 0), last = (l == nl - 1);
      a.in = src;
      a.inv_scale = first ? isc : nullptr;
      if (last) {
        a.out = out;
        a.scale = sc;
        a.out_mul = L->out_scale;
        a.addto = addto;
        if (first) conv<1, 1, false>(L, a);
        else conv<kC, 1, false>(L, a);
      } else {
        a.out = dst;
        if (first) conv<1, kC, true>(L, a);
        else conv<kC, kC, true>(L, a);
        advance();
      }
    }
  }
  CEDA_HIP_CHECK(hipGetLastError());
}

// weight gradient of layer l from input x (FP64 images scaled by isc, or
// NHWC16) and output gradient dy; written to the gradient, or added to it
template <int CIN, int COUT>
void wgrad(LearnedPrec* L, int l, int nb, int nx, int ny, const void* x, const double* isc, const void* dy,
           int accum)
{
  WGradArgs w{nb, nx, ny, L->dil[l], x, isc, dy, L->wpart};
  // about 8 pixel blocks per wavefront, at most kWGradWaves partials
  const int64_t nblk = (int64_t)nb * ny * ((nx + 15) / 16);
  const int nw = (int)std::min<int64_t>(kWGradWaves, std::max<int64_t>(kWavesPB, ((nblk + 7) / 8 + 3) / 4 * 4));
  hipLaunchKernelGGL((k_wgrad<CIN, COUT>), dim3(nw / kWavesPB), dim3(kBlock), 0, L->ex->stream, w);
  hipLaunchKernelGGL(k_wreduce, dim3((kPW + 63) / 64), dim3(kBlock), 0, L->ex->stream, L->wpart, nw,
                     L->cin[l], L->cout[l], L->d_g + L->woff[l], L->sym ? nullptr : L->d_g + L->boff[l],
                     accum);
}

// backward pass of the symmetric network from the output gradient L->gy.
// Each layer's weights are used twice, so its gradient has two parts: from
// the adjoint layer u_l = W_l^T u_{l+1} with output gradient g,
//   dW[co][ci][t] = sum_p u_{l+1}[p][co] g[p + off_t][ci]
// (the weight-gradient GEMM with the roles of input and output gradient
// exchanged), written first; then from the layer h_{l+1} = W_l h_l, added.
// The gradient of an adjoint layer's input is the forward convolution of its
// output gradient, that of a layer's input the adjoint convolution.
void lp_backward_sym(LearnedPrec* L, int nb, int nx, int ny, const double* r)
{
  const int nl = L->nl;
  ConvArgs a{};
  a.nb = nb;
  a.nx = nx;
  a.ny = ny;
  const void* g = L->gy;  // FP64 for the adjoint of layer 0
  unsigned short* gout = L->gA;
  for (int l = 0; l < nl; l++) {
    const void* uin = (l == nl - 1) ? (const void*)L->act[nl] : (const void*)L->uact[l + 1];
    if (l == 0) wgrad<1, kC>(L, l, nb, nx, ny, g, nullptr, uin, 0);
    else wgrad<kC, kC>(L, l, nb, nx, ny, g, nullptr, uin, 0);
    a.dil = L->dil[l];
    a.apack = L->d_afw + (int64_t)l * kPack;
    a.in = g;
    a.out = gout;
    if (l == 0) conv<1, kC, false>(L, a);
    else conv<kC, kC, false>(L, a);
    g = gout;
    gout = (gout == L->gA) ? L->gB : L->gA;
  }
  // g is now the gradient of h_nl = u_nl
  for (int l = nl - 1; l >= 0; l--) {
    if (l == 0) {
      wgrad<1, kC>(L, l, nb, nx, ny, r, L->isc, g, 1);
      break;
    }
    wgrad<kC, kC>(L, l, nb, nx, ny, L->act[l], nullptr, g, 1);
    a.dil = L->dil[l];
    a.apack = L->d_abw + (int64_t)l * kPack;
    a.in = g;
    a.out = gout;
    conv<kC, kC, false>(L, a);
    g = gout;
    gout = (gout == L->gA) ? L->gB : L->gA;
  }
  CEDA_HIP_CHECK(hipGetLastError());
}

// loss (device scalar *dloss) and parameter gradient (L->d_g) of a batch
void lp_lossgrad(LearnedPrec* L, int nb, int nx, int ny, double rx, double ry, const double* r,
                 const double* e, double* dloss)
{
  hipStream_t s = L->ex->stream;
  const int64_t npix = (int64_t)nx * ny, tot = npix * nb;
  lp_workspace(L, nb, nx, ny, true);
  lp_scale(L, nb, npix, r);
  lp_forward(L, nb, nx, ny, r, L->u, 0, true);
  const unsigned gg = grid_for(tot);
  hipLaunchKernelGGL(k_diff, dim3(gg), dim3(kBlock), 0, s, npix, nb, L->u, e, L->isc, L->dd);
  hipLaunchKernelGGL(k_applyA, dim3(gg), dim3(kBlock), 0, s, nb, nx, ny, rx, ry, L->dd, L->ad);
  const int P = rms_parts(npix);
  hipLaunchKernelGGL(k_dot2, dim3(P, nb), dim3(kBlock), 0, s, npix, P, L->dd, L->ad, e, r, L->part);
  hipLaunchKernelGGL(k_loss_final, dim3((nb + 63) / 64), dim3(64), 0, s, nb, P, L->part, L->isc,
                     L->out_scale, L->lossb, L->coef);
  hipLaunchKernelGGL(k_mean, dim3(1), dim3(1), 0, s, nb, L->lossb, dloss);
  // dL/d(pre-activation of the last layer) = coef_b * A d
  hipLaunchKernelGGL(k_scale_img, dim3(gg), dim3(kBlock), 0, s, npix, nb, L->coef, L->ad, L->gy);
  if (L->sym) {
    lp_backward_sym(L, nb, nx, ny, r);
    return;
  }
  // backward through the layers
  const int last = L->nl - 1;
  const void* dy = L->gy;  // FP64 for the last layer
  unsigned short* gout = L->gA;
  for (int l = last; l >= 0; l--) {
    const bool first = (l == 0), isl = (l == last);
    const void* x = first ? (const void*)r : (const void*)L->act[l];
    if (first && isl) wgrad<1, 1>(L, l, nb, nx, ny, x, L->isc, dy, 0);
    else if (first) wgrad<1, kC>(L, l, nb, nx, ny, x, L->isc, dy, 0);
    else if (isl) wgrad<kC, 1>(L, l, nb, nx, ny, x, L->isc, dy, 0);
    else wgrad<kC, kC>(L, l, nb, nx, ny, x, L->isc, dy, 0);
    if (first) break;
    ConvArgs a{};
    a.nb = nb;
    a.nx = nx;
    a.ny = ny;
    a.dil = L->dil[l];
    a.in = dy;
    a.out = gout;
    a.apack = L->d_abw + (int64_t)l * kPack;
    a.mask = L->act[l];
    if (isl) conv<1, kC, false>(L, a);
    else conv<kC, kC, false>(L, a);
    dy = gout;
    gout = (gout == L->gA) ? L->gB : L->gA;
  }
  CEDA_HIP_CHECK(hipGetLastError());
}

void adam_step(LearnedPrec* L, double lr)
{
  L->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)L->adam_t), c2 = 1.0f - powf(b2, (float)L->adam_t);
  hipLaunchKernelGGL(k_adam, dim3((L->np + 255) / 256), dim3(256), 0, L->ex->stream, L->np, L->d_p, L->d_g,
                     L->d_m, L->d_v, (float)lr, b1, b2, 1e-8f, c1, c2);
  lp_pack(L);
}

LearnedPrec* lp_create(int nlayer, const int* dil, SUNContext ctx, bool sym)
{
  LearnedPrec* L = new LearnedPrec();
  L->nl = nlayer;
  L->sym = sym;
  L->ex = ceda::context_exec(ctx);
  int off = 0;
  for (int l = 0; l < nlayer; l++) {
    L->dil[l] = dil[l];
    L->cin[l] = (l == 0) ? 1 : kC;
    L->cout[l] = (l == nlayer - 1 && !sym) ? 1 : kC;
    L->woff[l] = off;
    off += L->cout[l] * L->cin[l] * 9;
    L->boff[l] = off;
    off += L->cout[l];
  }
  L->np = off;
  void* na = nullptr;  // the fixed workspace's address on this device, for LearnedPrec_ApplyNoAlloc
  (void)na_scratch(&na);
  CEDA_HIP_CHECK(ceda::dev_malloc(&L->d_p, sizeof(float) * off));
  CEDA_HIP_CHECK(ceda::dev_malloc(&L->d_g, sizeof(float) * off));
  CEDA_HIP_CHECK(ceda::dev_malloc(&L->d_m, sizeof(float) * off));
  CEDA_HIP_CHECK(ceda::dev_malloc(&L->d_v, sizeof(float) * off));
  CEDA_HIP_CHECK(ceda::set_sync(L->d_p, 0, sizeof(float) * off));
  CEDA_HIP_CHECK(ceda::set_sync(L->d_g, 0, sizeof(float) * off));
  CEDA_HIP_CHECK(ceda::set_sync(L->d_m, 0, sizeof(float) * off));
  CEDA_HIP_CHECK(ceda::set_sync(L->d_v, 0, sizeof(float) * off));
  CEDA_HIP_CHECK(ceda::dev_malloc(&L->d_afw, sizeof(unsigned short) * kPack * nlayer));
  CEDA_HIP_CHECK(ceda::dev_malloc(&L->d_abw, sizeof(unsigned short) * kPack * nlayer));
  CEDA_HIP_CHECK(ceda::set_sync(L->d_afw, 0, sizeof(unsigned short) * kPack * nlayer));
  CEDA_HIP_CHECK(ceda::set_sync(L->d_abw, 0, sizeof(unsigned short) * kPack * nlayer));
  lp_pack(L);
  CEDA_HIP_CHECK(hipStreamSynchronize(L->ex->stream));
  return L;
}

}  // namespace

extern "C" {

LearnedPrec* LearnedPrec_Create(int nlayer, const int* dil, SUNContext ctx)
{
  if (nlayer < 2 || nlayer > kMaxL || !dil || !ctx) return nullptr;
  for (int l = 0; l < nlayer; l++)
    if (dil[l] < 1 || dil[l] > 64) return nullptr;
  return lp_create(nlayer, dil, ctx, false);
}

LearnedPrec* LearnedPrec_CreateSymmetric(int nlayer, const int* dil, SUNContext ctx)
{
  if (nlayer < 1 || nlayer > kMaxL - 1 || !dil || !ctx) return nullptr;
  for (int l = 0; l < nlayer; l++)
    if (dil[l] < 1 || dil[l] > 64) return nullptr;
  return lp_create(nlayer, dil, ctx, true);
}

int LearnedPrec_IsSymmetric(const LearnedPrec* L) { return L ? (L->sym ? 1 : 0) : -1; }

void LearnedPrec_Destroy(LearnedPrec* L)
{
  if (!L) return;
  CEDA_HIP_CHECK(hipStreamSynchronize(L->ex->stream));
  void* ps[] = {L->d_p, L->d_g, L->d_m, L->d_v, L->d_afw, L->d_abw, L->gA, L->gB, L->u, L->dd,
                L->ad, L->gy, L->sc, L->isc, L->part, L->lossb, L->coef, L->wpart};
  for (void* p : ps)
    if (p) CEDA_HIP_CHECK(ceda::dev_free(p));
  for (int l = 0; l < kMaxL; l++) {
    if (L->act[l]) CEDA_HIP_CHECK(ceda::dev_free(L->act[l]));
    if (L->uact[l]) CEDA_HIP_CHECK(ceda::dev_free(L->uact[l]));
  }
  delete L;
}

int LearnedPrec_NumParams(const LearnedPrec* L) { return L ? L->np : -1; }

int LearnedPrec_SetParams(LearnedPrec* L, const float* p)
{
  if (!L || !p) return -1;
  CEDA_HIP_CHECK(hipMemcpyAsync(L->d_p, p, sizeof(float) * L->np, hipMemcpyHostToDevice, L->ex->stream));
  lp_pack(L);
  CEDA_HIP_CHECK(hipStreamSynchronize(L->ex->stream));
  return 0;
}

int LearnedPrec_GetParams(const LearnedPrec* L, float* p)
{
  if (!L || !p) return -1;
  CEDA_HIP_CHECK(hipMemcpyAsync(p, L->d_p, sizeof(float) * L->np, hipMemcpyDeviceToHost, L->ex->stream));
  CEDA_HIP_CHECK(hipStreamSynchronize(L->ex->stream));
  return 0;
}

int LearnedPrec_InitIdentity(LearnedPrec* L, double noise, unsigned long seed)
{
  if (!L || noise < 0.0) return -1;
  std::vector<float> p(L->np, 0.0f);
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  for (int l = 0; l < L->nl; l++) {
    const int nw = L->cout[l] * L->cin[l] * 9;
    for (int k = 0; k < nw; k++) p[L->woff[l] + k] = (float)(noise * U(rng));
    auto W = [&](int co, int ci, int t) -> float& { return p[L->woff[l] + (co * L->cin[l] + ci) * 9 + t]; };
    const bool first = (l == 0), last = (l == L->nl - 1);
    if (L->sym) W(0, 0, 4) += 1.0f;  // K passes channel 0 through: K^T K = I
    else if (first && last) W(0, 0, 4) += 1.0f;
    else if (first) {
      W(0, 0, 4) += 1.0f;   // relu(x)
      W(1, 0, 4) -= 1.0f;   // relu(-x)
    } else if (last) {
      W(0, 0, 4) += 1.0f;
      W(0, 1, 4) -= 1.0f;
    } else {
      W(0, 0, 4) += 1.0f;
      W(1, 1, 4) += 1.0f;
    }
  }
  L->adam_t = 0;
  CEDA_HIP_CHECK(hipMemsetAsync(L->d_m, 0, sizeof(float) * L->np, L->ex->stream));
  CEDA_HIP_CHECK(hipMemsetAsync(L->d_v, 0, sizeof(float) * L->np, L->ex->stream));
  return LearnedPrec_SetParams(L, p.data());
}

int LearnedPrec_SetOutputScale(LearnedPrec* L, double s)
{
  if (!L || !(s > 0.0)) return -1;
  L->out_scale = s;
  return 0;
}

double LearnedPrec_GetOutputScale(const LearnedPrec* L) { return L ? L->out_scale : 0.0; }

int LearnedPrec_SetConvKernel(int kind)
{
  if (kind < 0 || kind > 2) return -1;
  g_conv_chain = kind;
  return 0;
}

int LearnedPrec_SetPairTuning(int rows, int ahead)
{
  if (rows < 1 || rows > 1024 || ahead < 1 || ahead > 3) return -1;
  g_pair_rows = rows;
  g_pair_ahead = ahead;
  return 0;
}

int LearnedPrec_SetChainTasks(int min_tasks)
{
  if (min_tasks < 1 || min_tasks > (1 << 20)) return -1;
  g_chain_tasks = min_tasks;
  return 0;
}

int LearnedPrec_Apply(LearnedPrec* L, int nb, int nx, int ny, const double* in, double* out, int addto)
{
  if (!L || nb < 1 || nx < 1 || ny < 1 || !in || !out) return -1;
  lp_workspace(L, nb, nx, ny, false);
  // the application keeps no activations: the two gradient buffers of
  // training serve as the ping-pong pair
  lp_apply_in(L, nb, nx, ny, in, out, addto, L->gA, L->gB, L->sc, L->isc, L->part);
  return 0;
}

int LearnedPrec_ApplyNoAlloc(LearnedPrec* L, int nb, int nx, int ny, const double* in, double* out, int addto)
{
  if (!L || nb < 1 || nx < 1 || ny < 1 || !in || !out) return -1;
  if (nb > kNAMaxImages || act_elems(nb, nx, ny) > kNAActElems) {
    fprintf(stderr, "LearnedPrec_ApplyNoAlloc: %d image(s) of %d x %d exceed the fixed workspace (at most %d "
                    "images and %lld pixels in all, rows padded to multiples of 16)\n",
            nb, nx, ny, kNAMaxImages, (long long)kNAMaxPixels);
    return -2;
  }
  void* base = nullptr;
  if (na_scratch(&base)) return -1;
  unsigned short* buf0 = static_cast<unsigned short*>(base);
  unsigned short* buf1 = buf0 + kNAActElems;
  double* sc = reinterpret_cast<double*>(buf1 + kNAActElems);
  double* isc = sc + kNAMaxImages;
  double* part = isc + kNAMaxImages;
  lp_apply_in(L, nb, nx, ny, in, out, addto, buf0, buf1, sc, isc, part);
  return 0;
}

// the network applied with a given output scale (the network's own one is
// left as it is)
static int lp_apply_scaled(LearnedPrec* L, int nx, int ny, const double* in, double* out, double out_scale)
{
  const double keep = L->out_scale;
  L->out_scale = out_scale;
  const int r = LearnedPrec_Apply(L, 1, nx, ny, in, out, 0);
  L->out_scale = keep;
  return r;
}

int LearnedPrec_LossGradXY(LearnedPrec* L, int nb, int nx, int ny, double ratio_x, double ratio_y,
                           const double* r, const double* e, double* loss, float* grad)
{
  if (!L || nb < 1 || nx < 1 || ny < 1 || !(ratio_x >= 0.0) || !(ratio_y >= 0.0) || !r || !e || !loss)
    return -1;
  double* dl = nullptr;
  CEDA_HIP_CHECK(ceda::dev_malloc(&dl, sizeof(double)));
  lp_lossgrad(L, nb, nx, ny, ratio_x, ratio_y, r, e, dl);
  CEDA_HIP_CHECK(hipMemcpyAsync(loss, dl, sizeof(double), hipMemcpyDeviceToHost, L->ex->stream));
  if (grad)
    CEDA_HIP_CHECK(hipMemcpyAsync(grad, L->d_g, sizeof(float) * L->np, hipMemcpyDeviceToHost, L->ex->stream));
  CEDA_HIP_CHECK(hipStreamSynchronize(L->ex->stream));
  CEDA_HIP_CHECK(ceda::dev_free(dl));
  return 0;
}

int LearnedPrec_LossGrad(LearnedPrec* L, int nb, int nx, int ny, double ratio, const double* r,
                         const double* e, double* loss, float* grad)
{
  return LearnedPrec_LossGradXY(L, nb, nx, ny, ratio, ratio, r, e, loss, grad);
}

int LearnedPrec_MakeTrainingSetXY(int nsamp, int n, double ratio_x, double ratio_y, double reltol,
                                  unsigned long seed, double* r, double* e, int* max_iters, SUNContext ctx)
{
  if (nsamp < 1 || n < 2 || n > 1024 || !(ratio_x >= 0.0) || !(ratio_y >= 0.0) || !(reltol > 0.0) || !r ||
      !e || !ctx)
    return -1;
  hipStream_t s = ceda::context_exec(ctx)->stream;
  const int64_t tot = (int64_t)nsamp * n * n;
  std::vector<double> S((size_t)n * n);
  for (int x = 0; x < n; x++)
    for (int k = 0; k < n; k++)  // (k+1)(x+1) exact: S[x][k] == S[k][x] bitwise
      S[(size_t)x * n + k] = sin(M_PI * ((double)(k + 1) * (double)(x + 1)) / (double)(n + 1));
  double *dS = nullptr, *tmp = nullptr, *p = nullptr, *p2 = nullptr;
  CGState* st = nullptr;
  CEDA_HIP_CHECK(ceda::dev_malloc(&dS, sizeof(double) * n * n));
  CEDA_HIP_CHECK(ceda::dev_malloc(&tmp, sizeof(double) * tot));
  CEDA_HIP_CHECK(ceda::dev_malloc(&p, sizeof(double) * tot));
  CEDA_HIP_CHECK(ceda::dev_malloc(&p2, sizeof(double) * tot));
  CEDA_HIP_CHECK(ceda::dev_malloc(&st, sizeof(CGState) * nsamp));
  CEDA_HIP_CHECK(hipMemcpyAsync(dS, S.data(), sizeof(double) * n * n, hipMemcpyHostToDevice, s));
  const unsigned gg = grid_for(tot);
  // fields: even samples smooth sine sums, odd samples (I - rx Dxx - ry Dyy) applied to
  // one (rough: the high-frequency content a Krylov residual carries)
  hipLaunchKernelGGL(k_coeffs, dim3(gg), dim3(kBlock), 0, s, nsamp, n, (uint64_t)seed, p);
  hipLaunchKernelGGL(k_synth, dim3(gg), dim3(kBlock), 0, s, nsamp, n, 0, dS, p, tmp);
  hipLaunchKernelGGL(k_synth, dim3(gg), dim3(kBlock), 0, s, nsamp, n, 1, dS, tmp, r);
  hipLaunchKernelGGL(k_applyA, dim3(gg), dim3(kBlock), 0, s, nsamp, n, n, ratio_x, ratio_y, r, tmp);
  hipLaunchKernelGGL(k_take_odd, dim3(gg), dim3(kBlock), 0, s, (int64_t)n * n, nsamp, tmp, r);
  // Jacobi-PCG on every pair
  const double idg = 1.0 / (1.0 + (2.0 * ratio_x + 2.0 * ratio_y));  // inverse Jacobi diagonal
  hipLaunchKernelGGL(k_cg_init, dim3(nsamp), dim3(kBlock), 0, s, n, idg, r, e, tmp, p, st);
  std::vector<CGState> hst(nsamp);
  const int maxit = 20000;
  int it = 0;
  bool all = false;
  while (!all && it < maxit) {
    for (int k = 0; k < 50; k++, it++)
      if (n % 2 == 0)
        hipLaunchKernelGGL(k_cg_iter<true>, dim3(nsamp), dim3(kBlock), 0, s, n, ratio_x, ratio_y, idg,
                           reltol * reltol, e, tmp, p, p2, st);
      else
        hipLaunchKernelGGL(k_cg_iter<false>, dim3(nsamp), dim3(kBlock), 0, s, n, ratio_x, ratio_y, idg,
                           reltol * reltol, e, tmp, p, p2, st);
    CEDA_HIP_CHECK(hipMemcpyAsync(hst.data(), st, sizeof(CGState) * nsamp, hipMemcpyDeviceToHost, s));
    CEDA_HIP_CHECK(hipStreamSynchronize(s));
    all = true;
    for (const CGState& c : hst) all = all && c.done;
  }
  int mx = 0;
  for (const CGState& c : hst) mx = std::max(mx, c.iters);
  if (max_iters) *max_iters = mx;
  CEDA_HIP_CHECK(hipGetLastError());
  CEDA_HIP_CHECK(hipStreamSynchronize(s));
  void* ps[] = {dS, tmp, p, p2, st};
  for (void* q : ps) CEDA_HIP_CHECK(ceda::dev_free(q));
  return all ? 0 : 1;
}

int LearnedPrec_MakeTrainingSet(int nsamp, int n, double ratio, double reltol, unsigned long seed,
                                double* r, double* e, int* max_iters, SUNContext ctx)
{
  return LearnedPrec_MakeTrainingSetXY(nsamp, n, ratio, ratio, reltol, seed, r, e, max_iters, ctx);
}

int LearnedPrec_TrainXY(LearnedPrec* L, int nsamp, int n, double ratio_x, double ratio_y, const double* r,
                        const double* e, int steps, int batch, double lr0, unsigned long seed, double* loss_hist)
{
  if (!L || nsamp < 1 || n < 2 || steps < 1 || batch < 1 || !(ratio_x >= 0.0) || !(ratio_y >= 0.0) || !r ||
      !e || !(lr0 > 0.0))
    return -1;
  hipStream_t s = L->ex->stream;
  const int64_t npix = (int64_t)n * n;
  double *rb = nullptr, *eb = nullptr, *dloss = nullptr;
  int* didx = nullptr;
  CEDA_HIP_CHECK(ceda::dev_malloc(&rb, sizeof(double) * npix * batch));
  CEDA_HIP_CHECK(ceda::dev_malloc(&eb, sizeof(double) * npix * batch));
  CEDA_HIP_CHECK(ceda::dev_malloc(&dloss, sizeof(double) * steps));
  CEDA_HIP_CHECK(ceda::dev_malloc(&didx, sizeof(int) * batch * steps));
  std::vector<int> idx((size_t)batch * steps);
  std::mt19937_64 rng(seed);
  std::uniform_int_distribution<int> U(0, nsamp - 1);
  for (int& v : idx) v = U(rng);
  CEDA_HIP_CHECK(hipMemcpyAsync(didx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, s));
  const unsigned gg = grid_for(npix * batch);
  for (int it = 0; it < steps; it++) {
    hipLaunchKernelGGL(k_gather, dim3(gg), dim3(kBlock), 0, s, npix, batch, didx + (int64_t)it * batch, r, e, rb, eb);
    lp_lossgrad(L, batch, n, n, ratio_x, ratio_y, rb, eb, dloss + it);
    const double lr = lr0 * 0.5 * (1.0 + cos(M_PI * (double)it / (double)steps));
    adam_step(L, lr);
  }
  if (loss_hist)
    CEDA_HIP_CHECK(hipMemcpyAsync(loss_hist, dloss, sizeof(double) * steps, hipMemcpyDeviceToHost, s));
  CEDA_HIP_CHECK(hipStreamSynchronize(s));
  void* ps[] = {rb, eb, dloss, didx};
  for (void* q : ps) CEDA_HIP_CHECK(ceda::dev_free(q));
  return 0;
}

int LearnedPrec_Train(LearnedPrec* L, int nsamp, int n, double ratio, const double* r, const double* e,
                      int steps, int batch, double lr0, unsigned long seed, double* loss_hist)
{
  return LearnedPrec_TrainXY(L, nsamp, n, ratio, ratio, r, e, steps, batch, lr0, seed, loss_hist);
}

int DiffusionProblem_SetLearnedPrec(DiffusionProblem* P, LearnedPrec* lp)
{
  if (!P) return -1;
  // the network's kernels run on its context's stream; the preconditioner's
  // vectors are on the problem's: one stream keeps them ordered
  if (lp && lp->ex != P->exec) {
    fprintf(stderr, "DiffusionProblem_SetLearnedPrec: the network and the problem belong to different "
                    "SUNContexts (streams); create the network on the problem's context\n");
    return -1;
  }
  P->learned = lp;
  P->learned_out_scale = lp ? lp->out_scale : 1.0;
  return 0;
}

int LearnedPSetup(sunrealtype t, N_Vector u, N_Vector f, sunbooleantype jok, sunbooleantype* jcurPtr,
                  sunrealtype gamma, void* user_data)
{
  DiffusionProblem* P = static_cast<DiffusionProblem*>(user_data);
  if (!P || !P->learned) return -1;
  const int r = PSetup(t, u, f, jok, jcurPtr, gamma, P);
  if (r) return r;
  const double d = 1.0 + gamma * (2.0 * P->p.kx / (P->dx * P->dx) + 2.0 * P->p.ky / (P->dy * P->dy));
  // kept with the problem: problems sharing one network keep their own scale
  P->learned_out_scale = 1.0 / d;
  return 0;
}

int LearnedPSolve(sunrealtype t, N_Vector u, N_Vector f, N_Vector r, N_Vector z, sunrealtype gamma,
                  sunrealtype delta, int lr, void* user_data)
{
  DiffusionProblem* P = static_cast<DiffusionProblem*>(user_data);
  if (!P || !P->learned || N_VGetLength(r) != P->n || N_VGetLength(z) != P->n) return -1;
  return lp_apply_scaled(P->learned, (int)P->p.nx, (int)P->p.ny, N_VGetDeviceArrayPointer(r),
                         N_VGetDeviceArrayPointer(z), P->learned_out_scale);
}

}  // extern "C"
