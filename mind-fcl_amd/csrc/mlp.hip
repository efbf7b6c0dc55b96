// Learned GJK warm start for gfx950.
//
// A 39-128-128-128-4 multilayer perceptron with GELU (erf form) maps the
// relative pose of a pair plus a 16-float descriptor of each shape to an
// initial GJK search direction (3, in shape a's frame) and a signed-distance
// estimate (1).  Its output only seeds GJK through the reference's guess hook
// (GJKSolver::gjk_guess, include/fcl/narrowphase/detail/gjk_solver.h:140-163);
// the exact algorithm still decides every result.
//
// Every product runs on the MFMA units as BF16 x BF16 -> FP32 through one
// hand-written kernel, C[M][N] = A[M][K] . B[N][K]^T with both operands
// K-contiguous (v_mfma_f32_32x32x16_bf16: one 32x32 tile per wavefront,
// split-K with FP32 atomics for the weight gradients).  Activations are
// sample-major; the epilogues also write the transposed copies that the
// weight-gradient products read.  Training keeps FP32 master weights with
// Adam; BF16 copies of W and W^T feed the forward and backward products.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "ctx.h"

namespace mfcl {

#define MFCL_LDS __attribute__((address_space(3)))

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kBlk = 256;

// ------------------------------------------------------------------ GEMM
enum Epi : int {
  kEpiStore = 0,     // C (FP32, row-major ldc); atomic add when split over K
  kEpiBiasGelu = 1,  // Z = C + b (FP32), H = gelu(Z) (BF16), H^T (BF16)
  kEpiBias = 2,      // Y = C + b (FP32)
  kEpiGeluGrad = 3,  // dZ = C * gelu'(Z) -> BF16 and its transpose
  kEpiGelu = 4,      // H = gelu(C + b) (BF16 only: inference)
};

struct EpiArgs {
  float* c;          // kEpiStore / kEpiBias / kEpiBiasGelu(Z)
  int ldc;
  const float* bias;
  const float* z;    // kEpiGeluGrad: saved pre-activations, row-major ldc
  __bf16* h;         // BF16 output, row-major ldc
  __bf16* ht;        // BF16 transposed output, row-major ldt
  int ldt;
  int atomic;        // kEpiStore: accumulate with atomics
};

__device__ __forceinline__ float geluF(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float geluGradF(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

// One wavefront per 32x32 output tile; blockIdx.y selects the K chunk.
// Register layout of the 32x32 accumulator (16 FP32 per lane):
//   acc[r] = C[m0 + 8 (r / 4) + 4 (lane / 32) + r % 4][n0 + lane % 32].
template <int EPI>
__global__ void __launch_bounds__(kBlk) gemmNT(const __bf16* __restrict__ A, int lda, const __bf16* __restrict__ B,
                                               int ldb, int M, int N, int K, int kchunk, EpiArgs e) {
  const int lane = threadIdx.x & 63;
  const int tiles_n = N / 32;
  const int tile = blockIdx.x * (kBlk / 64) + threadIdx.x / 64;
  if (tile >= (M / 32) * tiles_n) return;
  const int m0 = (tile / tiles_n) * 32, n0 = (tile % tiles_n) * 32;
  const int kb = blockIdx.y * kchunk;
  const int ke = kb + kchunk < K ? kb + kchunk : K;
  const int h = lane >> 5, r32 = lane & 31;
  const __bf16* pa = A + static_cast<size_t>(m0 + r32) * lda + h * 8;
  const __bf16* pb = B + static_cast<size_t>(n0 + r32) * ldb + h * 8;
  f32x16 acc = {};
  for (int k = kb; k < ke; k += 16) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(pa + k);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(pb + k);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  const int n = n0 + r32;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int m = m0 + 8 * (r / 4) + 4 * h + (r % 4);
    const float c = acc[r];
    const size_t o = static_cast<size_t>(m) * e.ldc + n;
    if constexpr (EPI == kEpiStore) {
      if (e.atomic) atomicAdd(&e.c[o], c);
      else e.c[o] = c;
    } else if constexpr (EPI == kEpiBias) {
      e.c[o] = c + e.bias[n];
    } else if constexpr (EPI == kEpiBiasGelu) {
      const float z = c + e.bias[n];
      e.c[o] = z;
      const __bf16 hv = static_cast<__bf16>(geluF(z));
      e.h[o] = hv;
      e.ht[static_cast<size_t>(n) * e.ldt + m] = hv;
    } else if constexpr (EPI == kEpiGelu) {
      e.h[o] = static_cast<__bf16>(geluF(c + e.bias[n]));
    } else {  // kEpiGeluGrad
      const __bf16 g = static_cast<__bf16>(c * geluGradF(e.z[o]));
      e.h[o] = g;
      e.ht[static_cast<size_t>(n) * e.ldt + m] = g;
    }
  }
}

template <int EPI>
static void launchGemm(const __bf16* A, int lda, const __bf16* B, int ldb, int M, int N, int K, int ksplit,
                       const EpiArgs& e, hipStream_t s) {
  const int tiles = (M / 32) * (N / 32);
  const int kchunk = ((K / ksplit + 15) / 16) * 16;
  const int kz = (K + kchunk - 1) / kchunk;
  dim3 grid((tiles + kBlk / 64 - 1) / (kBlk / 64), kz);
  gemmNT<EPI><<<grid, kBlk, 0, s>>>(A, lda, B, ldb, M, N, K, kchunk, e);
}

// ---------------------------------------------------------------- features
// Per pair: relative translation of b in a's frame (3), relative rotation
// quaternion (4, w >= 0), descriptors of a and b (16 + 16); columns 39..63
// are zero.  Written as BF16 rows of 64 (and optionally the transpose).
__global__ void __launch_bounds__(kBlk) featureKernel(const uint32_t* pa, const uint32_t* pb, const float* posea,
                                                      const float* poseb, const float* desc, uint32_t n_shapes,
                                                      uint64_t n, uint64_t n_pad, __bf16* X, __bf16* XT,
                                                      float* xf) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_pad) return;
  float f[kMlpInPad];
  for (int k = 0; k < kMlpInPad; ++k) f[k] = 0.f;
  if (i < n && pa[i] < n_shapes && pb[i] < n_shapes) {
    const float* A = posea + 7 * i;
    const float* B = poseb + 7 * i;
    float qa[4] = {A[3], A[4], A[5], A[6]}, qb[4] = {B[3], B[4], B[5], B[6]};
    const float na = rsqrtf(qa[0] * qa[0] + qa[1] * qa[1] + qa[2] * qa[2] + qa[3] * qa[3]);
    const float nb = rsqrtf(qb[0] * qb[0] + qb[1] * qb[1] + qb[2] * qb[2] + qb[3] * qb[3]);
    for (int k = 0; k < 4; ++k) { qa[k] *= na; qb[k] *= nb; }
    // R_a^T (t_b - t_a): rotate by the conjugate of q_a
    const float dx = B[0] - A[0], dy = B[1] - A[1], dz = B[2] - A[2];
    const float w = qa[0], x = -qa[1], y = -qa[2], z = -qa[3];
    const float tx = 2.f * (y * dz - z * dy), ty = 2.f * (z * dx - x * dz), tz = 2.f * (x * dy - y * dx);
    f[0] = dx + w * tx + (y * tz - z * ty);
    f[1] = dy + w * ty + (z * tx - x * tz);
    f[2] = dz + w * tz + (x * ty - y * tx);
    // q_a^* q_b
    float rw = w * qb[0] - x * qb[1] - y * qb[2] - z * qb[3];
    float rx = w * qb[1] + x * qb[0] + y * qb[3] - z * qb[2];
    float ry = w * qb[2] - x * qb[3] + y * qb[0] + z * qb[1];
    float rz = w * qb[3] + x * qb[2] - y * qb[1] + z * qb[0];
    if (rw < 0) { rw = -rw; rx = -rx; ry = -ry; rz = -rz; }
    f[3] = rw; f[4] = rx; f[5] = ry; f[6] = rz;
    const float* da = desc + 16 * static_cast<size_t>(pa[i]);
    const float* db = desc + 16 * static_cast<size_t>(pb[i]);
    for (int k = 0; k < 16; ++k) { f[7 + k] = da[k]; f[23 + k] = db[k]; }
  }
  for (int k = 0; k < kMlpInPad; ++k) {
    const __bf16 v = static_cast<__bf16>(f[k]);
    if (X) X[i * kMlpInPad + k] = v;
    if (XT) XT[static_cast<size_t>(k) * n_pad + i] = v;
    if (xf) xf[i * kMlpInPad + k] = f[k];
  }
}

// ------------------------------------------------------------------ loss
// L = mean over valid rows of (1 - cos(y[0:3], t[0:3])) + wd (y3 - t3)^2;
// rows with t = 0 direction are masked.  Writes dY (BF16, rows of 32; only
// the first 4 columns non-zero) and its transpose.
__global__ void __launch_bounds__(kBlk) lossKernel(const float* Y, const float* T, uint64_t n, uint64_t n_pad,
                                                   float inv_count, float wd, __bf16* dY, __bf16* dYT,
                                                   float* loss_sum) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  float g[4] = {0, 0, 0, 0};
  float l = 0.f;
  if (i < n) {
    const float* y = Y + i * kMlpOutPad;
    const float* t = T + 4 * i;
    const float tn = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    if (tn > 0.f) {
      const float yn = fmaxf(sqrtf(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]), 1e-12f);
      const float c = (y[0] * t[0] + y[1] * t[1] + y[2] * t[2]) / (yn * tn);
      l = 1.f - c;
      for (int k = 0; k < 3; ++k) g[k] = -(t[k] / tn - c * y[k] / yn) / yn * inv_count;
      const float e = y[3] - t[3];
      l += wd * e * e;
      g[3] = 2.f * wd * e * inv_count;
    }
  }
  if (i < n_pad) {
    for (int k = 0; k < kMlpOutPad; ++k) {
      const __bf16 v = static_cast<__bf16>(k < 4 ? g[k] : 0.f);
      dY[i * kMlpOutPad + k] = v;
      dYT[static_cast<size_t>(k) * n_pad + i] = v;
    }
  }
  // block reduction of the loss, one atomic per block
  __shared__ float red[kBlk / 64];
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int w = 0; w < kBlk / 64; ++w) s += red[w];
    atomicAdd(loss_sum, s * inv_count);
  }
}

// Row sums of a BF16 matrix [rows][n] (bias gradients from dZ^T).
__global__ void __launch_bounds__(kBlk) rowSumKernel(const __bf16* AT, uint64_t n, float* out) {
  const int row = blockIdx.x;
  float s = 0.f;
  for (uint64_t k = threadIdx.x; k < n; k += blockDim.x) s += static_cast<float>(AT[row * n + k]);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  __shared__ float red[kBlk / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < kBlk / 64; ++w) t += red[w];
    out[row] += t;
  }
}

__global__ void fillKernel(float* p, uint64_t n, float v) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Adam on the FP32 master parameters; padding entries have zero gradients
// and stay zero.
__global__ void adamKernel(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of the weights: W (forward operand) and W^T (backward operand).
__global__ void castKernel(const float* W, int rows, int cols, __bf16* Wb, __bf16* WTb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  const int r = i / cols, c = i % cols;
  const __bf16 v = static_cast<__bf16>(W[i]);
  Wb[i] = v;
  if (WTb) WTb[c * rows + r] = v;
}

// Output -> unit guess direction (shape a frame) and distance estimate.
// policy 1: pairs predicted to penetrate keep the solver's default guess
// (UnitX, gjk_solver-inl.h:77-78), whose final GJK simplex seeds EPA.
__global__ void guessKernel(const float* Y, uint64_t n, int policy, float* guess3, float* dist) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* y = Y + i * kMlpOutPad;
  float nn = sqrtf(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
  float g0 = 1.f, g1 = 0.f, g2 = 0.f;
  if (nn > 1e-20f && !(policy == 1 && y[3] < 0.f)) { g0 = y[0] / nn; g1 = y[1] / nn; g2 = y[2] / nn; }
  guess3[3 * i] = g0;
  guess3[3 * i + 1] = g1;
  guess3[3 * i + 2] = g2;
  if (dist) dist[i] = y[3];
}

// ------------------------------------------------------- fused inference
// Features, the four layers and the normalisation in one kernel: weights
// live in LDS (copied once per block from a packed image), and each
// wavefront carries 32 pairs through the network with the activations kept
// in registers.  Formulation: C = W . X^T (rows = units, columns = pairs),
// so a layer's accumulator tile holds, in lane l, pair l % 32 and the 16
// units 8 (r / 4) + 4 (l / 32) + r % 4 of the tile; the next layer takes
// those registers directly as its B operand (8 per K step) by contracting
// over the units in that order -- the packed weights are laid out with the
// same permutation of K, so no transpose through LDS is needed.
//
// Packed image (bf16x8 per lane): [layer][out tile][K step][lane], with
//   element i of lane l = W[32 t + l % 32][k(j, l / 32, i)],
//   layer 1: k = 16 j + 8 h + i (features in natural order),
//   layers 2-4: k = 32 (j / 2) + 16 (j % 2) + 8 (i / 4) + 4 h + i % 4.
constexpr int kPackL1 = 4 * 4, kPackL2 = 4 * 8, kPackL3 = 4 * 8, kPackL4 = 1 * 8;  // (tile, K step) blocks
constexpr int kPackBlocks = kPackL1 + kPackL2 + kPackL3 + kPackL4;
constexpr int kPackBytes = kPackBlocks * 64 * 16;
constexpr int kPackBiasFloats = 3 * kHid + kMlpOutPad;

__device__ __forceinline__ int permK(int j, int h, int i) { return 32 * (j / 2) + 16 * (j % 2) + 8 * (i / 4) + 4 * h + (i % 4); }

__global__ void packKernel(const float* P, __bf16* pack, float* bias) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // one (block, lane, element)
  if (idx < kPackBlocks * 64 * 8) {
    const int e = idx % 8, l = (idx / 8) % 64, b = idx / (8 * 64);
    const int h = l / 32, m = l % 32;
    float v;
    if (b < kPackL1) {
      const int t = b / 4, j = b % 4;
      v = P[kOffW1 + (32 * t + m) * kMlpInPad + 16 * j + 8 * h + e];
    } else if (b < kPackL1 + kPackL2) {
      const int q = b - kPackL1, t = q / 8, j = q % 8;
      v = P[kOffW2 + (32 * t + m) * kHid + permK(j, h, e)];
    } else if (b < kPackL1 + kPackL2 + kPackL3) {
      const int q = b - kPackL1 - kPackL2, t = q / 8, j = q % 8;
      v = P[kOffW3 + (32 * t + m) * kHid + permK(j, h, e)];
    } else {
      const int j = b - kPackL1 - kPackL2 - kPackL3;
      v = P[kOffW4 + m * kHid + permK(j, h, e)];
    }
    pack[idx] = static_cast<__bf16>(v);
  }
  if (idx < kPackBiasFloats) {
    const int off = idx < kHid ? kOffB1 + idx
                  : idx < 2 * kHid ? kOffB2 + idx - kHid
                  : idx < 3 * kHid ? kOffB3 + idx - 2 * kHid : kOffB4 + idx - 3 * kHid;
    bias[idx] = P[off];
  }
}

// one hidden layer: 4 output tiles, KS K steps, B operands in x[KS]
template <int KS>
__device__ __forceinline__ void denseGelu(const MFCL_LDS bf16x8* W, const bf16x8* x, const float* bias, int lane,
                                          bf16x8* out) {
  const int h = lane >> 5;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x16 acc = {};
#pragma unroll
    for (int j = 0; j < KS; ++j) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[(t * KS + j) * 64 + lane], x[j], acc, 0, 0, 0);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      bf16x8 o;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int r = 8 * s + i;
        const int unit = 32 * t + 8 * (r / 4) + 4 * h + (r % 4);
        o[i] = static_cast<__bf16>(geluF(acc[r] + bias[unit]));
      }
      out[2 * t + s] = o;
    }
    // keep the scheduler from overlapping the tiles (register pressure)
    __builtin_amdgcn_sched_barrier(0);
  }
}

// BLK threads per block share the block's LDS copy of the weights (one
// block per CU): more wavefronts per block hide the feature loads and the
// MFMA chains behind each other, within BLK's register budget.
template <int BLK>
__global__ void __launch_bounds__(BLK) mlpGuessFused(const uint32_t* pa, const uint32_t* pb, const float* posea,
                                                      const float* poseb, const float* desc, uint32_t n_shapes,
                                                      uint64_t n, const uint4* pack_g, const float* bias_g, int policy,
                                                      float* guess3, float* dist) {
  extern __shared__ __align__(16) uint4 lds_pack[];
  for (int k = threadIdx.x; k < kPackBytes / 16; k += blockDim.x) lds_pack[k] = pack_g[k];
  float* bias = reinterpret_cast<float*>(lds_pack + kPackBytes / 16);
  for (int k = threadIdx.x; k < kPackBiasFloats; k += blockDim.x) bias[k] = bias_g[k];
  __syncthreads();
  const MFCL_LDS bf16x8* W = (const MFCL_LDS bf16x8*)(lds_pack);
  const MFCL_LDS bf16x8* W1 = W;
  const MFCL_LDS bf16x8* W2 = W1 + kPackL1 * 64;
  const MFCL_LDS bf16x8* W3 = W2 + kPackL2 * 64;
  const MFCL_LDS bf16x8* W4 = W3 + kPackL3 * 64;
  const int lane = threadIdx.x & 63, h = lane >> 5;
  const uint64_t ntiles = (n + 31) / 32;
  const uint64_t wave = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  for (uint64_t tile = wave; tile < ntiles; tile += nwaves) {
    const uint64_t i = tile * 32 + (lane & 31);
    // features (the same formulas as featureKernel); lanes h = 0 / 1 keep
    // columns 16 j + 0..7 / 16 j + 8..15
    float t0 = 0, t1 = 0, t2 = 0, r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    float4 da[4] = {}, db[4] = {};
    if (i < n && pa[i] < n_shapes && pb[i] < n_shapes) {
      const float* A = posea + 7 * i;
      const float* B = poseb + 7 * i;
      float qa[4] = {A[3], A[4], A[5], A[6]}, qb[4] = {B[3], B[4], B[5], B[6]};
      const float na = rsqrtf(qa[0] * qa[0] + qa[1] * qa[1] + qa[2] * qa[2] + qa[3] * qa[3]);
      const float nb = rsqrtf(qb[0] * qb[0] + qb[1] * qb[1] + qb[2] * qb[2] + qb[3] * qb[3]);
#pragma unroll
      for (int k = 0; k < 4; ++k) { qa[k] *= na; qb[k] *= nb; }
      const float dx = B[0] - A[0], dy = B[1] - A[1], dz = B[2] - A[2];
      const float w = qa[0], x = -qa[1], y = -qa[2], z = -qa[3];
      const float tx = 2.f * (y * dz - z * dy), ty = 2.f * (z * dx - x * dz), tz = 2.f * (x * dy - y * dx);
      t0 = dx + w * tx + (y * tz - z * ty);
      t1 = dy + w * ty + (z * tx - x * tz);
      t2 = dz + w * tz + (x * ty - y * tx);
      r0 = w * qb[0] - x * qb[1] - y * qb[2] - z * qb[3];
      r1 = w * qb[1] + x * qb[0] + y * qb[3] - z * qb[2];
      r2 = w * qb[2] - x * qb[3] + y * qb[0] + z * qb[1];
      r3 = w * qb[3] + x * qb[2] - y * qb[1] + z * qb[0];
      if (r0 < 0) { r0 = -r0; r1 = -r1; r2 = -r2; r3 = -r3; }
      const float4* pda = reinterpret_cast<const float4*>(desc + 16 * static_cast<size_t>(pa[i]));
      const float4* pdb = reinterpret_cast<const float4*>(desc + 16 * static_cast<size_t>(pb[i]));
#pragma unroll
      for (int q = 0; q < 4; ++q) { da[q] = pda[q]; db[q] = pdb[q]; }
    }
    // feature k: 0-2 t, 3-6 r, 7-22 da, 23-38 db, 39-63 zero
    auto feat = [&](int k) -> float {
      if (k < 3) return k == 0 ? t0 : (k == 1 ? t1 : t2);
      if (k < 7) return k == 3 ? r0 : (k == 4 ? r1 : (k == 5 ? r2 : r3));
      if (k < 23) { const float4 v = da[(k - 7) / 4]; const int c = (k - 7) % 4; return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w)); }
      if (k < 39) { const float4 v = db[(k - 23) / 4]; const int c = (k - 23) % 4; return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w)); }
      return 0.f;
    };
    bf16x8 x0[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 8; ++e) x0[j][e] = static_cast<__bf16>(h ? feat(16 * j + 8 + e) : feat(16 * j + e));
    bf16x8 a1[8], a2[8];
    denseGelu<4>(W1, x0, bias, lane, a1);
    denseGelu<8>(W2, a1, bias + kHid, lane, a2);
    denseGelu<8>(W3, a2, bias + 2 * kHid, lane, a1);
    f32x16 acc = {};
#pragma unroll
    for (int j = 0; j < 8; ++j) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W4[j * 64 + lane], a1[j], acc, 0, 0, 0);
    // outputs 0..3 of pair lane % 32 are rows 0..3: registers 0..3 of the lanes with h = 0
    if (h == 0 && i < n) {
      const float* b4 = bias + 3 * kHid;
      const float y0 = acc[0] + b4[0], y1 = acc[1] + b4[1], y2 = acc[2] + b4[2], y3 = acc[3] + b4[3];
      const float nn = sqrtf(y0 * y0 + y1 * y1 + y2 * y2);
      float g0 = 1.f, g1 = 0.f, g2 = 0.f;
      if (nn > 1e-20f && !(policy == 1 && y3 < 0.f)) { g0 = y0 / nn; g1 = y1 / nn; g2 = y2 / nn; }
      guess3[3 * i] = g0;
      guess3[3 * i + 1] = g1;
      guess3[3 * i + 2] = g2;
      if (dist) dist[i] = y3;
    }
  }
}

// ------------------------------------------------------------- host side
namespace {

uint64_t pad32(uint64_t n) { return (n + 31) / 32 * 32; }

int ensureMlp(mfcl_ctx* ctx) {
  if (ctx->mlp.params) return MFCL_OK;
  ctx->err = "MLP parameters not initialised (mfcl_mlp_init or mfcl_mlp_set_params)";
  return MFCL_E_STATE;
}

// Refreshes the BF16 weight copies from the FP32 master parameters.
void castWeights(mfcl_ctx* ctx, hipStream_t s) {
  auto& M = ctx->mlp;
  const int T = 256;
  packKernel<<<(kPackBlocks * 64 * 8 + T - 1) / T, T, 0, s>>>(M.params, reinterpret_cast<__bf16*>(M.pack),
                                                               reinterpret_cast<float*>(static_cast<char*>(M.pack) +
                                                                                        kPackBytes));
  castKernel<<<(kHid * kMlpInPad + T - 1) / T, T, 0, s>>>(M.params + kOffW1, kHid, kMlpInPad, M.w1, M.w1t);
  castKernel<<<(kHid * kHid + T - 1) / T, T, 0, s>>>(M.params + kOffW2, kHid, kHid, M.w2, M.w2t);
  castKernel<<<(kHid * kHid + T - 1) / T, T, 0, s>>>(M.params + kOffW3, kHid, kHid, M.w3, M.w3t);
  castKernel<<<(kMlpOutPad * kHid + T - 1) / T, T, 0, s>>>(M.params + kOffW4, kMlpOutPad, kHid, M.w4, M.w4t);
}

struct Acts {
  __bf16 *x, *xt, *h1, *h1t, *h2, *h2t, *h3, *h3t, *dy, *dyt, *dz3, *dz3t, *dz2, *dz2t, *dz1, *dz1t;
  float *z1, *z2, *z3, *y, *xf;
};

// Carves the activation workspace for n_pad rows (training needs all of it,
// inference only x, h1..h3 and y).
size_t actsBytes(uint64_t np, bool train) {
  const size_t bf_in = np * kMlpInPad * 2, bf_h = np * kHid * 2, bf_o = np * kMlpOutPad * 2;
  const size_t f_h = np * kHid * 4, f_o = np * kMlpOutPad * 4, f_in = np * kMlpInPad * 4;
  size_t total = bf_in + 3 * bf_h + f_o + 256 * 8;
  if (train) total += bf_in + 3 * bf_h + 2 * bf_o + 6 * bf_h + 3 * f_h + f_in + 256 * 16;
  return total;
}

int carveActs(mfcl_ctx* ctx, uint64_t np, bool train, Acts& a) {
  const size_t bf_in = np * kMlpInPad * 2, bf_h = np * kHid * 2, bf_o = np * kMlpOutPad * 2;
  const size_t f_h = np * kHid * 4, f_o = np * kMlpOutPad * 4, f_in = np * kMlpInPad * 4;
  if (actsBytes(np, train) > ctx->mlp.ws.bytes) {
    ctx->err = "MLP workspace too small for " + std::to_string(np) +
               " rows: call mfcl_reserve(ctx, MFCL_WS_MLP, rows) first";
    return MFCL_E_STATE;
  }
  char* p = static_cast<char*>(ctx->mlp.ws.p);
  auto take = [&](size_t b) {
    char* q = p;
    p += (b + 255) & ~static_cast<size_t>(255);
    return q;
  };
  std::memset(&a, 0, sizeof(a));
  a.x = reinterpret_cast<__bf16*>(take(bf_in));
  a.h1 = reinterpret_cast<__bf16*>(take(bf_h));
  a.h2 = reinterpret_cast<__bf16*>(take(bf_h));
  a.h3 = reinterpret_cast<__bf16*>(take(bf_h));
  a.y = reinterpret_cast<float*>(take(f_o));
  if (train) {
    a.xt = reinterpret_cast<__bf16*>(take(bf_in));
    a.h1t = reinterpret_cast<__bf16*>(take(bf_h));
    a.h2t = reinterpret_cast<__bf16*>(take(bf_h));
    a.h3t = reinterpret_cast<__bf16*>(take(bf_h));
    a.dy = reinterpret_cast<__bf16*>(take(bf_o));
    a.dyt = reinterpret_cast<__bf16*>(take(bf_o));
    a.dz3 = reinterpret_cast<__bf16*>(take(bf_h));
    a.dz3t = reinterpret_cast<__bf16*>(take(bf_h));
    a.dz2 = reinterpret_cast<__bf16*>(take(bf_h));
    a.dz2t = reinterpret_cast<__bf16*>(take(bf_h));
    a.dz1 = reinterpret_cast<__bf16*>(take(bf_h));
    a.dz1t = reinterpret_cast<__bf16*>(take(bf_h));
    a.z1 = reinterpret_cast<float*>(take(f_h));
    a.z2 = reinterpret_cast<float*>(take(f_h));
    a.z3 = reinterpret_cast<float*>(take(f_h));
    a.xf = reinterpret_cast<float*>(take(f_in));
  }
  return MFCL_OK;
}

void forward(mfcl_ctx* ctx, const Acts& a, uint64_t np, bool train, hipStream_t s) {
  auto& M = ctx->mlp;
  const int N = static_cast<int>(np);
  EpiArgs e{};
  if (train) {
    e = EpiArgs{a.z1, kHid, M.params + kOffB1, nullptr, a.h1, a.h1t, N, 0};
    launchGemm<kEpiBiasGelu>(a.x, kMlpInPad, M.w1, kMlpInPad, N, kHid, kMlpInPad, 1, e, s);
    e = EpiArgs{a.z2, kHid, M.params + kOffB2, nullptr, a.h2, a.h2t, N, 0};
    launchGemm<kEpiBiasGelu>(a.h1, kHid, M.w2, kHid, N, kHid, kHid, 1, e, s);
    e = EpiArgs{a.z3, kHid, M.params + kOffB3, nullptr, a.h3, a.h3t, N, 0};
    launchGemm<kEpiBiasGelu>(a.h2, kHid, M.w3, kHid, N, kHid, kHid, 1, e, s);
  } else {
    e = EpiArgs{nullptr, kHid, M.params + kOffB1, nullptr, a.h1, nullptr, 0, 0};
    launchGemm<kEpiGelu>(a.x, kMlpInPad, M.w1, kMlpInPad, N, kHid, kMlpInPad, 1, e, s);
    e.bias = M.params + kOffB2; e.h = a.h2;
    launchGemm<kEpiGelu>(a.h1, kHid, M.w2, kHid, N, kHid, kHid, 1, e, s);
    e.bias = M.params + kOffB3; e.h = a.h3;
    launchGemm<kEpiGelu>(a.h2, kHid, M.w3, kHid, N, kHid, kHid, 1, e, s);
  }
  e = EpiArgs{a.y, kMlpOutPad, M.params + kOffB4, nullptr, nullptr, nullptr, 0, 0};
  launchGemm<kEpiBias>(a.h3, kHid, M.w4, kHid, N, kMlpOutPad, kHid, 1, e, s);
}

// Gradients of the mean loss into M.grad (zeroed here).
void backward(mfcl_ctx* ctx, const Acts& a, uint64_t np, hipStream_t s) {
  auto& M = ctx->mlp;
  const int N = static_cast<int>(np);
  fillKernel<<<(kMlpParams + 255) / 256, 256, 0, s>>>(M.grad, kMlpParams, 0.f);
  const int ks = N >= 4096 ? N / 2048 : 1;  // split the batch dimension of the weight gradients
  EpiArgs e{};
  // layer 4
  e = EpiArgs{M.grad + kOffW4, kHid, nullptr, nullptr, nullptr, nullptr, 0, 1};
  launchGemm<kEpiStore>(a.dyt, N, a.h3t, N, kMlpOutPad, kHid, N, ks, e, s);
  rowSumKernel<<<kMlpOutPad, kBlk, 0, s>>>(a.dyt, np, M.grad + kOffB4);
  e = EpiArgs{nullptr, kHid, nullptr, a.z3, a.dz3, a.dz3t, N, 0};
  launchGemm<kEpiGeluGrad>(a.dy, kMlpOutPad, M.w4t, kMlpOutPad, N, kHid, kMlpOutPad, 1, e, s);
  // layer 3
  e = EpiArgs{M.grad + kOffW3, kHid, nullptr, nullptr, nullptr, nullptr, 0, 1};
  launchGemm<kEpiStore>(a.dz3t, N, a.h2t, N, kHid, kHid, N, ks, e, s);
  rowSumKernel<<<kHid, kBlk, 0, s>>>(a.dz3t, np, M.grad + kOffB3);
  e = EpiArgs{nullptr, kHid, nullptr, a.z2, a.dz2, a.dz2t, N, 0};
  launchGemm<kEpiGeluGrad>(a.dz3, kHid, M.w3t, kHid, N, kHid, kHid, 1, e, s);
  // layer 2
  e = EpiArgs{M.grad + kOffW2, kHid, nullptr, nullptr, nullptr, nullptr, 0, 1};
  launchGemm<kEpiStore>(a.dz2t, N, a.h1t, N, kHid, kHid, N, ks, e, s);
  rowSumKernel<<<kHid, kBlk, 0, s>>>(a.dz2t, np, M.grad + kOffB2);
  e = EpiArgs{nullptr, kHid, nullptr, a.z1, a.dz1, a.dz1t, N, 0};
  launchGemm<kEpiGeluGrad>(a.dz2, kHid, M.w2t, kHid, N, kHid, kHid, 1, e, s);
  // layer 1
  e = EpiArgs{M.grad + kOffW1, kMlpInPad, nullptr, nullptr, nullptr, nullptr, 0, 1};
  launchGemm<kEpiStore>(a.dz1t, N, a.xt, N, kHid, kMlpInPad, N, ks, e, s);
  rowSumKernel<<<kHid, kBlk, 0, s>>>(a.dz1t, np, M.grad + kOffB1);
}

// Copies FP32 features [n][64] (host or device) into the BF16 workspace.
__global__ void featCastKernel(const float* xf, uint64_t n, uint64_t np, __bf16* X, __bf16* XT) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= np) return;
  for (int k = 0; k < kMlpInPad; ++k) {
    const __bf16 v = static_cast<__bf16>(i < n ? xf[i * kMlpInPad + k] : 0.f);
    X[i * kMlpInPad + k] = v;
    if (XT) XT[static_cast<size_t>(k) * np + i] = v;
  }
}

}  // namespace

int mlpReserve(mfcl_ctx* ctx, uint64_t rows) {
  if (rows > (1ull << 26)) {
    ctx->err = "MLP rows out of range (<= 2^26)";
    return MFCL_E_INVALID;
  }
  return ensure(ctx, ctx->mlp.ws, actsBytes(pad32(rows ? rows : 1), true));
}

int mlpInit(mfcl_ctx* ctx, uint64_t seed) {
  std::vector<float> p(kMlpParams, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int rows, int cols, int fan_in, int fan_out, int valid_rows, int valid_cols) {
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    std::uniform_real_distribution<double> U(-a, a);
    for (int r = 0; r < rows; ++r)
      for (int c = 0; c < cols; ++c)
        p[off + r * cols + c] = (r < valid_rows && c < valid_cols) ? static_cast<float>(U(rng)) : 0.f;
  };
  fill(kOffW1, kHid, kMlpInPad, kMlpIn, kHid, kHid, kMlpIn);
  fill(kOffW2, kHid, kHid, kHid, kHid, kHid, kHid);
  fill(kOffW3, kHid, kHid, kHid, kHid, kHid, kHid);
  fill(kOffW4, kMlpOutPad, kHid, kHid, kMlpOut, kMlpOut, kHid);
  return mlpSetParams(ctx, p.data(), kMlpParams, true);
}

int mlpSetParams(mfcl_ctx* ctx, const float* host, int n, bool reset_opt) {
  if (n != kMlpParams) {
    ctx->err = "expected " + std::to_string(kMlpParams) + " parameters";
    return MFCL_E_INVALID;
  }
  auto& M = ctx->mlp;
  if (!M.params) {
    const size_t fb = static_cast<size_t>(kMlpParams) * 4;
    MFCL_HIP_TRY(ctx, hipMalloc(&M.params, fb));
    MFCL_HIP_TRY(ctx, hipMalloc(&M.grad, fb));
    MFCL_HIP_TRY(ctx, hipMalloc(&M.m, fb));
    MFCL_HIP_TRY(ctx, hipMalloc(&M.v, fb));
    MFCL_HIP_TRY(ctx, hipMalloc(&M.loss, 256));
    MFCL_HIP_TRY(ctx, hipMalloc(&M.pack, kPackBytes + kPackBiasFloats * 4));
    const size_t wb = (static_cast<size_t>(kHid) * kMlpInPad + 2 * kHid * kHid + kMlpOutPad * kHid) * 2;
    char* w = nullptr;
    MFCL_HIP_TRY(ctx, hipMalloc(&w, 2 * wb));
    M.wbuf = w;
    __bf16* q = reinterpret_cast<__bf16*>(w);
    M.w1 = q; q += kHid * kMlpInPad;
    M.w2 = q; q += kHid * kHid;
    M.w3 = q; q += kHid * kHid;
    M.w4 = q; q += kMlpOutPad * kHid;
    M.w1t = q; q += kHid * kMlpInPad;
    M.w2t = q; q += kHid * kHid;
    M.w3t = q; q += kHid * kHid;
    M.w4t = q;
  }
  MFCL_HIP_TRY(ctx, hipMemcpyAsync(M.params, host, static_cast<size_t>(kMlpParams) * 4, hipMemcpyHostToDevice,
                                   ctx->stream));
  if (reset_opt) {
    fillKernel<<<(kMlpParams + 255) / 256, 256, 0, ctx->stream>>>(M.m, kMlpParams, 0.f);
    fillKernel<<<(kMlpParams + 255) / 256, 256, 0, ctx->stream>>>(M.v, kMlpParams, 0.f);
    M.step = 0;
  }
  castWeights(ctx, ctx->stream);
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return MFCL_OK;
}

int mlpGetParams(mfcl_ctx* ctx, float* host, int n, float* grad_host) {
  if (int rc = ensureMlp(ctx)) return rc;
  if (n != kMlpParams) {
    ctx->err = "expected " + std::to_string(kMlpParams) + " parameters";
    return MFCL_E_INVALID;
  }
  MFCL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (host) MFCL_HIP_TRY(ctx, hipMemcpy(host, ctx->mlp.params, static_cast<size_t>(n) * 4, hipMemcpyDeviceToHost));
  if (grad_host)
    MFCL_HIP_TRY(ctx, hipMemcpy(grad_host, ctx->mlp.grad, static_cast<size_t>(n) * 4, hipMemcpyDeviceToHost));
  return MFCL_OK;
}

int mlpFeatures(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea, const float* poseb,
                uint64_t n, float* xf_out, hipStream_t s) {
  if (!ctx->d_desc) {
    ctx->err = "no shapes uploaded";
    return MFCL_E_STATE;
  }
  if (n == 0) return MFCL_OK;
  featureKernel<<<static_cast<unsigned>((n + kBlk - 1) / kBlk), kBlk, 0, s>>>(
      pa, pb, posea, poseb, ctx->d_desc, ctx->n_shapes, n, n, nullptr, nullptr, xf_out);
  MFCL_HIP_TRY(ctx, hipGetLastError());
  return MFCL_OK;
}

int mlpGuess(mfcl_ctx* ctx, const uint32_t* pa, const uint32_t* pb, const float* posea, const float* poseb,
             uint64_t n, float* guess3, float* dist, hipStream_t s) {
  if (int rc = ensureMlp(ctx)) return rc;
  if (!ctx->d_desc) {
    ctx->err = "no shapes uploaded";
    return MFCL_E_STATE;
  }
  if (n == 0) return MFCL_OK;
  if (ctx->timing) MFCL_HIP_TRY(ctx, hipEventRecord(ctx->ev[8], s));
  const size_t lds = kPackBytes + kPackBiasFloats * 4;
  const uint64_t ntiles = (n + 31) / 32;
  const int waves = ctx->tune.mlp_waves;
  // persistent: at most one block per CU (the weights fill most of its LDS)
  const uint64_t want = (ntiles + waves - 1) / waves;
  const uint64_t cap = static_cast<uint64_t>(ctx->n_cu);
  const unsigned grid = static_cast<unsigned>(want < cap ? want : cap);
  const uint4* pk = static_cast<const uint4*>(ctx->mlp.pack);
  const float* pb4 = reinterpret_cast<const float*>(static_cast<char*>(ctx->mlp.pack) + kPackBytes);
  const int pol = ctx->tune.guess_policy;
  if (waves == 16)
    mlpGuessFused<1024><<<grid, 1024, lds, s>>>(pa, pb, posea, poseb, ctx->d_desc, ctx->n_shapes, n, pk, pb4, pol,
                                                guess3, dist);
  else if (waves == 8)
    mlpGuessFused<512><<<grid, 512, lds, s>>>(pa, pb, posea, poseb, ctx->d_desc, ctx->n_shapes, n, pk, pb4, pol,
                                              guess3, dist);
  else
    mlpGuessFused<256><<<grid, 256, lds, s>>>(pa, pb, posea, poseb, ctx->d_desc, ctx->n_shapes, n, pk, pb4, pol,
                                              guess3, dist);
  if (ctx->timing) MFCL_HIP_TRY(ctx, hipEventRecord(ctx->ev[9], s));
  MFCL_HIP_TRY(ctx, hipGetLastError());
  return MFCL_OK;
}

int mlpForwardFeat(mfcl_ctx* ctx, const float* xf, uint64_t n, float* y4_out, hipStream_t s) {
  if (int rc = ensureMlp(ctx)) return rc;
  const uint64_t np = pad32(n);
  Acts a;
  if (int rc = carveActs(ctx, np, false, a)) return rc;
  featCastKernel<<<static_cast<unsigned>((np + kBlk - 1) / kBlk), kBlk, 0, s>>>(xf, n, np, a.x, nullptr);
  forward(ctx, a, np, false, s);
  MFCL_HIP_TRY(ctx, hipMemcpy2DAsync(y4_out, 4 * sizeof(float), a.y, kMlpOutPad * sizeof(float), 4 * sizeof(float),
                                     n, hipMemcpyDeviceToDevice, s));
  MFCL_HIP_TRY(ctx, hipGetLastError());
  return MFCL_OK;
}

int mlpTrainStep(mfcl_ctx* ctx, const float* xf, const float* target4, uint64_t n, const float* hyper,
                 int apply_update, float* loss_out, hipStream_t s) {
  if (int rc = ensureMlp(ctx)) return rc;
  if (n == 0 || n > (1ull << 26)) {
    ctx->err = "training batch size out of range (1 .. 2^26)";
    return MFCL_E_INVALID;
  }
  auto& M = ctx->mlp;
  const uint64_t np = pad32(n);
  Acts a;
  if (int rc = carveActs(ctx, np, true, a)) return rc;
  featCastKernel<<<static_cast<unsigned>((np + kBlk - 1) / kBlk), kBlk, 0, s>>>(xf, n, np, a.x, a.xt);
  forward(ctx, a, np, true, s);
  fillKernel<<<1, 64, 0, s>>>(M.loss, 1, 0.f);
  const float wd = hyper ? hyper[5] : 1.0f;
  lossKernel<<<static_cast<unsigned>((np + kBlk - 1) / kBlk), kBlk, 0, s>>>(
      a.y, target4, n, np, 1.0f / static_cast<float>(n), wd, a.dy, a.dyt, M.loss);
  backward(ctx, a, np, s);
  if (apply_update) {
    const float lr = hyper ? hyper[0] : 1e-3f, b1 = hyper ? hyper[1] : 0.9f, b2 = hyper ? hyper[2] : 0.999f;
    const float eps = hyper ? hyper[3] : 1e-8f;
    M.step += 1;
    const float c1 = 1.f - std::pow(b1, static_cast<float>(M.step)), c2 = 1.f - std::pow(b2, static_cast<float>(M.step));
    adamKernel<<<(kMlpParams + 255) / 256, 256, 0, s>>>(M.params, M.grad, M.m, M.v, kMlpParams, lr, b1, b2, eps, c1,
                                                         c2);
    castWeights(ctx, s);
  }
  if (loss_out) MFCL_HIP_TRY(ctx, hipMemcpyAsync(loss_out, M.loss, 4, hipMemcpyDeviceToHost, s));
  MFCL_HIP_TRY(ctx, hipGetLastError());
  return MFCL_OK;
}

void mlpFree(mfcl_ctx* ctx) {
  auto& M = ctx->mlp;
  for (void* p : {static_cast<void*>(M.params), static_cast<void*>(M.grad), static_cast<void*>(M.m),
                  static_cast<void*>(M.v), static_cast<void*>(M.loss), static_cast<void*>(M.wbuf), M.ws.p, M.pack})
    if (p) (void)hipFree(p);
  M = MlpState{};
}

}  // namespace mfcl
