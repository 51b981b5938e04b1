// Neural radiance cache on CDNA4 (gfx950): a 81 -> 64 -> 64 -> 64 -> 64 -> 3 ReLU MLP whose
// inference, backpropagation and Adam update are hand-written here, with BF16 operands on the
// v_mfma_f32_16x16x32_bf16 matrix cores and FP32 accumulation. The reference renderer has no
// radiance cache; this follows Müller et al., "Real-time Neural Radiance Caching for Path
// Tracing" (2021): relative L2 loss, online training, paths terminated into the cache.
//
// Data layout (padded so every matrix product is a whole number of 16 x 16 x 32 MFMA tiles):
//   inputs   N x 96 BF16 (81 features, zero padded), row-major
//   forward  W0 [64][96], W1..3 [64][64], W4 [16][64] (rows 3.. zero), [out][in] row-major, the
//            input columns of W1..4 permuted by nrc_chain_column (training and inference)
//   backward Wt1..3 [64][64] ([in][out], the out columns permuted likewise), Wt4 [64][8] ([in][out],
//            outputs 3.. zero)
//   biases   b0..b3 [64], b4 [16] (FP32)
// MFMA operand layout (16x16x32): lane l feeds A[l % 16][8 (l / 16) + j] and
// B[8 (l / 16) + j][l % 16] for j = 0..7, and receives D[4 (l / 16) + i][l % 16] for i = 0..3.
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nrc.hpp"

namespace pbrt {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bit pattern, round to nearest even (one v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ uint16_t bf16_bits(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }

// sin(pi f) and cos(pi f) for f = 2^k p (exact, |2 f| < 2^24): f = q / 2 + r with q = rint(2 f) and
// |r| <= 1/4 (both exact), the Taylor series of sin(pi r) to r^9 and of cos(pi r) to r^10
// (truncation below 3e-9 at |r| = 1/4), then the quadrant q mod 4 as a swap and two sign flips.
__device__ __forceinline__ void sincospi_red(float f, float* s_out, float* c_out) {
  const float q = rintf(f * 2.0f);
  const float r = f - q * 0.5f;
  const float r2 = r * r;
  // coefficients (-1)^k pi^(2k+1) / (2k+1)! and (-1)^k pi^(2k) / (2k)!
  float sp = __builtin_fmaf(r2, 8.21458866111282287e-02f, -5.99264529320792077e-01f);
  sp = __builtin_fmaf(r2, sp, 2.55016403987734544e+00f);
  sp = __builtin_fmaf(r2, sp, -5.16771278004997003e+00f);
  sp = __builtin_fmaf(r2, sp, 3.14159265358979324e+00f);
  const float sv = r * sp;
  float cp = __builtin_fmaf(r2, -2.58068913900140508e-02f, 2.35330630358893123e-01f);
  cp = __builtin_fmaf(r2, cp, -1.33526276885458950e+00f);
  cp = __builtin_fmaf(r2, cp, 4.05871212641676822e+00f);
  cp = __builtin_fmaf(r2, cp, -4.93480220054467931e+00f);
  const float cv = __builtin_fmaf(r2, cp, 1.0f);
  // quadrant 0: (s, c); 1: (c, -s); 2: (-s, -c); 3: (-c, s)
  const uint32_t qi = (uint32_t)(int)q;
  const bool swap = (qi & 1u) != 0u;
  const float a = swap ? cv : sv, b = swap ? sv : cv;
  *s_out = __uint_as_float(__float_as_uint(a) ^ ((qi & 2u) << 30));
  *c_out = __uint_as_float(__float_as_uint(b) ^ (((qi + 1u) & 2u) << 30));
}
__device__ __forceinline__ float bf16_float(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

union Pack8 {
  uint16_t u[8];
  uint4 q;
  bf16x8 v;
};

// 8 consecutive BF16 values (16-byte aligned)
__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
  Pack8 x;
  x.q = *reinterpret_cast<const uint4*>(p);
  return x.v;
}
// 8 values `stride` elements apart
__device__ __forceinline__ bf16x8 gather8(const uint16_t* p, int stride) {
  Pack8 x;
#pragma unroll
  for (int j = 0; j < 8; ++j) x.u[j] = p[j * stride];
  return x.v;
}
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// LDS written by some lanes of a wave and then read by others of the same wave: order the accesses
// (no code motion across, stores visible before the loads).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
constexpr int IN = kNrcIn, INP = kNrcInPad, H = kNrcHidden, OUTP = 16, OUTK = 32;

// Inference with the activations in registers. Each product takes the weights as its A operand and
// 16 batch rows as its B operand, so lane (r, g) receives features 16 nb + 4 g + i of row r; with
// the hidden layers' weight columns in the chained order (nrc_chain_column) those four-feature
// groups, rounded to BF16 and rectified, are the lane's B operand of the next layer as they stand.
// The accumulators start at the biases. A workgroup stages the forward weights in LDS once (rows
// padded by 16 bytes: the 16 rows a product reads fall in distinct banks); a wave runs RT 16-row
// tiles through the five layers, every weight operand read from LDS once per RT products.
constexpr int kS0 = INP + 8, kSH = H + 8;  // LDS row strides (BF16)
struct ChainLds {
  uint16_t w0[64 * kS0];
  uint16_t w1[64 * kSH], w2[64 * kSH], w3[64 * kSH];
  uint16_t w4[16 * kSH];
  float b[4][H], b4[16];  // biases of layers 0..3 and of the output layer
};

__device__ __forceinline__ void chain_stage(const NrcForward& P, ChainLds& L) {
  for (int c = threadIdx.x; c < 64 * INP / 8; c += blockDim.x) {
    const int row = c / (INP / 8), k8 = c % (INP / 8);
    *reinterpret_cast<uint4*>(L.w0 + row * kS0 + k8 * 8) = *reinterpret_cast<const uint4*>(P.w0 + row * INP + k8 * 8);
  }
  for (int c = threadIdx.x; c < 64 * H / 8; c += blockDim.x) {
    const int row = c / (H / 8), k8 = c % (H / 8);
    *reinterpret_cast<uint4*>(L.w1 + row * kSH + k8 * 8) = *reinterpret_cast<const uint4*>(P.w1 + row * H + k8 * 8);
    *reinterpret_cast<uint4*>(L.w2 + row * kSH + k8 * 8) = *reinterpret_cast<const uint4*>(P.w2 + row * H + k8 * 8);
    *reinterpret_cast<uint4*>(L.w3 + row * kSH + k8 * 8) = *reinterpret_cast<const uint4*>(P.w3 + row * H + k8 * 8);
    if (row < 16)
      *reinterpret_cast<uint4*>(L.w4 + row * kSH + k8 * 8) = *reinterpret_cast<const uint4*>(P.w4 + row * H + k8 * 8);
  }
  if (threadIdx.x < 4 * H / 4) {
    const int l = threadIdx.x / (H / 4), c = threadIdx.x % (H / 4);
    const float* b = l == 0 ? P.b0 : l == 1 ? P.b1 : l == 2 ? P.b2 : P.b3;
    *reinterpret_cast<float4*>(&L.b[l][4 * c]) = *reinterpret_cast<const float4*>(b + 4 * c);
  } else if (threadIdx.x < 4 * H / 4 + 4) {
    const int c = threadIdx.x - 4 * H / 4;
    *reinterpret_cast<float4*>(&L.b4[4 * c]) = *reinterpret_cast<const float4*>(P.b4 + 4 * c);
  }
  __syncthreads();
}

typedef short i16x2 __attribute__((ext_vector_type(2)));

// max(x, 0) of two values as their BF16 patterns: rounding keeps the sign, and a BF16 pattern orders
// as a 16-bit signed integer the way the value does against zero (one packed integer max)
__device__ __forceinline__ uint32_t relu_bf16x2(float lo, float hi) {
  const uint32_t w = (uint32_t)bf16_bits(lo) | ((uint32_t)bf16_bits(hi) << 16);
  const i16x2 v = __builtin_elementwise_max(__builtin_bit_cast(i16x2, w), i16x2{0, 0});
  return __builtin_bit_cast(uint32_t, v);
}

template <int RT>
__device__ __forceinline__ void chain_act(const f32x4 (&acc)[RT][4], bf16x8 (&h)[RT][2]) {
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      uint4 w;
      w.x = relu_bf16x2(acc[t][2 * kk][0], acc[t][2 * kk][1]);
      w.y = relu_bf16x2(acc[t][2 * kk][2], acc[t][2 * kk][3]);
      w.z = relu_bf16x2(acc[t][2 * kk + 1][0], acc[t][2 * kk + 1][1]);
      w.w = relu_bf16x2(acc[t][2 * kk + 1][2], acc[t][2 * kk + 1][3]);
      Pack8 p;
      p.q = w;
      h[t][kk] = p.v;
    }
}

// accumulators of the four 16-feature blocks, started at the biases of the lane's features
template <int RT>
__device__ __forceinline__ void chain_bias(const float* __restrict__ b, int g, f32x4 (&acc)[RT][4]) {
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    const float4 v = *reinterpret_cast<const float4*>(b + nb * 16 + 4 * g);
#pragma unroll
    for (int t = 0; t < RT; ++t) acc[t][nb] = f32x4{v.x, v.y, v.z, v.w};
  }
}

template <int RT>
__device__ __forceinline__ void chain_hidden(const bf16x8 (&h)[RT][2], const uint16_t* W, const float* __restrict__ b,
                                             int r, int g, f32x4 (&acc)[RT][4]) {
  chain_bias<RT>(b, g, acc);
#pragma unroll
  for (int nb = 0; nb < 4; ++nb)
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const bf16x8 a = load8(W + (nb * 16 + r) * kSH + kk * 32 + g * 8);
#pragma unroll
      for (int t = 0; t < RT; ++t) acc[t][nb] = mfma(a, h[t][kk], acc[t][nb]);
    }
}

// The network on RT tiles whose layer-0 B operands (features 32 kk + 8 g .. + 7 of row r) are x0;
// y[t][i] is output i of row r of tile t in the lanes with g = 0 (i < 3).
template <int RT>
__device__ __forceinline__ void chain_forward(const NrcForward& P, const ChainLds& L, const bf16x8 (&x0)[RT][INP / 32],
                                              int r, int g, f32x4 (&y)[RT]) {
  // an opaque zero offset keeps the LDS weight reads inside the caller's loop (hoisted, they would
  // hold 150 registers for the whole kernel)
  int z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  f32x4 acc[RT][4];
  bf16x8 h[RT][2];
  chain_bias<RT>(L.b[0], g, acc);
#pragma unroll
  for (int nb = 0; nb < 4; ++nb)
#pragma unroll
    for (int kk = 0; kk < INP / 32; ++kk) {
      const bf16x8 a = load8(L.w0 + z + (nb * 16 + r) * kS0 + kk * 32 + g * 8);
#pragma unroll
      for (int t = 0; t < RT; ++t) acc[t][nb] = mfma(a, x0[t][kk], acc[t][nb]);
    }
  chain_act<RT>(acc, h);
  chain_hidden<RT>(h, L.w1 + z, L.b[1], r, g, acc);
  chain_act<RT>(acc, h);
  chain_hidden<RT>(h, L.w2 + z, L.b[2], r, g, acc);
  chain_act<RT>(acc, h);
  chain_hidden<RT>(h, L.w3 + z, L.b[3], r, g, acc);
  chain_act<RT>(acc, h);
  const float4 b4 = *reinterpret_cast<const float4*>(L.b4 + 4 * g);
#pragma unroll
  for (int t = 0; t < RT; ++t) y[t] = f32x4{b4.x, b4.y, b4.z, b4.w};
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
    const bf16x8 a = load8(L.w4 + z + r * kSH + kk * 32 + g * 8);
#pragma unroll
    for (int t = 0; t < RT; ++t) y[t] = mfma(a, h[t][kk], y[t]);
  }
}

constexpr int kChainTiles = 4;

// Outputs of the RT tiles from the lanes with g = 0: 16 lanes x 12 contiguous bytes per tile.
template <int RT>
__device__ __forceinline__ void chain_store(const f32x4 (&y)[RT], float* __restrict__ out, uint32_t row0, uint32_t n,
                                            int r, int g) {
  if (g != 0) return;
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    const uint32_t row = row0 + t * 16 + r;
    if (row < n) {
      out[(size_t)row * 3 + 0] = y[t][0];
      out[(size_t)row * 3 + 1] = y[t][1];
      out[(size_t)row * 3 + 2] = y[t][2];
    }
  }
}

// The radiance of the RT tiles' rows, clamped at 0, times the querying path's throughput, added to
// the path's radiance: beta[row] holds the throughput (xyz) and the path index (w, as bits).
template <int RT>
__device__ __forceinline__ void chain_apply(const f32x4 (&y)[RT], const float4* __restrict__ beta, float4* L,
                                            uint32_t row0, uint32_t n, int r, int g) {
  if (g != 0) return;
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    const uint32_t row = row0 + t * 16 + r;
    if (row < n) {
      const float4 bt = beta[row];
      const uint32_t j = __float_as_uint(bt.w);
      float4 l = L[j];
      l.x = l.x + bt.x * fmaxf(y[t][0], 0.f);
      l.y = l.y + bt.y * fmaxf(y[t][1], 0.f);
      l.z = l.z + bt.z * fmaxf(y[t][2], 0.f);
      L[j] = l;
    }
  }
}

// Inference: rows of `x` (N x 96 BF16) -> out (N x 3 FP32), kChainTiles 16-row tiles per wave,
// grid-stride.
__global__ __launch_bounds__(256) void k_nrc_infer(NrcForward P, const uint16_t* __restrict__ x, float* __restrict__ out,
                                                   uint32_t n, const uint32_t* __restrict__ n_dev) {
  constexpr int RT = kChainTiles;
  __shared__ __attribute__((aligned(16))) ChainLds L;
  chain_stage(P, L);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  if (n_dev) n = min(n, *n_dev);  // a batch counted on the device, bounded by the buffers' capacity
  const uint32_t groups = (n + RT * 16 - 1) / (RT * 16);
  for (uint32_t grp = blockIdx.x * 4 + wave; grp < groups; grp += gridDim.x * 4) {
    const uint32_t row0 = grp * RT * 16;
    bf16x8 x0[RT][INP / 32];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
      const uint32_t row = row0 + t * 16 + r;
#pragma unroll
      for (int kk = 0; kk < INP / 32; ++kk) {
        Pack8 z;
        z.q = make_uint4(0, 0, 0, 0);
        x0[t][kk] = row < n ? load8(x + (size_t)row * INP + kk * 32 + g * 8) : z.v;
      }
    }
    f32x4 y[RT];
    chain_forward<RT>(P, L, x0, r, g, y);
    chain_store<RT>(y, out, row0, n, r, g);
  }
}

// sin and cos of 2^k pi p for k = k0 .. k0 + 3, as features (sin, cos) pairs
__device__ __forceinline__ void sincos4(float p, int k0, Pack8& x) {
  float f = ldexpf(p, k0);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float sv, cv;
    sincospi_red(f, &sv, &cv);
    x.u[2 * i] = bf16_bits(sv);
    x.u[2 * i + 1] = bf16_bits(cv);
    f = f * 2.0f;
  }
}

// Input features KK * 32 + 8 g .. + 7 of the sample with raw inputs q (position, outgoing
// direction, normal, albedo): the same values k_nrc_encode stores. Branch-free across the lane
// groups g = 0..3 of a wave.
template <int KK>
__device__ __forceinline__ bf16x8 features8(const float* q, float3 lo, float3 ie, int g) {
  Pack8 x;
  if (KK < 2) {  // 24 features per axis (sin, cos of 2^k pi p, k = 0..11): 8 never straddle an axis
    const int j0 = KK * 32 + g * 8, a = j0 / 24;
    const float p = a == 0 ? (q[0] - lo.x) * ie.x : a == 1 ? (q[1] - lo.y) * ie.y : (q[2] - lo.z) * ie.z;
    sincos4(p, (j0 % 24) / 2, x);
  } else {  // features 64 + 8 g: z at k = 8..11 for g = 0; direction, normal, albedo, padding after
    sincos4((q[2] - lo.z) * ie.z, 8, x);
    if (g > 0) {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const float v = g == 1 ? q[3 + i] : (g == 2 && i == 0) ? q[11] : 0.0f;
        x.u[i] = bf16_bits(v);
      }
    }
  }
  return x.v;
}

// Inference straight from raw inputs (n x 12 floats): each lane's layer-0 B operands are computed
// in registers (no encoded copy in HBM); otherwise k_nrc_infer, with the same arithmetic per row.
// APPLY: instead of storing the outputs, add them to the querying paths' radiance (chain_apply) and
// count the queries into `total`.
template <bool APPLY>
__global__ __launch_bounds__(256) void k_nrc_query(NrcForward P, const float* __restrict__ raw, float3 lo, float3 ie,
                                                   float* __restrict__ out, uint32_t n,
                                                   const uint32_t* __restrict__ n_dev, const float4* __restrict__ beta,
                                                   float4* radiance, unsigned long long* total) {
  constexpr int RT = kChainTiles;
  __shared__ __attribute__((aligned(16))) ChainLds L;
  chain_stage(P, L);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  if (n_dev) n = min(n, *n_dev);
  if (APPLY && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(total, (unsigned long long)n);
  const uint32_t groups = (n + RT * 16 - 1) / (RT * 16);
  for (uint32_t grp = blockIdx.x * 4 + wave; grp < groups; grp += gridDim.x * 4) {
    const uint32_t row0 = grp * RT * 16;
    bf16x8 x0[RT][INP / 32];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
      const uint32_t row = row0 + t * 16 + r;
      float q[12];
      if (row < n) {
        const float4* src = reinterpret_cast<const float4*>(raw + (size_t)row * 12);
        const float4 a = src[0], b = src[1], c = src[2];
        q[0] = a.x; q[1] = a.y; q[2] = a.z; q[3] = a.w;
        q[4] = b.x; q[5] = b.y; q[6] = b.z; q[7] = b.w;
        q[8] = c.x; q[9] = c.y; q[10] = c.z; q[11] = c.w;
      } else {
#pragma unroll
        for (int i = 0; i < 12; ++i) q[i] = lo.x;  // any finite value: the row is never stored
      }
      x0[t][0] = features8<0>(q, lo, ie, g);
      x0[t][1] = features8<1>(q, lo, ie, g);
      x0[t][2] = features8<2>(q, lo, ie, g);
    }
    f32x4 y[RT];
    chain_forward<RT>(P, L, x0, r, g, y);
    if (APPLY)
      chain_apply<RT>(y, beta, radiance, row0, n, r, g);
    else
      chain_store<RT>(y, out, row0, n, r, g);
  }
}

// Training: forward, relative-L2 loss gradient and backward over 64-row tiles (4 waves x 16 rows) in
// the feature-major form of the inference chain. Every product of the forward and backward passes
// takes a weight matrix as its A operand and the wave's 16 batch rows as its B operand, so lane
// (r, g) holds features 16 nb + 4 g + i of row r from the first layer to the output and back: the
// backward products take W^T of layers 1..3 with its output columns in the chained order, and the
// ReLU masks come from the activations the lane kept. The weight gradients dW = dz^T a contract over
// the batch rows, so the waves also store each a and dz transposed ([feature][row], rows padded by
// 16 bytes) in LDS; after one barrier every wave accumulates its share of the dW tiles from 16-byte
// row reads, in MFMA registers across all the tiles the workgroup processes, and writes one
// partial per workgroup (padded layout, FP32). The bias gradients are dz^T 1 on the same operands
// (BF16 dz, as for the weights). Deterministic: fixed tile order, fixed reduction order.
constexpr int kST = 64 + 8;  // row stride (BF16) of the transposed tiles: 64 batch rows + 16 bytes
constexpr int kBw4 = 8;      // columns of the layer-4 backward weights the kernel reads
struct TrainLds {
  ChainLds fw;                // forward weights, chained
  uint16_t bw[3][64 * kSH];   // backward weights of layers 1..3: [in][chained out]
  uint16_t bw4[64 * kBw4];    // backward weights of layer 4: [in][out], outputs 0..7
  uint16_t xT[INP * kST];     // inputs [feature][row]
  uint16_t hT[4][64 * kST];   // activations of layers 0..3 [feature][row]
  uint16_t dzT[4][64 * kST];  // gradients at the pre-activations of layers 0..3
  uint16_t doT[4 * kST];      // gradients at the outputs (features 0..2; 3 is zero)
  float loss[4];
};
static_assert(sizeof(TrainLds) <= 160 * 1024, "training LDS above 160 KiB");
static_assert(4 * 5 * 64 * 4 <= sizeof(TrainLds::hT), "bias partials overlay the activations");

__device__ __forceinline__ uint32_t dpp_swap_pairs(uint32_t v) {  // lanes 2k <-> 2k + 1
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ float row16_sum(float v) {  // the sum over the 16 lanes of a DPP row
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x128, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x124, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x122, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x121, 0xF, 0xF, false));
  return v;
}

// Where a lane stores its words transposed. Lanes of rows 2k and 2k + 1 swap words; the even lane
// then writes feature f of both rows and the odd lane feature f + 1, each one 32-bit word of
// T[feature][row]: byte selector `sel` of v_perm over (swapped word, own word), and element offset
// `off` within a feature-row pair.
struct PairT {
  uint32_t sel;
  int off;
  __device__ explicit PairT(int row)
      : sel((row & 1) ? 0x03020706u : 0x05040100u), off((row & 1) * kST + (row & ~1)) {}
};
// Stores word w (features f and f + 1 of the lane's row, low half first) transposed. Every lane of
// a swapped pair takes part.
__device__ __forceinline__ void store_pair_t(uint16_t* T, int f, const PairT& pt, uint32_t w) {
  const uint32_t o = dpp_swap_pairs(w);
  *reinterpret_cast<uint32_t*>(T + f * kST + pt.off) = __builtin_amdgcn_perm(o, w, pt.sel);
}
// The two B operands of a hidden layer (chained features of the lane's row) into T
__device__ __forceinline__ void store_chain_t(uint16_t* T, const bf16x8 (&h)[2], const PairT& pt, int g) {
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
    Pack8 p;
    p.v = h[kk];
    store_pair_t(T, 32 * kk + 4 * g, pt, p.q.x);
    store_pair_t(T, 32 * kk + 4 * g + 2, pt, p.q.y);
    store_pair_t(T, 32 * kk + 16 + 4 * g, pt, p.q.z);
    store_pair_t(T, 32 * kk + 16 + 4 * g + 2, pt, p.q.w);
  }
}

// Weights into LDS with every global load issued before the first LDS store: chunk c = i + 256 k
// (16 bytes) of a [rows][cols] BF16 matrix goes to row c / (cols / 8) of `stride` elements.
template <int COLS>
__device__ __forceinline__ void stage_put(uint16_t* dst, int stride, int c, uint4 v) {
  *reinterpret_cast<uint4*>(dst + (c / (COLS / 8)) * stride + (c % (COLS / 8)) * 8) = v;
}
__device__ __forceinline__ uint4 stage_get(const uint16_t* __restrict__ src, int c) {
  return *reinterpret_cast<const uint4*>(src + c * 8);
}

__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u)); }

// dz of a hidden layer, rounded to BF16, from da (accumulator layout) and the layer's kept
// activations: ReLU'(a) is a > 0 of the BF16 activation. The result is the chained B operand of the
// next backward product.
__device__ __forceinline__ void relu_grad(const f32x4 (&da)[4], const bf16x8 (&h)[2], bf16x8 (&dz)[2]) {
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
    Pack8 a, d;
    a.v = h[kk];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int nb = 2 * kk + j / 4, i = j % 4;
      d.u[j] = bf16_bits((int16_t)a.u[j] > 0 ? da[nb][i] : 0.f);
    }
    dz[kk] = d.v;
  }
}

// da of the layer below: W^T (LDS, [in][chained out], row stride kSH) times the chained dz operands
__device__ __forceinline__ void back_chain(const uint16_t* Wt, const bf16x8 (&dz)[2], int r, int g, f32x4 (&da)[4]) {
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    da[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) da[nb] = mfma(load8(Wt + (nb * 16 + r) * kSH + kk * 32 + g * 8), dz[kk], da[nb]);
  }
}

__device__ __forceinline__ bf16x8 ones8() {  // eight BF16 1.0
  return __builtin_bit_cast(bf16x8, make_uint4(0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u));
}

// acc[i] += dz^T a over the 64 rows for the dW tiles (rows 16 m of dzT, rows 16 i of aT), i < NT,
// and db += dz^T 1 (every column the bias gradient of features 16 m + 4 g + e). Operands are read
// before the products.
template <int NT>
__device__ __forceinline__ void dw_tiles(const uint16_t* dzT, const uint16_t* aT, int m, int r, int g,
                                         f32x4 (&acc)[NT], f32x4& db) {
  bf16x8 a[2], b[2][NT];
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
    a[kk] = load8(dzT + (16 * m + r) * kST + kk * 32 + g * 8);
#pragma unroll
    for (int i = 0; i < NT; ++i) b[kk][i] = load8(aT + (16 * i + r) * kST + kk * 32 + g * 8);
  }
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
    for (int i = 0; i < NT; ++i) acc[i] = mfma(a[kk], b[kk][i], acc[i]);
    db = mfma(a[kk], ones8(), db);
  }
}

__global__ __launch_bounds__(256) void k_nrc_train(NrcDevice P, const uint16_t* __restrict__ x,
                                                   const float* __restrict__ target, uint32_t n,
                                                   const uint32_t* __restrict__ n_dev, float* __restrict__ partial) {
  __shared__ __attribute__((aligned(16))) TrainLds L;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  if (n_dev) n = min(n, *n_dev);
  const float inv_count = n ? 1.0f / (3.0f * (float)n) : 0.0f;  // mean over samples and channels
  // weight-gradient tiles: layer l's tiles (16 m + .., 16 i + ..) with m = wave
  f32x4 gw0[6], gw1[4], gw2[4], gw3[4], gw4[1];
#pragma unroll
  for (int i = 0; i < 6; ++i) gw0[i] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < 4; ++i) gw1[i] = gw2[i] = gw3[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  gw4[0] = f32x4{0.f, 0.f, 0.f, 0.f};
  f32x4 gb[5];  // bias gradients of layers 0..3 (features 16 wave + 4 g + e) and 4 (wave 0, 4 g + e)
#pragma unroll
  for (int l = 0; l < 5; ++l) gb[l] = f32x4{0.f, 0.f, 0.f, 0.f};
  float loss_acc = 0.f;
  const bf16x8 zero = zero8();
  const NrcForward& F = P.chain;
  const int tr = wave * 16 + r;  // the lane's row within the tile
  const PairT pt(tr);
  const uint32_t tiles = (n + 63) / 64;
  // the lane's inputs (features 32 kk + 8 g .. + 7) and, in lane group 0, targets of tile t's row
  bf16x8 xb[INP / 32];
  float tg[3];
  auto fetch = [&](uint32_t t) {
    const uint32_t row = t * 64 + tr;
    const bool live = row < n;
#pragma unroll
    for (int kk = 0; kk < INP / 32; ++kk) xb[kk] = live ? load8(x + (size_t)row * INP + kk * 32 + g * 8) : zero;
#pragma unroll
    for (int i = 0; i < 3; ++i) tg[i] = live && g == 0 ? target[(size_t)row * 3 + i] : 0.f;
  };
  {
    const int i = threadIdx.x;
    static_assert(64 * INP / 8 == 3 * 256 && 64 * H / 8 == 2 * 256 && 16 * H / 8 == 128 && 64 * kBw4 / 8 == 64, "");
    uint4 v0[3], v[6][2], v4 = make_uint4(0u, 0u, 0u, 0u), u4 = make_uint4(0u, 0u, 0u, 0u);
    const uint16_t* hid[6] = {F.w1, F.w2, F.w3, P.wt1, P.wt2, P.wt3};
#pragma unroll
    for (int k = 0; k < 3; ++k) v0[k] = stage_get(F.w0, i + 256 * k);
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
      for (int k = 0; k < 2; ++k) v[m][k] = stage_get(hid[m], i + 256 * k);
    if (i < 128) v4 = stage_get(F.w4, i);
    if (i < 64) u4 = stage_get(P.wt4, i);
    fetch(blockIdx.x);  // the first tile's rows, in flight with the weights
    uint16_t* dst[6] = {L.fw.w1, L.fw.w2, L.fw.w3, L.bw[0], L.bw[1], L.bw[2]};
#pragma unroll
    for (int k = 0; k < 3; ++k) stage_put<INP>(L.fw.w0, kS0, i + 256 * k, v0[k]);
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
      for (int k = 0; k < 2; ++k) stage_put<H>(dst[m], kSH, i + 256 * k, v[m][k]);
    if (i < 128) stage_put<H>(L.fw.w4, kSH, i, v4);
    if (i < 64) stage_put<kBw4>(L.bw4, kBw4, i, u4);
    if (i >= 128 && i < 128 + H) {  // biases: 64 float4 of layers 0..3, 4 of the output layer
      const int c = i - 128, l = c / (H / 4), k = c % (H / 4);
      const float* b = l == 0 ? F.b0 : l == 1 ? F.b1 : l == 2 ? F.b2 : F.b3;
      *reinterpret_cast<float4*>(&L.fw.b[l][4 * k]) = *reinterpret_cast<const float4*>(b + 4 * k);
    } else if (i >= 192 && i < 196) {
      *reinterpret_cast<float4*>(&L.fw.b4[4 * (i - 192)]) = *reinterpret_cast<const float4*>(F.b4 + 4 * (i - 192));
    }
  }
  __syncthreads();  // the weights are in LDS
  for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint32_t row = t * 64 + tr;
    const bool live = row < n;
    // forward, activations kept in registers and stored transposed
#pragma unroll
    for (int kk = 0; kk < INP / 32; ++kk) {
      Pack8 p;
      p.v = xb[kk];
      store_pair_t(L.xT, 32 * kk + 8 * g, pt, p.q.x);
      store_pair_t(L.xT, 32 * kk + 8 * g + 2, pt, p.q.y);
      store_pair_t(L.xT, 32 * kk + 8 * g + 4, pt, p.q.z);
      store_pair_t(L.xT, 32 * kk + 8 * g + 6, pt, p.q.w);
    }
    f32x4 acc[1][4];
    bf16x8 h0[1][2], h1[1][2], h2[1][2], h3[1][2];
    chain_bias<1>(L.fw.b[0], g, acc);
#pragma unroll
    for (int nb = 0; nb < 4; ++nb)
#pragma unroll
      for (int kk = 0; kk < INP / 32; ++kk)
        acc[0][nb] = mfma(load8(L.fw.w0 + (nb * 16 + r) * kS0 + kk * 32 + g * 8), xb[kk], acc[0][nb]);
    chain_act<1>(acc, h0);
    store_chain_t(L.hT[0], h0[0], pt, g);
    chain_hidden<1>(h0, L.fw.w1, L.fw.b[1], r, g, acc);
    chain_act<1>(acc, h1);
    store_chain_t(L.hT[1], h1[0], pt, g);
    chain_hidden<1>(h1, L.fw.w2, L.fw.b[2], r, g, acc);
    chain_act<1>(acc, h2);
    store_chain_t(L.hT[2], h2[0], pt, g);
    chain_hidden<1>(h2, L.fw.w3, L.fw.b[3], r, g, acc);
    chain_act<1>(acc, h3);
    store_chain_t(L.hT[3], h3[0], pt, g);
    const float4 b4 = *reinterpret_cast<const float4*>(L.fw.b4 + 4 * g);
    f32x4 y = f32x4{b4.x, b4.y, b4.z, b4.w};
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) y = mfma(load8(L.fw.w4 + r * kSH + kk * 32 + g * 8), h3[0][kk], y);
    // relative L2 loss and its gradient at the outputs (lanes g = 0 hold outputs 0..3 of row r)
    Pack8 dout;
    dout.q = make_uint4(0, 0, 0, 0);
    if (g == 0 && live) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const float tv = tg[i];
        const float yv = y[i];
        const float den = yv * yv + 0.01f;
        const float diff = yv - tv;
        loss_acc += diff * diff / den * inv_count;
        dout.u[i] = bf16_bits(2.0f * diff / den * inv_count);
      }
    }
    if (g == 0) {  // both lanes of every swapped pair are in lane group 0
      store_pair_t(L.doT, 0, pt, dout.q.x);
      store_pair_t(L.doT, 2, pt, dout.q.y);
    }
    // backward through layer 4: da3 = W4^T dout (K = outputs 0..7 in lane group 0)
    f32x4 da[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      const bf16x8 a = g == 0 ? load8(L.bw4 + (nb * 16 + r) * kBw4) : zero;
      da[nb] = mfma(a, dout.v, f32x4{0.f, 0.f, 0.f, 0.f});
    }
    bf16x8 dz[2];
    relu_grad(da, h3[0], dz);
    store_chain_t(L.dzT[3], dz, pt, g);
    back_chain(L.bw[2], dz, r, g, da);
    relu_grad(da, h2[0], dz);
    store_chain_t(L.dzT[2], dz, pt, g);
    back_chain(L.bw[1], dz, r, g, da);
    relu_grad(da, h1[0], dz);
    store_chain_t(L.dzT[1], dz, pt, g);
    back_chain(L.bw[0], dz, r, g, da);
    relu_grad(da, h0[0], dz);
    store_chain_t(L.dzT[0], dz, pt, g);
    __syncthreads();
    if (t + gridDim.x < tiles) fetch(t + gridDim.x);  // in flight during the products below
    // weight gradients over the tile's 64 rows: wave w owns output rows 16 w .. 16 w + 15 of
    // layers 0..3 and input columns 16 w .. of layer 4
    {
      bf16x8 a[2], b[2];
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        a[kk] = r < 4 ? load8(L.doT + r * kST + kk * 32 + g * 8) : zero;
        b[kk] = load8(L.hT[3] + (16 * wave + r) * kST + kk * 32 + g * 8);
      }
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        gw4[0] = mfma(a[kk], b[kk], gw4[0]);
        if (wave == 0) gb[4] = mfma(a[kk], ones8(), gb[4]);
      }
    }
    dw_tiles<4>(L.dzT[3], L.hT[2], wave, r, g, gw3, gb[3]);
    dw_tiles<4>(L.dzT[2], L.hT[1], wave, r, g, gw2, gb[2]);
    dw_tiles<4>(L.dzT[1], L.hT[0], wave, r, g, gw1, gb[1]);
    dw_tiles<6>(L.dzT[0], L.xT, wave, r, g, gw0, gb[0]);
    __syncthreads();
  }
  // this workgroup's partial in the padded layout: dW element (16 m + 4 g + e, 16 i + r) in lane
  // (r, g), element e of tile i
  float* out = partial + (size_t)blockIdx.x * kNrcPartialStride;
  auto put = [&](int off, int ncols, int m0, int n0, f32x4 v) {
#pragma unroll
    for (int e = 0; e < 4; ++e) out[off + (m0 + g * 4 + e) * ncols + n0 + r] = v[e];
  };
#pragma unroll
  for (int i = 0; i < 6; ++i) put(kNrcOffW0, INP, wave * 16, i * 16, gw0[i]);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    put(kNrcOffW1, H, wave * 16, i * 16, gw1[i]);
    put(kNrcOffW2, H, wave * 16, i * 16, gw2[i]);
    put(kNrcOffW3, H, wave * 16, i * 16, gw3[i]);
  }
  put(kNrcOffW4, H, 0, wave * 16, gw4[0]);
  // biases: features 16 wave + 4 g + e of layers 0..3 from lanes r = 0; layer 4's 64 entries from
  // wave 0 (features 0..15, then zeros)
  if (r == 0) {
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
      for (int e = 0; e < 4; ++e) out[kNrcOffB + l * H + 16 * wave + 4 * g + e] = gb[l][e];
    if (wave == 0)
#pragma unroll
      for (int e = 0; e < 4; ++e) out[kNrcOffB + 4 * H + 4 * g + e] = gb[4][e];
  }
  if (wave == 1 && lane >= 16) out[kNrcOffB + 4 * H + lane] = 0.f;
  for (int off = 32; off > 0; off >>= 1) loss_acc += __shfl_xor(loss_acc, off);
  if (lane == 0) L.loss[wave] = loss_acc;
  __syncthreads();
  if (threadIdx.x == 0) out[kNrcOffLoss] = ((L.loss[0] + L.loss[1]) + L.loss[2]) + L.loss[3];
}

// Sum over the workgroups' partials of padded entry `pad`: wave w of the kReduceWaves takes blocks
// w, w + kReduceWaves, ... (sixteen loads in flight) in eight running sums and s[w] their sum;
// returns the sum of the waves in a fixed order (valid in every lane after the barrier inside).
// Every thread calls it.
constexpr int kReduceWaves = 16;
__device__ __forceinline__ float reduce_entry(const float* __restrict__ partial, uint32_t blocks, int pad, bool valid,
                                              float (&s)[kReduceWaves][64]) {
  constexpr uint32_t W = kReduceWaves;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (valid) {
    uint32_t b = wave;
    for (; b + 15 * W < blocks; b += 16 * W) {
      float v[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = partial[(size_t)(b + W * q) * kNrcPartialStride + pad];
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[q % 8] += v[q];
    }
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (b + W * q < blocks) acc[q % 8] += partial[(size_t)(b + W * q) * kNrcPartialStride + pad];
  }
  s[wave][lane] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  __syncthreads();
  float v = 0.f;
#pragma unroll
  for (int w = 0; w < kReduceWaves; ++w) v += s[w][lane];
  return v;
}

// grads (padded layout) = the sum of the workgroups' partials, loss likewise
__global__ __launch_bounds__(64 * kReduceWaves) void k_nrc_reduce(const float* __restrict__ partial, uint32_t blocks,
                                                                  float* __restrict__ grads) {
  __shared__ float s[kReduceWaves][64];
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const float v = reduce_entry(partial, blocks, i, i <= kNrcOffLoss, s);
  if (threadIdx.x < 64 && i <= kNrcOffLoss) grads[i] = v;
}

// Parameter (layer, o, k) into the forward set F; `chained`: hidden input columns in the order
// nrc_chain_column gives.
__device__ __forceinline__ void pack_forward(const NrcForward& F, int layer, int o, int k, bool is_bias, float w,
                                             bool chained) {
  if (is_bias) {
    float* b = layer == 0 ? F.b0 : layer == 1 ? F.b1 : layer == 2 ? F.b2 : layer == 3 ? F.b3 : F.b4;
    b[o] = w;
    return;
  }
  if (chained && layer > 0) k = nrc_chain_column(k);
  const uint16_t q = bf16_bits(w);
  switch (layer) {
    case 0: F.w0[o * INP + k] = q; break;
    case 1: F.w1[o * H + k] = q; break;
    case 2: F.w2[o * H + k] = q; break;
    case 3: F.w3[o * H + k] = q; break;
    default: F.w4[o * H + k] = q; break;
  }
}

// Parameter i (FP32 value w, average a) into the BF16 sets: the chained forward weights (training
// and inference), the average's (with ema_decay > 0) and the backward weights.
__device__ __forceinline__ void pack_param(const NrcDevice& P, int i, float w, float a) {
  int layer, o, k;
  const bool is_bias = nrc_locate(i, &layer, &o, &k);
  pack_forward(P.chain, layer, o, k, is_bias, w, true);
  if (P.ema_decay > 0.f) pack_forward(P.avg, layer, o, k, is_bias, a, true);
  if (is_bias) return;
  const uint16_t q = bf16_bits(w);
  switch (layer) {
    case 0: break;
    case 1: P.wt1[k * H + nrc_chain_column(o)] = q; break;
    case 2: P.wt2[k * H + nrc_chain_column(o)] = q; break;
    case 3: P.wt3[k * H + nrc_chain_column(o)] = q; break;
    default: P.wt4[k * kBw4 + o] = q; break;
  }
}

// One optimiser step: the gradient of each parameter (the sum of the partials, also stored in
// grads, as k_nrc_reduce would), Adam (Kingma & Ba 2015) on the FP32 parameters, the average and
// the BF16 copies. Threads 64 b .. 64 b + 63 of the flat order, plus the loss after the last one.
__global__ __launch_bounds__(64 * kReduceWaves) void k_nrc_step(NrcDevice P, const float* __restrict__ partial,
                                                                uint32_t blocks, float* __restrict__ grads, float lr,
                                                                float b1, float b2, float eps, float c1, float c2) {
  __shared__ float s[kReduceWaves][64];
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int pad = i < kNrcParams ? nrc_padded_index(i) : kNrcOffLoss;
  const float gr = reduce_entry(partial, blocks, pad, i <= kNrcParams, s);
  if (threadIdx.x >= 64 || i > kNrcParams) return;
  grads[pad] = gr;
  if (i == kNrcParams) return;
  const float m = b1 * P.m[i] + (1.0f - b1) * gr;
  const float v = b2 * P.v[i] + (1.0f - b2) * gr * gr;
  P.m[i] = m;
  P.v[i] = v;
  const float mh = m / c1, vh = v / c2;
  const float w = P.master[i] - lr * mh / (sqrtf(vh) + eps);
  P.master[i] = w;
  float a = 0.f;
  if (P.ema_decay > 0.f) {
    a = P.ema_decay * P.ema[i] + (1.0f - P.ema_decay) * w;
    P.ema[i] = a;
  }
  pack_param(P, i, w, a);
}

__global__ void k_nrc_pack(NrcDevice P) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kNrcParams; i += gridDim.x * blockDim.x)
    pack_param(P, i, P.master[i], P.ema[i]);
}

// 12 raw floats per sample (position, outgoing direction, normal, albedo) -> 96 BF16 features
// (81 used, zero padded); p = (x - lo) * (1 / (hi - lo)) in FP32.
__global__ void k_nrc_encode(const float* __restrict__ raw, uint32_t n, const uint32_t* __restrict__ n_dev, float3 lo,
                             float3 inv_ext, uint16_t* __restrict__ x) {
  if (n_dev) n = min(n, *n_dev);
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const float* q = raw + (size_t)s * 12;
    uint16_t* o = x + (size_t)s * INP;
    const float p[3] = {(q[0] - lo.x) * inv_ext.x, (q[1] - lo.y) * inv_ext.y, (q[2] - lo.z) * inv_ext.z};
    int j = 0;
    for (int a = 0; a < 3; ++a) {
      float f = p[a];  // 2^k p is exact in FP32
      for (int k = 0; k < kNrcFreq; ++k) {
        float sv, cv;
        sincospi_red(f, &sv, &cv);
        o[j++] = bf16_bits(sv);
        o[j++] = bf16_bits(cv);
        f = f * 2.0f;
      }
    }
    for (int i = 3; i < 12; ++i) o[j++] = bf16_bits(q[i]);
    for (; j < INP; ++j) o[j] = 0;
  }
}

// The MFMA operand layout this file assumes, exercised on one 16x16x32 product.
__global__ void k_mfma_probe(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B, float* __restrict__ D) {
  const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
  const bf16x8 a = load8(A + r * 32 + g * 8);
  const bf16x8 b = gather8(B + (g * 8) * 16 + r, 16);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = mfma(a, b, acc);
  for (int i = 0; i < 4; ++i) D[(g * 4 + i) * 16 + r] = acc[i];
}

}  // namespace

hipError_t nrc_launch_infer(const NrcForward& P, const uint16_t* x, float* out, uint32_t n, const uint32_t* n_dev,
                            uint32_t cu, void* s) {
  if (!n) return hipSuccess;
  const uint32_t groups = (n + kChainTiles * 16 - 1) / (kChainTiles * 16);
  uint32_t blocks = (groups + 3) / 4;
  const uint32_t cap = cu * 3;  // 39 KB of LDS each: three workgroups per CU
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(k_nrc_infer, dim3(blocks), dim3(256), 0, (hipStream_t)s, P, x, out, n, n_dev);
  return hipGetLastError();
}

hipError_t nrc_launch_query(const NrcForward& P, const float* raw, const float lo[3], const float hi[3], float* out,
                            uint32_t n, const uint32_t* n_dev, uint32_t cu, void* s) {
  if (!n) return hipSuccess;
  const float3 l = make_float3(lo[0], lo[1], lo[2]);
  const float3 ie = make_float3(1.0f / (hi[0] - lo[0]), 1.0f / (hi[1] - lo[1]), 1.0f / (hi[2] - lo[2]));
  const uint32_t groups = (n + kChainTiles * 16 - 1) / (kChainTiles * 16);
  uint32_t blocks = (groups + 3) / 4;
  const uint32_t cap = cu * 3;
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(k_nrc_query<false>, dim3(blocks), dim3(256), 0, (hipStream_t)s, P, raw, l, ie, out, n, n_dev,
                     nullptr, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t nrc_launch_query_apply(const NrcForward& P, const float* raw, const float4* beta, float4* radiance,
                                  unsigned long long* total, const float lo[3], const float hi[3], uint32_t n,
                                  const uint32_t* n_dev, uint32_t cu, void* s) {
  if (!n) return hipSuccess;
  const float3 l = make_float3(lo[0], lo[1], lo[2]);
  const float3 ie = make_float3(1.0f / (hi[0] - lo[0]), 1.0f / (hi[1] - lo[1]), 1.0f / (hi[2] - lo[2]));
  const uint32_t groups = (n + kChainTiles * 16 - 1) / (kChainTiles * 16);
  uint32_t blocks = (groups + 3) / 4;
  const uint32_t cap = cu * 3;
  if (blocks > cap) blocks = cap;
  hipLaunchKernelGGL(k_nrc_query<true>, dim3(blocks), dim3(256), 0, (hipStream_t)s, P, raw, l, ie, nullptr, n, n_dev,
                     beta, radiance, total);
  return hipGetLastError();
}

hipError_t nrc_launch_train(const NrcDevice& P, const uint16_t* x, const float* target, uint32_t n,
                            const uint32_t* n_dev, uint32_t blocks, float* partial, float* grads, void* s) {
  hipLaunchKernelGGL(k_nrc_train, dim3(blocks), dim3(256), 0, (hipStream_t)s, P, x, target, n, n_dev, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_nrc_reduce, dim3((kNrcOffLoss + 1 + 63) / 64), dim3(64 * kReduceWaves), 0, (hipStream_t)s,
                     partial, blocks, grads);
  return hipGetLastError();
}

hipError_t nrc_launch_train_step(const NrcDevice& P, const uint16_t* x, const float* target, uint32_t n,
                                 const uint32_t* n_dev, uint32_t blocks, float* partial, float* grads, float lr,
                                 float b1, float b2, float eps, int step, void* s) {
  hipLaunchKernelGGL(k_nrc_train, dim3(blocks), dim3(256), 0, (hipStream_t)s, P, x, target, n, n_dev, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_nrc_step, dim3((kNrcParams + 1 + 63) / 64), dim3(64 * kReduceWaves), 0, (hipStream_t)s, P,
                     partial, blocks, grads, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t nrc_launch_pack(const NrcDevice& P, void* s) {
  hipLaunchKernelGGL(k_nrc_pack, dim3((kNrcParams + 255) / 256), dim3(256), 0, (hipStream_t)s, P);
  return hipGetLastError();
}

hipError_t nrc_launch_encode(const float* raw, uint32_t n, const uint32_t* n_dev, const float lo[3], const float hi[3],
                             uint16_t* x, void* s) {
  if (!n) return hipSuccess;
  const float3 l = make_float3(lo[0], lo[1], lo[2]);
  const float3 ie = make_float3(1.0f / (hi[0] - lo[0]), 1.0f / (hi[1] - lo[1]), 1.0f / (hi[2] - lo[2]));
  uint32_t blocks = (n + 255) / 256;
  if (blocks > 65535) blocks = 65535;
  hipLaunchKernelGGL(k_nrc_encode, dim3(blocks), dim3(256), 0, (hipStream_t)s, raw, n, n_dev, l, ie, x);
  return hipGetLastError();
}

hipError_t nrc_launch_probe(const uint16_t* A, const uint16_t* B, float* D, void* s) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, (hipStream_t)s, A, B, D);
  return hipGetLastError();
}

}  // namespace pbrt
