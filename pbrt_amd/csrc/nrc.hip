// Neural radiance cache on CDNA4 (gfx950): a 81 -> 64 -> 64 -> 64 -> 64 -> 3 ReLU MLP whose
// inference, backpropagation and Adam update are hand-written here, with BF16 operands on the
// v_mfma_f32_16x16x32_bf16 matrix cores and FP32 accumulation. The reference renderer has no
// radiance cache; this follows Müller et al., "Real-time Neural Radiance Caching for Path
// Tracing" (2021): relative L2 loss, online training, paths terminated into the cache.
//
// Data layout (padded so every matrix product is a whole number of 16 x 16 x 32 MFMA tiles):
//   inputs   N x 96 BF16 (81 features, zero padded), row-major
//   forward  Wf0 [64][96], Wf1..3 [64][64], Wf4 [16][64] (rows 3.. zero), [out][in] row-major
//   inference the same, with the input columns of Wf1..4 permuted by nrc_chain_column
//   backward Wt1..3 [64][64], Wt4 [64][32] ([in][out], out padded with zeros)
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
// sum over the four lanes l, l ^ 16, l ^ 32, l ^ 48 (same MFMA output column), fixed order
__device__ __forceinline__ float col_sum(float s) {
  s += __shfl_xor(s, 16);
  s += __shfl_xor(s, 32);
  return s;
}

constexpr int IN = kNrcIn, INP = kNrcInPad, H = kNrcHidden, OUTP = 16, OUTK = 32;

// One layer of the forward pass for the 16 rows of a wave: out[r][c] = act(sum_k in[r][k] W[c][k]
// + b[c]), `in` in LDS (row stride in_stride), W the forward weights (row stride w_stride).
template <int K, int NOUT, bool RELU>
__device__ __forceinline__ void layer_fwd(const uint16_t* in, int in_stride, const uint16_t* W, int w_stride,
                                          const float* __restrict__ b, uint16_t* out, int out_stride, int lane,
                                          f32x4* keep) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int nb = 0; nb < NOUT / 16; ++nb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < K / 32; ++kk) {
      const bf16x8 a = load8(in + r * in_stride + kk * 32 + g * 8);
      const bf16x8 bb = load8(W + (nb * 16 + r) * w_stride + kk * 32 + g * 8);
      acc = mfma(a, bb, acc);
    }
    const int col = nb * 16 + r;
    const float bias = b[col];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float v = acc[i] + bias;
      if (RELU) v = v > 0.f ? v : 0.f;
      acc[i] = v;
      if (out) out[(g * 4 + i) * out_stride + col] = bf16_bits(v);
    }
    if (keep) keep[nb] = acc;
  }
}

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
  chain_bias<RT>(P.b0, g, acc);
#pragma unroll
  for (int nb = 0; nb < 4; ++nb)
#pragma unroll
    for (int kk = 0; kk < INP / 32; ++kk) {
      const bf16x8 a = load8(L.w0 + z + (nb * 16 + r) * kS0 + kk * 32 + g * 8);
#pragma unroll
      for (int t = 0; t < RT; ++t) acc[t][nb] = mfma(a, x0[t][kk], acc[t][nb]);
    }
  chain_act<RT>(acc, h);
  chain_hidden<RT>(h, L.w1 + z, P.b1, r, g, acc);
  chain_act<RT>(acc, h);
  chain_hidden<RT>(h, L.w2 + z, P.b2, r, g, acc);
  chain_act<RT>(acc, h);
  chain_hidden<RT>(h, L.w3 + z, P.b3, r, g, acc);
  chain_act<RT>(acc, h);
  const float4 b4 = *reinterpret_cast<const float4*>(P.b4 + 4 * g);
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

// Training: forward, relative-L2 loss gradient and backward over 64-row tiles (4 waves x 16 rows),
// weight gradients accumulated in MFMA registers across every tile the block processes, then one
// partial (padded layout, FP32) per block. Deterministic: fixed tile order, fixed reduction order.
// Rows of activations, gradients and weights are padded by 16 bytes in LDS, so that the 16 rows a
// product reads fall in distinct banks; the weights (forward [out][in], backward [in][out]) are
// staged once per workgroup.
constexpr int kSX = INP + 8, kSO = OUTK + 8;  // LDS row strides (BF16) of the inputs and dout
struct TrainLds {
  uint16_t x[64 * kSX];
  uint16_t h[4][64 * kSH];   // post-ReLU activations of layers 0..3
  uint16_t dz[2][64 * kSH];  // BF16 gradients at a layer's pre-activation (ping-pong)
  uint16_t dout[64 * kSO];
  uint16_t f0[64 * kS0], f1[64 * kSH], f2[64 * kSH], f3[64 * kSH], f4[16 * kSH];  // forward weights
  uint16_t t1[64 * kSH], t2[64 * kSH], t3[64 * kSH], t4[64 * kSO];                // backward weights
  float db[4][5][H];       // per-wave bias gradient partials
  float loss[4];
};
static_assert(sizeof(TrainLds) <= 160 * 1024, "training LDS above 160 KiB");

// Copies a rows x cols BF16 matrix (cols a multiple of 8) into LDS rows of `stride` elements.
__device__ __forceinline__ void stage_rows(uint16_t* dst, int stride, const uint16_t* __restrict__ src, int rows, int cols) {
  for (int c = threadIdx.x; c < rows * cols / 8; c += blockDim.x) {
    const int row = c / (cols / 8), k8 = c % (cols / 8);
    *reinterpret_cast<uint4*>(dst + row * stride + k8 * 8) = *reinterpret_cast<const uint4*>(src + row * cols + k8 * 8);
  }
}

// dW tile (rows m0.. of the output dim, cols n0.. of the input dim) += dz^T act over the tile's 64 rows
__device__ __forceinline__ f32x4 dw_tile(const uint16_t* dz, int dz_stride, const uint16_t* act, int act_stride, int m0,
                                         int n0, int lane, f32x4 acc) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
    const bf16x8 a = gather8(dz + (kk * 32 + g * 8) * dz_stride + m0 + r, dz_stride);
    const bf16x8 b = gather8(act + (kk * 32 + g * 8) * act_stride + n0 + r, act_stride);
    acc = mfma(a, b, acc);
  }
  return acc;
}

// Backward through W_l for the wave's 16 rows: da = dz W (K = the layer's output dim), times
// ReLU'(a) from the stored activation; writes BF16 dz of the previous layer and sums its FP32
// values into the bias partial.
template <int KOUT>
__device__ __forceinline__ void back_rows(const uint16_t* dz_in, int dz_stride, const uint16_t* Wt, int wt_stride,
                                          const uint16_t* act, uint16_t* dz_out, float* dbp, int row0, int lane) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int nb = 0; nb < H / 16; ++nb) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < KOUT / 32; ++kk) {
      const bf16x8 a = load8(dz_in + (row0 + r) * dz_stride + kk * 32 + g * 8);
      const bf16x8 b = load8(Wt + (nb * 16 + r) * wt_stride + kk * 32 + g * 8);
      acc = mfma(a, b, acc);
    }
    const int col = nb * 16 + r;
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = row0 + g * 4 + i;
      const float v = bf16_float(act[row * kSH + col]) > 0.f ? acc[i] : 0.f;
      s += v;
      dz_out[row * kSH + col] = bf16_bits(v);
    }
    s = col_sum(s);
    if (g == 0) dbp[col] += s;
  }
}

__global__ __launch_bounds__(256) void k_nrc_train(NrcDevice P, const uint16_t* __restrict__ x,
                                                   const float* __restrict__ target, uint32_t n,
                                                   const uint32_t* __restrict__ n_dev, float* __restrict__ partial) {
  __shared__ __attribute__((aligned(16))) TrainLds L;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (n_dev) n = min(n, *n_dev);
  const float inv_count = n ? 1.0f / (3.0f * (float)n) : 0.0f;  // mean over samples and channels
  const int r = lane & 15, g = lane >> 4;
  for (int i = threadIdx.x; i < 4 * 5 * H; i += 256) (&L.db[0][0][0])[i] = 0.f;
  for (int i = threadIdx.x; i < 64 * kSO; i += 256) L.dout[i] = 0;
  stage_rows(L.f0, kS0, P.wf0, 64, INP);
  stage_rows(L.f1, kSH, P.wf1, 64, H);
  stage_rows(L.f2, kSH, P.wf2, 64, H);
  stage_rows(L.f3, kSH, P.wf3, 64, H);
  stage_rows(L.f4, kSH, P.wf4, 16, H);
  stage_rows(L.t1, kSH, P.wt1, 64, H);
  stage_rows(L.t2, kSH, P.wt2, 64, H);
  stage_rows(L.t3, kSH, P.wt3, 64, H);
  stage_rows(L.t4, kSO, P.wt4, 64, OUTK);
  if (threadIdx.x < 4) L.loss[threadIdx.x] = 0.f;
  // weight-gradient accumulators, tile t of a layer owned by wave t % 4
  f32x4 gw0[6], gw1[4], gw2[4], gw3[4], gw4;
#pragma unroll
  for (int i = 0; i < 6; ++i) gw0[i] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < 4; ++i) gw1[i] = gw2[i] = gw3[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  gw4 = f32x4{0.f, 0.f, 0.f, 0.f};
  float loss_acc = 0.f;
  __syncthreads();
  const uint32_t tiles = (n + 63) / 64;
  for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint32_t base = t * 64;
    // stage the 64 input rows (16-byte chunks), zeros past the end
    for (int c = threadIdx.x; c < 64 * INP / 8; c += 256) {
      const int row = c / (INP / 8), k8 = c % (INP / 8);
      uint4 v = make_uint4(0, 0, 0, 0);
      if (base + row < n) v = *reinterpret_cast<const uint4*>(x + (size_t)(base + row) * INP + k8 * 8);
      *reinterpret_cast<uint4*>(L.x + row * kSX + k8 * 8) = v;
    }
    __syncthreads();
    const int row0 = wave * 16;
    // forward (the wave's own rows)
    layer_fwd<INP, H, true>(L.x + row0 * kSX, kSX, L.f0, kS0, P.b0, L.h[0] + row0 * kSH, kSH, lane, nullptr);
    wave_sync();
    layer_fwd<H, H, true>(L.h[0] + row0 * kSH, kSH, L.f1, kSH, P.b1, L.h[1] + row0 * kSH, kSH, lane, nullptr);
    wave_sync();
    layer_fwd<H, H, true>(L.h[1] + row0 * kSH, kSH, L.f2, kSH, P.b2, L.h[2] + row0 * kSH, kSH, lane, nullptr);
    wave_sync();
    layer_fwd<H, H, true>(L.h[2] + row0 * kSH, kSH, L.f3, kSH, P.b3, L.h[3] + row0 * kSH, kSH, lane, nullptr);
    wave_sync();
    f32x4 y;
    layer_fwd<H, OUTP, false>(L.h[3] + row0 * kSH, kSH, L.f4, kSH, P.b4, nullptr, 0, lane, &y);
    // relative L2 loss and its gradient at the outputs
    {
      float s = 0.f;
      if (r < 3) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int row = row0 + g * 4 + i;
          float dz = 0.f;
          if (base + row < n) {
            const float tv = target[(size_t)(base + row) * 3 + r];
            const float yv = y[i];
            const float den = yv * yv + 0.01f;
            const float diff = yv - tv;
            loss_acc += diff * diff / den * inv_count;
            dz = 2.0f * diff / den * inv_count;
          }
          s += dz;
          L.dout[row * kSO + r] = bf16_bits(dz);
        }
      }
      s = col_sum(s);
      if (g == 0 && r < 3) L.db[wave][4][r] += s;
    }
    __syncthreads();
    // layer 4: dW4 (one 16-row tile x four 16-col tiles: wave w owns col tile w), then dz3 = dout W4 * relu'
    gw4 = dw_tile(L.dout, kSO, L.h[3], kSH, 0, wave * 16, lane, gw4);
    back_rows<OUTK>(L.dout, kSO, L.t4, kSO, L.h[3], L.dz[0], L.db[wave][3], row0, lane);
    __syncthreads();
    // layer 3
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int tile = wave + 4 * i;
      gw3[i] = dw_tile(L.dz[0], kSH, L.h[2], kSH, (tile / 4) * 16, (tile % 4) * 16, lane, gw3[i]);
    }
    back_rows<H>(L.dz[0], kSH, L.t3, kSH, L.h[2], L.dz[1], L.db[wave][2], row0, lane);
    __syncthreads();
    // layer 2
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int tile = wave + 4 * i;
      gw2[i] = dw_tile(L.dz[1], kSH, L.h[1], kSH, (tile / 4) * 16, (tile % 4) * 16, lane, gw2[i]);
    }
    back_rows<H>(L.dz[1], kSH, L.t2, kSH, L.h[1], L.dz[0], L.db[wave][1], row0, lane);
    __syncthreads();
    // layer 1
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int tile = wave + 4 * i;
      gw1[i] = dw_tile(L.dz[0], kSH, L.h[0], kSH, (tile / 4) * 16, (tile % 4) * 16, lane, gw1[i]);
    }
    back_rows<H>(L.dz[0], kSH, L.t1, kSH, L.h[0], L.dz[1], L.db[wave][0], row0, lane);
    __syncthreads();
    // layer 0: dW0 = dz0^T x (4 row tiles x 6 col tiles)
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const int tile = wave + 4 * i;
      gw0[i] = dw_tile(L.dz[1], kSH, L.x, kSX, (tile / 6) * 16, (tile % 6) * 16, lane, gw0[i]);
    }
    __syncthreads();
  }
  // write this block's partial in the padded layout
  float* out = partial + (size_t)blockIdx.x * kNrcPartialStride;
  auto put = [&](int off, int ncols, int m0, int n0, f32x4 v) {
#pragma unroll
    for (int i = 0; i < 4; ++i) out[off + (m0 + g * 4 + i) * ncols + n0 + r] = v[i];
  };
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const int tile = wave + 4 * i;
    put(kNrcOffW0, INP, (tile / 6) * 16, (tile % 6) * 16, gw0[i]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int tile = wave + 4 * i;
    put(kNrcOffW1, H, (tile / 4) * 16, (tile % 4) * 16, gw1[i]);
    put(kNrcOffW2, H, (tile / 4) * 16, (tile % 4) * 16, gw2[i]);
    put(kNrcOffW3, H, (tile / 4) * 16, (tile % 4) * 16, gw3[i]);
  }
  put(kNrcOffW4, H, 0, wave * 16, gw4);
  // loss: reduce over the wave's lanes in a fixed order
  for (int off = 32; off > 0; off >>= 1) loss_acc += __shfl_xor(loss_acc, off);
  if (lane == 0) L.loss[wave] = loss_acc;
  __syncthreads();
  for (int i = threadIdx.x; i < 5 * H; i += 256) {
    const int l = i / H, c = i % H;  // bias of layer l (0..3 hidden, 4 output), summed over the waves
    const float s = ((L.db[0][l][c] + L.db[1][l][c]) + L.db[2][l][c]) + L.db[3][l][c];
    out[kNrcOffB + l * H + c] = s;
  }
  if (threadIdx.x == 0) out[kNrcOffLoss] = ((L.loss[0] + L.loss[1]) + L.loss[2]) + L.loss[3];
}

// grads (padded layout) = sum over the workgroups' partials, loss likewise: eight running sums
// over blocks b = 8k + r (independent loads in flight), added in a fixed order at the end.
__global__ __launch_bounds__(64) void k_nrc_reduce(const float* __restrict__ partial, uint32_t blocks,
                                                   float* __restrict__ grads) {
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  if (i > (uint32_t)kNrcOffLoss) return;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  uint32_t b = 0;
  for (; b + 8 <= blocks; b += 8) {
#pragma unroll
    for (int r = 0; r < 8; ++r) acc[r] += partial[(size_t)(b + r) * kNrcPartialStride + i];
  }
  for (int r = 0; b < blocks; ++b, ++r) acc[r] += partial[(size_t)b * kNrcPartialStride + i];
  grads[i] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

// Adam (Kingma & Ba 2015) on the flat FP32 parameters, then the BF16 forward / backward copies.
__global__ void k_nrc_adam(NrcDevice P, const float* __restrict__ grads, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kNrcParams; i += gridDim.x * blockDim.x) {
    const float gr = grads[nrc_padded_index(i)];
    const float m = b1 * P.m[i] + (1.0f - b1) * gr;
    const float v = b2 * P.v[i] + (1.0f - b2) * gr * gr;
    P.m[i] = m;
    P.v[i] = v;
    const float mh = m / c1, vh = v / c2;
    const float w = P.master[i] - lr * mh / (sqrtf(vh) + eps);
    P.master[i] = w;
    if (P.ema_decay > 0.f) P.ema[i] = P.ema_decay * P.ema[i] + (1.0f - P.ema_decay) * w;
  }
}

// Parameter (layer, o, k) into the forward set F; `chained`: hidden input columns in the order
// nrc_chain_column gives (the inference sets).
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

__global__ void k_nrc_pack(NrcDevice P) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kNrcParams; i += gridDim.x * blockDim.x) {
    const float w = P.master[i];
    int layer, o, k;
    const bool is_bias = nrc_locate(i, &layer, &o, &k);
    pack_forward(P.train(), layer, o, k, is_bias, w, false);
    pack_forward(P.chain, layer, o, k, is_bias, w, true);  // w0 and biases: the same words again
    if (P.ema_decay > 0.f) pack_forward(P.avg, layer, o, k, is_bias, P.ema[i], true);
    if (is_bias) continue;
    const uint16_t q = bf16_bits(w);
    switch (layer) {
      case 0: break;
      case 1: P.wt1[k * H + o] = q; break;
      case 2: P.wt2[k * H + o] = q; break;
      case 3: P.wt3[k * H + o] = q; break;
      default: P.wt4[k * OUTK + o] = q; break;
    }
  }
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
  hipLaunchKernelGGL(k_nrc_reduce, dim3((kNrcOffLoss + 1 + 63) / 64), dim3(64), 0, (hipStream_t)s, partial, blocks,
                     grads);
  return hipGetLastError();
}

hipError_t nrc_launch_adam(const NrcDevice& P, const float* grads, float lr, float b1, float b2, float eps, int step,
                           void* s) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_nrc_adam, dim3((kNrcParams + 255) / 256), dim3(256), 0, (hipStream_t)s, P, grads, lr, b1, b2,
                     eps, c1, c2);
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
