// Neural radiance cache: sizes, padded device layouts and the launch helpers of nrc.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pbrt_gpu.h"
#include "context.hpp"

namespace pbrt {

constexpr int kNrcIn = 81;      // 3 axes x 12 frequencies x (sin, cos) + direction + normal + albedo
constexpr int kNrcInPad = 96;   // three K = 32 MFMA steps
constexpr int kNrcHidden = 64;
constexpr int kNrcOut = 3;
constexpr int kNrcFreq = 12;
constexpr int kNrcLayers = 5;
constexpr int kNrcParams = (kNrcIn * 64 + 64) + 3 * (64 * 64 + 64) + (64 * 3 + 3);  // 17923

// Padded FP32 gradient layout (also one block's partial): W0 [64][96], W1..W3 [64][64], W4 [16][64],
// biases [5][64], the loss, padded to a multiple of 64 floats.
constexpr int kNrcOffW0 = 0;
constexpr int kNrcOffW1 = kNrcOffW0 + 64 * kNrcInPad;
constexpr int kNrcOffW2 = kNrcOffW1 + 64 * 64;
constexpr int kNrcOffW3 = kNrcOffW2 + 64 * 64;
constexpr int kNrcOffW4 = kNrcOffW3 + 64 * 64;
constexpr int kNrcOffB = kNrcOffW4 + 16 * 64;
constexpr int kNrcOffLoss = kNrcOffB + 5 * 64;
constexpr int kNrcPartialStride = (kNrcOffLoss + 1 + 63) / 64 * 64;

// Position of flat parameter i (per layer: W [out][in] row-major, then b [out]). Returns true for
// a bias (then *k is 0).
__host__ __device__ inline bool nrc_locate(int i, int* layer, int* o, int* k) {
  int off = 0;
  for (int l = 0; l < kNrcLayers; ++l) {
    const int ni = l == 0 ? kNrcIn : kNrcHidden, no = l == kNrcLayers - 1 ? kNrcOut : kNrcHidden;
    if (i < off + ni * no) {
      *layer = l;
      *o = (i - off) / ni;
      *k = (i - off) % ni;
      return false;
    }
    off += ni * no;
    if (i < off + no) {
      *layer = l;
      *o = i - off;
      *k = 0;
      return true;
    }
    off += no;
  }
  *layer = -1;
  *o = *k = 0;
  return false;
}

__host__ __device__ inline int nrc_padded_index(int i) {
  int l, o, k;
  if (nrc_locate(i, &l, &o, &k)) return kNrcOffB + l * 64 + o;
  switch (l) {
    case 0: return kNrcOffW0 + o * kNrcInPad + k;
    case 1: return kNrcOffW1 + o * 64 + k;
    case 2: return kNrcOffW2 + o * 64 + k;
    case 3: return kNrcOffW3 + o * 64 + k;
    default: return kNrcOffW4 + o * 64 + k;
  }
}

// Column of hidden input feature k (0..63) in the chained layout of the inference weights of
// layers 1..4. A 16x16x32 MFMA with the weights as A operand and a layer's activations as B
// operand leaves features 16 nb + 4 g + i (i = 0..3, nb = 0..3) of its batch row in lane group g;
// the next layer's B operand of lane group g holds columns 32 kk + 8 g + j, so feature
// 16 (2 kk + j / 4) + 4 g + j % 4 goes to column 32 kk + 8 g + j and the activations never leave
// the lane.
__host__ __device__ constexpr int nrc_chain_column(int k) {
  return (k >> 5) * 32 + ((k >> 2) & 3) * 8 + ((k >> 4) & 1) * 4 + (k & 3);
}

// The weights a forward pass reads: BF16 [out][in] matrices ([64][96], [64][64] x 3, [16][64]) and
// FP32 biases ([64] x 4, [16]).
struct NrcForward {
  uint16_t *w0, *w1, *w2, *w3, *w4;
  float *b0, *b1, *b2, *b3, *b4;
};

// Device views of the network (one allocation). BF16 values are stored as their 16-bit patterns.
struct NrcDevice {
  NrcForward chain;                // the parameters in BF16: layer 0 [64][96], the hidden layers' input
                                   // columns in the chained order (training and inference)
  uint16_t *wt1, *wt2, *wt3, *wt4;  // backward, [in][out]: [64][64] x 3 with the output columns in the
                                   // chained order, [64][8] (outputs 0..2 used)
  float *master, *m, *v;           // FP32 parameters and Adam moments, flat order
  float* ema;                      // exponential moving average of the parameters, flat order
  NrcForward avg;                  // inference weights of the average, chained likewise
  float ema_decay;                 // 0: inference reads the parameters themselves
  __host__ __device__ NrcForward infer() const { return ema_decay > 0.f ? avg : chain; }
};

// Bytes of one NrcDevice allocation and the carving of it.
size_t nrc_device_bytes();
NrcDevice nrc_carve(void* base);

// `n` is the batch size, or with `n_dev` set the capacity bound of a batch counted on the device.
hipError_t nrc_launch_infer(const NrcForward& F, const uint16_t* x, float* out, uint32_t n, const uint32_t* n_dev,
                            uint32_t cu, void* stream);
// Inference from raw inputs (n x 12 floats, 16-byte aligned rows), features computed in registers.
hipError_t nrc_launch_query(const NrcForward& F, const float* raw, const float lo[3], const float hi[3], float* out,
                            uint32_t n, const uint32_t* n_dev, uint32_t cu, void* stream);
// The same queries (at most n, the count at n_dev) with each output, clamped at 0, times beta[k].xyz
// added to radiance[bits of beta[k].w]; adds the count to *total.
hipError_t nrc_launch_query_apply(const NrcForward& F, const float* raw, const float4* beta, float4* radiance,
                                  unsigned long long* total, const float lo[3], const float hi[3], uint32_t n,
                                  const uint32_t* n_dev, uint32_t cu, void* stream);
hipError_t nrc_launch_train(const NrcDevice& P, const uint16_t* x, const float* target, uint32_t n,
                            const uint32_t* n_dev, uint32_t blocks, float* partial, float* grads, void* stream);
// The training kernel, then the reduction of its partials fused with Adam (step 1-based) and the
// BF16 copies of the new parameters; grads receives the gradient as nrc_launch_train leaves it.
hipError_t nrc_launch_train_step(const NrcDevice& P, const uint16_t* x, const float* target, uint32_t n,
                                 const uint32_t* n_dev, uint32_t blocks, float* partial, float* grads, float lr,
                                 float b1, float b2, float eps, int step, void* stream);
// BF16 copies of the parameters (training) and, with ema_decay > 0, of their average (inference).
hipError_t nrc_launch_pack(const NrcDevice& P, void* stream);
hipError_t nrc_launch_encode(const float* raw, uint32_t n, const uint32_t* n_dev, const float lo[3], const float hi[3],
                             uint16_t* x, void* stream);
hipError_t nrc_launch_probe(const uint16_t* A, const uint16_t* B, float* D, void* stream);

}  // namespace pbrt

// The cache behind the C ABI's opaque handle.
struct pbrt_nrc {
  pbrt_ctx* ctx = nullptr;
  void* base = nullptr;       // NrcDevice allocation
  pbrt::NrcDevice dev{};
  float* partial = nullptr;   // per-block gradient partials
  float* grads = nullptr;     // reduced gradient, padded layout, loss at kNrcOffLoss
  uint32_t max_blocks = 0;
  float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {1.f, 1.f, 1.f};
  pbrt_nrc_optim optim{1e-3f, 0.9f, 0.99f, 1e-8f, 0.f};
  uint64_t steps = 0;
};

namespace pbrt {
// Gradient over a batch of at most `cap` samples (`n_dev`: the count on the device, or NULL for
// exactly cap), one Adam step and the BF16 refresh, all on `stream`.
hipError_t nrc_train_device(pbrt_nrc* n, const uint16_t* x, const float* target, uint32_t cap, const uint32_t* n_dev,
                            void* stream);
}  // namespace pbrt
