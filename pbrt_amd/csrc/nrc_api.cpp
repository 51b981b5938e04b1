// C ABI of the neural radiance cache (include/pbrt_gpu.h): parameter storage, initialisation,
// and the train / infer entry points over the kernels of nrc.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "context.hpp"
#include "nrc.hpp"
#include "rng.hpp"

using namespace pbrt;

namespace pbrt {

namespace {
constexpr size_t kAlign = 256;
size_t up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
}  // namespace

// Two forward sets (the parameters and their average), each BF16 weights [64][96], [64][64] x 3,
// [16][64] and FP32 biases [64] x 4, [16]; the backward weights; master, m, v and ema.
size_t nrc_device_bytes() {
  size_t b = 0;
  b += 2 * (up(64 * kNrcInPad * 2) + 3 * up(64 * 64 * 2) + up(16 * 64 * 2));  // forward weights
  b += 2 * (4 * up(64 * 4) + up(16 * 4));                                     // and biases
  b += 3 * up(64 * 64 * 2) + up(64 * 8 * 2);                                  // backward weights
  b += 4 * up((size_t)kNrcParams * 4);                                        // master, m, v, ema
  return b;
}

NrcDevice nrc_carve(void* base) {
  char* p = (char*)base;
  auto take = [&](size_t bytes) {
    void* r = p;
    p += up(bytes);
    return r;
  };
  auto forward_set = [&](NrcForward& f) {
    f.w0 = (uint16_t*)take(64 * kNrcInPad * 2);
    f.w1 = (uint16_t*)take(64 * 64 * 2);
    f.w2 = (uint16_t*)take(64 * 64 * 2);
    f.w3 = (uint16_t*)take(64 * 64 * 2);
    f.w4 = (uint16_t*)take(16 * 64 * 2);
    f.b0 = (float*)take(64 * 4);
    f.b1 = (float*)take(64 * 4);
    f.b2 = (float*)take(64 * 4);
    f.b3 = (float*)take(64 * 4);
    f.b4 = (float*)take(16 * 4);
  };
  NrcDevice d{};
  forward_set(d.chain);
  forward_set(d.avg);
  d.wt1 = (uint16_t*)take(64 * 64 * 2);
  d.wt2 = (uint16_t*)take(64 * 64 * 2);
  d.wt3 = (uint16_t*)take(64 * 64 * 2);
  d.wt4 = (uint16_t*)take(64 * 8 * 2);
  d.master = (float*)take((size_t)kNrcParams * 4);
  d.m = (float*)take((size_t)kNrcParams * 4);
  d.v = (float*)take((size_t)kNrcParams * 4);
  d.ema = (float*)take((size_t)kNrcParams * 4);
  d.ema_decay = 0.f;
  return d;
}

hipError_t nrc_train_device(pbrt_nrc* n, const uint16_t* x, const float* target, uint32_t cap, const uint32_t* n_dev,
                            void* stream) {
  const uint32_t tiles = (cap + 63) / 64;
  const uint32_t blocks = tiles < n->max_blocks ? tiles : n->max_blocks;
  const pbrt_nrc_optim& o = n->optim;
  const hipError_t e = nrc_launch_train_step(n->dev, x, target, cap, n_dev, blocks ? blocks : 1, n->partial, n->grads,
                                             o.lr, o.beta1, o.beta2, o.eps, (int)(n->steps + 1), stream);
  if (e == hipSuccess) ++n->steps;
  return e;
}

}  // namespace pbrt

namespace {

void check(pbrt_nrc* n) {
  if (!n || !n->ctx) throw Error(PBRT_ERR_INVALID_ARG, "null radiance cache");
  PBRT_HIP(hipSetDevice(n->ctx->device));
}
hipStream_t stream_of(pbrt_nrc* n, void* s) { return s ? (hipStream_t)s : n->ctx->stream; }

// Encoded rows are read as 16-byte vectors; targets, outputs and raw vertices as floats.
void check_batch(const void* x, const void* f, uint32_t count) {
  if (count > (1u << 31)) throw Error(PBRT_ERR_INVALID_ARG, "batch larger than 2^31 samples");
  if (((uintptr_t)x & 15u) != 0) throw Error(PBRT_ERR_INVALID_ARG, "encoded features must be 16-byte aligned");
  if (((uintptr_t)f & 3u) != 0) throw Error(PBRT_ERR_INVALID_ARG, "float buffers must be 4-byte aligned");
}

// Uploads master parameters, clears the moments and refreshes the BF16 copies (synchronous).
void upload_params(pbrt_nrc* n, const float* params) {
  hipStream_t s = n->ctx->stream;
  PBRT_HIP(hipMemcpyAsync(n->dev.master, params, (size_t)kNrcParams * 4, hipMemcpyHostToDevice, s));
  PBRT_HIP(hipMemsetAsync(n->dev.m, 0, (size_t)kNrcParams * 4, s));
  PBRT_HIP(hipMemsetAsync(n->dev.v, 0, (size_t)kNrcParams * 4, s));
  PBRT_HIP(hipMemcpyAsync(n->dev.ema, params, (size_t)kNrcParams * 4, hipMemcpyHostToDevice, s));
  PBRT_HIP(nrc_launch_pack(n->dev, s));
  PBRT_HIP(hipStreamSynchronize(s));
  n->steps = 0;
}

void run_gradients(pbrt_nrc* n, const uint16_t* x, const float* target, uint32_t count, hipStream_t s) {
  if (count == 0) throw Error(PBRT_ERR_INVALID_ARG, "empty training batch");
  if (!x || !target) throw Error(PBRT_ERR_INVALID_ARG, "null batch");
  check_batch(x, target, count);
  const uint32_t tiles = (count + 63) / 64;
  const uint32_t blocks = tiles < n->max_blocks ? tiles : n->max_blocks;
  PBRT_HIP(nrc_launch_train(n->dev, x, target, count, nullptr, blocks, n->partial, n->grads, s));
}

}  // namespace

extern "C" {

int pbrt_nrc_create(pbrt_ctx* ctx, uint64_t seed, pbrt_nrc** out) {
  return guarded([&] {
    if (!ctx || !out) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    PBRT_HIP(hipSetDevice(ctx->device));
    auto* n = new pbrt_nrc();
    n->ctx = ctx;
    try {
      PBRT_HIP(hipMalloc(&n->base, nrc_device_bytes()));
      PBRT_HIP(hipMemset(n->base, 0, nrc_device_bytes()));  // zero padding of every BF16 layout
      n->dev = nrc_carve(n->base);
      n->max_blocks = (uint32_t)ctx->cu_count;
      PBRT_HIP(hipMalloc(&n->partial, (size_t)n->max_blocks * kNrcPartialStride * 4));
      PBRT_HIP(hipMalloc(&n->grads, (size_t)kNrcPartialStride * 4));
      // He-uniform hidden weights U(-sqrt(6 / fan_in), +sqrt(6 / fan_in)); zero biases and a zero
      // output layer, so an untrained cache answers 0 rather than noise
      std::vector<float> p(kNrcParams, 0.f);
      for (int i = 0; i < kNrcParams; ++i) {
        int l, o, k;
        if (nrc_locate(i, &l, &o, &k) || l == kNrcLayers - 1) continue;
        const double fan_in = l == 0 ? kNrcIn : kNrcHidden;
        const double u = to_unit(hash4(seed, (uint64_t)i, 0x6e7263ull, 0));
        p[i] = (float)((2.0 * u - 1.0) * std::sqrt(6.0 / fan_in));
      }
      upload_params(n, p.data());
    } catch (...) {
      if (n->base) (void)hipFree(n->base);
      if (n->partial) (void)hipFree(n->partial);
      if (n->grads) (void)hipFree(n->grads);
      delete n;
      throw;
    }
    *out = n;
  });
}

int pbrt_nrc_destroy(pbrt_nrc* n) {
  return guarded([&] {
    if (!n) return;
    PBRT_HIP(hipSetDevice(n->ctx->device));
    PBRT_HIP(hipStreamSynchronize(n->ctx->stream));
    if (n->base) PBRT_HIP(hipFree(n->base));
    if (n->partial) PBRT_HIP(hipFree(n->partial));
    if (n->grads) PBRT_HIP(hipFree(n->grads));
    delete n;
  });
}

int pbrt_nrc_set_bounds(pbrt_nrc* n, const float lo[3], const float hi[3]) {
  return guarded([&] {
    check(n);
    if (!lo || !hi) throw Error(PBRT_ERR_INVALID_ARG, "null bounds");
    for (int a = 0; a < 3; ++a)
      if (!(hi[a] > lo[a]) || !std::isfinite(lo[a]) || !std::isfinite(hi[a]))
        throw Error(PBRT_ERR_INVALID_ARG, "bounds must be finite with hi > lo on every axis");
    std::memcpy(n->lo, lo, sizeof n->lo);
    std::memcpy(n->hi, hi, sizeof n->hi);
  });
}

int pbrt_nrc_set_optim(pbrt_nrc* n, const pbrt_nrc_optim* o) {
  return guarded([&] {
    check(n);
    if (!o || !(o->lr > 0.f) || !(o->beta1 >= 0.f && o->beta1 < 1.f) || !(o->beta2 >= 0.f && o->beta2 < 1.f) ||
        !(o->eps > 0.f) || !(o->ema >= 0.f && o->ema < 1.f))
      throw Error(PBRT_ERR_INVALID_ARG, "optimiser needs lr > 0, betas and ema in [0, 1), eps > 0");
    PBRT_HIP(hipStreamSynchronize(n->ctx->stream));
    const bool was = n->dev.ema_decay > 0.f;
    n->optim = *o;
    n->dev.ema_decay = o->ema;
    if (o->ema > 0.f && !was) {  // the average starts from the current parameters
      PBRT_HIP(hipMemcpyAsync(n->dev.ema, n->dev.master, (size_t)kNrcParams * 4, hipMemcpyDeviceToDevice,
                              n->ctx->stream));
      PBRT_HIP(nrc_launch_pack(n->dev, n->ctx->stream));
      PBRT_HIP(hipStreamSynchronize(n->ctx->stream));
    }
  });
}

int pbrt_nrc_set_params(pbrt_nrc* n, const float* params, uint32_t count) {
  return guarded([&] {
    check(n);
    if (!params || count != (uint32_t)kNrcParams) throw Error(PBRT_ERR_INVALID_ARG, "expected PBRT_NRC_PARAMS floats");
    upload_params(n, params);
  });
}

int pbrt_nrc_get_params(pbrt_nrc* n, float* params, uint32_t count) {
  return guarded([&] {
    check(n);
    if (!params || count != (uint32_t)kNrcParams) throw Error(PBRT_ERR_INVALID_ARG, "expected PBRT_NRC_PARAMS floats");
    PBRT_HIP(hipStreamSynchronize(n->ctx->stream));
    PBRT_HIP(hipMemcpy(params, n->dev.master, (size_t)kNrcParams * 4, hipMemcpyDeviceToHost));
  });
}

int pbrt_nrc_encode(pbrt_nrc* n, const float* raw, uint32_t count, uint16_t* x, void* stream) {
  return guarded([&] {
    check(n);
    if (count && (!raw || !x)) throw Error(PBRT_ERR_INVALID_ARG, "null buffer");
    check_batch(x, raw, count);
    PBRT_HIP(nrc_launch_encode(raw, count, nullptr, n->lo, n->hi, x, stream_of(n, stream)));
  });
}

int pbrt_nrc_infer(pbrt_nrc* n, const uint16_t* x, uint32_t count, float* out, void* stream) {
  return guarded([&] {
    check(n);
    if (count && (!x || !out)) throw Error(PBRT_ERR_INVALID_ARG, "null buffer");
    check_batch(x, out, count);
    PBRT_HIP(nrc_launch_infer(n->dev.infer(), x, out, count, nullptr, (uint32_t)n->ctx->cu_count, stream_of(n, stream)));
  });
}

int pbrt_nrc_query(pbrt_nrc* n, const float* raw, uint32_t count, float* out, void* stream) {
  return guarded([&] {
    check(n);
    if (count && (!raw || !out)) throw Error(PBRT_ERR_INVALID_ARG, "null buffer");
    check_batch(raw, out, count);
    PBRT_HIP(nrc_launch_query(n->dev.infer(), raw, n->lo, n->hi, out, count, nullptr, (uint32_t)n->ctx->cu_count,
                              stream_of(n, stream)));
  });
}

int pbrt_nrc_gradients(pbrt_nrc* n, const uint16_t* x, const float* target, uint32_t count, float* grads,
                       float* loss) {
  return guarded([&] {
    check(n);
    hipStream_t s = n->ctx->stream;
    run_gradients(n, x, target, count, s);
    std::vector<float> padded(kNrcPartialStride);
    PBRT_HIP(hipMemcpyAsync(padded.data(), n->grads, padded.size() * 4, hipMemcpyDeviceToHost, s));
    PBRT_HIP(hipStreamSynchronize(s));
    if (grads)
      for (int i = 0; i < kNrcParams; ++i) grads[i] = padded[nrc_padded_index(i)];
    if (loss) *loss = padded[kNrcOffLoss];
  });
}

int pbrt_nrc_train_step(pbrt_nrc* n, const uint16_t* x, const float* target, uint32_t count, float* loss_dev,
                        void* stream) {
  return guarded([&] {
    check(n);
    hipStream_t s = stream_of(n, stream);
    if (count == 0) throw Error(PBRT_ERR_INVALID_ARG, "empty training batch");
    if (!x || !target) throw Error(PBRT_ERR_INVALID_ARG, "null batch");
    check_batch(x, target, count);
    PBRT_HIP(nrc_train_device(n, x, target, count, nullptr, s));
    if (loss_dev) PBRT_HIP(hipMemcpyAsync(loss_dev, n->grads + kNrcOffLoss, 4, hipMemcpyDeviceToDevice, s));
  });
}

int pbrt_nrc_step_count(pbrt_nrc* n, uint64_t* steps) {
  return guarded([&] {
    check(n);
    if (!steps) throw Error(PBRT_ERR_INVALID_ARG, "null output");
    *steps = n->steps;
  });
}

int pbrt_nrc_mfma_check(pbrt_ctx* ctx, const uint16_t* a, const uint16_t* b, float* d) {
  return guarded([&] {
    if (!ctx || !a || !b || !d) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    PBRT_HIP(hipSetDevice(ctx->device));
    void* buf = nullptr;
    PBRT_HIP(hipMalloc(&buf, 2048 + 1024 + 1024));
    uint16_t* da = (uint16_t*)buf;
    uint16_t* db = da + 1024;
    float* dd = (float*)(db + 512);
    hipError_t e = hipMemcpy(da, a, 1024, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, 1024, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = nrc_launch_probe(da, db, dd, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(d, dd, 1024, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    PBRT_HIP(e);
  });
}

}  // extern "C"
