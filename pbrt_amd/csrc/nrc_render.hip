// The radiance cache inside the wavefront renderer: turning training-path records into (input,
// target) samples, choosing a frame's training batch, and the FP64 convergence measure against a
// reference film. (The cache's answers reach the paths that ended into it in k_nrc_query.)
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include "nrc_render.hpp"

namespace pbrt {
namespace {

// Slot (rank, v) for vertices v = 1..V of every training path: the radiance the path gathered over
// the window the cache stands for at a query vertex — next-event terms at vertices v .. v+W-1 and
// emission at v+1 .. v+W, summed in the order the path added them — divided by the throughput on
// arrival at v. Valid when v was recorded as eligible and the throughput is positive.
__global__ void k_nrc_targets(RenderParams p, PathBuffers b, NrcTrainBuffers t, uint32_t ranks) {
  const uint32_t V = p.query_vertex, W = p.max_depth - p.query_vertex, D = p.max_depth;
  const size_t cap = b.train_cap, fs = (size_t)D * cap;
  const uint32_t slots = ranks * V;
  const uint64_t npass = (uint64_t)p.width * p.height * p.pass_samples;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < slots; s += gridDim.x * blockDim.x) {
    const uint32_t r = s / V, v = s % V + 1;
    // rank r is path r * stride - offset of this pass; ranks outside the pass hold no record
    const uint64_t jr = (uint64_t)r * p.train_stride;
    const bool in_pass = jr >= p.train_offset && jr - p.train_offset < npass;
    const float4 beta = in_pass ? b.rec[kRecBeta * fs + (size_t)(v - 1) * cap + r] : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (beta.w == 3.0f && beta.x > 0.f && beta.y > 0.f && beta.z > 0.f) {
      float sx = 0.f, sy = 0.f, sz = 0.f;
      for (uint32_t m = v; m < v + W && m <= D; ++m) {
        const float4 ne = b.rec[kRecNee * fs + (size_t)(m - 1) * cap + r];
        sx = sx + ne.x;
        sy = sy + ne.y;
        sz = sz + ne.z;
        if (m + 1 <= D) {
          const float4 em = b.rec[kRecEmit * fs + (size_t)m * cap + r];
          sx = sx + em.x;
          sy = sy + em.y;
          sz = sz + em.z;
        }
      }
      const float tx = sx / beta.x, ty = sy / beta.y, tz = sz / beta.z;
      if (isfinite(tx) && isfinite(ty) && isfinite(tz)) out = make_float4(tx, ty, tz, 1.f);
    }
    t.slot_t[s] = out;
    for (int f = 0; f < 3; ++f) t.slot_raw[3 * (size_t)s + f] = b.rec[(kRecRaw + f) * fs + (size_t)(v - 1) * cap + r];
  }
}

// Valid slots in slot order, deterministic: per-workgroup counts, one workgroup scanning them,
// then each workgroup writing its valid slots from its offset in lane order (wave ballots).
constexpr int kCompactBlock = 256;
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kCompactBlock) void k_compact_count(NrcTrainBuffers t, uint32_t slots,
                                                                 uint32_t* __restrict__ block_count) {
  __shared__ uint32_t wsum[kCompactBlock / 64];
  const uint32_t s = blockIdx.x * kCompactBlock + threadIdx.x;
  const bool v = s < slots && t.slot_t[s].w != 0.f;
  const uint64_t m = __ballot(v);
  if (__lane_id() == 0) wsum[threadIdx.x / 64] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t c = 0;
    for (int w = 0; w < kCompactBlock / 64; ++w) c += wsum[w];
    block_count[blockIdx.x] = c;
  }
}

// exclusive scan of the workgroup counts in place; the kCntValid counter = total
__global__ __launch_bounds__(kScanThreads) void k_compact_scan(PathBuffers b, uint32_t* __restrict__ block_count,
                                                               uint32_t nblocks) {
  __shared__ uint32_t part[kScanThreads];
  const uint32_t tid = threadIdx.x;
  const uint32_t chunk = (nblocks + kScanThreads - 1) / kScanThreads;
  const uint32_t lo = min(nblocks, tid * chunk), hi = min(nblocks, lo + chunk);
  uint32_t c = 0;
  for (uint32_t k = lo; k < hi; ++k) c += block_count[k];
  part[tid] = c;
  __syncthreads();
  for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
    const uint32_t add = tid >= off ? part[tid - off] : 0u;
    __syncthreads();
    part[tid] += add;
    __syncthreads();
  }
  uint32_t run = part[tid] - c;
  for (uint32_t k = lo; k < hi; ++k) {
    const uint32_t x = block_count[k];
    block_count[k] = run;
    run += x;
  }
  if (tid == kScanThreads - 1) *counter(b.counts, kCntValid) = part[tid];
}

__global__ __launch_bounds__(kCompactBlock) void k_compact_scatter(NrcTrainBuffers t, uint32_t slots,
                                                                   const uint32_t* __restrict__ block_off) {
  __shared__ uint32_t wsum[kCompactBlock / 64];
  const uint32_t s = blockIdx.x * kCompactBlock + threadIdx.x;
  const bool v = s < slots && t.slot_t[s].w != 0.f;
  const uint64_t m = __ballot(v);
  const uint32_t lane = __lane_id(), w = threadIdx.x / 64;
  if (lane == 0) wsum[w] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t off = block_off[blockIdx.x];
  for (uint32_t k = 0; k < w; ++k) off += wsum[k];
  if (v) t.idx[off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = s;
}

// The pass's training batches, one per frame of the pass: b = min(batch, count / frames) samples
// for each of the `frames` optimisation steps, spread evenly over the slot order (which runs frame
// by frame), step k's samples at rows [k * batch, k * batch + b). the kCntBatch counter = b.
__global__ void k_nrc_gather(PathBuffers b, NrcTrainBuffers t, uint32_t batch, uint32_t frames) {
  const uint32_t total = *counter(b.counts, kCntValid);
  const uint32_t bsz = min(batch, total / frames);
  const uint32_t n = bsz * frames;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t s = t.idx[(uint32_t)(((uint64_t)i * total) / n)];
    const size_t row = (size_t)(i / bsz) * batch + i % bsz;
    for (int f = 0; f < 3; ++f) t.batch_raw[3 * row + f] = t.slot_raw[3 * (size_t)s + f];
    const float4 tt = t.slot_t[s];
    t.batch_t[3 * row] = tt.x;
    t.batch_t[3 * row + 1] = tt.y;
    t.batch_t[3 * row + 2] = tt.z;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *counter(b.counts, kCntBatch) = bsz;
}

// Sum over pixels and channels of (film / k - ref)^2 in FP64: per-workgroup partials in a fixed
// tree order, then one workgroup adds the partials in order and writes the RMSE.
constexpr int kRmseThreads = 256;
__global__ __launch_bounds__(kRmseThreads) void k_rmse_partial(const double* __restrict__ film, double inv_k,
                                                               const double* __restrict__ ref, uint32_t n,
                                                               double* __restrict__ partial) {
  __shared__ double sh[kRmseThreads];
  double acc = 0.0;
  for (uint32_t i = blockIdx.x * kRmseThreads + threadIdx.x; i < n; i += gridDim.x * kRmseThreads) {
    const double d = film[i] * inv_k - ref[i];
    acc += d * d;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kRmseThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_rmse_final(const double* __restrict__ partial, uint32_t blocks, uint32_t n, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (uint32_t i = 0; i < blocks; ++i) s += partial[i];
    *out = sqrt(s / (double)n);
  }
}

uint32_t grid_for(uint32_t n, uint32_t block, uint32_t cap) {
  uint32_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return g > cap ? cap : g;
}

}  // namespace

hipError_t launch_nrc_targets(const RenderParams& p, const PathBuffers& b, const NrcTrainBuffers& t, uint32_t ranks,
                              void* s) {
  const uint32_t slots = ranks * p.query_vertex;
  hipLaunchKernelGGL(k_nrc_targets, dim3(grid_for(slots, 256, 4096)), dim3(256), 0, (hipStream_t)s, p, b, t, ranks);
  return hipGetLastError();
}

hipError_t launch_nrc_select(const PathBuffers& b, const NrcTrainBuffers& t, uint32_t slots, uint32_t batch,
                             uint32_t frames, void* s) {
  const uint32_t nb = (slots + kCompactBlock - 1) / kCompactBlock;
  if (nb == 0 || nb > t.block_cap) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(kCompactBlock), 0, (hipStream_t)s, t, slots, t.block_count);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(kScanThreads), 0, (hipStream_t)s, b, t.block_count, nb);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_compact_scatter, dim3(nb), dim3(kCompactBlock), 0, (hipStream_t)s, t, slots, t.block_count);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_nrc_gather, dim3(grid_for(batch * frames, 256, 2048)), dim3(256), 0, (hipStream_t)s, b, t,
                     batch, frames);
  return hipGetLastError();
}

hipError_t launch_rmse(const double* film, double inv_k, const double* ref, uint32_t n, double* partial,
                       uint32_t partial_cap, double* out, void* s) {
  const uint32_t blocks = grid_for(n, kRmseThreads, partial_cap);
  hipLaunchKernelGGL(k_rmse_partial, dim3(blocks), dim3(kRmseThreads), 0, (hipStream_t)s, film, inv_k, ref, n, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_rmse_final, dim3(1), dim3(64), 0, (hipStream_t)s, partial, blocks, n, out);
  return hipGetLastError();
}

}  // namespace pbrt
