// The traversal walks that count the records they fetch (COUNT == 2 of traverse_kernels.hpp): per
// ray, the node records (64 bytes on the binary tree, 128 bytes on the 4-wide tree) and the
// triangle records (48 bytes) that the timed kernels' walk loads, for the algorithmic bytes of a
// launch. One launch shape per walk (128 threads, 16 LDS stack entries, per-lane refill): a ray's
// visiting order, hence its counts and its result, does not depend on the shape.
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "traverse.hpp"
#include "traverse_kernels.hpp"

namespace pbrt {

hipError_t launch_traversal_work(const LaunchArgs& a, int any_hit, uint32_t* work) {
  constexpr int kBlock = kWorkBlock, kDepth = kWorkLdsStack;
  if (a.block != (uint32_t)kBlock || a.lds_stack != (uint32_t)kDepth) return hipErrorInvalidValue;
  const uint32_t per_block_waves = kBlock / 64;
  const uint32_t g = a.cu_count * (a.waves_per_cu / per_block_waves > 0 ? a.waves_per_cu / per_block_waves : 1);
  const uint32_t need = (a.n + kBlock - 1) / kBlock;
  uint32_t grid = g < need ? g : need;
  if (grid == 0) grid = 1;
  const uint32_t spill_cap = a.spill_entries_per_thread;
  if ((uint64_t)spill_cap * grid * kBlock > a.spill_capacity_entries) return hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)a.stream;
  if (a.width == 4) {
    RootDesc root4 = a.root;
    root4.ref = a.root4;
    if (any_hit)
      hipLaunchKernelGGL((k_any4<kBlock, kDepth, false, 2>), dim3(grid), dim3(kBlock), 0, s, a.nodes4, a.tris, root4,
                         (const float4*)a.rays, a.occluded, a.n, a.spill, spill_cap, a.status, 0u, a.tile_w, work);
    else
      hipLaunchKernelGGL((k_closest4<kBlock, kDepth, false, 2>), dim3(grid), dim3(kBlock), 0, s, a.nodes4, a.tris,
                         root4, (const float4*)a.rays, (float4*)a.hits, a.n, a.spill, spill_cap, a.status, 0u,
                         a.tile_w, work);
    return hipGetLastError();
  }
  if (any_hit)
    hipLaunchKernelGGL((k_any<kBlock, kDepth, false, 2>), dim3(grid), dim3(kBlock), 0, s, a.nodes, a.tris, a.root,
                       (const float4*)a.rays, a.occluded, a.n, a.spill, spill_cap, a.status, 0u, a.tile_w, work);
  else
    hipLaunchKernelGGL((k_closest<kBlock, kDepth, false, 2>), dim3(grid), dim3(kBlock), 0, s, a.nodes, a.tris, a.root,
                       (const float4*)a.rays, (float4*)a.hits, a.n, work, a.spill, spill_cap, a.status, 0u, a.tile_w);
  return hipGetLastError();
}

}  // namespace pbrt
