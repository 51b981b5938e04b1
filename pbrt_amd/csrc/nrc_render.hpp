// Radiance-cache stages of the wavefront renderer (nrc_render.hip).
#pragma once
#include "scene.hpp"

namespace pbrt {

// Per-pass training buffers: one slot per (training path, vertex 1..query_vertex).
struct NrcTrainBuffers {
  float4* slot_raw;   // [slots][3] raw cache inputs
  float4* slot_t;     // [slots] target rgb, w = 1 when valid
  uint32_t* idx;      // [slots] valid slots in slot order
  float4* batch_raw;  // [frames * batch][3]
  float* batch_t;     // [frames * batch][3]
  uint32_t* block_count;  // [block_cap] compaction counts / offsets (256 slots per workgroup)
  uint32_t block_cap;
};

hipError_t launch_nrc_targets(const RenderParams& p, const PathBuffers& b, const NrcTrainBuffers& t, uint32_t ranks,
                              void* stream);
// Compacts the valid slots and gathers one batch of at most `batch` samples per frame of the pass.
hipError_t launch_nrc_select(const PathBuffers& b, const NrcTrainBuffers& t, uint32_t slots, uint32_t batch,
                             uint32_t frames, void* stream);
// RMSE of film / k against ref over n values, in FP64, written to *out (device).
hipError_t launch_rmse(const double* film, double inv_k, const double* ref, uint32_t n, double* partial,
                       uint32_t partial_cap, double* out, void* stream);

}  // namespace pbrt
