// Mesh BVH built on the device: the host builder's tree (bvh_build.cpp), flattened into the same
// node and triangle arrays, from vertices and indices already in device memory.
#pragma once
#include "common.hpp"

namespace pbrt {

struct GpuBuilt {
  DevNode* nodes = nullptr;  // device, n_nodes entries (hipMalloc'd; the caller frees)
  DevTri* tris = nullptr;    // device, nt entries
  uint64_t n_nodes = 0;
  RootDesc root{};
  pbrt_build_stats stats{};
};

// Builds on `stream` and returns when the arrays are complete. Throws Error on invalid input, on
// a leaf of more than 32 triangles (the device layout's bound, as flatten_mesh_bvh), and on a node
// of more than 1024 triangles that share one centroid (split by index on the host; not here).
void gpu_build_mesh_bvh(hipStream_t stream, const float* xyz_dev, uint32_t nv, const uint32_t* idx_dev, uint32_t nt,
                        const pbrt_bvh_params& p, GpuBuilt* out);

}  // namespace pbrt
