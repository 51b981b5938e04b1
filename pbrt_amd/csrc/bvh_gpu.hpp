// Mesh BVH built on the device: the host builder's tree (bvh_build.cpp), flattened into the same
// node and triangle arrays, from vertices and indices already in device memory.
#pragma once
#include <vector>

#include "common.hpp"

namespace pbrt {

struct GpuBuilt {
  DevNode* nodes = nullptr;  // device, n_nodes entries (hipMalloc'd; the caller frees)
  DevTri* tris = nullptr;    // device, nt entries
  uint64_t n_nodes = 0;
  RootDesc root{};
  pbrt_build_stats stats{};
};

// A mesh of a batch: vertices [vert_base, vert_base + nv) of the shared vertex array; its nt
// triangles' vertex indices (local to the mesh) follow the previous mesh's in the shared index array.
struct GpuMeshSpan {
  uint32_t vert_base, nv, nt;
};
struct GpuMeshOut {
  uint64_t node_base = 0, n_nodes = 0, tri_base = 0;
  RootDesc root{};  // references local to the mesh
};
struct GpuBuiltMany {
  DevNode* nodes = nullptr;  // device: every mesh's nodes, mesh m's from node_base, references local
  DevTri* tris = nullptr;    // device: every mesh's triangles, mesh m's from tri_base
  uint64_t n_nodes = 0;
  std::vector<GpuMeshOut> meshes;
  pbrt_build_stats stats{};  // over the batch (max_depth: the deepest mesh)
};

// Builds every mesh of a batch in one pass (levels shared across meshes); each mesh's arrays are
// those a build of that mesh alone gives. Arrays hipMalloc'd; the caller frees them.
void gpu_build_meshes(hipStream_t stream, const float* xyz_dev, const uint32_t* idx_dev,
                      const std::vector<GpuMeshSpan>& spans, const pbrt_bvh_params& p, GpuBuiltMany* out);

// Builds on `stream` and returns when the arrays are complete. Throws Error on invalid input, on
// a leaf of more than 32 triangles (the device layout's bound, as flatten_mesh_bvh), and on a node
// of more than 256 triangles that share one centroid (split by index on the host; not here).
void gpu_build_mesh_bvh(hipStream_t stream, const float* xyz_dev, uint32_t nv, const uint32_t* idx_dev, uint32_t nt,
                        const pbrt_bvh_params& p, GpuBuilt* out);

// Collapses every mesh's binary tree into 4-wide nodes on the device, by the host collapse's rule
// (collapse_bvh4: each node takes its binary node's two children and opens the interior child of
// largest surface area until it has four). blas: the meshes' binary nodes (mesh m's from
// node_base[m], references local to the mesh, root root_ref[m]), n_binary of them in all. The 4-wide
// nodes are laid out level by level over all meshes, each node's interior children consecutive;
// interior references index the returned array (node base 0 for every mesh), leaf references are
// unchanged. root4[m] receives mesh m's root reference (a leaf root stays a leaf). *out is
// hipMalloc'd (n_out nodes); the caller frees it. Returns when the array is complete.
void gpu_collapse4(hipStream_t stream, const DevNode* blas, const std::vector<uint64_t>& node_base,
                   const std::vector<int32_t>& root_ref, uint64_t n_binary, DevNode4** out, uint64_t* n_out,
                   std::vector<int32_t>* root4);

}  // namespace pbrt
