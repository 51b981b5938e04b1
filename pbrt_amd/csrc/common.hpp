// Shared host-side definitions of the MI355X ray path: error plumbing for the C ABI and the
// device data layout of a flattened mesh BVH.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pbrt_gpu.h"

namespace pbrt {

// Error carrying a C-ABI status code; converted at the boundary by `guarded`.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define PBRT_HIP(expr)                                                                                  \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess)                                                                               \
      throw ::pbrt::Error(e_ == hipErrorOutOfMemory ? PBRT_ERR_OUT_OF_MEMORY : PBRT_ERR_DEVICE,          \
                          std::string(#expr) + ": " + hipGetErrorString(e_));                           \
  } while (0)

// Runs `fn` and turns exceptions into status codes + a thread-local message.
template <class F>
int guarded(F&& fn) {
  try {
    fn();
    return PBRT_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return PBRT_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return PBRT_ERR_INTERNAL;
  }
}

// ---- device layout ----------------------------------------------------------------------------
// An interior node of the flattened BVH2 holds the boxes of its two children (64 bytes, one load
// of four 16-byte words): child c's x/y extents in xy[c] = (lo.x, hi.x, lo.y, hi.y), both z extents
// in z01 = (lo0.z, hi0.z, lo1.z, hi1.z), child references in ref.x / ref.y. A reference >= 0 is an
// interior node index; a negative reference is a leaf: ~ref = (first_triangle << 5) | (count - 1).
struct alignas(16) DevNode {
  float xy0[4];
  float xy1[4];
  float z01[4];
  int32_t ref[4];
};
static_assert(sizeof(DevNode) == 64, "node is 64 bytes");

// A 4-wide node collapsed from the binary tree (128 bytes, one cache line): the boxes of up to four
// children as six float4 (min x, max x, min y, max y, min z, max z of children 0..3), their references
// (as DevNode's), and the number of children in count[0] (2 to 4; unused slots come last).
struct alignas(16) DevNode4 {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  int32_t ref[4];
  int32_t count[4];
};
static_assert(sizeof(DevNode4) == 128, "4-wide node is 128 bytes");

// Collapses a flattened binary tree into 4-wide nodes (depth-first, parents before children): each
// node takes its binary node's two children and repeatedly opens the interior child of largest
// surface area until it has four. Leaves keep their references. root4 receives the new root
// reference (a leaf root stays a leaf).
void collapse_bvh4(const std::vector<DevNode>& n2, int32_t root_ref, std::vector<DevNode4>* n4, int32_t* root4);

// A triangle in leaf order, pre-gathered: p0 (w = original index bits), e1 = p1 - p0, e2 = p2 - p0.
struct alignas(16) DevTri {
  float p0[4];
  float e1[4];
  float e2[4];
};
static_assert(sizeof(DevTri) == 48, "triangle is 48 bytes");

constexpr int kLeafCountBits = 5;
constexpr uint32_t kMaxLeafTris = 1u << kLeafCountBits;
constexpr uint32_t kMaxLeafFirst = (1u << (31 - kLeafCountBits)) - 1;

inline int32_t leaf_ref(uint32_t first, uint32_t count) {
  return ~(int32_t)((first << kLeafCountBits) | (count - 1));
}

// Root description passed by value to the kernels.
struct RootDesc {
  float lo[3], hi[3];
  int32_t ref;
};

// Host-side build result (binned SAH over f64 centroids, exact FP32 boxes).
struct BuildNode {
  double lo[3], hi[3];
  uint32_t left, right;  // children (interior)
  uint32_t first, count; // range of `order` (leaf)
  bool leaf;
};

struct BuiltBVH {
  std::vector<BuildNode> nodes;  // preorder, root at 0
  std::vector<uint32_t> order;   // triangle indices in leaf order
  pbrt_build_stats stats{};
};

BuiltBVH build_mesh_bvh(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt, const pbrt_bvh_params& p);

// Flattens a built tree into the device layout (nodes, triangles) and describes its root.
void flatten_mesh_bvh(const BuiltBVH& b, const float* xyz, const uint32_t* idx, std::vector<DevNode>* nodes,
                      std::vector<DevTri>* tris, RootDesc* root);

}  // namespace pbrt
