// Launch interface of the traversal kernels (traverse.hip).
#pragma once
#include "common.hpp"

namespace pbrt {

struct LaunchArgs {
  const DevNode* nodes;
  const DevTri* tris;
  RootDesc root;
  const pbrt_ray* rays;
  pbrt_hit* hits;
  uint8_t* occluded;
  uint32_t* stats;
  uint32_t n;
  uint2* spill;
  uint32_t spill_entries_per_thread;
  uint64_t spill_capacity_entries;
  int* status;
  void* stream;
  uint32_t block, lds_stack, waves_per_cu, cu_count;
  uint32_t refill;  // 0: eager per-lane refill; N > 0: wave refill once N lanes wait
  uint32_t width;   // 2: the binary tree; 4: the 4-wide tree (closest and any hit; counters always binary)
  const DevNode4* nodes4;
  int32_t root4;    // root reference in the 4-wide tree (the root box is root's)
  uint32_t packet;  // 1: closest hit by wave packets on the binary tree (k_closest_packet)
  uint32_t tile_w;  // 0: rays in batch order; W: an image W rays wide, taken in 8 x 16 tiles (not packets)
  uint32_t max_threads;  // 0: the grid of waves_per_cu; else the persistent grid holds at most this many threads
};
// Pending subtrees the packet kernel's wave-wide stack holds: trees deeper than this need the
// per-lane kernels.
constexpr uint32_t kPacketStackDepth = 32;

// kind: 0 closest hit, 1 closest hit with per-ray work counters, 2 any hit.
hipError_t launch_traversal(const LaunchArgs& a, int kind);
// The walk of a.width (closest hit, or any hit when any_hit != 0) counting, per ray, the node and
// triangle records it fetches (work: 2 words per ray). Runs at one launch shape: a.block must be
// kWorkBlock and a.lds_stack kWorkLdsStack (the spill area sized for them); refill is not used.
constexpr uint32_t kWorkBlock = 128, kWorkLdsStack = 16;
hipError_t launch_traversal_work(const LaunchArgs& a, int any_hit, uint32_t* work);
// Threads of the persistent grid the launch would use (the spill area is sized from it).
uint32_t traversal_grid_threads(const LaunchArgs& a);
// Device-to-device copy of bytes (multiple of 16) by a plain streaming kernel.
hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, void* stream);

}  // namespace pbrt
