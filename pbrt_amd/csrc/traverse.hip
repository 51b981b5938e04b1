// Closest-hit and any-hit traversal of a flattened mesh BVH on CDNA4 (gfx950): the launches of the
// timed kernels and of the reference's instrumented walk (kernels in traverse_kernels.hpp), the
// wave-packet walk for coherent rays, and a plain copy kernel.
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <cfloat>

#include "common.hpp"
#include "device_common.hpp"
#include "traverse.hpp"
#include "traverse_kernels.hpp"

namespace pbrt {
namespace {
using namespace dev;

// Packet traversal for coherent rays: the 64 lanes of a wave take 64 consecutive rays and walk the
// binary tree together on one wave-wide stack. The node and triangle addresses are the same in
// every lane, so they are fetched once for the wave (scalar loads); each lane tests only the boxes
// and triangles its own ray reaches: a lane takes part in a subtree when its ray hits the subtree's
// box (within its own nearest distance so far) and, for a pending subtree, when the entry distance
// it recorded is not beyond its nearest hit — the per-ray rule of k_closest, so the hits are the
// same. When both children are hit, the wave goes first where most of the lanes that hit both are
// nearer. One wave per workgroup; the stack holds kPacketDepth pending subtrees (the tree's depth
// is checked on the host).
constexpr int kPacketDepth = (int)kPacketStackDepth;

__global__ __launch_bounds__(64) void k_closest_packet(const DevNode* __restrict__ nodes,
                                                       const DevTri* __restrict__ tris, RootDesc root,
                                                       const float4* __restrict__ rays, float4* __restrict__ hits,
                                                       uint32_t n, int* __restrict__ status) {
  __shared__ float s_entry[kPacketDepth][64];
  __shared__ int32_t s_ref[kPacketDepth];
  const uint32_t lane = threadIdx.x;
  const uint32_t npk = (n + 63) / 64;
  for (uint32_t pk = blockIdx.x; pk < npk; pk += gridDim.x) {
    const uint32_t i = pk * 64 + lane;
    const bool valid = i < n;
    RayState r{};
    Closest s{};
    s.min_t = FLT_MAX;
    s.best = 0xffffffffu;
    bool act = false;
    if (valid) {
      r = make_ray(rays[2 * (size_t)i], rays[2 * (size_t)i + 1]);
      guard_once(r, root.lo, root.hi);
      s.cur_far = r.far;
      float entry;
      act = box_test(r, root.lo[0], root.hi[0], root.lo[1], root.hi[1], root.lo[2], root.hi[2], r.far, &entry);
    }
    int32_t cur = root.ref;
    int sp = 0;
    bool go = __ballot(act) != 0;
    while (go) {
      if (cur >= 0) {
        const DevNode* nd = nodes + cur;
        const float4 a = *reinterpret_cast<const float4*>(nd->xy0);
        const float4 b = *reinterpret_cast<const float4*>(nd->xy1);
        const float4 z = *reinterpret_cast<const float4*>(nd->z01);
        const int2 c = *reinterpret_cast<const int2*>(nd->ref);
        float dl = 0.f, dr = 0.f;
        const bool hl = act && box_test(r, a.x, a.y, a.z, a.w, z.x, z.y, s.cur_far, &dl);
        const bool hr = act && box_test(r, b.x, b.y, b.z, b.w, z.z, z.w, s.cur_far, &dr);
        const uint64_t ml = __ballot(hl), mr = __ballot(hr);
        if (ml && mr) {
          const uint64_t both = ml & mr, right_nearer = __ballot(hl && hr && dr < dl);
          const bool right_first = 2 * __popcll(right_nearer) > __popcll(both);
          if (sp >= kPacketDepth) {  // unreachable: the host checks the tree's depth
            if (lane == 0) atomicOr(status, 1);
            break;
          }
          s_entry[sp][lane] = (right_first ? hl : hr) ? (right_first ? dl : dr) : INFINITY;
          if (lane == 0) s_ref[sp] = right_first ? c.x : c.y;
          ++sp;
          cur = right_first ? c.y : c.x;
          act = right_first ? hr : hl;
          continue;
        }
        if (ml) {
          cur = c.x;
          act = hl;
          continue;
        }
        if (mr) {
          cur = c.y;
          act = hr;
          continue;
        }
      } else {
        const uint32_t code = ~(uint32_t)cur;
        const uint32_t first = code >> kLeafCountBits, end = first + (code & (kMaxLeafTris - 1)) + 1;
        for (uint32_t k = first; k < end; ++k) {
          float t, u, v;
          uint32_t id;
          if (act && tri_test(r, tris, k, &t, &u, &v, &id)) {
            if (t >= r.near && t <= r.far && (t < s.min_t || (t == s.min_t && id < s.best))) {
              s.min_t = t; s.t = t; s.u = u; s.v = v; s.best = id;
              s.cur_far = fminf(r.far, s.min_t);
            }
          }
        }
      }
      // the next pending subtree that a lane still needs
      go = false;
      while (sp > 0) {
        --sp;
        const bool need = valid && s_entry[sp][lane] <= s.min_t;
        if (__ballot(need)) {
          cur = __builtin_amdgcn_readfirstlane(s_ref[sp]);
          act = need;
          go = true;
          break;
        }
      }
    }
    if (valid) {
      const bool found = s.best != 0xffffffffu;
      hits[i] = make_float4(__int_as_float(found ? (int32_t)s.best : -1), found ? s.t : 0.f, found ? s.u : 0.f,
                            found ? s.v : 0.f);
    }
  }
}

template <int BLOCK, int DEPTH>
void launch_shape(const LaunchArgs& a, uint32_t* grid) {
  const uint32_t per_block_waves = BLOCK / 64;
  const uint32_t g = a.cu_count * (a.waves_per_cu / per_block_waves > 0 ? a.waves_per_cu / per_block_waves : 1);
  const uint32_t need = (a.n + BLOCK - 1) / BLOCK;
  *grid = g < need ? g : need;
  if (a.max_threads && *grid > a.max_threads / BLOCK) *grid = a.max_threads / BLOCK;
  if (*grid == 0) *grid = 1;
}

template <int BLOCK, int DEPTH>
hipError_t launch_t(const LaunchArgs& a, int kind) {
  uint32_t grid;
  launch_shape<BLOCK, DEPTH>(a, &grid);
  const uint32_t spill_stride = grid * BLOCK;
  const uint32_t spill_cap = a.spill_entries_per_thread;
  if ((uint64_t)spill_cap * spill_stride > a.spill_capacity_entries) return hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)a.stream;
  const bool wave = a.refill > 0;
  if (a.width == 4 && kind != 1) {
    RootDesc root4 = a.root;
    root4.ref = a.root4;
    if (kind == 0 && !wave)
      hipLaunchKernelGGL((k_closest4<BLOCK, DEPTH, false, 0>), dim3(grid), dim3(BLOCK), 0, s, a.nodes4, a.tris, root4,
                         (const float4*)a.rays, (float4*)a.hits, a.n, a.spill, spill_cap, a.status, 0u, a.tile_w, nullptr);
    else if (kind == 0)
      hipLaunchKernelGGL((k_closest4<BLOCK, DEPTH, true, 0>), dim3(grid), dim3(BLOCK), 0, s, a.nodes4, a.tris, root4,
                         (const float4*)a.rays, (float4*)a.hits, a.n, a.spill, spill_cap, a.status, a.refill, a.tile_w,
                         nullptr);
    else if (!wave)
      hipLaunchKernelGGL((k_any4<BLOCK, DEPTH, false, 0>), dim3(grid), dim3(BLOCK), 0, s, a.nodes4, a.tris, root4,
                         (const float4*)a.rays, a.occluded, a.n, a.spill, spill_cap, a.status, 0u, a.tile_w, nullptr);
    else
      hipLaunchKernelGGL((k_any4<BLOCK, DEPTH, true, 0>), dim3(grid), dim3(BLOCK), 0, s, a.nodes4, a.tris, root4,
                         (const float4*)a.rays, a.occluded, a.n, a.spill, spill_cap, a.status, a.refill, a.tile_w,
                         nullptr);
    return hipGetLastError();
  }
  if (kind == 0 && !wave)
    hipLaunchKernelGGL((k_closest<BLOCK, DEPTH, false, 0>), dim3(grid), dim3(BLOCK), 0, s, a.nodes, a.tris,
                       a.root, (const float4*)a.rays, (float4*)a.hits, a.n, nullptr, a.spill, spill_cap, a.status, 0u, a.tile_w);
  else if (kind == 0)
    hipLaunchKernelGGL((k_closest<BLOCK, DEPTH, true, 0>), dim3(grid), dim3(BLOCK), 0, s, a.nodes, a.tris,
                       a.root, (const float4*)a.rays, (float4*)a.hits, a.n, nullptr, a.spill, spill_cap, a.status,
                       a.refill, a.tile_w);
  else if (kind == 1 && !wave)
    hipLaunchKernelGGL((k_closest<BLOCK, DEPTH, false, 1>), dim3(grid), dim3(BLOCK), 0, s, a.nodes, a.tris,
                       a.root, (const float4*)a.rays, (float4*)a.hits, a.n, a.stats, a.spill, spill_cap, a.status, 0u, a.tile_w);
  else if (kind == 1)
    hipLaunchKernelGGL((k_closest<BLOCK, DEPTH, true, 1>), dim3(grid), dim3(BLOCK), 0, s, a.nodes, a.tris,
                       a.root, (const float4*)a.rays, (float4*)a.hits, a.n, a.stats, a.spill, spill_cap, a.status,
                       a.refill, a.tile_w);
  else if (!wave)
    hipLaunchKernelGGL((k_any<BLOCK, DEPTH, false, 0>), dim3(grid), dim3(BLOCK), 0, s, a.nodes, a.tris, a.root,
                       (const float4*)a.rays, a.occluded, a.n, a.spill, spill_cap, a.status, 0u, a.tile_w, nullptr);
  else
    hipLaunchKernelGGL((k_any<BLOCK, DEPTH, true, 0>), dim3(grid), dim3(BLOCK), 0, s, a.nodes, a.tris, a.root,
                       (const float4*)a.rays, a.occluded, a.n, a.spill, spill_cap, a.status, a.refill, a.tile_w,
                       nullptr);
  return hipGetLastError();
}

template <int BLOCK>
hipError_t launch_b(const LaunchArgs& a, int kind) {
  switch (a.lds_stack) {
    case 8: return launch_t<BLOCK, 8>(a, kind);
    case 12: return launch_t<BLOCK, 12>(a, kind);
    case 16: return launch_t<BLOCK, 16>(a, kind);
    case 24: return launch_t<BLOCK, 24>(a, kind);
    case 32: return launch_t<BLOCK, 32>(a, kind);
    default: return hipErrorInvalidValue;
  }
}

// Plain streaming copy (16-byte words, grid-stride): the known-size reference against which
// hardware byte counters are calibrated.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, uint64_t n16) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * 256ull) dst[i] = src[i];
}

}  // namespace

hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, void* stream) {
  const uint64_t n16 = bytes / 16;
  if (n16 == 0) return hipSuccess;
  uint64_t blocks = (n16 + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_copy, dim3((uint32_t)blocks), dim3(256), 0, (hipStream_t)stream, (const float4*)src,
                     (float4*)dst, n16);
  return hipGetLastError();
}

uint32_t traversal_grid_threads(const LaunchArgs& a) {
  uint32_t grid = 0;
  switch (a.block) {
    case 64: launch_shape<64, 8>(a, &grid); return grid * 64;
    case 128: launch_shape<128, 8>(a, &grid); return grid * 128;
    case 256: launch_shape<256, 8>(a, &grid); return grid * 256;
    default: return 0;
  }
}

hipError_t launch_traversal(const LaunchArgs& a, int kind) {
  if (a.packet && kind == 0) {
    const uint32_t per_cu = a.waves_per_cu ? a.waves_per_cu : 16;
    uint32_t grid = a.cu_count * per_cu;
    const uint32_t packets = (a.n + 63) / 64;
    if (grid > packets) grid = packets;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(k_closest_packet, dim3(grid), dim3(64), 0, (hipStream_t)a.stream, a.nodes, a.tris, a.root,
                       (const float4*)a.rays, (float4*)a.hits, a.n, a.status);
    return hipGetLastError();
  }
  switch (a.block) {
    case 64: return launch_b<64>(a, kind);
    case 128: return launch_b<128>(a, kind);
    case 256: return launch_b<256>(a, kind);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace pbrt
