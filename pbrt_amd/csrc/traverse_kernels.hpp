// Traversal kernels of a flattened mesh BVH on CDNA4 (gfx950), shared by the translation units
// that instantiate them (traverse.hip: the timed kernels and the reference's instrumented walk;
// traverse_work.hip: the walks that count fetched records).
//
// One lane per ray, a persistent grid that strides over the ray batch. Each interior node is one
// 64-byte record holding both children's boxes (or a 128-byte record of four); the traversal stack
// lives in LDS (a short stack of LDS_DEPTH entries per lane, interleaved so that a wave's 64 lanes
// hit 64 consecutive 8-byte words) and continues into an HBM spill area past that depth.
//
// Arithmetic follows the reference statement by statement, in FP32, without contraction (the
// library is built with -ffp-contract=off) and with IEEE division:
//   slab test     src/geom/aabound.rs:323-355
//   Möller-Trumbore src/shapes/triangle_mesh/triangle_mesh.rs:183-217
//   closest hit   src/shapes/triangle_mesh/bvh.rs:450-526 (ordered, nearer child first; the right
//                 child leads only when both are hit and it is strictly nearer)
//   any hit       src/shapes/triangle_mesh/bvh.rs:381-420
// Ties are resolved by the order-independent rule (smallest t, then smallest triangle index): a
// pending node is dropped only when its entry distance is strictly beyond the nearest hit.
#pragma once
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <cfloat>

#include "common.hpp"
#include "device_common.hpp"

namespace pbrt {
namespace {
using namespace dev;

__device__ __forceinline__ void write_work(uint32_t* __restrict__ work, uint32_t ri, uint32_t nodes, uint32_t tris) {
  *reinterpret_cast<uint2*>(work + 2 * (size_t)ri) = make_uint2(nodes, tris);
}

// Per-ray state of the closest-hit traversal.
struct Closest {
  float min_t, t, u, v, cur_far;
  uint32_t best, n_vis, n_box, n_tri;
};

// Flat traversal loop: each iteration advances every live lane by one interior node, or by one
// triangle of the leaf it is in (a leaf's triangles take one iteration each, so the wave never
// runs a whole leaf loop for a few lanes). A lane whose ray is finished writes its result and
// takes its next ray (i, i + stride, i + 2*stride, ... per lane) instead of waiting for the rest of
// its wave: at once (WAVE_REFILL false), or (WAVE_REFILL true) together with the other waiting
// lanes once `refill` of them wait or no lane of the wave is still tracing. The two forms compile
// to different schedules; which is faster depends on the rays' coherence, and the autotuner picks.
// The order of node visits and triangle tests within a ray is the reference's, so results and work
// counters are the same either way.
//
// COUNT selects the per-ray counters written to `stats` (and, in the other kernels, `work`):
//   0  none (the timed kernels);
//   1  BVHTree::query_instrumented's (triangle_mesh/bvh.rs:18-47,354-361): nodes visited, box tests,
//      triangle tests (3 words per ray; binary tree only);
//   2  the records the walk fetches from HBM: node records (64 bytes binary, 128 bytes 4-wide) and
//      triangle records (48 bytes) (2 words per ray). The visiting order of a ray does not depend on
//      the launch shape, the refill policy or the ray order, so neither do these counts.
template <int BLOCK, int DEPTH, bool WAVE_REFILL, int COUNT>
__global__ __launch_bounds__(BLOCK) void k_closest(const DevNode* __restrict__ nodes, const DevTri* __restrict__ tris,
                                                   RootDesc root, const float4* __restrict__ rays,
                                                   float4* __restrict__ hits, uint32_t n, uint32_t* __restrict__ stats,
                                                   uint2* __restrict__ spill, uint32_t spill_cap,
                                                   int* __restrict__ status, uint32_t refill, uint32_t tile_w) {
  __shared__ uint2 lds[DEPTH][BLOCK];
  const __amdgpu_buffer_rsrc_t trs = tri_rsrc(tris);
  Stack<BLOCK, DEPTH> st{lds, spill, gridDim.x * BLOCK, spill_cap, threadIdx.x, xcd_block() * BLOCK + threadIdx.x, 0};
  const uint32_t stride = gridDim.x * BLOCK;
  uint32_t i = st.gtid;
  if (i >= n) return;
  RayState r;
  uint32_t ri = 0;  // the batch index of the ray in progress
  Closest s;
  int32_t cur = 0;
  uint32_t tk = 0, tend = 0;  // triangles of the current leaf still to test: [tk, tend)
  bool live = false, idle = false;
  auto start = [&](uint32_t k) {
    ri = tiled_ray(k, tile_w);
    r = make_ray(rays[2 * (size_t)ri], rays[2 * (size_t)ri + 1]);
    guard_once(r, root.lo, root.hi);
    s.min_t = FLT_MAX;
    s.t = s.u = s.v = 0.f;
    s.cur_far = r.far;  // min(far, min_t): the box tests' narrowed interval
    s.best = 0xffffffffu;
    s.n_vis = 0;
    s.n_box = 1;
    s.n_tri = 0;
    st.sp = 0;
    tk = tend = 0;
    float entry;
    live = box_test(r, root.lo[0], root.hi[0], root.lo[1], root.hi[1], root.lo[2], root.hi[2], r.far, &entry);
    cur = root.ref;
  };
  start(i);
  for (;;) {
    bool pop_now = false;
    if (live && cur >= 0) {
      // interior node: test both children over [near, min(far, min_t)]
      if (COUNT == 1) { ++s.n_vis; s.n_box += 2; }
      if (COUNT == 2) ++s.n_vis;
      const DevNode* nd = nodes + cur;
      const float4 a = *reinterpret_cast<const float4*>(nd->xy0);
      const float4 b = *reinterpret_cast<const float4*>(nd->xy1);
      const float4 z = *reinterpret_cast<const float4*>(nd->z01);
      const int2 c = *reinterpret_cast<const int2*>(nd->ref);
      float dl, dr;
      const bool hl = box_test(r, a.x, a.y, a.z, a.w, z.x, z.y, s.cur_far, &dl);
      const bool hr = box_test(r, b.x, b.y, b.z, b.w, z.z, z.w, s.cur_far, &dr);
      if (hl && hr) {
        const bool right_first = dr < dl;
        if (!st.push(right_first ? c.x : c.y, right_first ? dl : dr)) {
          atomicOr(status, 1);
          live = false;
        }
        cur = right_first ? c.y : c.x;
      } else if (hl) {
        cur = c.x;
      } else if (hr) {
        cur = c.y;
      } else {
        pop_now = true;
      }
    }
    if (live && !pop_now && cur < 0) {
      if (tk == tend) {  // a leaf reached: ~cur = (first << 5) | (count - 1)
        if (COUNT == 1) ++s.n_vis;
        const uint32_t code = ~(uint32_t)cur;
        tk = code >> kLeafCountBits;
        tend = tk + (code & (kMaxLeafTris - 1)) + 1;
      }
      if (COUNT) ++s.n_tri;
      float t, u, v;
      uint32_t id;
      if (tri_test(r, trs, tk, &t, &u, &v, &id)) {
        if (t >= r.near && t <= r.far && (t < s.min_t || (t == s.min_t && id < s.best))) {
          s.min_t = t; s.t = t; s.u = u; s.v = v; s.best = id;
          s.cur_far = fminf(r.far, s.min_t);
        }
      }
      ++tk;
      pop_now = tk == tend;
    }
    if (live && pop_now) {
      // the next pending node whose entry is not beyond the nearest hit
      live = false;
      while (st.sp > 0) {
        const uint2 e = st.pop();
        if (__uint_as_float(e.y) > s.min_t) continue;
        cur = (int32_t)e.x;
        live = true;
        break;
      }
    }
    if (!live && !idle) {
      const bool found = s.best != 0xffffffffu;
      hits[ri] = make_float4(__int_as_float(found ? (int32_t)s.best : -1), found ? s.t : 0.f, found ? s.u : 0.f,
                            found ? s.v : 0.f);
      if (COUNT == 1) {
        stats[3 * (size_t)ri] = s.n_vis;
        stats[3 * (size_t)ri + 1] = s.n_box;
        stats[3 * (size_t)ri + 2] = s.n_tri;
      }
      if (COUNT == 2) write_work(stats, ri, s.n_vis, s.n_tri);
      i += stride;
      if (i >= n) break;
      if (WAVE_REFILL)
        idle = true;
      else
        start(i);
    }
    if (WAVE_REFILL) {
      const uint64_t waiting = __ballot(idle), active = __ballot(1);
      if (idle && ((uint32_t)__popcll(waiting) >= refill || waiting == active)) {
        start(i);
        idle = false;
      }
    }
  }
}

// Any hit, unordered (bvh.rs:381-420): the reference pushes left then right, so the right child is
// examined first; the first triangle hit within [near, far] ends the ray. Same flat loop.
template <int BLOCK, int DEPTH, bool WAVE_REFILL, int COUNT>
__global__ __launch_bounds__(BLOCK) void k_any(const DevNode* __restrict__ nodes, const DevTri* __restrict__ tris,
                                               RootDesc root, const float4* __restrict__ rays,
                                               uint8_t* __restrict__ occluded, uint32_t n, uint2* __restrict__ spill,
                                               uint32_t spill_cap, int* __restrict__ status, uint32_t refill, uint32_t tile_w,
                                               uint32_t* __restrict__ work) {
  __shared__ uint2 lds[DEPTH][BLOCK];
  const __amdgpu_buffer_rsrc_t trs = tri_rsrc(tris);
  Stack<BLOCK, DEPTH> st{lds, spill, gridDim.x * BLOCK, spill_cap, threadIdx.x, xcd_block() * BLOCK + threadIdx.x, 0};
  const uint32_t stride = gridDim.x * BLOCK;
  uint32_t i = st.gtid;
  if (i >= n) return;
  RayState r;
  uint32_t ri = 0;  // the batch index of the ray in progress
  int32_t cur = 0;
  uint32_t tk = 0, tend = 0;
  bool live = false, hit = false, idle = false;
  uint32_t n_node = 0, n_tri = 0;  // COUNT == 2: node and triangle records fetched
  auto start = [&](uint32_t k) {
    ri = tiled_ray(k, tile_w);
    r = make_ray(rays[2 * (size_t)ri], rays[2 * (size_t)ri + 1]);
    guard_once(r, root.lo, root.hi);
    st.sp = 0;
    hit = false;
    n_node = n_tri = 0;
    tk = tend = 0;
    float entry;
    live = box_test(r, root.lo[0], root.hi[0], root.lo[1], root.hi[1], root.lo[2], root.hi[2], r.far, &entry);
    cur = root.ref;
  };
  start(i);
  for (;;) {
    bool pop_now = false;
    if (live && cur >= 0) {
      if (COUNT == 2) ++n_node;
      const DevNode* nd = nodes + cur;
      const float4 a = *reinterpret_cast<const float4*>(nd->xy0);
      const float4 b = *reinterpret_cast<const float4*>(nd->xy1);
      const float4 z = *reinterpret_cast<const float4*>(nd->z01);
      const int2 c = *reinterpret_cast<const int2*>(nd->ref);
      float dl, dr;
      const bool hl = box_test(r, a.x, a.y, a.z, a.w, z.x, z.y, r.far, &dl);
      const bool hr = box_test(r, b.x, b.y, b.z, b.w, z.z, z.w, r.far, &dr);
      if (hl && hr) {
        if (!st.push(c.x, 0.f)) {
          atomicOr(status, 1);
          live = false;
        }
        cur = c.y;
      } else if (hr) {
        cur = c.y;
      } else if (hl) {
        cur = c.x;
      } else {
        pop_now = true;
      }
    }
    if (live && !pop_now && cur < 0) {
      if (tk == tend) {
        const uint32_t code = ~(uint32_t)cur;
        tk = code >> kLeafCountBits;
        tend = tk + (code & (kMaxLeafTris - 1)) + 1;
      }
      if (COUNT == 2) ++n_tri;
      float t, u, v;
      uint32_t id;
      if (tri_test(r, trs, tk, &t, &u, &v, &id) && t >= r.near && t <= r.far) {
        hit = true;
        live = false;
      }
      ++tk;
      pop_now = tk == tend;
    }
    if (live && pop_now) {
      live = st.sp > 0;
      if (live) cur = (int32_t)st.pop().x;
    }
    if (!live && !idle) {
      occluded[ri] = hit ? 1 : 0;
      if (COUNT == 2) write_work(work, ri, n_node, n_tri);
      i += stride;
      if (i >= n) break;
      if (WAVE_REFILL)
        idle = true;
      else
        start(i);
    }
    if (WAVE_REFILL) {
      const uint64_t waiting = __ballot(idle), active = __ballot(1);
      if (idle && ((uint32_t)__popcll(waiting) >= refill || waiting == active)) {
        start(i);
        idle = false;
      }
    }
  }
}

// Closest hit over the 4-wide tree: the flat loop of k_closest with one 4-wide node per iteration
// (nearest child next, the others pushed farthest first). Hits are those of k_closest (the tie rule
// makes them independent of the visiting order); work counters are taken on the binary tree only.
template <int BLOCK, int DEPTH, bool WAVE_REFILL, int COUNT>
__global__ __launch_bounds__(BLOCK) void k_closest4(const DevNode4* __restrict__ nodes, const DevTri* __restrict__ tris,
                                                    RootDesc root, const float4* __restrict__ rays,
                                                    float4* __restrict__ hits, uint32_t n, uint2* __restrict__ spill,
                                                    uint32_t spill_cap, int* __restrict__ status, uint32_t refill, uint32_t tile_w,
                                                    uint32_t* __restrict__ work) {
  __shared__ uint2 lds[DEPTH][BLOCK];
  const __amdgpu_buffer_rsrc_t trs = tri_rsrc(tris);
  const __amdgpu_buffer_rsrc_t nrs = node4_rsrc(nodes);
  Stack<BLOCK, DEPTH> st{lds, spill, gridDim.x * BLOCK, spill_cap, threadIdx.x, xcd_block() * BLOCK + threadIdx.x, 0};
  const uint32_t stride = gridDim.x * BLOCK;
  uint32_t i = st.gtid;
  if (i >= n) return;
  RayState r;
  uint32_t ri = 0;  // the batch index of the ray in progress
  Closest s;
  int32_t cur = 0;
  uint32_t tk = 0, tend = 0;
  bool live = false, idle = false;
  auto start = [&](uint32_t k) {
    ri = tiled_ray(k, tile_w);
    r = make_ray(rays[2 * (size_t)ri], rays[2 * (size_t)ri + 1]);
    guard_once(r, root.lo, root.hi);
    s.min_t = FLT_MAX;
    s.t = s.u = s.v = 0.f;
    s.cur_far = r.far;
    s.best = 0xffffffffu;
    s.n_vis = s.n_tri = 0;
    st.sp = 0;
    tk = tend = 0;
    float entry;
    live = box_test(r, root.lo[0], root.hi[0], root.lo[1], root.hi[1], root.lo[2], root.hi[2], r.far, &entry);
    cur = root.ref;
  };
  start(i);
  for (;;) {
    bool pop_now = false;
    if (live && cur >= 0) {
      if (COUNT == 2) ++s.n_vis;
      int32_t r0, r1, r2, r3;
      float k0, k1, k2, k3;
      const int nh = node4_step(r, nrs, (uint32_t)cur, s.cur_far, r0, r1, r2, r3, k0, k1, k2, k3);
      if (nh == 0) {
        pop_now = true;
      } else {
        bool ok = true;
        if (nh > 3) ok = st.push(r3, k3);
        if (nh > 2) ok = ok && st.push(r2, k2);
        if (nh > 1) ok = ok && st.push(r1, k1);
        if (!ok) {
          atomicOr(status, 1);
          live = false;
        }
        cur = r0;
      }
    }
    if (live && !pop_now && cur < 0) {
      if (tk == tend) {
        const uint32_t code = ~(uint32_t)cur;
        tk = code >> kLeafCountBits;
        tend = tk + (code & (kMaxLeafTris - 1)) + 1;
      }
      if (COUNT == 2) ++s.n_tri;
      float t, u, v;
      uint32_t id;
      if (tri_test(r, trs, tk, &t, &u, &v, &id)) {
        if (t >= r.near && t <= r.far && (t < s.min_t || (t == s.min_t && id < s.best))) {
          s.min_t = t; s.t = t; s.u = u; s.v = v; s.best = id;
          s.cur_far = fminf(r.far, s.min_t);
        }
      }
      ++tk;
      pop_now = tk == tend;
    }
    if (live && pop_now) {
      live = false;
      while (st.sp > 0) {
        const uint2 e = st.pop();
        if (__uint_as_float(e.y) > s.min_t) continue;
        cur = (int32_t)e.x;
        live = true;
        break;
      }
    }
    if (!live && !idle) {
      const bool found = s.best != 0xffffffffu;
      hits[ri] = make_float4(__int_as_float(found ? (int32_t)s.best : -1), found ? s.t : 0.f, found ? s.u : 0.f,
                            found ? s.v : 0.f);
      if (COUNT == 2) write_work(work, ri, s.n_vis, s.n_tri);
      i += stride;
      if (i >= n) break;
      if (WAVE_REFILL)
        idle = true;
      else
        start(i);
    }
    if (WAVE_REFILL) {
      const uint64_t waiting = __ballot(idle), active = __ballot(1);
      if (idle && ((uint32_t)__popcll(waiting) >= refill || waiting == active)) {
        start(i);
        idle = false;
      }
    }
  }
}

// Any hit over the 4-wide tree (the result does not depend on the visiting order).
template <int BLOCK, int DEPTH, bool WAVE_REFILL, int COUNT>
__global__ __launch_bounds__(BLOCK) void k_any4(const DevNode4* __restrict__ nodes, const DevTri* __restrict__ tris,
                                                RootDesc root, const float4* __restrict__ rays,
                                                uint8_t* __restrict__ occluded, uint32_t n, uint2* __restrict__ spill,
                                                uint32_t spill_cap, int* __restrict__ status, uint32_t refill, uint32_t tile_w,
                                                uint32_t* __restrict__ work) {
  __shared__ uint2 lds[DEPTH][BLOCK];
  const __amdgpu_buffer_rsrc_t trs = tri_rsrc(tris);
  const __amdgpu_buffer_rsrc_t nrs = node4_rsrc(nodes);
  Stack<BLOCK, DEPTH> st{lds, spill, gridDim.x * BLOCK, spill_cap, threadIdx.x, xcd_block() * BLOCK + threadIdx.x, 0};
  const uint32_t stride = gridDim.x * BLOCK;
  uint32_t i = st.gtid;
  if (i >= n) return;
  RayState r;
  uint32_t ri = 0;  // the batch index of the ray in progress
  int32_t cur = 0;
  uint32_t tk = 0, tend = 0;
  bool live = false, hit = false, idle = false;
  uint32_t n_node = 0, n_tri = 0;  // COUNT == 2: node and triangle records fetched
  auto start = [&](uint32_t k) {
    ri = tiled_ray(k, tile_w);
    r = make_ray(rays[2 * (size_t)ri], rays[2 * (size_t)ri + 1]);
    guard_once(r, root.lo, root.hi);
    st.sp = 0;
    hit = false;
    n_node = n_tri = 0;
    tk = tend = 0;
    float entry;
    live = box_test(r, root.lo[0], root.hi[0], root.lo[1], root.hi[1], root.lo[2], root.hi[2], r.far, &entry);
    cur = root.ref;
  };
  start(i);
  for (;;) {
    bool pop_now = false;
    if (live && cur >= 0) {
      if (COUNT == 2) ++n_node;
      int32_t r0, r1, r2, r3;
      float k0, k1, k2, k3;
      const int nh = node4_step(r, nrs, (uint32_t)cur, r.far, r0, r1, r2, r3, k0, k1, k2, k3);
      if (nh == 0) {
        pop_now = true;
      } else {
        bool ok = true;
        if (nh > 3) ok = st.push(r3, 0.f);
        if (nh > 2) ok = ok && st.push(r2, 0.f);
        if (nh > 1) ok = ok && st.push(r1, 0.f);
        if (!ok) {
          atomicOr(status, 1);
          live = false;
        }
        cur = r0;
      }
    }
    if (live && !pop_now && cur < 0) {
      if (tk == tend) {
        const uint32_t code = ~(uint32_t)cur;
        tk = code >> kLeafCountBits;
        tend = tk + (code & (kMaxLeafTris - 1)) + 1;
      }
      if (COUNT == 2) ++n_tri;
      float t, u, v;
      uint32_t id;
      if (tri_test(r, trs, tk, &t, &u, &v, &id) && t >= r.near && t <= r.far) {
        hit = true;
        live = false;
      }
      ++tk;
      pop_now = tk == tend;
    }
    if (live && pop_now) {
      live = st.sp > 0;
      if (live) cur = (int32_t)st.pop().x;
    }
    if (!live && !idle) {
      occluded[ri] = hit ? 1 : 0;
      if (COUNT == 2) write_work(work, ri, n_node, n_tri);
      i += stride;
      if (i >= n) break;
      if (WAVE_REFILL)
        idle = true;
      else
        start(i);
    }
    if (WAVE_REFILL) {
      const uint64_t waiting = __ballot(idle), active = __ballot(1);
      if (idle && ((uint32_t)__popcll(waiting) >= refill || waiting == active)) {
        start(i);
        idle = false;
      }
    }
  }
}

}  // namespace
}  // namespace pbrt
