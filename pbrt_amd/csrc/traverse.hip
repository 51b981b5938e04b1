// Closest-hit and any-hit traversal kernels for a flattened mesh BVH on CDNA4 (gfx950).
//
// One lane per ray, a persistent grid that strides over the ray batch. Each interior node is one
// 64-byte record holding both children's boxes; the traversal stack lives in LDS (a short stack of
// LDS_DEPTH entries per lane, interleaved so that a wave's 64 lanes hit 64 consecutive 8-byte
// words) and continues into an HBM spill area past that depth.
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
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <cfloat>

#include "common.hpp"
#include "traverse.hpp"

namespace pbrt {
namespace {

// 2 * gamma(3) for FP32 (aabound.rs:18-39 in the FP32 analogue): u = 2^-24.
__device__ __forceinline__ float slab_widening() {
  const float u = FLT_EPSILON * 0.5f;
  const float g3 = 3.0f * u / (1.0f - 3.0f * u);
  return 2.0f * g3;
}

struct RayState {
  float ox, oy, oz, ix, iy, iz;  // origin, reciprocal direction
  float dx, dy, dz;
  float near, far;
  bool nx, ny, nz;    // reciprocal < 0: the slab's near plane is its max
  bool parallel;      // a direction component is exactly zero
};

__device__ __forceinline__ RayState make_ray(float4 a, float4 b) {
  RayState r;
  r.ox = a.x; r.oy = a.y; r.oz = a.z; r.near = a.w;
  r.dx = b.x; r.dy = b.y; r.dz = b.z;
  r.far = fminf(b.w, FLT_MAX);
  r.ix = 1.0f / b.x;
  r.iy = 1.0f / b.y;
  r.iz = 1.0f / b.z;
  r.nx = r.ix < 0.0f;
  r.ny = r.iy < 0.0f;
  r.nz = r.iz < 0.0f;
  r.parallel = (b.x == 0.0f) | (b.y == 0.0f) | (b.z == 0.0f);
  return r;
}

// Slab test for a ray with no zero direction component. Computing both slab distances through the
// plane selected by the reciprocal's sign equals the reference's compute-then-swap; taking the
// max/min over all three axes at the end equals its per-axis early exit (a rejection at an earlier
// axis implies one at the end). fmaxf/fminf ignore a NaN operand, as Rust's f64::max/min do.
__device__ __forceinline__ bool box_fast(const RayState& r, float lx, float hx, float ly, float hy, float lz, float hz,
                                         float tmax_in, float* entry) {
  const float W = slab_widening();
  const float px0 = r.nx ? hx : lx, px1 = r.nx ? lx : hx;
  const float py0 = r.ny ? hy : ly, py1 = r.ny ? ly : hy;
  const float pz0 = r.nz ? hz : lz, pz1 = r.nz ? lz : hz;
  float t0x = (px0 - r.ox) * r.ix, t1x = (px1 - r.ox) * r.ix;
  float t0y = (py0 - r.oy) * r.iy, t1y = (py1 - r.oy) * r.iy;
  float t0z = (pz0 - r.oz) * r.iz, t1z = (pz1 - r.oz) * r.iz;
  t0x = t0x - fabsf(t0x) * W; t1x = t1x + fabsf(t1x) * W;
  t0y = t0y - fabsf(t0y) * W; t1y = t1y + fabsf(t1y) * W;
  t0z = t0z - fabsf(t0z) * W; t1z = t1z + fabsf(t1z) * W;
  const float tmin = fmaxf(fmaxf(fmaxf(t0x, r.near), t0y), t0z);
  const float tmax = fminf(fminf(fminf(t1x, tmax_in), t1y), t1z);
  *entry = tmin;
  return !(tmax < tmin);
}

// The reference's statement order, including its branch for a direction component of exactly
// zero (inside the slab iff the origin is). Taken only by rays that have such a component.
__device__ __forceinline__ bool slab_literal(float o, float d, float lo, float hi, float W, float* tmin, float* tmax) {
  if (d == 0.0f) return !(o < lo || o > hi);
  const float inv = 1.0f / d;
  float t0 = (lo - o) * inv;
  float t1 = (hi - o) * inv;
  if (inv < 0.0f) {
    const float s = t0;
    t0 = t1;
    t1 = s;
  }
  t0 = t0 - fabsf(t0) * W;
  t1 = t1 + fabsf(t1) * W;
  *tmin = fmaxf(t0, *tmin);
  *tmax = fminf(t1, *tmax);
  return !(*tmax < *tmin);
}

__device__ __forceinline__ bool box_literal(const RayState& r, float lx, float hx, float ly, float hy, float lz,
                                            float hz, float tmax, float* entry) {
  const float W = slab_widening();
  float tmin = r.near;
  if (!slab_literal(r.ox, r.dx, lx, hx, W, &tmin, &tmax)) return false;
  if (!slab_literal(r.oy, r.dy, ly, hy, W, &tmin, &tmax)) return false;
  if (!slab_literal(r.oz, r.dz, lz, hz, W, &tmin, &tmax)) return false;
  *entry = tmin;
  return true;
}

__device__ __forceinline__ bool box_test(const RayState& r, float lx, float hx, float ly, float hy, float lz, float hz,
                                         float tmax, float* entry) {
  if (__builtin_expect(r.parallel, 0)) return box_literal(r, lx, hx, ly, hy, lz, hz, tmax, entry);
  return box_fast(r, lx, hx, ly, hy, lz, hz, tmax, entry);
}

// Möller–Trumbore with the edges pre-computed in FP32 on the host (the same subtraction the
// reference performs first).
__device__ __forceinline__ bool tri_test(const RayState& r, const DevTri* __restrict__ tris, uint32_t k, float* t_out,
                                         float* u_out, float* v_out, uint32_t* id_out) {
  const float4 a = *reinterpret_cast<const float4*>(tris[k].p0);
  const float4 e1 = *reinterpret_cast<const float4*>(tris[k].e1);
  const float4 e2 = *reinterpret_cast<const float4*>(tris[k].e2);
  // pv = cross(d, e2)
  const float pvx = r.dy * e2.z - r.dz * e2.y;
  const float pvy = r.dz * e2.x - r.dx * e2.z;
  const float pvz = r.dx * e2.y - r.dy * e2.x;
  const float det = e1.x * pvx + e1.y * pvy + e1.z * pvz;
  if (fabsf(det) < 1e-8f) return false;
  const float inv_det = 1.0f / det;
  const float sx = r.ox - a.x, sy = r.oy - a.y, sz = r.oz - a.z;
  const float u = inv_det * (sx * pvx + sy * pvy + sz * pvz);
  if (u < 0.0f || u > 1.0f) return false;
  // q = cross(s, e1)
  const float qx = sy * e1.z - sz * e1.y;
  const float qy = sz * e1.x - sx * e1.z;
  const float qz = sx * e1.y - sy * e1.x;
  const float v = inv_det * (r.dx * qx + r.dy * qy + r.dz * qz);
  if (v < 0.0f || u + v > 1.0f) return false;
  *t_out = inv_det * (e2.x * qx + e2.y * qy + e2.z * qz);
  *u_out = u;
  *v_out = v;
  *id_out = __float_as_uint(a.w);
  return true;
}

template <int BLOCK, int DEPTH>
struct Stack {
  uint2 (*lds)[BLOCK];
  uint2* __restrict__ spill;
  uint32_t spill_stride, spill_cap, lane, gtid;
  int sp;
  __device__ __forceinline__ bool push(int32_t ref, float d) {
    const uint2 e = make_uint2((uint32_t)ref, __float_as_uint(d));
    if (sp < DEPTH) {
      lds[sp][lane] = e;
    } else {
      const uint32_t k = (uint32_t)(sp - DEPTH);
      if (k >= spill_cap) return false;
      spill[(size_t)k * spill_stride + gtid] = e;
    }
    ++sp;
    return true;
  }
  __device__ __forceinline__ uint2 pop() {
    --sp;
    if (sp < DEPTH) return lds[sp][lane];
    return spill[(size_t)(sp - DEPTH) * spill_stride + gtid];
  }
};

// Per-ray state of the closest-hit traversal.
struct Closest {
  float min_t, t, u, v, cur_far;
  uint32_t best, n_vis, n_box, n_tri;
};

// Flat traversal loop: each iteration advances every live lane by one interior node and/or one
// leaf, and a lane whose ray is finished writes its result and takes its next ray at once (rays
// i, i + stride, i + 2*stride, ... per lane) instead of waiting for the rest of its wave. The order
// of node visits within a ray is the reference's, so results and work counters are unchanged.
template <int BLOCK, int DEPTH, bool STATS>
__global__ __launch_bounds__(BLOCK) void k_closest(const DevNode* __restrict__ nodes, const DevTri* __restrict__ tris,
                                                   RootDesc root, const float4* __restrict__ rays,
                                                   float4* __restrict__ hits, uint32_t n, uint32_t* __restrict__ stats,
                                                   uint2* __restrict__ spill, uint32_t spill_cap,
                                                   int* __restrict__ status) {
  __shared__ uint2 lds[DEPTH][BLOCK];
  Stack<BLOCK, DEPTH> st{lds, spill, gridDim.x * BLOCK, spill_cap, threadIdx.x, blockIdx.x * BLOCK + threadIdx.x, 0};
  const uint32_t stride = gridDim.x * BLOCK;
  uint32_t i = st.gtid;
  if (i >= n) return;
  RayState r;
  Closest s;
  int32_t cur = 0;
  bool live = false;
  auto start = [&](uint32_t k) {
    r = make_ray(rays[2 * (size_t)k], rays[2 * (size_t)k + 1]);
    s.min_t = FLT_MAX;
    s.t = s.u = s.v = 0.f;
    s.cur_far = r.far;  // min(far, min_t): the box tests' narrowed interval
    s.best = 0xffffffffu;
    s.n_vis = 0;
    s.n_box = 1;
    s.n_tri = 0;
    st.sp = 0;
    float entry;
    live = box_test(r, root.lo[0], root.hi[0], root.lo[1], root.hi[1], root.lo[2], root.hi[2], r.far, &entry);
    cur = root.ref;
  };
  start(i);
  for (;;) {
    bool pop_now = false;
    if (live && cur >= 0) {
      // interior node: test both children over [near, min(far, min_t)]
      if (STATS) { ++s.n_vis; s.n_box += 2; }
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
      // leaf: ~cur = (first << 5) | (count - 1)
      if (STATS) ++s.n_vis;
      const uint32_t code = ~(uint32_t)cur;
      const uint32_t first = code >> kLeafCountBits;
      const uint32_t cnt = (code & (kMaxLeafTris - 1)) + 1;
      for (uint32_t k = first; k < first + cnt; ++k) {
        if (STATS) ++s.n_tri;
        float t, u, v;
        uint32_t id;
        if (tri_test(r, tris, k, &t, &u, &v, &id)) {
          if (t >= r.near && t <= r.far && (t < s.min_t || (t == s.min_t && id < s.best))) {
            s.min_t = t; s.t = t; s.u = u; s.v = v; s.best = id;
            s.cur_far = fminf(r.far, s.min_t);
          }
        }
      }
      pop_now = true;
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
    if (!live) {
      const bool found = s.best != 0xffffffffu;
      hits[i] = make_float4(__int_as_float(found ? (int32_t)s.best : -1), found ? s.t : 0.f, found ? s.u : 0.f,
                            found ? s.v : 0.f);
      if (STATS) {
        stats[3 * (size_t)i] = s.n_vis;
        stats[3 * (size_t)i + 1] = s.n_box;
        stats[3 * (size_t)i + 2] = s.n_tri;
      }
      i += stride;
      if (i >= n) break;
      start(i);
    }
  }
}

// Any hit, unordered (bvh.rs:381-420): the reference pushes left then right, so the right child is
// examined first; the first triangle hit within [near, far] ends the ray. Same flat loop.
template <int BLOCK, int DEPTH>
__global__ __launch_bounds__(BLOCK) void k_any(const DevNode* __restrict__ nodes, const DevTri* __restrict__ tris,
                                               RootDesc root, const float4* __restrict__ rays,
                                               uint8_t* __restrict__ occluded, uint32_t n, uint2* __restrict__ spill,
                                               uint32_t spill_cap, int* __restrict__ status) {
  __shared__ uint2 lds[DEPTH][BLOCK];
  Stack<BLOCK, DEPTH> st{lds, spill, gridDim.x * BLOCK, spill_cap, threadIdx.x, blockIdx.x * BLOCK + threadIdx.x, 0};
  const uint32_t stride = gridDim.x * BLOCK;
  uint32_t i = st.gtid;
  if (i >= n) return;
  RayState r;
  int32_t cur = 0;
  bool live = false, hit = false;
  auto start = [&](uint32_t k) {
    r = make_ray(rays[2 * (size_t)k], rays[2 * (size_t)k + 1]);
    st.sp = 0;
    hit = false;
    float entry;
    live = box_test(r, root.lo[0], root.hi[0], root.lo[1], root.hi[1], root.lo[2], root.hi[2], r.far, &entry);
    cur = root.ref;
  };
  start(i);
  for (;;) {
    bool pop_now = false;
    if (live && cur >= 0) {
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
      const uint32_t code = ~(uint32_t)cur;
      const uint32_t first = code >> kLeafCountBits;
      const uint32_t cnt = (code & (kMaxLeafTris - 1)) + 1;
      for (uint32_t k = first; k < first + cnt; ++k) {
        float t, u, v;
        uint32_t id;
        if (tri_test(r, tris, k, &t, &u, &v, &id) && t >= r.near && t <= r.far) {
          hit = true;
          break;
        }
      }
      pop_now = true;
      if (hit) live = false;
    }
    if (live && pop_now) {
      live = st.sp > 0;
      if (live) cur = (int32_t)st.pop().x;
    }
    if (!live) {
      occluded[i] = hit ? 1 : 0;
      i += stride;
      if (i >= n) break;
      start(i);
    }
  }
}

template <int BLOCK, int DEPTH>
void launch_shape(const LaunchArgs& a, uint32_t* grid) {
  const uint32_t per_block_waves = BLOCK / 64;
  const uint32_t g = a.cu_count * (a.waves_per_cu / per_block_waves > 0 ? a.waves_per_cu / per_block_waves : 1);
  const uint32_t need = (a.n + BLOCK - 1) / BLOCK;
  *grid = g < need ? g : need;
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
  if (kind == 0)
    hipLaunchKernelGGL((k_closest<BLOCK, DEPTH, false>), dim3(grid), dim3(BLOCK), 0, s, a.nodes, a.tris, a.root,
                       (const float4*)a.rays, (float4*)a.hits, a.n, nullptr, a.spill, spill_cap, a.status);
  else if (kind == 1)
    hipLaunchKernelGGL((k_closest<BLOCK, DEPTH, true>), dim3(grid), dim3(BLOCK), 0, s, a.nodes, a.tris, a.root,
                       (const float4*)a.rays, (float4*)a.hits, a.n, a.stats, a.spill, spill_cap, a.status);
  else
    hipLaunchKernelGGL((k_any<BLOCK, DEPTH>), dim3(grid), dim3(BLOCK), 0, s, a.nodes, a.tris, a.root,
                       (const float4*)a.rays, a.occluded, a.n, a.spill, spill_cap, a.status);
  return hipGetLastError();
}

template <int BLOCK>
hipError_t launch_b(const LaunchArgs& a, int kind) {
  switch (a.lds_stack) {
    case 8: return launch_t<BLOCK, 8>(a, kind);
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
  switch (a.block) {
    case 64: return launch_b<64>(a, kind);
    case 128: return launch_b<128>(a, kind);
    case 256: return launch_b<256>(a, kind);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace pbrt
