// Device helpers shared by the traversal kernels: ray state, the widened slab test, Möller–Trumbore
// and the LDS short stack with its HBM spill. Arithmetic follows the reference statement by
// statement in FP32 (the library is built with -ffp-contract=off, IEEE division and square root):
//   slab test        src/geom/aabound.rs:323-355
//   Möller-Trumbore  src/shapes/triangle_mesh/triangle_mesh.rs:183-217
#pragma once
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <cfloat>

#include "common.hpp"

namespace pbrt {
namespace dev {

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

// The batch index of the k-th ray a kernel takes. tile_w = 0: batch order. tile_w > 0: the batch is
// an image of tile_w rays per row (tile_w a multiple of 8, the ray count a multiple of 16 tile_w)
// taken in tiles of 8 x 16 rays, band of 16 rows after band, so that the 64 lanes of a wave start
// on an 8 x 8 patch of the image (the next wave on the patch below) rather than on a 64-ray strip
// of one row. Results are stored at the ray's own index.
__device__ __forceinline__ uint32_t tiled_ray(uint32_t k, uint32_t tile_w) {
  if (tile_w == 0) return k;
  const uint32_t band = k / (16 * tile_w), in_band = k - band * 16 * tile_w;
  const uint32_t tile = in_band >> 7, slot = in_band & 127;
  return (band * 16 + (slot >> 3)) * tile_w + (tile << 3) + (slot & 7);
}

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

// The 4-wide node step shared by the mesh and scene kernels: the children's boxes tested over
// [near, tmax], hit children ordered by entry distance (a box the ray misses sorts last with an
// infinite key: a hit's entry is finite, at most tmax). Returns the number of children hit; the
// nearest is in r0 / k0, the others in r1..r3 / k1..k3 in increasing distance.
__device__ __forceinline__ int node4_step(const RayState& r, const DevNode4* __restrict__ nodes, int32_t cur, float tmax,
                                          int32_t& r0, int32_t& r1, int32_t& r2, int32_t& r3, float& k0, float& k1,
                                          float& k2, float& k3) {
  const float4* p = reinterpret_cast<const float4*>(nodes + cur);
  const float4 lx = p[0], hx = p[1], ly = p[2], hy = p[3], lz = p[4], hz = p[5];
  const int4 rf = reinterpret_cast<const int4*>(p)[6];
  const int cnt = reinterpret_cast<const int4*>(p)[7].x;
  float d0, d1, d2, d3;
  const bool h0 = box_test(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, tmax, &d0);
  const bool h1 = box_test(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, tmax, &d1);
  const bool h2 = cnt > 2 && box_test(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, tmax, &d2);
  const bool h3 = cnt > 3 && box_test(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, tmax, &d3);
  k0 = h0 ? d0 : INFINITY;
  k1 = h1 ? d1 : INFINITY;
  k2 = h2 ? d2 : INFINITY;
  k3 = h3 ? d3 : INFINITY;
  r0 = rf.x;
  r1 = rf.y;
  r2 = rf.z;
  r3 = rf.w;
  auto cswap = [](float& ka, int32_t& ra, float& kb, int32_t& rb) {
    const bool sw = kb < ka;
    const float tk = sw ? kb : ka;
    kb = sw ? ka : kb;
    ka = tk;
    const int32_t tr = sw ? rb : ra;
    rb = sw ? ra : rb;
    ra = tr;
  };
  cswap(k0, r0, k1, r1);
  cswap(k2, r2, k3, r3);
  cswap(k0, r0, k2, r2);
  cswap(k1, r1, k3, r3);
  cswap(k1, r1, k2, r2);
  return (int)h0 + (int)h1 + (int)h2 + (int)h3;
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

}  // namespace dev
}  // namespace pbrt
