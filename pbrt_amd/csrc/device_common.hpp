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
  bool parallel;      // a direction component is exactly zero, or its reciprocal overflows
  bool once;          // box tests may widen each interval once (set by guard_once; see box_once)
  uint32_t sx, sy, sz;  // byte offset (0 or 16) of the near-plane word of each axis in a 4-wide node
                        // (0 for a parallel ray, whose literal test takes the planes in order)
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

// The logical index of this workgroup in a grid of gridDim.x: the hardware hands workgroups to the
// eight XCDs in turn (workgroup b to XCD b % 8), so the logical index numbers each XCD's workgroups
// consecutively (XCD x takes logical workgroups offset_x .. offset_x + count_x - 1). With rays
// assigned by logical index, the rays in flight on one XCD are neighbours in the batch, and its L2
// holds the part of the tree they share. A bijection of [0, gridDim.x).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t b = blockIdx.x, g = gridDim.x;
  const uint32_t q = g / kXcds, rem = g % kXcds, x = b % kXcds, j = b / kXcds;
  return x * q + min(x, rem) + j;
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
  r.parallel = (b.x == 0.0f) | (b.y == 0.0f) | (b.z == 0.0f) | (fabsf(r.ix) == INFINITY) | (fabsf(r.iy) == INFINITY) |
               (fabsf(r.iz) == INFINITY);
  r.sx = r.parallel ? 0u : (r.nx ? 16u : 0u);
  r.sy = r.parallel ? 0u : (r.ny ? 16u : 0u);
  r.sz = r.parallel ? 0u : (r.nz ? 16u : 0u);
  r.once = false;
  return r;
}

// Marks a ray whose slab distances through every plane of the box [lo, hi] (a tree's root box, which
// holds every child box of the tree) are finite: with o the origin and i the reciprocal, each
// |(p - o) * i| is at most max(|lo - o|, |hi - o|) * |i| in FP32 (rounding is monotonic), so when
// that bound is finite no plane product overflows and the box tests may take box_once.
__device__ __forceinline__ void guard_once(RayState& r, const float lo[3], const float hi[3]) {
  const float bx = fmaxf(fabsf(lo[0] - r.ox), fabsf(hi[0] - r.ox)) * fabsf(r.ix);
  const float by = fmaxf(fabsf(lo[1] - r.oy), fabsf(hi[1] - r.oy)) * fabsf(r.iy);
  const float bz = fmaxf(fabsf(lo[2] - r.oz), fabsf(hi[2] - r.oz)) * fabsf(r.iz);
  r.once = !r.parallel && bx <= FLT_MAX && by <= FLT_MAX && bz <= FLT_MAX;
}

// Slab tests of a non-parallel ray against planes already ordered by the reciprocal's sign (px0 the
// near plane of x, px1 the far one; likewise y and z). Computing both slab
// distances through the plane selected by the sign equals the reference's compute-then-swap
// (aabound.rs:335-340), and taking the max / min over the three axes at the end equals its
// per-axis early exit (a rejection at an earlier axis implies one at the end).
// fmaxf / fminf ignore a NaN operand, as Rust's f64::max / min do.
//
// box_planes: the reference's widening of every plane distance, t0 - |t0| W and t1 + |t1| W
// (aabound.rs:343-344).
__device__ __forceinline__ bool box_planes(const RayState& r, float px0, float px1, float py0, float py1, float pz0,
                                           float pz1, float tmax_in, float* entry) {
  const float W = slab_widening();
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

// box_once: the same interval with the largest entry and the smallest exit widened once. Both maps
// are non-decreasing in FP32 on [-inf, inf), so this equals box_planes whenever no plane distance is
// infinite (an overflowed product, where the per-plane map gives NaN, which max / min ignore): the
// rays guard_once marks.
__device__ __forceinline__ bool box_once(const RayState& r, float px0, float px1, float py0, float py1, float pz0,
                                         float pz1, float tmax_in, float* entry) {
  const float W = slab_widening();
  const float t0x = (px0 - r.ox) * r.ix, t1x = (px1 - r.ox) * r.ix;
  const float t0y = (py0 - r.oy) * r.iy, t1y = (py1 - r.oy) * r.iy;
  const float t0z = (pz0 - r.oz) * r.iz, t1z = (pz1 - r.oz) * r.iz;
  const float a = fmaxf(fmaxf(t0x, t0y), t0z), b = fminf(fminf(t1x, t1y), t1z);
  const float tmin = fmaxf(a - fabsf(a) * W, r.near);
  const float tmax = fminf(b + fabsf(b) * W, tmax_in);
  *entry = tmin;
  return !(tmax < tmin);
}

// Slab test for a non-parallel ray with the planes in box order (lo, hi per axis), in the
// reference's per-plane form: both distances through the plane the reciprocal's sign selects (its
// compute-then-swap), each widened, then the max / min over the axes (its per-axis early exit).
// The binary walks take it for the rays guard_once does not mark.
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

// The binary walks' box test: the widen-once form for the rays guard_once marks (planes chosen by
// the reciprocal's sign), else the per-plane form, or the literal one for a parallel ray.
__device__ __forceinline__ bool box_test(const RayState& r, float lx, float hx, float ly, float hy, float lz, float hz,
                                         float tmax, float* entry) {
  if (__builtin_expect(r.once, 1))
    return box_once(r, r.nx ? hx : lx, r.nx ? lx : hx, r.ny ? hy : ly, r.ny ? ly : hy, r.nz ? hz : lz,
                    r.nz ? lz : hz, tmax, entry);
  if (__builtin_expect(r.parallel, 0)) return box_literal(r, lx, hx, ly, hy, lz, hz, tmax, entry);
  return box_fast(r, lx, hx, ly, hy, lz, hz, tmax, entry);
}

// Möller–Trumbore with the edges pre-computed in FP32 on the host (the same subtraction the
// reference performs first): a = (p0, original index bits), e1 = p1 - p0, e2 = p2 - p0.
__device__ __forceinline__ bool tri_test_words(const RayState& r, const float4 a, const float4 e1, const float4 e2,
                                               float* t_out, float* u_out, float* v_out, uint32_t* id_out) {
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

__device__ __forceinline__ bool tri_test(const RayState& r, const DevTri* __restrict__ tris, uint32_t k, float* t_out,
                                         float* u_out, float* v_out, uint32_t* id_out) {
  return tri_test_words(r, *reinterpret_cast<const float4*>(tris[k].p0), *reinterpret_cast<const float4*>(tris[k].e1),
                        *reinterpret_cast<const float4*>(tris[k].e2), t_out, u_out, v_out, id_out);
}

// A buffer resource over a 4-wide node array (base in scalar registers; raw loads, no range
// limit): the node loads then take one 32-bit byte offset each instead of a 64-bit address. A
// 4-wide array stays below 4 GiB (the host refuses larger ones).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t node4_rsrc(const DevNode4* nodes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<DevNode4*>(nodes), (short)0, -1, 0x00020000);
}

__device__ __forceinline__ float4 load_f4(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
}

// The same over a mesh's triangle array (one mesh's triangles stay below 4 GiB: the leaf code
// bounds a mesh to 2^26 triangles, 3.2 GB): triangle k at byte offset 48 k.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tri_rsrc(const DevTri* tris) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<DevTri*>(tris), (short)0, -1, 0x00020000);
}

__device__ __forceinline__ bool tri_test(const RayState& r, __amdgpu_buffer_rsrc_t tris, uint32_t k, float* t_out,
                                         float* u_out, float* v_out, uint32_t* id_out) {
  const uint32_t o = k * 48u;
  return tri_test_words(r, load_f4(tris, o), load_f4(tris, o + 16u), load_f4(tris, o + 32u), t_out, u_out, v_out,
                        id_out);
}

// The words of a 4-wide node a ray's step reads: the near and far plane words of each axis, chosen
// by the ray's sign through the load offset (in box order for a parallel ray), the references and
// the child count.
struct Node4Words {
  float4 x0, x1, y0, y1, z0, z1;
  int4 rf;
  int cnt;
};

// From the mesh kernels' buffer resource (node4_rsrc; node cur at byte offset 128 cur).
__device__ __forceinline__ Node4Words load_node4(const RayState& r, __amdgpu_buffer_rsrc_t nodes, uint32_t cur) {
  const uint32_t o = cur << 7;
  Node4Words w;
  w.x0 = load_f4(nodes, o | r.sx);
  w.x1 = load_f4(nodes, o | (r.sx ^ 16u));
  w.y0 = load_f4(nodes, o | (32u | r.sy));
  w.y1 = load_f4(nodes, o | (32u | (r.sy ^ 16u)));
  w.z0 = load_f4(nodes, o | (64u | r.sz));
  w.z1 = load_f4(nodes, o | (64u | (r.sz ^ 16u)));
  const float4 rc = load_f4(nodes, o | 96u);
  w.rf = make_int4(__float_as_int(rc.x), __float_as_int(rc.y), __float_as_int(rc.z), __float_as_int(rc.w));
  w.cnt = (int)__builtin_amdgcn_raw_buffer_load_b32(nodes, o | 112u, 0, 0);
  return w;
}

// From a node array by address (the scene walks, whose array base differs per lane's object: a
// buffer resource needs a wave-uniform base).
__device__ __forceinline__ Node4Words load_node4(const RayState& r, const DevNode4* __restrict__ nodes, uint32_t cur) {
  const char* nb = reinterpret_cast<const char*>(nodes + cur);
  Node4Words w;
  w.x0 = *reinterpret_cast<const float4*>(nb + r.sx);
  w.x1 = *reinterpret_cast<const float4*>(nb + (r.sx ^ 16u));
  w.y0 = *reinterpret_cast<const float4*>(nb + 32 + r.sy);
  w.y1 = *reinterpret_cast<const float4*>(nb + 32 + (r.sy ^ 16u));
  w.z0 = *reinterpret_cast<const float4*>(nb + 64 + r.sz);
  w.z1 = *reinterpret_cast<const float4*>(nb + 64 + (r.sz ^ 16u));
  w.rf = *reinterpret_cast<const int4*>(nb + 96);
  w.cnt = reinterpret_cast<const int4*>(nb + 112)->x;
  return w;
}

// The 4-wide node step shared by the mesh and scene kernels: node `cur` (from a buffer resource or
// an array), its children's boxes tested over [near, tmax], hit children ordered by entry distance
// (a box the ray misses sorts last with an infinite key: a hit's entry is finite, at most tmax).
// Returns the number of children hit; the nearest is in r0 / k0, the others in r1..r3 / k1..k3 in
// increasing distance.
template <class Nodes>
__device__ __forceinline__ int node4_step(const RayState& r, Nodes nodes, uint32_t cur, float tmax, int32_t& r0,
                                          int32_t& r1, int32_t& r2, int32_t& r3, float& k0, float& k1, float& k2,
                                          float& k3) {
  const Node4Words w = load_node4(r, nodes, cur);
  const float4 &x0 = w.x0, &x1 = w.x1, &y0 = w.y0, &y1 = w.y1, &z0 = w.z0, &z1 = w.z1;
  const int4& rf = w.rf;
  const int cnt = w.cnt;
  float d0, d1, d2, d3;
  bool h0, h1, h2, h3;
  if (__builtin_expect(r.once, 1)) {
    // all four slots tested without branches (an unused slot holds a zero box; its result is
    // dropped by the child count)
    h0 = box_once(r, x0.x, x1.x, y0.x, y1.x, z0.x, z1.x, tmax, &d0);
    h1 = box_once(r, x0.y, x1.y, y0.y, y1.y, z0.y, z1.y, tmax, &d1);
    h2 = box_once(r, x0.z, x1.z, y0.z, y1.z, z0.z, z1.z, tmax, &d2) & (cnt > 2);
    h3 = box_once(r, x0.w, x1.w, y0.w, y1.w, z0.w, z1.w, tmax, &d3) & (cnt > 3);
  } else {
    auto test = [&](float a0, float a1, float b0, float b1, float c0, float c1, float* d) {
      if (r.parallel) return box_literal(r, a0, a1, b0, b1, c0, c1, tmax, d);
      return box_planes(r, a0, a1, b0, b1, c0, c1, tmax, d);
    };
    h0 = test(x0.x, x1.x, y0.x, y1.x, z0.x, z1.x, &d0);
    h1 = test(x0.y, x1.y, y0.y, y1.y, z0.y, z1.y, &d1);
    h2 = cnt > 2 && test(x0.z, x1.z, y0.z, y1.z, z0.z, z1.z, &d2);
    h3 = cnt > 3 && test(x0.w, x1.w, y0.w, y1.w, z0.w, z1.w, &d3);
  }
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
  // The empty asm statements keep the LDS and the spill accesses apart: merged by the optimiser, the
  // two become one generic (flat) access through a selected pointer, which goes through the vector
  // memory path and waits on both the LDS and the vector-memory counters. As separate LDS and
  // global instructions the headline traversal batches ran 3-4 % faster.
  __device__ __forceinline__ bool push(int32_t ref, float d) {
    const uint2 e = make_uint2((uint32_t)ref, __float_as_uint(d));
    if (sp < DEPTH) {
      lds[sp][lane] = e;
      __asm__ volatile("" ::: "memory");
    } else {
      const uint32_t k = (uint32_t)(sp - DEPTH);
      if (k >= spill_cap) return false;
      spill[(size_t)k * spill_stride + gtid] = e;
      __asm__ volatile("" ::: "memory");
    }
    ++sp;
    return true;
  }
  __device__ __forceinline__ uint2 pop() {
    --sp;
    uint2 e;
    if (sp < DEPTH) {
      e = lds[sp][lane];
      __asm__ volatile("" ::: "memory");
    } else {
      e = spill[(size_t)(sp - DEPTH) * spill_stride + gtid];
      __asm__ volatile("" ::: "memory");
    }
    return e;
  }
};

}  // namespace dev
}  // namespace pbrt
