// Scene-level traversal and the wavefront path tracer on CDNA4 (gfx950).
//
// Scene traversal restates the reference's two levels in FP32:
//   top level    the median-split object BVH, ordered closest hit (src/bvh.rs:153-230) with the
//                visitor of Scene::find_nearest (src/scene.rs:186-220) and unordered any hit
//                (bvh.rs:239-275); box tests narrowed to [near, nearest hit]
//   objects      Transformed (src/objects/transformed.rs:26-77, src/geom/transform.rs:98-116),
//                Simple taking the shape's first intersection (src/objects/simple.rs:27-42)
//   shapes       Sphere (src/shapes/sphere.rs:108-187), AABox (src/shapes/aabox.rs:66-193),
//                Rectangle (src/shapes/rectangle.rs:26-58), TriangleMesh's BVH (nested, sharing the
//                lane's stack) and shading quantities (triangle_mesh.rs:73-132, 219-258)
// Ties between objects at equal distance go to the smaller (object, triangle) index.
//
// The path tracer is PathIntegrator::li (src/integrators/path.rs:48-115) turned into a wavefront:
// generate (pinhole camera, src/cameras/pin_hole.rs, pixel jitter of renderers/mt.rs:123-127) ->
// per bounce: extend (closest hit) -> shade (emission after specular bounces, next-event estimation
// with one uniformly chosen light, material scatter) -> shadow (any hit) -> accumulate (per-pixel
// FP64 sums in a fixed sample order). Every random number is drawn from a counter keyed by (seed,
// pixel * spp + sample, bounce, slot), so the result does not depend on scheduling.
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "device_common.hpp"
#include "rng.hpp"
#include "scene.hpp"

namespace pbrt {
namespace {
using namespace dev;

constexpr float kPi = (float)3.14159265358979323846;

// ---- small FP32 vector algebra in the reference's operation order (src/geom/vector3.rs) ---------
struct V {
  float x, y, z;
};
__device__ __forceinline__ V vadd(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V vsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V vmul(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V vcross(V u, V v) {
  return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
__device__ __forceinline__ V vnorm(V v) {
  const float n = vdot(v, v);
  const float inv = 1.0f / sqrtf(n);
  return vmul(v, inv);
}
__device__ __forceinline__ float vlen(V v) { return sqrtf(vdot(v, v)); }

// matrix4.rs:187-213 on the first three rows (affine: w = 1)
__device__ __forceinline__ V xpoint(const float* M, V v) {
  return {M[0] * v.x + M[1] * v.y + M[2] * v.z + M[3], M[4] * v.x + M[5] * v.y + M[6] * v.z + M[7],
          M[8] * v.x + M[9] * v.y + M[10] * v.z + M[11]};
}
__device__ __forceinline__ V xdir(const float* M, V v) {
  return {M[0] * v.x + M[1] * v.y + M[2] * v.z, M[4] * v.x + M[5] * v.y + M[6] * v.z,
          M[8] * v.x + M[9] * v.y + M[10] * v.z};
}
__device__ __forceinline__ V xnormal(const float* M, V v) {
  return {M[0] * v.x + M[4] * v.y + M[8] * v.z, M[1] * v.x + M[5] * v.y + M[9] * v.z,
          M[2] * v.x + M[6] * v.y + M[10] * v.z};
}

__device__ __forceinline__ RayState ray_from(V o, V d, float near, float far) {
  return make_ray(make_float4(o.x, o.y, o.z, near), make_float4(d.x, d.y, d.z, far));
}

// ---- analytic shapes: distance of the first intersection the shape records -------------------------
// sphere.rs:108-144
__device__ __forceinline__ bool sphere_t(V o, V d, float r, float near, float far, float* t_out) {
  const V l = vmul(o, -1.0f);
  const float tca = vdot(l, d);
  if (tca < 0.0f) return false;
  const float r2 = r * r;
  const float d2 = vdot(l, l) - tca * tca;
  if (d2 > r2) return false;
  const float thc = sqrtf(r2 - d2);
  const float t0 = tca - thc, t1 = tca + thc;
  const float tmin = fminf(t0, t1), tmax = fmaxf(t0, t1);
  if (tmax < near || tmin > far) return false;
  if (tmin > near) {
    *t_out = tmin;
    return true;
  }
  if (tmax > near && tmax < far) {
    *t_out = tmax;
    return true;
  }
  return false;
}
// aabox.rs:66-193; *face: index into the normal table (x-, x+, y-, y+, z-, z+)
__device__ __forceinline__ bool box_t(V o, V d, const float* ext, float near, float far, float* t_out, int* face) {
  const V mn{0.0f - ext[0] * 0.5f, 0.0f - ext[1] * 0.5f, 0.0f - ext[2] * 0.5f};
  const V mx{ext[0] * 0.5f, ext[1] * 0.5f, ext[2] * 0.5f};
  const V inv{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
  float tmin, tmax;
  int nmin, nmax;
  if (inv.x >= 0.0f) { tmin = (mn.x - o.x) * inv.x; tmax = (mx.x - o.x) * inv.x; nmin = 0; nmax = 1; }
  else { tmin = (mx.x - o.x) * inv.x; tmax = (mn.x - o.x) * inv.x; nmin = 1; nmax = 0; }
  float tymin, tymax;
  int nymin, nymax;
  if (inv.y >= 0.0f) { tymin = (mn.y - o.y) * inv.y; tymax = (mx.y - o.y) * inv.y; nymin = 2; nymax = 3; }
  else { tymin = (mx.y - o.y) * inv.y; tymax = (mn.y - o.y) * inv.y; nymin = 3; nymax = 2; }
  if (tmin > tymax || tymin > tmax) return false;
  if (tymin > tmin) { tmin = tymin; nmin = nymin; }
  if (tymax < tmax) { tmax = tymax; nmax = nymax; }
  float tzmin, tzmax;
  int nzmin, nzmax;
  if (inv.z >= 0.0f) { tzmin = (mn.z - o.z) * inv.z; tzmax = (mx.z - o.z) * inv.z; nzmin = 4; nzmax = 5; }
  else { tzmin = (mx.z - o.z) * inv.z; tzmax = (mn.z - o.z) * inv.z; nzmin = 5; nzmax = 4; }
  if (tmin > tzmax || tzmin > tmax) return false;
  if (tzmin > tmin) { tmin = tzmin; nmin = nzmin; }
  if (tzmax < tmax) { tmax = tzmax; nmax = nzmax; }
  if (tmax < near || tmin > far) return false;
  if (tmin > near) {
    *t_out = tmin;
    *face = nmin;
    return true;
  }
  if (tmax < far) {
    *t_out = tmax;
    *face = nmax;
    return true;
  }
  return false;
}
// rectangle.rs:26-58
__device__ __forceinline__ bool rect_t(V o, V d, float w, float h, float near, float far, float* t_out) {
  if (d.y == 0.0f) return false;
  const float t = (o.y * -1.0f) / d.y;
  if (t < near || t > far) return false;
  V p = vadd(o, vmul(d, t));
  if (fabsf(p.x) > w / 2.0f || fabsf(p.z) > h / 2.0f) return false;
  *t_out = t;
  return true;
}

// ---- CSG (shapes/csg/*.rs) ----------------------------------------------------------------------
// Every intersection a sphere records, in order (sphere.rs:133-143): at most two distances.
__device__ __forceinline__ int sphere_all(V o, V d, float r, float near, float far, float* t) {
  const V l = vmul(o, -1.0f);
  const float tca = vdot(l, d);
  if (tca < 0.0f) return 0;
  const float r2 = r * r;
  const float d2 = vdot(l, l) - tca * tca;
  if (d2 > r2) return 0;
  const float thc = sqrtf(r2 - d2);
  const float t0 = tca - thc, t1 = tca + thc;
  const float tmin = fminf(t0, t1), tmax = fmaxf(t0, t1);
  if (tmax < near || tmin > far) return 0;
  int n = 0;
  if (tmin > near) t[n++] = tmin;
  if (tmax > near && tmax < far) t[n++] = tmax;
  return n;
}
// Every intersection a box records, in order (aabox.rs:155-190), with its face.
__device__ __forceinline__ int box_all(V o, V d, const float* ext, float near, float far, float* t, int* face) {
  const V mn{0.0f - ext[0] * 0.5f, 0.0f - ext[1] * 0.5f, 0.0f - ext[2] * 0.5f};
  const V mx{ext[0] * 0.5f, ext[1] * 0.5f, ext[2] * 0.5f};
  const V inv{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
  float tmin, tmax;
  int nmin, nmax;
  if (inv.x >= 0.0f) { tmin = (mn.x - o.x) * inv.x; tmax = (mx.x - o.x) * inv.x; nmin = 0; nmax = 1; }
  else { tmin = (mx.x - o.x) * inv.x; tmax = (mn.x - o.x) * inv.x; nmin = 1; nmax = 0; }
  float tymin, tymax;
  int nymin, nymax;
  if (inv.y >= 0.0f) { tymin = (mn.y - o.y) * inv.y; tymax = (mx.y - o.y) * inv.y; nymin = 2; nymax = 3; }
  else { tymin = (mx.y - o.y) * inv.y; tymax = (mn.y - o.y) * inv.y; nymin = 3; nymax = 2; }
  if (tmin > tymax || tymin > tmax) return 0;
  if (tymin > tmin) { tmin = tymin; nmin = nymin; }
  if (tymax < tmax) { tmax = tymax; nmax = nymax; }
  float tzmin, tzmax;
  int nzmin, nzmax;
  if (inv.z >= 0.0f) { tzmin = (mn.z - o.z) * inv.z; tzmax = (mx.z - o.z) * inv.z; nzmin = 4; nzmax = 5; }
  else { tzmin = (mx.z - o.z) * inv.z; tzmax = (mn.z - o.z) * inv.z; nzmin = 5; nzmax = 4; }
  if (tmin > tzmax || tzmin > tmax) return 0;
  if (tzmin > tmin) { tmin = tzmin; nmin = nzmin; }
  if (tzmax < tmax) { tmax = tzmax; nmax = nzmax; }
  if (tmax < near || tmin > far) return 0;
  int n = 0;
  if (tmin > near) {
    face[n] = nmin;
    t[n++] = tmin;
  }
  if (tmax < far) {
    face[n] = nmax;
    t[n++] = tmax;
  }
  return n;
}
// Shape::contain_point of an element at a point of the CSG frame (sphere.rs:146-148,
// aabox.rs:195-202; both closed), through transform_point_to_local.
__device__ __forceinline__ bool csg_contains(const DevCsgElem& e, V p) {
  const V lp = xpoint(e.minv, p);
  if (e.shape == kShapeSphere) {
    const float r = e.param[0];
    return vdot(lp, lp) <= r * r;
  }
  const float hx = e.param[0] * 0.5f, hy = e.param[1] * 0.5f, hz = e.param[2] * 0.5f;
  return lp.x >= 0.0f - hx && lp.x <= hx && lp.y >= 0.0f - hy && lp.y <= hy && lp.z >= 0.0f - hz && lp.z <= hz;
}
// An element's intersections: distances, faces and the points in the CSG frame.
__device__ __forceinline__ int csg_elem_hits(const DevCsgElem& e, V o, V d, float near, float far, float* t, int* face,
                                             V* p) {
  const V lo = xpoint(e.minv, o), ld = vnorm(xdir(e.minv, d));
  int n;
  if (e.shape == kShapeSphere) {
    n = sphere_all(lo, ld, e.param[0], near, far, t);
    face[0] = face[1] = 0;
  } else {
    n = box_all(lo, ld, e.param, near, far, t, face);
  }
  for (int i = 0; i < n; ++i) p[i] = xpoint(e.m, vadd(lo, vmul(ld, t[i])));
  return n;
}
// The first intersection of a CSG shape on a ray of its frame (union.rs:22-42, intersection.rs:
// 23-43, substraction.rs:22-70): the nearest kept candidate, the earliest recorded among equal
// distances (the reference sorts them with a stable sort). *elem: the element it belongs to;
// *face: its box face, | 8 when the normal is flipped.
__device__ __forceinline__ bool csg_first(const SceneView& S, const DevObject& ob, V o, V d, float near, float far,
                                         float* t_out, int32_t* elem, int32_t* face_out) {
  const DevCsgElem* E = S.csg + ob.mesh;
  const int op = (int)ob.param[0];
  const uint32_t cnt = (uint32_t)ob.param[1];
  bool found = false;
  float best = 0.0f;
  float t[2];
  int face[2];
  V p[2];
  auto consider = [&](float tt, uint32_t k, int f) {
    if (!found || tt < best) {
      found = true;
      best = tt;
      *t_out = tt;
      *elem = (int32_t)k;
      *face_out = f;
    }
  };
  if (op == kCsgSubtraction) {
    if (cnt == 0) return false;
    int n = csg_elem_hits(E[0], o, d, near, far, t, face, p);
    for (int i = 0; i < n; ++i) {
      bool covered = false;
      for (uint32_t m = 1; m < cnt && !covered; ++m) covered = csg_contains(E[m], p[i]);
      if (!covered) consider(t[i], 0, face[i]);
    }
    for (uint32_t k = 1; k < cnt; ++k) {
      n = csg_elem_hits(E[k], o, d, near, far, t, face, p);
      for (int i = 0; i < n; ++i) {
        if (!csg_contains(E[0], p[i])) continue;
        bool covered = false;
        for (uint32_t m = 1; m < cnt && !covered; ++m) covered = m != k && csg_contains(E[m], p[i]);
        if (!covered) consider(t[i], k, face[i] | 8);
      }
    }
    return found;
  }
  const bool uni = op == kCsgUnion;
  for (uint32_t k = 0; k < cnt; ++k) {
    const int n = csg_elem_hits(E[k], o, d, near, far, t, face, p);
    for (int i = 0; i < n; ++i) {
      // union: inside no other element; intersection: inside every other element
      bool keep = true;
      for (uint32_t m = 0; m < cnt && keep; ++m)
        if (m != k && csg_contains(E[m], p[i]) == uni) keep = false;
      if (keep) consider(t[i], k, face[i]);
    }
  }
  return found;
}

// A pushed traversal entry: node reference and entry distance.
template <int BLOCK, int DEPTH>
__device__ __forceinline__ bool push_or_flag(Stack<BLOCK, DEPTH>& st, int32_t ref, float d, int* status) {
  if (st.push(ref, d)) return true;
  atomicOr(status, 1);
  return false;
}

// BVHTree::query on an object's mesh (nested on the lane's stack above `base`): closest (t, prim)
// over [near, far] with the order-independent tie rule. Each iteration advances the lane by one
// interior node or by one triangle of its leaf, so a wave never runs a whole leaf loop for the
// few lanes in leaves (the mesh kernels' flat loop, inside the scene walk).
template <int BLOCK, int DEPTH>
__device__ bool mesh_closest(const SceneView& S, const DevObject& ob, const RayState& ray, float far,
                             Stack<BLOCK, DEPTH>& st, int* status, float* t_out, float* u_out, float* v_out,
                             uint32_t* prim_out) {
  RayState r = ray;
  guard_once(r, ob.root_lo, ob.root_hi);
  float entry;
  if (!box_test(r, ob.root_lo[0], ob.root_hi[0], ob.root_lo[1], ob.root_hi[1], ob.root_lo[2], ob.root_hi[2], far,
                &entry))
    return false;
  const int base = st.sp;
  const DevNode* nodes = S.blas + ob.node_base;
  const DevTri* tris = S.tris + ob.tri_base;
  float min_t = FLT_MAX, cur_far = far;
  uint32_t best = 0xffffffffu;
  int32_t cur = ob.root_ref;
  uint32_t tk = 0, tend = 0;  // triangles of the current leaf still to test: [tk, tend)
  bool live = true;
  while (live) {
    bool pop_now = false;
    if (cur >= 0) {
      const DevNode* nd = nodes + cur;
      const float4 a = *reinterpret_cast<const float4*>(nd->xy0);
      const float4 b = *reinterpret_cast<const float4*>(nd->xy1);
      const float4 z = *reinterpret_cast<const float4*>(nd->z01);
      const int2 c = *reinterpret_cast<const int2*>(nd->ref);
      float dl, dr;
      const bool hl = box_test(r, a.x, a.y, a.z, a.w, z.x, z.y, cur_far, &dl);
      const bool hr = box_test(r, b.x, b.y, b.z, b.w, z.z, z.w, cur_far, &dr);
      if (hl && hr) {
        const bool right_first = dr < dl;
        if (!push_or_flag(st, right_first ? c.x : c.y, right_first ? dl : dr, status)) break;
        cur = right_first ? c.y : c.x;
      } else if (hl) {
        cur = c.x;
      } else if (hr) {
        cur = c.y;
      } else {
        pop_now = true;
      }
    }
    if (!pop_now && cur < 0) {
      if (tk == tend) {  // a leaf reached: ~cur = (first << 5) | (count - 1)
        const uint32_t code = ~(uint32_t)cur;
        tk = code >> kLeafCountBits;
        tend = tk + (code & (kMaxLeafTris - 1)) + 1;
      }
      float t, u, v;
      uint32_t id;
      if (tri_test(r, tris, tk, &t, &u, &v, &id) && t >= r.near && t <= far &&
          (t < min_t || (t == min_t && id < best))) {
        min_t = t;
        best = id;
        *t_out = t;
        *u_out = u;
        *v_out = v;
        cur_far = fminf(far, min_t);
      }
      ++tk;
      pop_now = tk == tend;
    }
    if (pop_now) {
      live = false;
      while (st.sp > base) {
        const uint2 e = st.pop();
        if (__uint_as_float(e.y) > min_t) continue;
        cur = (int32_t)e.x;
        live = true;
        break;
      }
    }
  }
  st.sp = base;
  *prim_out = best;
  return best != 0xffffffffu;
}

template <int BLOCK, int DEPTH>
__device__ bool mesh_any(const SceneView& S, const DevObject& ob, const RayState& ray, Stack<BLOCK, DEPTH>& st,
                         int* status) {
  RayState r = ray;
  guard_once(r, ob.root_lo, ob.root_hi);
  float entry;
  if (!box_test(r, ob.root_lo[0], ob.root_hi[0], ob.root_lo[1], ob.root_hi[1], ob.root_lo[2], ob.root_hi[2], r.far,
                &entry))
    return false;
  const int base = st.sp;
  const DevNode* nodes = S.blas + ob.node_base;
  const DevTri* tris = S.tris + ob.tri_base;
  int32_t cur = ob.root_ref;
  uint32_t tk = 0, tend = 0;  // triangles of the current leaf still to test: [tk, tend)
  bool hit = false, live = true;
  while (live && !hit) {
    bool pop_now = false;
    if (cur >= 0) {
      const DevNode* nd = nodes + cur;
      const float4 a = *reinterpret_cast<const float4*>(nd->xy0);
      const float4 b = *reinterpret_cast<const float4*>(nd->xy1);
      const float4 z = *reinterpret_cast<const float4*>(nd->z01);
      const int2 c = *reinterpret_cast<const int2*>(nd->ref);
      float dl, dr;
      const bool hl = box_test(r, a.x, a.y, a.z, a.w, z.x, z.y, r.far, &dl);
      const bool hr = box_test(r, b.x, b.y, b.z, b.w, z.z, z.w, r.far, &dr);
      if (hl && hr) {
        if (!push_or_flag(st, c.x, 0.f, status)) break;
        cur = c.y;
      } else if (hr) {
        cur = c.y;
      } else if (hl) {
        cur = c.x;
      } else {
        pop_now = true;
      }
    }
    if (!pop_now && cur < 0) {
      if (tk == tend) {
        const uint32_t code = ~(uint32_t)cur;
        tk = code >> kLeafCountBits;
        tend = tk + (code & (kMaxLeafTris - 1)) + 1;
      }
      float t, u, v;
      uint32_t id;
      if (tri_test(r, tris, tk, &t, &u, &v, &id) && t >= r.near && t <= r.far) hit = true;
      ++tk;
      pop_now = !hit && tk == tend;
    }
    if (pop_now) {
      live = st.sp > base;
      if (live) cur = (int32_t)st.pop().x;
    }
  }
  st.sp = base;
  return hit;
}

// The same walks over the mesh's 4-wide tree (scenes committed with width 4): one 4-wide node per
// step, the nearest child next and the others pushed farthest first; same hits.
template <int BLOCK, int DEPTH>
__device__ bool mesh_closest4(const SceneView& S, const DevObject& ob, const RayState& ray, float far,
                              Stack<BLOCK, DEPTH>& st, int* status, float* t_out, float* u_out, float* v_out,
                              uint32_t* prim_out) {
  RayState r = ray;
  guard_once(r, ob.root_lo, ob.root_hi);
  float entry;
  if (!box_test(r, ob.root_lo[0], ob.root_hi[0], ob.root_lo[1], ob.root_hi[1], ob.root_lo[2], ob.root_hi[2], far,
                &entry))
    return false;
  const int base = st.sp;
  const DevNode4* nodes = S.blas4 + ob.node_base4;
  const DevTri* tris = S.tris + ob.tri_base;
  float min_t = FLT_MAX, cur_far = far;
  uint32_t best = 0xffffffffu;
  int32_t cur = ob.root_ref4;
  uint32_t tk = 0, tend = 0;  // triangles of the current leaf still to test: [tk, tend)
  bool live = true;
  while (live) {
    bool pop_now = false;
    if (cur >= 0) {
      int32_t r0, r1, r2, r3;
      float k0, k1, k2, k3;
      const int nh = node4_step(r, nodes, (uint32_t)cur, cur_far, r0, r1, r2, r3, k0, k1, k2, k3);
      if (nh == 0) {
        pop_now = true;
      } else {
        if (nh > 3 && !push_or_flag(st, r3, k3, status)) break;
        if (nh > 2 && !push_or_flag(st, r2, k2, status)) break;
        if (nh > 1 && !push_or_flag(st, r1, k1, status)) break;
        cur = r0;
      }
    }
    if (!pop_now && cur < 0) {
      if (tk == tend) {  // a leaf reached: ~cur = (first << 5) | (count - 1)
        const uint32_t code = ~(uint32_t)cur;
        tk = code >> kLeafCountBits;
        tend = tk + (code & (kMaxLeafTris - 1)) + 1;
      }
      float t, u, v;
      uint32_t id;
      if (tri_test(r, tris, tk, &t, &u, &v, &id) && t >= r.near && t <= far &&
          (t < min_t || (t == min_t && id < best))) {
        min_t = t;
        best = id;
        *t_out = t;
        *u_out = u;
        *v_out = v;
        cur_far = fminf(far, min_t);
      }
      ++tk;
      pop_now = tk == tend;
    }
    if (pop_now) {
      live = false;
      while (st.sp > base) {
        const uint2 e = st.pop();
        if (__uint_as_float(e.y) > min_t) continue;
        cur = (int32_t)e.x;
        live = true;
        break;
      }
    }
  }
  st.sp = base;
  *prim_out = best;
  return best != 0xffffffffu;
}

template <int BLOCK, int DEPTH>
__device__ bool mesh_any4(const SceneView& S, const DevObject& ob, const RayState& ray, Stack<BLOCK, DEPTH>& st,
                          int* status) {
  RayState r = ray;
  guard_once(r, ob.root_lo, ob.root_hi);
  float entry;
  if (!box_test(r, ob.root_lo[0], ob.root_hi[0], ob.root_lo[1], ob.root_hi[1], ob.root_lo[2], ob.root_hi[2], r.far,
                &entry))
    return false;
  const int base = st.sp;
  const DevNode4* nodes = S.blas4 + ob.node_base4;
  const DevTri* tris = S.tris + ob.tri_base;
  int32_t cur = ob.root_ref4;
  uint32_t tk = 0, tend = 0;  // triangles of the current leaf still to test: [tk, tend)
  bool hit = false, live = true;
  while (live && !hit) {
    bool pop_now = false;
    if (cur >= 0) {
      int32_t r0, r1, r2, r3;
      float k0, k1, k2, k3;
      const int nh = node4_step(r, nodes, (uint32_t)cur, r.far, r0, r1, r2, r3, k0, k1, k2, k3);
      if (nh == 0) {
        pop_now = true;
      } else {
        if (nh > 3 && !push_or_flag(st, r3, 0.f, status)) break;
        if (nh > 2 && !push_or_flag(st, r2, 0.f, status)) break;
        if (nh > 1 && !push_or_flag(st, r1, 0.f, status)) break;
        cur = r0;
      }
    }
    if (!pop_now && cur < 0) {
      if (tk == tend) {
        const uint32_t code = ~(uint32_t)cur;
        tk = code >> kLeafCountBits;
        tend = tk + (code & (kMaxLeafTris - 1)) + 1;
      }
      float t, u, v;
      uint32_t id;
      if (tri_test(r, tris, tk, &t, &u, &v, &id) && t >= r.near && t <= r.far) hit = true;
      ++tk;
      pop_now = !hit && tk == tend;
    }
    if (pop_now) {
      live = st.sp > base;
      if (live) cur = (int32_t)st.pop().x;
    }
  }
  st.sp = base;
  return hit;
}

// Object::intersect for leaf-order object k on the world ray; far = the nearest hit so far.
template <int BLOCK, int DEPTH, bool WIDE>
__device__ bool object_closest(const SceneView& S, uint32_t k, V wo, V wd, float near, float far,
                               Stack<BLOCK, DEPTH>& st, int* status, float* t, float* u, float* v, int32_t* prim,
                               int32_t* face) {
  const DevObject& ob = S.objs[k];
  V o = wo, d = wd;
  if (ob.has_xform) {  // transform.rs:98-103: the local direction is renormalised
    o = xpoint(ob.minv, wo);
    d = vnorm(xdir(ob.minv, wd));
  }
  *prim = -1;
  *u = 0.f;
  *v = 0.f;
  switch (ob.shape) {
    case kShapeCsg: return csg_first(S, ob, o, d, near, far, t, prim, face);
    case kShapeSphere: return sphere_t(o, d, ob.param[0], near, far, t);
    case kShapeBox: return box_t(o, d, ob.param, near, far, t, face);
    case kShapeRect: return rect_t(o, d, ob.param[0], ob.param[1], near, far, t);
    default: {
      const RayState r = ray_from(o, d, near, far);
      uint32_t id;
      if constexpr (WIDE) {
        if (!mesh_closest4(S, ob, r, far, st, status, t, u, v, &id)) return false;
      } else {
        if (!mesh_closest(S, ob, r, far, st, status, t, u, v, &id)) return false;
      }
      *prim = (int32_t)id;
      return true;
    }
  }
}

template <int BLOCK, int DEPTH, bool WIDE>
__device__ bool object_any(const SceneView& S, uint32_t k, V wo, V wd, float near, float far, Stack<BLOCK, DEPTH>& st,
                           int* status) {
  const DevObject& ob = S.objs[k];
  V o = wo, d = wd;
  if (ob.has_xform) {
    o = xpoint(ob.minv, wo);
    d = vnorm(xdir(ob.minv, wd));
  }
  float t;
  switch (ob.shape) {
    case kShapeCsg: {
      int32_t elem, face;
      return csg_first(S, ob, o, d, near, far, &t, &elem, &face);
    }
    case kShapeSphere: return sphere_t(o, d, ob.param[0], near, far, &t);
    case kShapeBox: {
      int face;
      return box_t(o, d, ob.param, near, far, &t, &face);
    }
    case kShapeRect: return rect_t(o, d, ob.param[0], ob.param[1], near, far, &t);
    default:
      if constexpr (WIDE)
        return mesh_any4(S, ob, ray_from(o, d, near, far), st, status);
      else
        return mesh_any(S, ob, ray_from(o, d, near, far), st, status);
  }
}

// Closest hit over the scene (bvh.rs:153-230 + scene.rs:186-220).
// WIDE: the meshes' 4-wide trees are walked (the scene has them); the binary trees otherwise.
template <int BLOCK, int DEPTH, bool WIDE>
__device__ bool scene_closest(const SceneView& S, V o, V d, float near, float far_in, Stack<BLOCK, DEPTH>& st,
                              int* status, SceneHitRec* out) {
  out->obj = -1;
  out->prim = -1;
  out->orig = -1;
  if (!S.has_tlas) return false;
  const float far = fminf(far_in, FLT_MAX);
  RayState r = ray_from(o, d, near, far);
  guard_once(r, S.root.lo, S.root.hi);
  float min_t = far;
  int32_t best_orig = INT_MAX, best_prim = INT_MAX;
  bool found = false;
  float entry;
  st.sp = 0;
  if (!box_test(r, S.root.lo[0], S.root.hi[0], S.root.lo[1], S.root.hi[1], S.root.lo[2], S.root.hi[2], min_t, &entry))
    return false;
  int32_t cur = S.root.ref;
  bool live = true;
  while (live) {
    bool pop_now = false;
    if (cur >= 0) {
      const DevNode* nd = S.tlas + cur;
      const float4 a = *reinterpret_cast<const float4*>(nd->xy0);
      const float4 b = *reinterpret_cast<const float4*>(nd->xy1);
      const float4 z = *reinterpret_cast<const float4*>(nd->z01);
      const int2 c = *reinterpret_cast<const int2*>(nd->ref);
      float dl, dr;
      const bool hl = box_test(r, a.x, a.y, a.z, a.w, z.x, z.y, min_t, &dl);
      const bool hr = box_test(r, b.x, b.y, b.z, b.w, z.z, z.w, min_t, &dr);
      if (hl && hr) {
        const bool right_first = dr < dl;
        if (!push_or_flag(st, right_first ? c.x : c.y, right_first ? dl : dr, status)) break;
        cur = right_first ? c.y : c.x;
      } else if (hl) {
        cur = c.x;
      } else if (hr) {
        cur = c.y;
      } else {
        pop_now = true;
      }
    }
    if (!pop_now && cur < 0) {
      const uint32_t code = ~(uint32_t)cur;
      const uint32_t first = code >> kLeafCountBits;
      const uint32_t cnt = (code & (kMaxLeafTris - 1)) + 1;
      for (uint32_t k = first; k < first + cnt; ++k) {
        float t, u, v;
        int32_t prim, face = 0;
        if (!object_closest<BLOCK, DEPTH, WIDE>(S, k, o, d, near, min_t, st, status, &t, &u, &v, &prim, &face))
          continue;
        const int32_t orig = S.objs[k].orig;
        const bool take = t < min_t || (t == min_t && (orig < best_orig || (orig == best_orig && prim < best_prim)));
        if (!take) continue;
        min_t = t;
        best_orig = orig;
        best_prim = prim;
        found = true;
        out->t = t;
        out->u = u;
        out->v = v;
        out->obj = (int32_t)k;
        out->prim = prim;
        out->orig = orig;
        out->face = face;
      }
      pop_now = true;
    }
    if (pop_now) {
      live = false;
      while (st.sp > 0) {
        const uint2 e = st.pop();
        if (__uint_as_float(e.y) > min_t) continue;
        cur = (int32_t)e.x;
        live = true;
        break;
      }
    }
  }
  return found;
}

// Any hit over the scene (bvh.rs:239-275 + scene.rs:157-168).
template <int BLOCK, int DEPTH, bool WIDE>
__device__ bool scene_any(const SceneView& S, V o, V d, float near, float far_in, Stack<BLOCK, DEPTH>& st,
                          int* status) {
  if (!S.has_tlas) return false;
  const float far = fminf(far_in, FLT_MAX);
  RayState r = ray_from(o, d, near, far);
  guard_once(r, S.root.lo, S.root.hi);
  float entry;
  st.sp = 0;
  if (!box_test(r, S.root.lo[0], S.root.hi[0], S.root.lo[1], S.root.hi[1], S.root.lo[2], S.root.hi[2], far, &entry))
    return false;
  int32_t cur = S.root.ref;
  for (;;) {
    bool pop_now = false;
    if (cur >= 0) {
      const DevNode* nd = S.tlas + cur;
      const float4 a = *reinterpret_cast<const float4*>(nd->xy0);
      const float4 b = *reinterpret_cast<const float4*>(nd->xy1);
      const float4 z = *reinterpret_cast<const float4*>(nd->z01);
      const int2 c = *reinterpret_cast<const int2*>(nd->ref);
      float dl, dr;
      const bool hl = box_test(r, a.x, a.y, a.z, a.w, z.x, z.y, far, &dl);
      const bool hr = box_test(r, b.x, b.y, b.z, b.w, z.z, z.w, far, &dr);
      if (hl && hr) {
        if (!push_or_flag(st, c.y, 0.f, status)) return false;
        cur = c.x;  // the reference pushes right then left: the left child is examined first
      } else if (hl) {
        cur = c.x;
      } else if (hr) {
        cur = c.y;
      } else {
        pop_now = true;
      }
    }
    if (!pop_now && cur < 0) {
      const uint32_t code = ~(uint32_t)cur;
      const uint32_t first = code >> kLeafCountBits;
      const uint32_t cnt = (code & (kMaxLeafTris - 1)) + 1;
      for (uint32_t k = first; k < first + cnt; ++k)
        if (object_any<BLOCK, DEPTH, WIDE>(S, k, o, d, near, far, st, status)) return true;
      pop_now = true;
    }
    if (pop_now) {
      if (st.sp == 0) return false;
      cur = (int32_t)st.pop().x;
    }
  }
}

template <int DEPTH, bool WIDE>
__global__ __launch_bounds__(kSceneBlock) void k_scene_trace(SceneView S, const float4* __restrict__ rays,
                                                             SceneHitRec* __restrict__ hits,
                                                             uint8_t* __restrict__ occluded, uint32_t n,
                                                             uint2* __restrict__ spill, uint32_t spill_cap,
                                                             int* __restrict__ status, int any_hit) {
  __shared__ uint2 lds[DEPTH][kSceneBlock];
  Stack<kSceneBlock, DEPTH> st{lds, spill, gridDim.x * kSceneBlock, spill_cap, threadIdx.x,
                               blockIdx.x * kSceneBlock + threadIdx.x, 0};
  for (uint32_t i = st.gtid; i < n; i += gridDim.x * kSceneBlock) {
    const float4 a = rays[2 * (size_t)i], b = rays[2 * (size_t)i + 1];
    const V o{a.x, a.y, a.z}, d{b.x, b.y, b.z};
    if (any_hit) {
      occluded[i] = scene_any<kSceneBlock, DEPTH, WIDE>(S, o, d, a.w, b.w, st, status) ? 1 : 0;
    } else {
      SceneHitRec h;
      scene_closest<kSceneBlock, DEPTH, WIDE>(S, o, d, a.w, b.w, st, status, &h);
      hits[i] = h;
    }
  }
}

// ---- shading ---------------------------------------------------------------------------------------
struct Isect {
  V p, n, wo, dpdu;
  float d;
  float u, v;  // surface coordinates (textures); 0 unless asked for
};

// The world-space intersection record of a hit (triangle_mesh.rs:73-132, sphere.rs:152-187,
// aabox.rs:31-60 + 155-190, rectangle.rs:45-54, transform.rs:105-116).
// Sphere surface coordinates of a local hit (sphere.rs:159-166): u = phi / 2 pi, v = theta / pi.
__device__ __forceinline__ void sphere_uv(V hit, float r, float* u, float* v) {
  float phi = atan2f(hit.y, hit.x);
  if (phi < 0.0f) phi += 2.0f * kPi;
  *u = phi / (2.0f * kPi);
  float zr = hit.z / r;
  zr = zr < -1.0f ? -1.0f : (zr > 1.0f ? 1.0f : zr);
  *v = acosf(zr) / kPi;
}

// uv: also the surface coordinates (sphere.rs:159-166, the box face table's du / dv, rectangle.rs:
// 50-51, the default triangle UVs of triangle_mesh.rs:219-224).
__device__ __forceinline__ Isect interaction(const SceneView& S, const SceneHitRec& h, V wo_ray, V wd, bool uv) {
  const DevObject& ob = S.objs[h.obj];
  V o = wo_ray, d = wd;
  if (ob.has_xform) {
    o = xpoint(ob.minv, wo_ray);
    d = vnorm(xdir(ob.minv, wd));
  }
  Isect is;
  is.d = h.t;
  is.u = is.v = 0.0f;
  is.wo = vmul(d, -1.0f);
  const V hit = vadd(o, vmul(d, h.t));
  switch (ob.shape) {
    case kShapeCsg: {
      // the element's record (sphere.rs:152-187 or the box face table) through its transform to
      // the CSG frame (transform.rs:105-116), the normal flipped for a subtracted element
      const DevCsgElem& e = S.csg[ob.mesh + h.prim];
      const V lo = xpoint(e.minv, o), ld = vnorm(xdir(e.minv, d));
      const V lh = vadd(lo, vmul(ld, h.t));
      V n, du;
      if (e.shape == kShapeSphere) {
        n = vnorm(lh);
        du = vmul(V{-lh.y, lh.x, 0.0f}, 2.0f * kPi);
        if (uv) sphere_uv(lh, e.param[0], &is.u, &is.v);
      } else {
        const int face = h.face & 7, axis = face >> 1;
        const float sg = (face & 1) ? 1.0f : -1.0f;
        n = V{axis == 0 ? sg : 0.0f, axis == 1 ? sg : 0.0f, axis == 2 ? sg : 0.0f};
        du = V{axis == 1 ? sg : 0.0f, axis == 2 ? sg : 0.0f, axis == 0 ? sg : 0.0f};
        if (uv) {
          is.u = vdot(lh, du);
          is.v = vdot(lh, V{axis == 2 ? 1.0f : 0.0f, axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f});
        }
      }
      is.p = xpoint(e.m, lh);
      is.n = xnormal(e.minv, n);
      if (h.face & 8) is.n = vmul(is.n, -1.0f);
      is.wo = xdir(e.m, vmul(ld, -1.0f));
      is.dpdu = xdir(e.m, du);
      break;
    }
    case kShapeSphere: {
      is.p = hit;
      is.n = vnorm(hit);
      is.dpdu = vmul(V{-hit.y, hit.x, 0.0f}, 2.0f * kPi);
      if (uv) sphere_uv(hit, ob.param[0], &is.u, &is.v);
      break;
    }
    case kShapeBox: {
      // the face table of aabox.rs:31-60: face 2a + s has normal (2s - 1) e_a and dpdu
      // (2s - 1) e_(a+2 mod 3); selected per component (a table indexed by the face would be
      // placed in private memory)
      const int face = h.face, axis = face >> 1;
      const float sg = (face & 1) ? 1.0f : -1.0f;
      is.p = hit;
      is.n = V{axis == 0 ? sg : 0.0f, axis == 1 ? sg : 0.0f, axis == 2 ? sg : 0.0f};
      is.dpdu = V{axis == 1 ? sg : 0.0f, axis == 2 ? sg : 0.0f, axis == 0 ? sg : 0.0f};
      if (uv) {  // dv = e_(a+1 mod 3)
        is.u = vdot(hit, is.dpdu);
        is.v = vdot(hit, V{axis == 2 ? 1.0f : 0.0f, axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f});
      }
      break;
    }
    case kShapeRect: {
      V p = hit;
      p.y = 0.0f;
      is.p = p;
      is.n = V{0.0f, 1.0f, 0.0f};
      is.dpdu = V{1.0f, 0.0f, 0.0f};
      is.u = p.x;
      is.v = p.z;
      break;
    }
    default: {
      const float* vb = S.verts + 3 * (size_t)ob.vert_base;
      const uint32_t* ib = S.idx + 3 * (size_t)(ob.idx_base + (uint32_t)h.prim);
      const V p0{vb[3 * ib[0]], vb[3 * ib[0] + 1], vb[3 * ib[0] + 2]};
      const V p1{vb[3 * ib[1]], vb[3 * ib[1] + 1], vb[3 * ib[1] + 2]};
      const V p2{vb[3 * ib[2]], vb[3 * ib[2] + 1], vb[3 * ib[2] + 2]};
      // compute_texture_derivatives with the default UVs (0,0), (1,0), (1,1): regular case, det = -1
      const float duv02x = -1.0f, duv02y = -1.0f, duv12x = 0.0f, duv12y = -1.0f;
      const V dp02 = vsub(p0, p2), dp12 = vsub(p1, p2);
      const float det = duv12x * duv02y - duv12y * duv02x;
      const float inv_det = 1.0f / det;
      const V dpdu = vmul(vsub(vmul(dp12, duv02y), vmul(dp02, duv12y)), inv_det);
      const V dpdv = vmul(vsub(vmul(dp02, duv12x), vmul(dp12, duv02x)), inv_det);
      is.p = hit;
      is.n = vnorm(vcross(dpdv, dpdu));
      is.dpdu = dpdu;
      if (uv) {  // barycentric interpolation of (0,0), (1,0), (1,1)
        const float w = 1.0f - h.u - h.v;
        is.u = w * 0.0f + h.u * 1.0f + h.v * 1.0f;
        is.v = w * 0.0f + h.u * 0.0f + h.v * 1.0f;
      }
      break;
    }
  }
  if (ob.has_xform) {
    is.n = xnormal(ob.minv, is.n);
    is.p = xpoint(ob.m, is.p);
    is.wo = xdir(ob.m, is.wo);
    is.dpdu = xdir(ob.m, is.dpdu);
  }
  return is;
}

// intersectable.rs:92-144
__device__ __forceinline__ V to_local(const Isect& i, V v) {
  const V ss = vnorm(i.dpdu);
  const V ts = vcross(i.n, ss);
  return {vdot(ss, v), vdot(ts, v), vdot(i.n, v)};
}
__device__ __forceinline__ V to_world(const Isect& i, V v) {
  const V ns = i.n, ss = vnorm(i.dpdu);
  const V ts = vcross(ns, ss);
  return {ss.x * v.x + ts.x * v.y + ns.x * v.z, ss.y * v.x + ts.y * v.y + ns.y * v.z,
          ss.z * v.x + ts.z * v.y + ns.z * v.z};
}

// One uniform of path `path` (= pixel * spp + sample). sobol_spp = 0: the counter-based draw;
// sobol_spp = spp (the Sobol sampler): the stratified pairs' draws come from the pixel's scrambled
// Sobol points at its sample index (rng.hpp), the others stay counter-based.
__device__ __forceinline__ float urand(uint64_t seed, uint64_t path, uint32_t bounce, uint32_t slot, uint32_t sobol_spp) {
  uint32_t axis = 0;
  const int pair = sobol_spp ? sobol_pair(bounce, slot, &axis) : -1;
  if (pair >= 0)
    return (float)(sobol_draw32(seed, path / sobol_spp, (uint32_t)(path % sobol_spp), (uint32_t)pair, axis) >> 8) *
           0x1p-24f;
  return to_unit_f32(hash4(seed, path, bounce, slot));
}
// The pixel jitter of the camera ray (renderers/mt.rs:123-127), under the same rule, in FP32.
__device__ __forceinline__ float jitter(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t spp, uint32_t slot,
                                       uint32_t sobol) {
  if (sobol) return (float)(sobol_draw32(seed, pixel, sample, 0, slot) >> 8) * 0x1p-24f;
  return to_unit_f32(hash4(seed, (uint64_t)pixel * spp + sample, 0, slot));
}
enum : uint32_t {
  kSlotJx = 0, kSlotJy = 1, kSlotLight = 2, kSlotLu = 3, kSlotLv = 4, kSlotS0 = 5, kSlotS1 = 6, kSlotCoin = 7,
  kSlotMetal0 = 8,  // Metal's point in the unit ball: attempt k uses slots kSlotMetal0 + 3k .. + 2
  kSlotRoulette = 8 + 3 * 64  // Russian roulette's draw, past the 64 Metal attempts
};
constexpr uint32_t kMetalAttempts = 64;

// CheckerBoard::shade (textures/checker_board.rs:17-29) at (u, v): true for the first colour.
// Rust's f64 % is fmod; signum is +-1 by the sign bit.
__device__ __forceinline__ bool checker_first(float u, float v, float scale) {
  const float su = fmodf(fabsf(u) * scale, 1.0f), sv = fmodf(fabsf(v) * scale, 1.0f);
  const bool use_color = (su < 0.5f && sv < 0.5f) || (su >= 0.5f && sv >= 0.5f);
  const bool positive = copysignf(1.0f, u) * copysignf(1.0f, v) >= 0.0f;
  return use_color != positive;
}

// BackgroundInfiniteLight::le (lights/background_infinite_light.rs:24-28), summed over the
// background lights.
__device__ __forceinline__ V background_le(const SceneView& S, V d) {
  V sum{0.f, 0.f, 0.f};
  const V u = vnorm(d);
  const float factor = 0.5f * (u.y + 1.0f);
#pragma unroll 1
  for (uint32_t i = 0; i < S.n_lights; ++i) {
    const DevLight& L = S.lights[i];
    if (L.kind != kLightBackground) continue;
    sum = {sum.x + (L.pos[0] * (1.0f - factor) + L.rgb[0] * factor),
           sum.y + (L.pos[1] * (1.0f - factor) + L.rgb[1] * factor),
           sum.z + (L.pos[2] * (1.0f - factor) + L.rgb[2] * factor)};
  }
  return sum;
}

// GGX microfacet reflection with Smith masking and Schlick Fresnel (albedo as F0)
__device__ __forceinline__ float ggx_d(float cos_h, float a) {
  const float a2 = a * a;
  const float c2 = cos_h * cos_h;
  const float den = c2 * (a2 - 1.0f) + 1.0f;
  return a2 / (kPi * den * den);
}
__device__ __forceinline__ float ggx_g1(float cos_v, float a) {
  const float c = fabsf(cos_v);
  const float a2 = a * a;
  return 2.0f * c / (c + sqrtf(a2 + (1.0f - a2) * c * c));
}
__device__ __forceinline__ V ggx_f(const DevMaterial& m, const Isect& is, V wo_w, V wi_w) {
  const V wo = to_local(is, wo_w), wi = to_local(is, wi_w);
  if (wo.z <= 0.0f || wi.z <= 0.0f) return {0.f, 0.f, 0.f};
  const V h = vnorm(vadd(wo, wi));
  const float a = m.roughness;
  const float D = ggx_d(h.z, a);
  const float G = ggx_g1(wo.z, a) * ggx_g1(wi.z, a);
  const float c = 1.0f - vdot(wo, h);
  const float c5 = c * c * c * c * c;
  const float s = D * G / (4.0f * wo.z * wi.z);
  return {(m.rgb[0] + (1.0f - m.rgb[0]) * c5) * s, (m.rgb[1] + (1.0f - m.rgb[1]) * c5) * s,
          (m.rgb[2] + (1.0f - m.rgb[2]) * c5) * s};
}

// Material::f (lambertian.rs:41-52)
__device__ __forceinline__ V mat_f(const DevMaterial& m, const Isect& is, V wo_w, V wi_w) {
  if (m.kind == kMatLambertian) {
    const V wi = to_local(is, wi_w), wo = to_local(is, wo_w);
    if (wo.z * wi.z > 0.0f) return {m.rgb[0] / kPi, m.rgb[1] / kPi, m.rgb[2] / kPi};
    return {0.f, 0.f, 0.f};
  }
  if (m.kind == kMatGGX) return ggx_f(m, is, wo_w, wi_w);
  return {0.f, 0.f, 0.f};
}

struct ScatterOut {
  V att, o, d;
  float pdf;
  bool specular;
};

// Material::scatter: lambertian.rs:24-38 (+ pdfs/cosine.rs:14-24), dielectric.rs:31-182. Each
// branch computes into locals and the output is written once (stores to different fields from
// different branches would otherwise keep the output in private memory).
__device__ __forceinline__ bool scatter(const DevMaterial& m, const Isect& is, uint64_t seed, uint64_t path,
                                        uint32_t bounce, uint32_t sobol, ScatterOut* s) {
  V att, org, dir;
  float pdf;
  bool spec;
  if (m.kind == kMatLambertian) {
    const float two_pi = 2.0f * kPi;
    const float phi = two_pi * urand(seed, path, bounce, kSlotS0, sobol);
    const float r2 = urand(seed, path, bounce, kSlotS1, sobol);
    const float r2s = sqrtf(r2);
    const V lwi{cosf(phi) * r2s, sinf(phi) * r2s, sqrtf(1.0f - r2)};
    const V wi = to_world(is, lwi);
    org = vadd(is.p, vmul(is.n, 0.001f));
    dir = vnorm(wi);
    att = {m.rgb[0] / kPi, m.rgb[1] / kPi, m.rgb[2] / kPi};
    pdf = fabsf(lwi.z / kPi);
    spec = false;
  } else if (m.kind == kMatDielectric) {
    const V lwo = vnorm(to_local(is, is.wo));
    const V lrefl{-lwo.x, -lwo.y, lwo.z};
    const bool leaving = lwo.z <= 0.0f;
    const V lnorm{0.f, 0.f, leaving ? -1.f : 1.f};
    const V wnorm = leaving ? vmul(is.n, -1.0f) : is.n;
    const float ni = leaving ? m.ior : 1.0f, nt = leaving ? 1.0f : m.ior;
    const float eta = ni / nt;
    // refract (dielectric.rs:129-157)
    const float ci = vdot(lwo, lnorm);
    const float s2i = 1.0f - ci * ci;
    const float s2t = eta * eta * s2i;
    const float disc = 1.0f - s2t;
    bool reflect = true;
    V lrefr{0.f, 0.f, 0.f};
    if (disc > 0.0f) {
      const float ct = sqrtf(disc);
      lrefr = vnorm(vadd(vmul(vmul(lwo, -1.0f), eta), vmul(lnorm, eta * ci - ct)));
      // fresnel (dielectric.rs:168-182)
      const float e = ni / nt;
      const float fci = lwo.z * lnorm.z;
      const float fct = sqrtf(1.0f - e * e * (1.0f - fci * fci));
      const float ro = (ni * fci - nt * fct) / (ni * fci + nt * fct);
      const float rp = (nt * fci - ni * fct) / (nt * fci + ni * fct);
      const float refl = (ro * ro + rp * rp) * 0.5f;
      reflect = urand(seed, path, bounce, kSlotCoin, sobol) < refl;
    }
    const V ldir = reflect ? lrefl : lrefr;
    org = vadd(is.p, vmul(wnorm, reflect ? 0.001f : -0.001f));
    dir = vnorm(to_world(is, ldir));
    const float c = fabsf(ldir.z);
    att = {m.rgb[0] / c, m.rgb[1] / c, m.rgb[2] / c};
    pdf = 1.0f;
    spec = true;
  } else if (m.kind == kMatMetal) {
    // metal.rs:22-54: the mirror direction plus fuzz times a point of the unit ball (utils.rs:19-27,
    // by rejection; attempt k draws three uniforms of its own)
    const V lwo = vnorm(to_local(is, is.wo));
    const V lrefl{-lwo.x, -lwo.y, lwo.z};
    V q{0.f, 0.f, 0.f};
#pragma unroll 1
    for (uint32_t k = 0; k < kMetalAttempts; ++k) {
      const V c{urand(seed, path, bounce, kSlotMetal0 + 3 * k, sobol) * 2.0f - 1.0f,
                urand(seed, path, bounce, kSlotMetal0 + 3 * k + 1, sobol) * 2.0f - 1.0f,
                urand(seed, path, bounce, kSlotMetal0 + 3 * k + 2, sobol) * 2.0f - 1.0f};
      if (vdot(c, c) < 1.0f) {
        q = c;
        break;
      }
    }
    const V lt = vnorm(vadd(lrefl, vmul(q, m.roughness)));
    if (!(lt.z > 0.0f)) return false;
    org = vadd(is.p, vmul(is.n, 0.001f));
    dir = vnorm(to_world(is, lt));
    const float c = fabsf(lt.z);
    att = {m.rgb[0] / c, m.rgb[1] / c, m.rgb[2] / c};
    pdf = 1.0f;
    spec = true;
  } else if (m.kind == kMatGGX) {
    const V wo = to_local(is, is.wo);
    if (wo.z <= 0.0f) return false;
    const float a = m.roughness;
    const float u1 = urand(seed, path, bounce, kSlotS0, sobol), u2 = urand(seed, path, bounce, kSlotS1, sobol);
    const float phi = 2.0f * kPi * u1;
    const float cos_t = sqrtf((1.0f - u2) / (1.0f + (a * a - 1.0f) * u2));
    const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    const V h{sin_t * cosf(phi), sin_t * sinf(phi), cos_t};
    const float woh = vdot(wo, h);
    const V wi = vsub(vmul(h, 2.0f * woh), wo);
    if (wi.z <= 0.0f) return false;
    pdf = ggx_d(h.z, a) * h.z / (4.0f * woh);
    att = ggx_f(m, is, is.wo, to_world(is, wi));
    org = vadd(is.p, vmul(is.n, 0.001f));
    dir = vnorm(to_world(is, wi));
    spec = false;
  } else {
    return false;
  }
  s->att = att;
  s->o = org;
  s->d = dir;
  s->pdf = pdf;
  s->specular = spec;
  return true;
}

struct LiOut {
  V li, wi, so, sd;
  float pdf, dist;
};

// Light::sample_li: point_light.rs:23-35 + lights.rs:38-46; area light: uniform point on the
// rectangle, one-sided emission, solid-angle pdf d^2 / (A cos) (ideas/area_light.md §3).
__device__ __forceinline__ bool sample_li(const SceneView& S, const DevLight& L, const Isect& is, uint64_t seed, uint64_t path,
                          uint32_t bounce, uint32_t sobol, LiOut* s) {
  if (L.kind == kLightPoint) {
    const V pos{L.pos[0], L.pos[1], L.pos[2]};
    const V wi = vsub(pos, is.p);
    const float d2 = vdot(wi, wi);
    s->li = {L.rgb[0] / d2, L.rgb[1] / d2, L.rgb[2] / d2};
    s->wi = vnorm(wi);
    s->pdf = 1.0f;
    s->so = is.p;
    s->sd = vnorm(vnorm(vsub(pos, is.p)));
    s->dist = vlen(vsub(pos, is.p));
    return true;
  }
  if (L.kind == kLightBackground) {
    // background_infinite_light.rs:31-44: a direction uniform on the sphere (pdfs/sphere.rs:14-23),
    // VisibilityTester::towards_infinity (lights.rs:57-62)
    const float phi = 2.0f * kPi * urand(seed, path, bounce, kSlotLu, sobol);
    const float cos_t = 1.0f - 2.0f * urand(seed, path, bounce, kSlotLv, sobol);
    const float sin_t = sqrtf(1.0f - cos_t * cos_t);
    const V wi{cosf(phi) * sin_t, sinf(phi) * sin_t, cos_t};
    const float factor = 0.5f * (vnorm(wi).y + 1.0f);
    s->li = {L.pos[0] * (1.0f - factor) + L.rgb[0] * factor, L.pos[1] * (1.0f - factor) + L.rgb[1] * factor,
             L.pos[2] * (1.0f - factor) + L.rgb[2] * factor};
    s->wi = wi;
    s->pdf = 1.0f / (4.0f * kPi);
    s->so = is.p;
    s->sd = vnorm(wi);
    s->dist = FLT_MAX;
    return true;
  }
  const DevObject& o = S.objs[L.object];
  const float w = o.param[0], h = o.param[1];
  const float u1 = urand(seed, path, bounce, kSlotLu, sobol), u2 = urand(seed, path, bounce, kSlotLv, sobol);
  const V pl_local{(u1 - 0.5f) * w, 0.0f, (u2 - 0.5f) * h};
  const V pl = o.has_xform ? xpoint(o.m, pl_local) : pl_local;
  const V nl = o.has_xform ? vnorm(xnormal(o.minv, V{0.f, 1.f, 0.f})) : V{0.f, 1.f, 0.f};
  const V dv = vsub(pl, is.p);
  const float dist2 = vdot(dv, dv);
  const float dist = sqrtf(dist2);
  const V wi = vmul(dv, 1.0f / dist);
  const float cos_l = -vdot(nl, wi);
  if (!(cos_l > 0.0f)) return false;
  const float sx = o.has_xform ? vlen(xdir(o.m, V{1.f, 0.f, 0.f})) : 1.0f;
  const float sz = o.has_xform ? vlen(xdir(o.m, V{0.f, 0.f, 1.f})) : 1.0f;
  const float area = w * sx * h * sz;
  s->li = {L.rgb[0], L.rgb[1], L.rgb[2]};
  s->wi = wi;
  s->pdf = dist2 / (area * cos_l);
  s->so = is.p;
  s->sd = wi;
  s->dist = dist - 1e-4f;
  return true;
}

// Multiple importance sampling (the reference's plan after the area light, ideas/area_light.md:
// 158-165): the densities, in solid angle, of the two strategies that can produce a direction.
// BSDF: the density with which scatter() produces wi_w (0 for specular and non-scattering materials).
__device__ __forceinline__ float bsdf_pdf(const DevMaterial& m, const Isect& is, V wo_w, V wi_w) {
  if (m.kind == kMatLambertian) {
    const float c = vdot(is.n, wi_w);
    return c > 0.0f ? c / kPi : 0.0f;
  }
  if (m.kind == kMatGGX) {
    const V wo = to_local(is, wo_w), wi = to_local(is, wi_w);
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    const V h = vnorm(vadd(wo, wi));
    return ggx_d(h.z, m.roughness) * h.z / (4.0f * vdot(wo, h));
  }
  return 0.0f;
}
// Light: the density, with the choice of light, with which next-event estimation picks direction d
// towards the area light on object k met at distance t (sample_li's pdf); 0 if no area light sits
// on k or d reaches its back.
__device__ __forceinline__ float light_pdf(const SceneView& S, uint32_t k, V d, float t) {
  for (uint32_t i = 0; i < S.n_lights; ++i) {
    const DevLight& L = S.lights[i];
    if (L.kind == kLightPoint || L.object != k) continue;
    const DevObject& o = S.objs[k];
    const V nl = o.has_xform ? vnorm(xnormal(o.minv, V{0.f, 1.f, 0.f})) : V{0.f, 1.f, 0.f};
    const float cos_l = -vdot(nl, d);
    if (!(cos_l > 0.0f)) return 0.0f;
    const float sx = o.has_xform ? vlen(xdir(o.m, V{1.f, 0.f, 0.f})) : 1.0f;
    const float sz = o.has_xform ? vlen(xdir(o.m, V{0.f, 0.f, 1.f})) : 1.0f;
    const float area = o.param[0] * sx * o.param[1] * sz;
    return (1.0f / (float)S.n_lights) * (t * t / (area * cos_l));
  }
  return 0.0f;
}
// Power heuristic a^2 / (a^2 + b^2) as 1 / (1 + (b / a)^2): an infinite density (a light sample at
// a grazing angle) gives 1 or 0 rather than inf / inf; 0 when a is 0, 1/2 when both are infinite.
__device__ __forceinline__ float mis_power(float a, float b) {
  if (!(a > 0.0f)) return 0.0f;
  const float r = b / a;
  if (r != r) return 0.5f;
  return 1.0f / (1.0f + r * r);
}

// ---- wavefront kernels -------------------------------------------------------------------------------
// PinHoleCamera::get_ray (pin_hole.rs:42-50) with Matrix4::transform_point / transform_direction
// (matrix4.rs:187-206: the w-divide only when w != 1) and the normalisations of get_ray and
// Ray::new, in FP32: the matrices are built in f64 on the host and rounded once.
__device__ __forceinline__ void cam_point(const float* M, float x, float y, float z, float* r) {
  float rx = M[0] * x + M[1] * y + M[2] * z + M[3];
  float ry = M[4] * x + M[5] * y + M[6] * z + M[7];
  float rz = M[8] * x + M[9] * y + M[10] * z + M[11];
  const float w = M[12] * x + M[13] * y + M[14] * z + M[15];
  if (w != 1.0f) {
    const float iw = 1.0f / w;
    rx = rx * iw;
    ry = ry * iw;
    rz = rz * iw;
  }
  r[0] = rx;
  r[1] = ry;
  r[2] = rz;
}
__device__ __forceinline__ void fnorm3(float* v) {
  const float n = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const float inv = 1.0f / sqrtf(n);
  v[0] = v[0] * inv;
  v[1] = v[1] * inv;
  v[2] = v[2] * inv;
}
// The camera ray through pixel coordinate (x, y) (the pixel's corner plus its jitter).
__device__ __forceinline__ void camera_ray(const float* r2s, const float* c2w, float x, float y, float* o, float* d) {
  float cv[3];
  cam_point(r2s, x, y, 0.0f, cv);
  fnorm3(cv);
  cam_point(c2w, 0.0f, 0.0f, 0.0f, o);
  d[0] = c2w[0] * cv[0] + c2w[1] * cv[1] + c2w[2] * cv[2];
  d[1] = c2w[4] * cv[0] + c2w[5] * cv[1] + c2w[6] * cv[2];
  d[2] = c2w[8] * cv[0] + c2w[9] * cv[1] + c2w[10] * cv[2];
  fnorm3(d);
}

// One camera path per (pixel, sample of this pass): PinHoleCamera::get_ray (pin_hole.rs:42-50) in
// FP32 through the pixel jittered by two uniforms (renderers/mt.rs:123-127).
template <bool NRC>
__global__ void k_generate(RenderParams p, PathBuffers b) {
  const uint32_t npix = p.width * p.height;
  const uint32_t n = npix * p.pass_samples;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    uint32_t rank;
    if (NRC && is_training_path(p, j, &rank)) {
      for (uint32_t f = 0; f < 3; ++f)
        for (uint32_t m = 0; m < p.max_depth; ++m)
          b.rec[((size_t)f * p.max_depth + m) * b.train_cap + rank] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const uint32_t pixel = j % npix, sample = p.pass_first_sample + j / npix;
    const float jx = jitter(p.seed, pixel, sample, p.spp, kSlotJx, p.sampler),
                jy = jitter(p.seed, pixel, sample, p.spp, kSlotJy, p.sampler);
    float o[3], d[3];
    camera_ray(p.r2s, p.c2w, (float)(pixel % p.width) + jx, (float)(pixel / p.width) + jy, o, d);
    b.ray_o[j] = make_float4(o[0], o[1], o[2], p.near);
    b.ray_d[j] = make_float4(d[0], d[1], d[2], p.far);
    b.beta[j] = make_float4(1.f, 1.f, 1.f, 1.f);
    b.L[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    b.info[j] = make_uint4(pixel, sample, 0u, 1u);
    b.queue[0][j] = j;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *counter(b.counts, kCntQueue0) = n;
    *counter(b.counts, kCntQueue1) = 0;
    *counter(b.counts, kCntShadow) = 0;
    if (NRC) *counter(b.counts, kCntQuery) = 0;
  }
}

__global__ void k_reset(PathBuffers b, int q) {
  if (threadIdx.x == 0) {
    *counter(b.counts, q ^ 1) = 0;
    *counter(b.counts, kCntShadow) = 0;
  }
}

// counter kCntRays holds a 64-bit count of every ray traced (extension and shadow)
__device__ __forceinline__ void count_rays(PathBuffers& b, uint32_t n) {
  if (blockIdx.x == 0 && threadIdx.x == 0)
    atomicAdd(reinterpret_cast<unsigned long long*>(counter(b.counts, kCntRays)), (unsigned long long)n);
}

// OCC5: compiled for five waves per SIMD (96 registers, a few spilled) instead of the compiler's
// own choice (101 registers, four waves): the walk is bound by the latency of dependent node loads,
// and the extra wave pays on mesh-heavy scenes, less so where CSG objects need the registers. The
// scene's launch settings pick the variant.
template <int DEPTH, bool OCC5, bool WIDE>
__global__ __launch_bounds__(kSceneBlock) __attribute__((amdgpu_waves_per_eu(OCC5 ? 5 : 1))) void k_extend(SceneView S, PathBuffers b, int q, uint2* __restrict__ spill,
                                                        uint32_t spill_cap, int* __restrict__ status) {
  __shared__ uint2 lds[DEPTH][kSceneBlock];
  Stack<kSceneBlock, DEPTH> st{lds, spill, gridDim.x * kSceneBlock, spill_cap, threadIdx.x,
                               blockIdx.x * kSceneBlock + threadIdx.x, 0};
  const uint32_t n = *counter(b.counts, q);
  count_rays(b, n);
  for (uint32_t i = st.gtid; i < n; i += gridDim.x * kSceneBlock) {
    const uint32_t j = b.queue[q][i];
    const float4 o = b.ray_o[j], d = b.ray_d[j];
    SceneHitRec h;
    scene_closest<kSceneBlock, DEPTH, WIDE>(S, V{o.x, o.y, o.z}, V{d.x, d.y, d.z}, o.w, d.w, st, status, &h);
    b.hit[j] = h;
  }
}

// Slots in a queue for the lanes of a wave that want one: one atomic per wave (the lowest
// requesting lane adds the wave's count), each lane's slot at its rank among the requesting lanes.
// Every lane of the wave must call it (a convergent point).
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool want) {
  const uint64_t mask = __ballot(want);
  if (mask == 0) return 0;
  const uint32_t lane = __lane_id();
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = (uint32_t)__shfl((int)base, (int)leader);
  return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

// What one shaded vertex appends to the queues. A vertex queues a shadow ray or a cache query,
// never both, so the two share w0..w2 (fewer registers live across the queue appends).
struct ShadeOut {
  bool shadow = false, next = false, query = false;
  float4 w0, w1, w2;  // shadow: origin, direction, contribution (path index bits in w); query: raw inputs
  float4 qb;          // query: throughput (path index bits in w)
};

// One path vertex (path.rs:55-112): emission after a specular bounce, the bounce cut, next-event
// estimation (its shadow ray queued with its contribution), then scattering into the next queue.
// With the radiance cache (NRC): a training path records the vertex; any other path ends into the
// cache at its first non-specular, non-emissive vertex at or past p.query_vertex, queueing the
// query in place of next-event estimation and scattering.
template <bool NRC>
__device__ __forceinline__ void shade_vertex(const SceneView& S, const RenderParams& p, const PathBuffers& b,
                                             uint32_t j, ShadeOut* out) {
  const SceneHitRec h = b.hit[j];
  const bool naive = p.integrator == PBRT_INTEGRATOR_NAIVE;
  if (p.integrator == PBRT_INTEGRATOR_NORMAL) {
    // normal.rs:17-30: the camera ray's hit normal mapped to (n + 1) / 2 (black on a miss); the path
    // ends at its first vertex
    if (h.obj >= 0) {
      const float4 o4 = b.ray_o[j], d4 = b.ray_d[j];
      const Isect is = interaction(S, h, V{o4.x, o4.y, o4.z}, V{d4.x, d4.y, d4.z}, false);
      b.L[j] = make_float4((is.n.x + 1.0f) * 0.5f, (is.n.y + 1.0f) * 0.5f, (is.n.z + 1.0f) * 0.5f, 0.f);
    }
    return;
  }
  if (h.obj < 0 && S.n_background == 0 && !naive) return;  // the path leaves the scene: it ends
  uint4 info = b.info[j];
  // the path's job (a batch of jobs) or the render's own parameters
  uint32_t spp = p.spp, max_depth = p.max_depth;
  uint64_t seed = p.seed;
  float near = p.near, far = p.far;
  if (p.n_jobs) {
    const JobDesc& J = p.jobs[info.w];
    spp = J.spp;
    max_depth = J.max_depth;
    seed = J.seed;
    near = J.near;
    far = J.far;
  }
  const uint64_t path = (uint64_t)info.x * spp + info.y;
  const uint32_t sobol = p.sampler ? spp : 0u;
  float4 beta4 = b.beta[j], L4 = b.L[j];
  const float4 o4 = b.ray_o[j], d4 = b.ray_d[j];
  const V rd{d4.x, d4.y, d4.z};
  // beta.w: 1 after a specular bounce (and for the camera ray), else minus the BSDF density of the
  // direction that led here (0 without multiple importance sampling)
  const bool specular = beta4.w > 0.0f;
  if (h.obj < 0 && naive) {
    // naive.rs:21-26: the integrator's own sky, white to (0.5, 0.7, 1.0) with the direction's y
    const V u = vnorm(rd);
    const float t = 0.5f * (u.y + 1.0f);
    const V sky{1.0f * (1.0f - t) + 0.5f * t, 1.0f * (1.0f - t) + 0.7f * t, 1.0f * (1.0f - t) + 1.0f * t};
    b.L[j] = make_float4(L4.x + sky.x * beta4.x, L4.y + sky.y * beta4.y, L4.z + sky.z * beta4.z, L4.w);
    return;
  }
  if (h.obj < 0) {
    // path.rs:101-108: the background's radiance after a specular bounce; with MIS also after a
    // non-specular one, weighted against next-event estimation's density
    V e{0.f, 0.f, 0.f};
    if (specular || p.mis) {
      const V le = background_le(S, rd);
      const float w = specular ? 1.0f : mis_power(-beta4.w, (1.0f / (float)S.n_lights) * (1.0f / (4.0f * kPi)));
      e = specular ? V{le.x * beta4.x, le.y * beta4.y, le.z * beta4.z}
                   : V{(le.x * beta4.x) * w, (le.y * beta4.y) * w, (le.z * beta4.z) * w};
      L4.x = L4.x + e.x;
      L4.y = L4.y + e.y;
      L4.z = L4.z + e.z;
      b.L[j] = L4;
    }
    uint32_t rank;
    if (NRC && info.z + 1 <= max_depth && is_training_path(p, j, &rank)) {
      // the emission of the vertex the path would have reached, for the training targets
      const size_t fs = (size_t)p.max_depth * b.train_cap;
      b.rec[kRecEmit * fs + (size_t)info.z * b.train_cap + rank] = make_float4(e.x, e.y, e.z, 0.f);
    }
    return;
  }
  const DevMaterial m0 = S.mats[S.objs[h.obj].material];
  const Isect is = interaction(S, h, V{o4.x, o4.y, o4.z}, rd, m0.tex == kTexChecker);
  DevMaterial m = m0;  // with its texture evaluated at the hit
  if (m0.tex == kTexChecker && !checker_first(is.u, is.v, m0.scale)) {
    m.rgb[0] = m0.rgb2[0];
    m.rgb[1] = m0.rgb2[1];
    m.rgb[2] = m0.rgb2[2];
  }
  V emitted{0.f, 0.f, 0.f};
  if ((specular || naive) && m.kind == kMatDiffuseLight) {  // naive.rs:36-39: at every hit
    emitted = {m.rgb[0] * beta4.x, m.rgb[1] * beta4.y, m.rgb[2] * beta4.z};
    L4.x = L4.x + emitted.x;
    L4.y = L4.y + emitted.y;
    L4.z = L4.z + emitted.z;
  } else if (p.mis && m.kind == kMatDiffuseLight) {
    // the emission a BSDF sample met, weighted against next-event estimation's density
    const float w = mis_power(-beta4.w, light_pdf(S, (uint32_t)h.obj, rd, h.t));
    emitted = {(m.rgb[0] * beta4.x) * w, (m.rgb[1] * beta4.y) * w, (m.rgb[2] * beta4.z) * w};
    L4.x = L4.x + emitted.x;
    L4.y = L4.y + emitted.y;
    L4.z = L4.z + emitted.z;
  }
  info.z += 1;
  const uint32_t bounce = info.z;
  b.L[j] = L4;
  b.info[j] = info;
  if (NRC) {
    const bool eligible = m.kind == kMatLambertian || m.kind == kMatGGX;
    uint32_t rank;
    const bool training = is_training_path(p, j, &rank);
    if (training || (eligible && bounce >= p.query_vertex && bounce < p.max_depth)) {
      const V wo = vmul(rd, -1.0f);
      const float4 raw0 = make_float4(is.p.x, is.p.y, is.p.z, wo.x);
      const float4 raw1 = make_float4(wo.y, wo.z, is.n.x, is.n.y);
      const float4 raw2 = make_float4(is.n.z, m.rgb[0], m.rgb[1], m.rgb[2]);
      if (training) {
        const size_t cap = b.train_cap, base = (size_t)(bounce - 1) * cap + rank, fs = (size_t)p.max_depth * cap;
        b.rec[kRecBeta * fs + base] = make_float4(beta4.x, beta4.y, beta4.z, eligible ? 3.0f : 1.0f);
        b.rec[kRecEmit * fs + base] = make_float4(emitted.x, emitted.y, emitted.z, 0.f);
        b.rec[(kRecRaw + 0) * fs + base] = raw0;
        b.rec[(kRecRaw + 1) * fs + base] = raw1;
        b.rec[(kRecRaw + 2) * fs + base] = raw2;
      } else {
        out->query = true;  // the cache stands for everything this path would gather from here on
        out->w0 = raw0;
        out->w1 = raw1;
        out->w2 = raw2;
        out->qb = make_float4(beta4.x, beta4.y, beta4.z, __uint_as_float(j));
        return;
      }
    }
  }
  if (bounce == max_depth) return;
  if (S.n_lights > 0 && !naive) {
    const float u = urand(seed, path, bounce, kSlotLight, sobol);
    const float fi = fminf(u * (float)S.n_lights, (float)S.n_lights - 1.0f);
    const DevLight Ld = S.lights[(uint32_t)fi];
    const float pl = 1.0f / (float)S.n_lights;
    LiOut ls;
    if (sample_li(S, Ld, is, seed, path, bounce, sobol, &ls)) {
      const V f0 = mat_f(m, is, vmul(rd, -1.0f), ls.wi);
      const float c = fabsf(vdot(ls.wi, is.n));
      const V f{f0.x * c, f0.y * c, f0.z * c};
      const float den = pl * ls.pdf;
      V contrib{((ls.li.x * f.x) * beta4.x) / den, ((ls.li.y * f.y) * beta4.y) / den,
                ((ls.li.z * f.z) * beta4.z) / den};
      if (p.mis && Ld.kind != kLightPoint) {
        const float w = mis_power(den, bsdf_pdf(m, is, vmul(rd, -1.0f), ls.wi));
        contrib = {contrib.x * w, contrib.y * w, contrib.z * w};
      }
      out->shadow = true;
      out->w0 = make_float4(ls.so.x, ls.so.y, ls.so.z, 1e-5f);
      out->w1 = make_float4(ls.sd.x, ls.sd.y, ls.sd.z, ls.dist);
      out->w2 = make_float4(contrib.x, contrib.y, contrib.z, __uint_as_float(j));
    }
  }
  ScatterOut sc;
  if (!scatter(m, is, seed, path, bounce, sobol, &sc)) return;
  const float c = fabsf(vdot(sc.d, is.n));
  beta4.x = beta4.x * ((sc.att.x * c) / sc.pdf);
  beta4.y = beta4.y * ((sc.att.y * c) / sc.pdf);
  beta4.z = beta4.z * ((sc.att.z * c) / sc.pdf);
  // a grazing specular event can give 1/|cos| = inf against cos = 0 in FP32: such a path carries
  // nothing further (the reference, in f64, would carry a NaN); end it
  if (!(isfinite(beta4.x) && isfinite(beta4.y) && isfinite(beta4.z))) return;
  if (!NRC && p.roulette) {
    // path.rs:93-100 (commented out in the reference): continue with probability q, the largest
    // throughput component capped at 1, the throughput divided by q. The path survives only when
    // the draw is below q, so a throughput of zero (q = 0) always ends it (never 0 / 0)
    const float q = fminf(1.0f, fmaxf(fmaxf(beta4.x, beta4.y), beta4.z));
    if (!(urand(seed, path, bounce, kSlotRoulette, sobol) < q)) return;
    beta4.x = beta4.x / q;
    beta4.y = beta4.y / q;
    beta4.z = beta4.z / q;
  }
  beta4.w = sc.specular ? 1.0f : (p.mis ? -sc.pdf : 0.0f);
  b.beta[j] = beta4;
  b.ray_o[j] = make_float4(sc.o.x, sc.o.y, sc.o.z, near);
  b.ray_d[j] = make_float4(sc.d.x, sc.d.y, sc.d.z, far);
  out->next = true;
}

// Queue slots for a whole workgroup of kShadeBlock threads: per-wave counts by ballot, one atomic
// per queue per workgroup, each lane's slot at its rank in (wave, lane) order. Every thread of the
// workgroup must call it.
constexpr int kShadeBlock = 256;
template <int K>
__device__ __forceinline__ void block_append(uint32_t* const* ctr, const bool* want, uint32_t* slot) {
  __shared__ uint32_t wcount[K][kShadeBlock / 64];
  __shared__ uint32_t base[K];
  const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
  uint64_t mask[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    mask[k] = __ballot(want[k]);
    if (lane == 0) wcount[k][wave] = (uint32_t)__popcll(mask[k]);
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    uint32_t total = 0;
    for (int w = 0; w < kShadeBlock / 64; ++w) total += wcount[k][w];
    base[k] = total ? atomicAdd(ctr[k], total) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    uint32_t off = base[k];
    for (uint32_t w = 0; w < wave; ++w) off += wcount[k][w];
    slot[k] = off + (uint32_t)__popcll(mask[k] & ((1ull << lane) - 1ull));
  }
  __syncthreads();  // the shared counts are reused by the next call
}

// The vertices of queue q: the grid strides over the queue a workgroup at a time, so every thread
// of a workgroup reaches the queue appends together. Four waves per SIMD (128 registers, no
// spill): shading is bound by the latency of its scattered loads, and the compiler's own choice
// (about 150 registers) allows three.
template <bool NRC>
__global__ __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_shade(SceneView S, RenderParams p, PathBuffers b, int q) {
  const uint32_t n = *counter(b.counts, q);
  for (uint32_t base = blockIdx.x * kShadeBlock; base < n; base += gridDim.x * kShadeBlock) {
    const uint32_t i = base + threadIdx.x;
    ShadeOut o;
    uint32_t j = 0;
    if (i < n) {
      j = b.queue[q][i];
      shade_vertex<NRC>(S, p, b, j, &o);
    }
    constexpr int K = NRC ? 3 : 2;
    uint32_t* const ctr[3] = {counter(b.counts, kCntShadow), counter(b.counts, q ^ 1), counter(b.counts, kCntQuery)};
    const bool want[3] = {o.shadow, o.next, o.query};
    uint32_t slot[3] = {0, 0, 0};
    block_append<K>(ctr, want, slot);
    if (o.shadow) {
      b.sh_o[slot[0]] = o.w0;
      b.sh_d[slot[0]] = o.w1;
      b.sh_c[slot[0]] = o.w2;
    }
    if (o.next) b.queue[q ^ 1][slot[1]] = j;
    if (NRC && o.query) {
      b.q_raw[3 * (size_t)slot[2] + 0] = o.w0;
      b.q_raw[3 * (size_t)slot[2] + 1] = o.w1;
      b.q_raw[3 * (size_t)slot[2] + 2] = o.w2;
      b.q_beta[slot[2]] = o.qb;
    }
  }
}

// VisibilityTester::unoccluded (lights.rs:67-69): add the contribution when nothing blocks it.
template <int DEPTH, bool NRC, bool WIDE>
__global__ __launch_bounds__(kSceneBlock) void k_shadow(SceneView S, RenderParams p, PathBuffers b,
                                                        uint2* __restrict__ spill, uint32_t spill_cap,
                                                        int* __restrict__ status) {
  __shared__ uint2 lds[DEPTH][kSceneBlock];
  Stack<kSceneBlock, DEPTH> st{lds, spill, gridDim.x * kSceneBlock, spill_cap, threadIdx.x,
                               blockIdx.x * kSceneBlock + threadIdx.x, 0};
  const uint32_t n = *counter(b.counts, kCntShadow);
  count_rays(b, n);
  for (uint32_t i = st.gtid; i < n; i += gridDim.x * kSceneBlock) {
    const float4 o = b.sh_o[i], d = b.sh_d[i], c = b.sh_c[i];
    if (scene_any<kSceneBlock, DEPTH, WIDE>(S, V{o.x, o.y, o.z}, V{d.x, d.y, d.z}, o.w, d.w, st, status)) continue;
    const uint32_t j = __float_as_uint(c.w);
    float4 L = b.L[j];
    L.x = L.x + c.x;
    L.y = L.y + c.y;
    L.z = L.z + c.z;
    b.L[j] = L;
    uint32_t rank;
    if (NRC && is_training_path(p, j, &rank)) {  // one shadow ray per vertex: the vertex's own record
      const uint32_t m = b.info[j].z;
      b.rec[((size_t)kRecNee * p.max_depth + (m - 1)) * b.train_cap + rank] = make_float4(c.x, c.y, c.z, 0.f);
    }
  }
}

// The job of path j (or pixel j) of a batch pass: the last job whose base is at most j.
__device__ __forceinline__ uint32_t job_of(const JobDesc* jobs, uint32_t n, uint32_t j, bool pixel) {
  uint32_t k = 0;
  for (uint32_t i = 1; i < n; ++i)
    if ((pixel ? jobs[i].pix_base : jobs[i].path_base) <= j) k = i;
  return k;
}

// k_generate for a batch of jobs: path j belongs to the job whose paths start at or before it;
// within the job, the same pixel / sample numbering and camera ray as a render of that job alone.
__global__ void k_generate_jobs(RenderParams p, PathBuffers b) {
  const uint32_t n = p.pass_paths;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t k = job_of(p.jobs, p.n_jobs, j, false);
    const JobDesc& J = p.jobs[k];
    const uint32_t npix = J.width * J.height, local = j - J.path_base;
    const uint32_t pixel = local % npix, sample = J.pass_first + local / npix;
    const float jx = jitter(J.seed, pixel, sample, J.spp, kSlotJx, p.sampler),
                jy = jitter(J.seed, pixel, sample, J.spp, kSlotJy, p.sampler);
    float o[3], d[3];
    camera_ray(J.r2s, J.c2w, (float)(pixel % J.width) + jx, (float)(pixel / J.width) + jy, o, d);
    b.ray_o[j] = make_float4(o[0], o[1], o[2], J.near);
    b.ray_d[j] = make_float4(d[0], d[1], d[2], J.far);
    b.beta[j] = make_float4(1.f, 1.f, 1.f, 1.f);
    b.L[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    b.info[j] = make_uint4(pixel, sample, 0u, k);
    b.queue[0][j] = j;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *counter(b.counts, kCntQueue0) = n;
    *counter(b.counts, kCntQueue1) = 0;
    *counter(b.counts, kCntShadow) = 0;
  }
}

// k_accumulate for a batch: every job's pixels, each job's samples of the pass in sample order.
__global__ void k_accumulate_jobs(RenderParams p, PathBuffers b) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < p.batch_pixels; g += gridDim.x * blockDim.x) {
    const JobDesc& J = p.jobs[job_of(p.jobs, p.n_jobs, g, true)];
    const uint32_t npix = J.width * J.height, px = g - J.pix_base;
    double* f = b.film + J.film_base + 3 * (size_t)px;
    double acc[3] = {f[0], f[1], f[2]};
    for (uint32_t s = 0; s < J.pass_samples; ++s) {
      const float4 L = b.L[(size_t)J.path_base + (size_t)s * npix + px];
      acc[0] += (double)L.x;
      acc[1] += (double)L.y;
      acc[2] += (double)L.z;
    }
    f[0] = acc[0];
    f[1] = acc[1];
    f[2] = acc[2];
  }
}

// Per-pixel FP64 sums over this pass's samples, in sample order.
__global__ void k_accumulate(RenderParams p, PathBuffers b) {
  const uint32_t npix = p.width * p.height;
  for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += gridDim.x * blockDim.x) {
    double acc[3] = {b.film[3 * (size_t)px], b.film[3 * (size_t)px + 1], b.film[3 * (size_t)px + 2]};
    for (uint32_t s = 0; s < p.pass_samples; ++s) {
      const float4 L = b.L[(size_t)s * npix + px];
      acc[0] += (double)L.x;
      acc[1] += (double)L.y;
      acc[2] += (double)L.z;
    }
    b.film[3 * (size_t)px] = acc[0];
    b.film[3 * (size_t)px + 1] = acc[1];
    b.film[3 * (size_t)px + 2] = acc[2];
  }
}

uint32_t blocks_for(uint32_t n, uint32_t block) {
  uint32_t g = (n + block - 1) / block;
  return g < 1 ? 1 : (g > 65535 ? 65535 : g);
}

}  // namespace

// The scene walks come in three LDS stack depths; the scene's launch settings pick one.
#define PBRT_SCENE_STACK(stack, CALL)            \
  switch (stack) {                               \
    case 8: {                                    \
      constexpr int D = 8;                       \
      CALL;                                      \
    } break;                                     \
    case 24: {                                   \
      constexpr int D = 24;                      \
      CALL;                                      \
    } break;                                     \
    default: {                                   \
      constexpr int D = 16;                      \
      CALL;                                      \
    } break;                                     \
  }

hipError_t launch_scene_trace(const SceneLaunch& a, int any_hit) {
  if (a.n == 0) return hipSuccess;
  if (a.view.blas4) {
    PBRT_SCENE_STACK(a.view.stack,
                     hipLaunchKernelGGL((k_scene_trace<D, true>), dim3(a.grid), dim3(kSceneBlock), 0,
                                        (hipStream_t)a.stream, a.view, (const float4*)a.rays, a.hits, a.occluded, a.n,
                                        a.spill, a.spill_per_thread, a.status, any_hit));
  } else {
    PBRT_SCENE_STACK(a.view.stack,
                     hipLaunchKernelGGL((k_scene_trace<D, false>), dim3(a.grid), dim3(kSceneBlock), 0,
                                        (hipStream_t)a.stream, a.view, (const float4*)a.rays, a.hits, a.occluded, a.n,
                                        a.spill, a.spill_per_thread, a.status, any_hit));
  }
  return hipGetLastError();
}

hipError_t launch_generate(const SceneView&, const RenderParams& p, const PathBuffers& b, void* stream) {
  const uint32_t n = p.width * p.height * p.pass_samples;
  hipLaunchKernelGGL(k_generate<false>, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, p, b);
  return hipGetLastError();
}

hipError_t launch_generate_nrc(const RenderParams& p, const PathBuffers& b, void* stream) {
  const uint32_t n = p.width * p.height * p.pass_samples;
  hipLaunchKernelGGL(k_generate<true>, dim3(blocks_for(n, 256)), dim3(256), 0, (hipStream_t)stream, p, b);
  return hipGetLastError();
}

hipError_t launch_generate_jobs(const RenderParams& p, const PathBuffers& b, void* stream) {
  hipLaunchKernelGGL(k_generate_jobs, dim3(blocks_for(p.pass_paths, 256)), dim3(256), 0, (hipStream_t)stream, p, b);
  return hipGetLastError();
}

hipError_t launch_accumulate_jobs(const RenderParams& p, const PathBuffers& b, void* stream) {
  hipLaunchKernelGGL(k_accumulate_jobs, dim3(blocks_for(p.batch_pixels, 256)), dim3(256), 0, (hipStream_t)stream, p,
                     b);
  return hipGetLastError();
}

hipError_t launch_reset_queue(const PathBuffers& b, int q, void* stream) {
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, (hipStream_t)stream, b, q);
  return hipGetLastError();
}

uint32_t scene_blocks_per_cu(uint32_t stack) {
  // the fewest workgroups of the traversal kernels that fit on one CU at once (the scene's grids are
  // sized from it)
  int a = 0, b = 0, c = 0;
  hipError_t e = hipSuccess;
  PBRT_SCENE_STACK(stack, {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_extend<D, true, false>, kSceneBlock, 0);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_shadow<D, false, false>, kSceneBlock, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_shadow<D, true, false>, kSceneBlock, 0);
  });
  if (e != hipSuccess) return 0;
  return (uint32_t)(a < b ? (a < c ? a : c) : (b < c ? b : c));
}

hipError_t launch_extend(const SceneView& v, const PathBuffers& b, int q, uint2* spill, uint32_t spill_per_thread,
                         uint32_t grid, int* status, void* stream) {
#define PBRT_EXTEND(OCC, WIDE)                                                                             \
  PBRT_SCENE_STACK(v.stack, hipLaunchKernelGGL((k_extend<D, OCC, WIDE>), dim3(grid), dim3(kSceneBlock), 0, \
                                               (hipStream_t)stream, v, b, q, spill, spill_per_thread, status))
  if (v.occ5 && v.blas4) {
    PBRT_EXTEND(true, true);
  } else if (v.occ5) {
    PBRT_EXTEND(true, false);
  } else if (v.blas4) {
    PBRT_EXTEND(false, true);
  } else {
    PBRT_EXTEND(false, false);
  }
#undef PBRT_EXTEND
  return hipGetLastError();
}

hipError_t launch_shade(const SceneView& v, const RenderParams& p, const PathBuffers& b, int q, void* stream) {
  hipLaunchKernelGGL(k_shade<false>, dim3(blocks_for(b.capacity, kShadeBlock)), dim3(kShadeBlock), 0,
                     (hipStream_t)stream, v, p, b, q);
  return hipGetLastError();
}

hipError_t launch_shade_nrc(const SceneView& v, const RenderParams& p, const PathBuffers& b, int q, void* stream,
                            uint32_t grid_cap) {
  uint32_t grid = blocks_for(b.capacity, kShadeBlock);
  if (grid_cap && grid > grid_cap) grid = grid_cap;
  hipLaunchKernelGGL(k_shade<true>, dim3(grid), dim3(kShadeBlock), 0, (hipStream_t)stream, v, p, b, q);
  return hipGetLastError();
}

hipError_t launch_shadow(const SceneView& v, const PathBuffers& b, uint2* spill, uint32_t spill_per_thread,
                         uint32_t grid, int* status, void* stream) {
  RenderParams none{};
  if (v.blas4) {
    PBRT_SCENE_STACK(v.stack, hipLaunchKernelGGL((k_shadow<D, false, true>), dim3(grid), dim3(kSceneBlock), 0,
                                                 (hipStream_t)stream, v, none, b, spill, spill_per_thread, status));
  } else {
    PBRT_SCENE_STACK(v.stack, hipLaunchKernelGGL((k_shadow<D, false, false>), dim3(grid), dim3(kSceneBlock), 0,
                                                 (hipStream_t)stream, v, none, b, spill, spill_per_thread, status));
  }
  return hipGetLastError();
}

hipError_t launch_shadow_nrc(const SceneView& v, const RenderParams& p, const PathBuffers& b, uint2* spill,
                             uint32_t spill_per_thread, uint32_t grid, int* status, void* stream) {
  if (v.blas4) {
    PBRT_SCENE_STACK(v.stack, hipLaunchKernelGGL((k_shadow<D, true, true>), dim3(grid), dim3(kSceneBlock), 0,
                                                 (hipStream_t)stream, v, p, b, spill, spill_per_thread, status));
  } else {
    PBRT_SCENE_STACK(v.stack, hipLaunchKernelGGL((k_shadow<D, true, false>), dim3(grid), dim3(kSceneBlock), 0,
                                                 (hipStream_t)stream, v, p, b, spill, spill_per_thread, status));
  }
  return hipGetLastError();
}

hipError_t launch_accumulate(const RenderParams& p, const PathBuffers& b, void* stream) {
  hipLaunchKernelGGL(k_accumulate, dim3(blocks_for(p.width * p.height, 256)), dim3(256), 0, (hipStream_t)stream, p, b);
  return hipGetLastError();
}

}  // namespace pbrt
