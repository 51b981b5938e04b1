// ORACLE — test infrastructure, not product code.
//
// CPU restatement of the reference's scene level and path integrator, templated on the scalar
// type like mesh_oracle.hpp (double: the reference's arithmetic; float: the device's):
//   shapes      Sphere (src/shapes/sphere.rs:108-187), AABox (src/shapes/aabox.rs:66-193),
//               Rectangle (src/shapes/rectangle.rs:26-58), TriangleMesh (triangle_mesh.rs:73-132)
//   objects     Simple (src/objects/simple.rs:27-42), Transformed (src/objects/transformed.rs:26-77)
//               with Transform (src/geom/transform.rs:98-116)
//   scene       median-split BVH over objects (src/bvh.rs:109-132, 291-388), closest hit
//               (bvh.rs:153-230 + scene.rs:186-249) and any hit (bvh.rs:239-275 + scene.rs:157-168)
//   shading     Lambertian (materials/lambertian.rs:23-52, pdfs/cosine.rs:9-25), Dielectric
//               (materials/dielectric.rs:31-182), DiffuseLight (materials/diffuse_light.rs:18-22),
//               PointLight (lights/point_light.rs:23-35), VisibilityTester (lights.rs:21-70),
//               frames (geom/intersectable.rs:92-144)
//   integrator  PathIntegrator::li (integrators/path.rs:31-115)
//   camera      PinHoleCamera (cameras/pin_hole.rs), pixel loop (renderers/mt.rs:112-130)
// Additions the reference lacks, restated from the plans it documents or new:
//   * a counter-based sampler keyed by (seed, pixel, sample, bounce, slot) (ideas/rng_graine.md);
//   * AreaLight on a rectangle with uniform area sampling and one-sided emission
//     (ideas/area_light.md:27-165), the specular guard kept;
//   * the GGX microfacet conductor/dielectric-free glossy reflector (new);
//   * multiple importance sampling of next-event estimation and BSDF sampling with the power
//     heuristic, as the reference plans it after the area light (ideas/area_light.md:158-165,
//     IDEAS.md:18-30): an option (mis), off by default.
#pragma once
#include <algorithm>
#include <cmath>
#include <limits>
#include <cstdint>
#include <vector>
#include <thread>
#include <atomic>

#include "camera_oracle.hpp"
#include "mesh_oracle.hpp"

namespace orc {

// ---- counter-based sampler (the same integer mixing as the device code) ---------------------------
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t hash4(uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  return mix64(mix64(mix64(mix64(a) ^ b) ^ c) ^ d);
}
inline double to_unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }
// One uniform for (seed, path, bounce, slot); path = pixel * spp + sample.
inline double rand_u(uint64_t seed, uint64_t path, uint32_t bounce, uint32_t slot) {
  return to_unit(hash4(seed, path, bounce, slot));
}
// The same draw at the working precision, in [0, 1): 53 bits in f64; in FP32 the top 24 bits,
// exactly representable (rounding 53 bits to FP32 could give 1.0, which the samplers exclude).
template <class R>
inline R uni(uint64_t seed, uint64_t path, uint32_t bounce, uint32_t slot);
template <>
inline double uni<double>(uint64_t seed, uint64_t path, uint32_t bounce, uint32_t slot) {
  return rand_u(seed, path, bounce, slot);
}
template <>
inline float uni<float>(uint64_t seed, uint64_t path, uint32_t bounce, uint32_t slot) {
  return (float)(hash4(seed, path, bounce, slot) >> 40) * 0x1p-24f;
}
// ---- the Sobol sampler's stratified draws -------------------------------------------------------
// Owen-scrambled Sobol (0,2)-sequences by hashing (Burley, "Practical Hash-based Owen Scrambling",
// JCGT 9(4), 2020): per pixel and per stratified pair of draws, the sample index is shuffled by a
// nested uniform scramble, the point is (radical inverse, second Sobol dimension) of the shuffled
// index, and each coordinate is scrambled the same way.
inline uint32_t bit_reverse(uint32_t x) {
  uint32_t r = 0;
  for (int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
  return r;
}
inline uint32_t owen_hash(uint32_t x, uint32_t seed) {  // Burley's Laine-Karras style permutation
  x += seed;
  x ^= x * 0x6c50b47cu;
  x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u;
  x ^= x * 0x8d22f6e6u;
  return x;
}
inline uint32_t owen_scramble(uint32_t x, uint32_t seed) { return bit_reverse(owen_hash(bit_reverse(x), seed)); }
// Sobol's second dimension: the generator matrix is Pascal's triangle mod 2 (columns v_k with
// v_0 = 2^31 and v_k = v_{k-1} xor (v_{k-1} >> 1)).
inline uint32_t sobol_second(uint32_t i) {
  uint32_t r = 0, v = 1u << 31;
  for (int k = 0; k < 32; ++k) {
    if ((i >> k) & 1u) r ^= v;
    v ^= v >> 1;
  }
  return r;
}
// The stratified pairs: the jitter (bounce 0, slots 0, 1) is pair 0; at bounces 1 and 2 the light
// sample (slots 3, 4) is pair 2b - 1 and the BSDF sample (slots 5, 6) pair 2b; -1 otherwise.
inline int stratified_pair(uint32_t bounce, uint32_t slot, uint32_t* axis) {
  if (bounce == 0 && (slot == 0 || slot == 1)) {
    *axis = slot;
    return 0;
  }
  if ((bounce == 1 || bounce == 2) && slot >= 3 && slot <= 6) {
    *axis = (slot == 4 || slot == 6) ? 1u : 0u;
    return (int)(2 * bounce - 1) + (slot >= 5 ? 1 : 0);
  }
  return -1;
}
inline uint32_t sobol_bits(uint64_t seed, uint64_t pixel, uint32_t sample, uint32_t pair, uint32_t axis) {
  const uint64_t h = hash4(seed, pixel, 0x50B0Cu + pair, 0);
  const uint32_t idx = owen_scramble(sample, (uint32_t)h);
  const uint32_t pt = axis == 0 ? bit_reverse(idx) : sobol_second(idx);
  return owen_scramble(pt, (uint32_t)(h >> 32) ^ (axis ? 0x9E3779B9u : 0u));
}

enum Slot : uint32_t {
  kJitterX = 0, kJitterY = 1, kLightPick = 2, kLightU = 3, kLightV = 4, kScatter0 = 5, kScatter1 = 6, kCoin = 7,
  kMetal0 = 8,  // Metal's point in the unit ball: attempt k uses slots kMetal0 + 3k .. + 2
  kRoulette = 8 + 3 * 64  // Russian roulette's draw, past the 64 Metal attempts
};
constexpr uint32_t kMetalAttempts = 64;

// ---- scene description (shared layout with the ctypes side) -------------------------------------
enum ShapeKind : int32_t { kMesh = 0, kSphere = 1, kBox = 2, kRect = 3, kCsg = 4 };
enum CsgOp : int32_t { kCsgUnion = 0, kCsgIntersection = 1, kCsgSubtraction = 2 };
enum MatKind : int32_t { kLambertian = 0, kDielectric = 1, kDiffuseLight = 2, kGGX = 3, kMetal = 4 };
enum LightKind : int32_t { kPointLight = 0, kAreaLight = 1, kBackgroundLight = 2 };
enum TexKind : int32_t { kPlainColor = 0, kCheckerBoard = 1 };

struct MaterialDesc {
  int32_t kind, tex;        // tex: kPlainColor (rgb) or kCheckerBoard (rgb, rgb2, scale)
  double rgb[3];
  double ior, roughness;    // roughness: GGX alpha, or Metal's fuzz
  double rgb2[3];
  double scale;
};
struct ObjectDesc {
  int32_t shape, mesh;      // shape kind; mesh index for kMesh
  int32_t material, has_xform;
  double param[3];          // sphere: r; box: extent xyz; rect: width, height
  double m[16], minv[16];   // local -> world and its inverse, row-major
};
// One element of a CSG shape (shapes/csg.rs: Elem { shape, transform }): a sphere or a box and its
// transform. A CSG object is an ObjectDesc with shape kCsg, mesh = its first element, param[0] =
// the operation and param[1] = the element count, elements in the order the operation's
// constructor receives them (Union::new(elements), csg/union.rs:13-17).
struct CsgElemDesc {
  int32_t shape, pad;
  double param[3];
  double m[16], minv[16];
};
struct LightDesc {
  int32_t kind, object;     // object: the emitting rectangle of an area light
  double pos[3], rgb[3];    // background light: pos = the colour looking down, rgb = looking up
};

template <class R>
struct Spec {
  R c[3];
};
template <class R> inline Spec<R> sadd(Spec<R> a, Spec<R> b) { return {{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}}; }
template <class R> inline Spec<R> smul(Spec<R> a, Spec<R> b) { return {{a.c[0] * b.c[0], a.c[1] * b.c[1], a.c[2] * b.c[2]}}; }
template <class R> inline Spec<R> sscale(Spec<R> a, R s) { return {{a.c[0] * s, a.c[1] * s, a.c[2] * s}}; }
template <class R> inline Spec<R> sdiv(Spec<R> a, R s) { return {{a.c[0] / s, a.c[1] / s, a.c[2] / s}}; }

// Intersection record (geom/intersectable.rs): point, distance, normal, -direction, uv, derivatives.
template <class R>
struct Isect {
  V3<R> p, n, wo, dpdu, dpdv;
  R d, u, v;
};

template <class R>
struct SceneHit {
  int64_t object = -1, prim = -1;
  Isect<R> is;
};

template <class R>
struct Xform {
  R m[4][4], inv[4][4];
  V3<R> point(const R (*M)[4], V3<R> v) const {  // matrix4.rs:187-199
    R x = M[0][0] * v.x + M[0][1] * v.y + M[0][2] * v.z + M[0][3];
    R y = M[1][0] * v.x + M[1][1] * v.y + M[1][2] * v.z + M[1][3];
    R z = M[2][0] * v.x + M[2][1] * v.y + M[2][2] * v.z + M[2][3];
    R w = M[3][0] * v.x + M[3][1] * v.y + M[3][2] * v.z + M[3][3];
    V3<R> r{x, y, z};
    if (w != R(1)) r = r * (R(1) / w);
    return r;
  }
  V3<R> dir(const R (*M)[4], V3<R> v) const {  // matrix4.rs:201-206
    return {M[0][0] * v.x + M[0][1] * v.y + M[0][2] * v.z, M[1][0] * v.x + M[1][1] * v.y + M[1][2] * v.z,
            M[2][0] * v.x + M[2][1] * v.y + M[2][2] * v.z};
  }
  V3<R> normal(const R (*M)[4], V3<R> v) const {  // matrix4.rs:208-213
    return {M[0][0] * v.x + M[1][0] * v.y + M[2][0] * v.z, M[0][1] * v.x + M[1][1] * v.y + M[2][1] * v.z,
            M[0][2] * v.x + M[1][2] * v.y + M[2][2] * v.z};
  }
  // transform.rs:98-103 — Ray::new renormalises the local direction.
  Ray<R> to_local(const Ray<R>& r) const { return {point(inv, r.o), normalized(dir(inv, r.d))}; }
  // transform.rs:105-116 — d, u, v unchanged; the normal through the inverse transpose, not renormalised.
  Isect<R> to_world(const Isect<R>& i) const {
    Isect<R> o = i;
    o.n = normal(inv, i.n);
    o.p = point(m, i.p);
    o.wo = dir(m, i.wo);
    o.dpdu = dir(m, i.dpdu);
    o.dpdv = dir(m, i.dpdv);
    return o;
  }
};

template <class R>
class SceneOracle {
 public:
  struct Obj {
    ObjectDesc d;
    Xform<R> x;
    AABB<double> box;  // world bounds (f64, as the reference)
  };
  struct CsgElem {
    CsgElemDesc d;
    Xform<R> x;
  };
  std::vector<MeshBVH<R>*> meshes;
  std::vector<Obj> objs;
  std::vector<CsgElem> csg;
  std::vector<MaterialDesc> mats;
  std::vector<LightDesc> lights;
  bool canonical;
  bool mis = false;  // multiple importance sampling (power heuristic) instead of the specular guard
  bool roulette = false;  // Russian roulette after each scattering (path.rs:93-100, commented out there)
  int integrator = 0;     // integrators.rs:23-28: 0 PathIntegrator, 1 NaiveIntegrator, 2 NormalIntegrator
  int sampler = 0;        // 0: every draw counter-based; 1: the stratified pairs from scrambled Sobol points
  mutable uint32_t draw_spp = 1;  // samples per pixel of the render in progress (path = pixel * spp + sample)

  // One uniform of path `path` at the working precision.
  R draw(uint64_t seed, uint64_t path, uint32_t bounce, uint32_t slot) const {
    uint32_t axis = 0;
    const int pair = sampler ? stratified_pair(bounce, slot, &axis) : -1;
    if (pair < 0) return uni<R>(seed, path, bounce, slot);
    const uint32_t bits = sobol_bits(seed, path / draw_spp, (uint32_t)(path % draw_spp), (uint32_t)pair, axis);
    if (sizeof(R) == 4) return (R)((float)(bits >> 8) * 0x1p-24f);
    return (R)((double)bits * 0x1p-32);
  }
  // The pixel jitter (renderers/mt.rs:123-127): f64 for the f64 camera, FP32 for the FP32 one.
  double jitter(uint64_t seed, uint64_t pixel, uint32_t sample, uint32_t spp, uint32_t slot) const {
    if (sampler) return (double)sobol_bits(seed, pixel, sample, 0, slot) * 0x1p-32;
    return rand_u(seed, pixel * spp + sample, 0, slot);
  }
  float jitter_f32(uint64_t seed, uint64_t pixel, uint32_t sample, uint32_t spp, uint32_t slot) const {
    if (sampler) return (float)(sobol_bits(seed, pixel, sample, 0, slot) >> 8) * 0x1p-24f;
    return uni<float>(seed, pixel * spp + sample, 0, slot);
  }
  // scene BVH in the reference's layout (bvh.rs:15-31)
  struct SNode {
    AABB<double> box;
    AABB<R> rbox;
    uint64_t left_first, count;
  };
  std::vector<SNode> nodes;
  std::vector<uint32_t> prims;  // object indices in leaf order

  ~SceneOracle() {
    for (auto* m : meshes) delete m;
  }

  // ---- shapes ------------------------------------------------------------------------------------
  // sphere.rs:108-144 (+ 152-187). Returns the first recorded intersection (Simple takes [0]).
  bool sphere(const Ray<R>& ray, R r, R near, R far, Isect<R>* out) const {
    Isect<R> all[2];
    if (sphere_all(ray, r, near, far, all) == 0) return false;
    *out = all[0];
    return true;
  }
  // Every intersection the sphere records, in order (at most 2).
  int sphere_all(const Ray<R>& ray, R r, R near, R far, Isect<R>* out) const {
    V3<R> l = ray.o * R(-1);
    R tca = dot(l, ray.d);
    if (tca < R(0)) return 0;
    R r2 = r * r;
    R d2 = dot(l, l) - tca * tca;
    if (d2 > r2) return 0;
    R thc = std::sqrt(r2 - d2);
    R t0 = tca - thc, t1 = tca + thc;
    R tmin = rmin(t0, t1), tmax = rmax(t0, t1);
    if (tmax < near || tmin > far) return 0;
    int n = 0;
    if (tmin > near) out[n++] = sphere_details(ray, r, tmin);
    if (tmax > near && tmax < far) out[n++] = sphere_details(ray, r, tmax);
    return n;
  }
  Isect<R> sphere_details(const Ray<R>& ray, R r, R t) const {
    const R PI = R(M_PI);
    V3<R> hit = ray.o + ray.d * t;
    V3<R> n = normalized(hit);
    R phi = std::atan2(hit.y, hit.x);
    if (phi < R(0)) phi += R(2) * PI;
    R u = phi / (R(2) * PI);
    R zr = hit.z / r;
    zr = zr < R(-1) ? R(-1) : (zr > R(1) ? R(1) : zr);
    R theta = std::acos(zr);
    R v = theta / PI;
    R z_radius = std::sqrt(hit.x * hit.x + hit.y * hit.y);
    R inv_z_r = R(1) / z_radius;
    R cos_phi = hit.x * inv_z_r, sin_phi = hit.y * inv_z_r;
    V3<R> dpdu = V3<R>{-hit.y, hit.x, R(0)} * (R(2) * PI);
    V3<R> dpdv = V3<R>{hit.z * cos_phi, hit.z * sin_phi, -r * std::sin(theta)} * PI;
    return {hit, n, ray.d * R(-1), dpdu, dpdv, t, u, v};
  }
  // aabox.rs:66-193 with the constant normal/derivative table of aabox.rs:31-60.
  bool box(const Ray<R>& ray, const R ext[3], R near, R far, Isect<R>* out) const {
    Isect<R> all[2];
    int face[2];
    if (box_all(ray, ext, near, far, all, face) == 0) return false;
    *out = all[0];
    return true;
  }
  // Every intersection the box records, in order (at most 2), with the face of each.
  int box_all(const Ray<R>& ray, const R ext[3], R near, R far, Isect<R>* out, int* face) const {
    static const R NT[6][9] = {
        {-1, 0, 0, 0, 0, -1, 0, 1, 0}, {1, 0, 0, 0, 0, 1, 0, 1, 0},  {0, -1, 0, -1, 0, 0, 0, 0, 1},
        {0, 1, 0, 1, 0, 0, 0, 0, 1},   {0, 0, -1, 0, -1, 0, 1, 0, 0}, {0, 0, 1, 0, 1, 0, 1, 0, 0}};
    V3<R> mn{R(0) - ext[0] * R(0.5), R(0) - ext[1] * R(0.5), R(0) - ext[2] * R(0.5)};
    V3<R> mx{ext[0] * R(0.5), ext[1] * R(0.5), ext[2] * R(0.5)};
    V3<R> inv{R(1) / ray.d.x, R(1) / ray.d.y, R(1) / ray.d.z};
    R tmin, tmax;
    int nmin, nmax;
    if (inv.x >= R(0)) { tmin = (mn.x - ray.o.x) * inv.x; tmax = (mx.x - ray.o.x) * inv.x; nmin = 0; nmax = 1; }
    else { tmin = (mx.x - ray.o.x) * inv.x; tmax = (mn.x - ray.o.x) * inv.x; nmin = 1; nmax = 0; }
    R tymin, tymax;
    int nymin, nymax;
    if (inv.y >= R(0)) { tymin = (mn.y - ray.o.y) * inv.y; tymax = (mx.y - ray.o.y) * inv.y; nymin = 2; nymax = 3; }
    else { tymin = (mx.y - ray.o.y) * inv.y; tymax = (mn.y - ray.o.y) * inv.y; nymin = 3; nymax = 2; }
    if (tmin > tymax || tymin > tmax) return 0;
    if (tymin > tmin) { tmin = tymin; nmin = nymin; }
    if (tymax < tmax) { tmax = tymax; nmax = nymax; }
    R tzmin, tzmax;
    int nzmin, nzmax;
    if (inv.z >= R(0)) { tzmin = (mn.z - ray.o.z) * inv.z; tzmax = (mx.z - ray.o.z) * inv.z; nzmin = 4; nzmax = 5; }
    else { tzmin = (mx.z - ray.o.z) * inv.z; tzmax = (mn.z - ray.o.z) * inv.z; nzmin = 5; nzmax = 4; }
    if (tmin > tzmax || tzmin > tmax) return 0;
    if (tzmin > tmin) { tmin = tzmin; nmin = nzmin; }
    if (tzmax < tmax) { tmax = tzmax; nmax = nzmax; }
    if (tmax < near || tmin > far) return 0;
    auto rec = [&](R t, int k) {
      V3<R> p = ray.o + ray.d * t;
      const R* e = NT[k];
      V3<R> du{e[3], e[4], e[5]}, dv{e[6], e[7], e[8]};
      return Isect<R>{p, {e[0], e[1], e[2]}, ray.d * R(-1), du, dv, t, dot(p, du), dot(p, dv)};
    };
    int n = 0;
    if (tmin > near) {
      face[n] = nmin;
      out[n++] = rec(tmin, nmin);
    }
    if (tmax < far) {
      face[n] = nmax;
      out[n++] = rec(tmax, nmax);
    }
    return n;
  }
  // rectangle.rs:26-58 (the constructor halves width and height, rectangle.rs:15-20).
  bool rect(const Ray<R>& ray, R w, R h, R near, R far, Isect<R>* out) const {
    if (ray.d.y == R(0)) return false;
    R d = (ray.o.y * R(-1)) / ray.d.y;
    if (d < near || d > far) return false;
    V3<R> p = ray.o + ray.d * d;
    p.y = R(0);
    if (std::fabs(p.x) > w / R(2) || std::fabs(p.z) > h / R(2)) return false;
    *out = {p, {0, 1, 0}, ray.d * R(-1), {1, 0, 0}, {0, 0, 1}, d, p.x, p.z};
    return true;
  }
  // triangle_mesh.rs:73-132 — shading quantities of the hit triangle with the default UVs.
  Isect<R> mesh_details(const MeshBVH<R>& m, const Ray<R>& ray, uint64_t tri, const TriHit<R>& h) const {
    const uint32_t* ix = &m.indices[3 * tri];
    V3<R> p0 = m.vertR(ix[0]), p1 = m.vertR(ix[1]), p2 = m.vertR(ix[2]);
    V3<R> hit = ray.o + ray.d * h.t;
    R w = R(1) - h.u - h.v;
    // interpolate_texture_coords with uv0 = (0,0), uv1 = (1,0), uv2 = (1,1) (triangle_mesh.rs:219-224)
    R uu = w * R(0) + h.u * R(1) + h.v * R(1);
    R vv = w * R(0) + h.u * R(0) + h.v * R(1);
    // compute_texture_derivatives (triangle_mesh.rs:226-258), regular case: det = -1
    const R duv02x = R(-1), duv02y = R(-1), duv12x = R(0), duv12y = R(-1);
    V3<R> dp02 = p0 - p2, dp12 = p1 - p2;
    R det = duv12x * duv02y - duv12y * duv02x;
    R inv_det = R(1) / det;
    V3<R> dpdu = (dp12 * duv02y - dp02 * duv12y) * inv_det;
    V3<R> dpdv = (dp02 * duv12x - dp12 * duv02x) * inv_det;
    V3<R> n = normalized(cross(dpdv, dpdu));
    return {hit, n, ray.d * R(-1), dpdu, dpdv, h.t, uu, vv};
  }

  // ---- CSG (shapes/csg/*.rs) -------------------------------------------------------------------
  // Shape::contain_point of an element at a point of the CSG frame: sphere.rs:146-148 (closed),
  // aabox.rs:195-202 (closed), through transform_point_to_local.
  bool elem_contains(const CsgElem& e, V3<R> p) const {
    const V3<R> lp = e.x.point(e.x.inv, p);
    if (e.d.shape == kSphere) {
      const R r = (R)e.d.param[0];
      return dot(lp, lp) <= r * r;
    }
    const R hx = (R)e.d.param[0] * R(0.5), hy = (R)e.d.param[1] * R(0.5), hz = (R)e.d.param[2] * R(0.5);
    return lp.x >= R(0) - hx && lp.x <= hx && lp.y >= R(0) - hy && lp.y <= hy && lp.z >= R(0) - hz && lp.z <= hz;
  }
  // The element's intersections in the CSG frame (transform_ray_to_local, shape.intersect,
  // transform_interaction_to_world), with the box face of each.
  int elem_hits(const CsgElem& e, const Ray<R>& ray, R near, R far, Isect<R>* out, int* face) const {
    const Ray<R> lr = e.x.to_local(ray);
    Isect<R> li[2];
    int n;
    if (e.d.shape == kSphere) {
      n = sphere_all(lr, (R)e.d.param[0], near, far, li);
      face[0] = face[1] = 0;
    } else {
      const R ext[3] = {(R)e.d.param[0], (R)e.d.param[1], (R)e.d.param[2]};
      n = box_all(lr, ext, near, far, li, face);
    }
    for (int i = 0; i < n; ++i) out[i] = e.x.to_world(li[i]);
    return n;
  }
  // The first intersection of a CSG shape: union.rs:22-42 (kept when inside no other element),
  // intersection.rs:23-43 (kept when inside every other element), substraction.rs:22-70 (the
  // base's hits outside every subtracted element; a subtracted element's hits inside the base and
  // outside every other subtracted element, normal flipped). The candidates are sorted by distance
  // with a stable sort, so the first is the nearest, the earliest recorded among equals. *prim is
  // the element the hit belongs to.
  bool csg_hit(const Obj& o, const Ray<R>& ray, R near, R far, Isect<R>* out, int64_t* prim) const {
    const int op = (int)o.d.param[0];
    const uint32_t base = (uint32_t)o.d.mesh, cnt = (uint32_t)o.d.param[1];
    bool found = false;
    R best = R(0);
    auto consider = [&](const Isect<R>& h, uint32_t k) {
      if (!found || h.d < best) {
        found = true;
        best = h.d;
        *out = h;
        *prim = (int64_t)k;
      }
    };
    auto in_subtracted = [&](V3<R> p, uint32_t exclude) {
      for (uint32_t k = 1; k < cnt; ++k)
        if (k != exclude && elem_contains(csg[base + k], p)) return true;
      return false;
    };
    Isect<R> h[2];
    int face[2];
    if (op == kCsgSubtraction) {
      if (cnt == 0) return false;
      const int n0 = elem_hits(csg[base], ray, near, far, h, face);
      for (int i = 0; i < n0; ++i)
        if (!in_subtracted(h[i].p, 0xffffffffu)) consider(h[i], 0);
      for (uint32_t k = 1; k < cnt; ++k) {
        const int n = elem_hits(csg[base + k], ray, near, far, h, face);
        for (int i = 0; i < n; ++i)
          if (elem_contains(csg[base], h[i].p) && !in_subtracted(h[i].p, k)) {
            h[i].n = h[i].n * R(-1);
            consider(h[i], k);
          }
      }
      return found;
    }
    for (uint32_t k = 0; k < cnt; ++k) {
      const int n = elem_hits(csg[base + k], ray, near, far, h, face);
      for (int i = 0; i < n; ++i) {
        // union: inside no other element; intersection: inside every other element
        bool keep = true;
        for (uint32_t m = 0; m < cnt && keep; ++m)
          if (m != k && elem_contains(csg[base + m], h[i].p) == (op == kCsgUnion)) keep = false;
        if (keep) consider(h[i], k);
      }
    }
    return found;
  }

  // Object::intersect on a local (or world) ray: the first intersection of the shape.
  bool shape_hit(const Obj& o, const Ray<R>& ray, R near, R far, Isect<R>* out, int64_t* prim) const {
    *prim = -1;
    switch (o.d.shape) {
      case kCsg: return csg_hit(o, ray, near, far, out, prim);
      case kSphere: return sphere(ray, (R)o.d.param[0], near, far, out);
      case kBox: {
        R e[3] = {(R)o.d.param[0], (R)o.d.param[1], (R)o.d.param[2]};
        return box(ray, e, near, far, out);
      }
      case kRect: return rect(ray, (R)o.d.param[0], (R)o.d.param[1], near, far, out);
      case kMesh: {
        const MeshBVH<R>& m = *meshes[o.d.mesh];
        Stats st;
        TriHit<R> h;
        uint64_t id;
        if (!m.traverse(ray, near, far, canonical, &st, &h, &id)) return false;
        *out = mesh_details(m, ray, id, h);
        *prim = (int64_t)id;
        return true;
      }
    }
    return false;
  }
  bool shape_any(const Obj& o, const Ray<R>& ray, R near, R far) const {
    if (o.d.shape == kMesh) return meshes[o.d.mesh]->intersect_p(ray, near, far);
    Isect<R> tmp;
    int64_t prim;
    return shape_hit(o, ray, near, far, &tmp, &prim);  // Intersectable::intersect_p default (intersectable.rs:35-37)
  }
  bool object_hit(uint32_t k, const Ray<R>& ray, R near, R far, Isect<R>* out, int64_t* prim) const {
    const Obj& o = objs[k];
    if (!o.d.has_xform) return shape_hit(o, ray, near, far, out, prim);
    Ray<R> lr = o.x.to_local(ray);
    Isect<R> li;
    if (!shape_hit(o, lr, near, far, &li, prim)) return false;
    *out = o.x.to_world(li);
    return true;
  }
  bool object_any(uint32_t k, const Ray<R>& ray, R near, R far) const {
    const Obj& o = objs[k];
    if (!o.d.has_xform) return shape_any(o, ray, near, far);
    return shape_any(o, o.x.to_local(ray), near, far);
  }

  // ---- build ---------------------------------------------------------------------------------------
  static AABB<double> local_bounds(const ObjectDesc& d, const MeshBVH<R>* m) {
    switch (d.shape) {
      case kSphere: return {{-d.param[0], -d.param[0], -d.param[0]}, {d.param[0], d.param[0], d.param[0]}};
      case kBox: {
        V3<double> h{d.param[0] * 0.5, d.param[1] * 0.5, d.param[2] * 0.5};
        return {{0.0 - h.x, 0.0 - h.y, 0.0 - h.z}, h};
      }
      case kRect: return {{-d.param[0] / 2.0, 0.0, -d.param[1] / 2.0}, {d.param[0] / 2.0, 0.0, d.param[1] / 2.0}};
      default: return vertex_bounds(m->verts64.data(), m->verts64.size() / 3);
    }
  }
  // CSG bounds in its own frame: union.rs:58-72 (union of the elements' boxes), intersection.rs:
  // 72-85 (their intersection), substraction.rs:94-104 (the base element's box).
  AABB<double> csg_bounds(const ObjectDesc& d) const {
    const int op = (int)d.param[0];
    const uint32_t base = (uint32_t)d.mesh, cnt = (uint32_t)d.param[1];
    if (cnt == 0) return {{0, 0, 0}, {0, 0, 0}};
    auto elem_box = [&](uint32_t k) {
      const CsgElemDesc& e = csg[base + k].d;
      ObjectDesc tmp{};
      tmp.shape = e.shape;
      tmp.has_xform = 1;
      for (int a = 0; a < 3; ++a) tmp.param[a] = e.param[a];
      for (int a = 0; a < 16; ++a) tmp.m[a] = e.m[a];
      return world_bounds(local_bounds(tmp, nullptr), tmp);
    };
    AABB<double> b = elem_box(0);
    if (op == kCsgSubtraction) return b;
    for (uint32_t k = 1; k < cnt; ++k) {
      const AABB<double> e = elem_box(k);
      if (op == kCsgUnion) {
        b.combine_with(e);
      } else {
        b.bmin = {rmax(b.bmin.x, e.bmin.x), rmax(b.bmin.y, e.bmin.y), rmax(b.bmin.z, e.bmin.z)};
        b.bmax = {rmin(b.bmax.x, e.bmax.x), rmin(b.bmax.y, e.bmax.y), rmin(b.bmax.z, e.bmax.z)};
      }
    }
    return b;
  }
  // aabound.rs 'transform': the eight corners through the transform.
  static AABB<double> world_bounds(const AABB<double>& b, const ObjectDesc& d) {
    if (!d.has_xform) return b;
    const double* M = d.m;
    const V3<double> c[8] = {{b.bmin.x, b.bmin.y, b.bmin.z}, {b.bmin.x, b.bmin.y, b.bmax.z}, {b.bmin.x, b.bmax.y, b.bmin.z},
                             {b.bmin.x, b.bmax.y, b.bmax.z}, {b.bmax.x, b.bmax.y, b.bmax.z}, {b.bmax.x, b.bmax.y, b.bmin.z},
                             {b.bmax.x, b.bmin.y, b.bmax.z}, {b.bmax.x, b.bmin.y, b.bmin.z}};
    AABB<double> r = AABB<double>::empty();
    for (const auto& v : c) {
      double x = M[0] * v.x + M[1] * v.y + M[2] * v.z + M[3];
      double y = M[4] * v.x + M[5] * v.y + M[6] * v.z + M[7];
      double z = M[8] * v.x + M[9] * v.y + M[10] * v.z + M[11];
      double w = M[12] * v.x + M[13] * v.y + M[14] * v.z + M[15];
      V3<double> p{x, y, z};
      if (w != 1.0) p = p * (1.0 / w);
      r.bmin = {rmin(r.bmin.x, p.x), rmin(r.bmin.y, p.y), rmin(r.bmin.z, p.z)};
      r.bmax = {rmax(r.bmax.x, p.x), rmax(r.bmax.y, p.y), rmax(r.bmax.z, p.z)};
    }
    return r;
  }

  // bvh.rs:291-335 — median split on the widest centroid axis, stable sort, 1 object per leaf.
  AABB<double> subdivide(uint64_t ni) {
    uint64_t first = nodes[ni].left_first, count = nodes[ni].count;
    if (count == 1) {
      nodes[ni].box = objs[prims[first]].box;
      return nodes[ni].box;
    }
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (uint64_t i = first; i < first + count; ++i) {
      const AABB<double>& b = objs[prims[i]].box;
      V3<double> c = (b.bmax + b.bmin) * 0.5;
      for (int a = 0; a < 3; ++a) {
        mn[a] = rmin(mn[a], c[a]);
        mx[a] = rmax(mx[a], c[a]);
      }
    }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
      if (mx[a] - mn[a] > mx[axis] - mn[axis]) axis = a;
    std::stable_sort(prims.begin() + first, prims.begin() + first + count, [&](uint32_t a, uint32_t b) {
      const AABB<double>&A = objs[a].box, &B = objs[b].box;
      return ((A.bmax + A.bmin) * 0.5)[axis] < ((B.bmax + B.bmin) * 0.5)[axis];
    });
    uint64_t lc = count / 2;
    uint64_t li = nodes.size();
    nodes.push_back({AABB<double>::empty(), {}, first, lc});
    nodes.push_back({AABB<double>::empty(), {}, first + lc, count - lc});
    nodes[ni].left_first = li;
    nodes[ni].count = 0;
    AABB<double> l = subdivide(li), r = subdivide(li + 1);
    AABB<double> b = AABB<double>::empty();
    b.combine_with(l);
    b.combine_with(r);
    nodes[ni].box = b;
    return b;
  }

  void build(const float* verts, const uint64_t* vert_off, const uint32_t* idx, const uint64_t* idx_off, uint32_t n_mesh,
             const ObjectDesc* o, uint32_t n_obj, const MaterialDesc* m, uint32_t n_mat, const LightDesc* l,
             uint32_t n_light, const CsgElemDesc* ce, uint32_t n_csg, BuildParams bp, bool canon) {
    canonical = canon;
    for (uint32_t i = 0; i < n_csg; ++i) {
      CsgElem e;
      e.d = ce[i];
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
          e.x.m[r][c] = (R)ce[i].m[4 * r + c];
          e.x.inv[r][c] = (R)ce[i].minv[4 * r + c];
        }
      csg.push_back(e);
    }
    // the meshes' trees are independent: built by up to 16 threads, each taking the next mesh
    meshes.assign(n_mesh, nullptr);
    {
      std::atomic<uint32_t> next{0};
      auto work = [&] {
        for (uint32_t i = next++; i < n_mesh; i = next++) {
          uint64_t nv = (vert_off[i + 1] - vert_off[i]), nt = (idx_off[i + 1] - idx_off[i]);
          meshes[i] = new MeshBVH<R>(verts + 3 * vert_off[i], nv, idx + 3 * idx_off[i], nt, bp);
        }
      };
      const uint32_t nthr = std::min<uint32_t>(n_mesh, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
      std::vector<std::thread> pool;
      for (uint32_t t = 1; t < nthr; ++t) pool.emplace_back(work);
      work();
      for (auto& th : pool) th.join();
    }
    mats.assign(m, m + n_mat);
    lights.assign(l, l + n_light);
    for (uint32_t i = 0; i < n_obj; ++i) {
      Obj ob;
      ob.d = o[i];
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
          ob.x.m[r][c] = (R)o[i].m[4 * r + c];
          ob.x.inv[r][c] = (R)o[i].minv[4 * r + c];
        }
      ob.box = world_bounds(o[i].shape == kCsg ? csg_bounds(o[i])
                                               : local_bounds(o[i], o[i].shape == kMesh ? meshes[o[i].mesh] : nullptr),
                            o[i]);
      objs.push_back(ob);
    }
    if (objs.empty()) return;  // scene.rs:278-284: an empty scene has no tree
    prims.resize(objs.size());
    for (uint32_t i = 0; i < objs.size(); ++i) prims[i] = i;
    nodes.push_back({AABB<double>::empty(), {}, 0, objs.size()});
    subdivide(0);
    for (auto& nd : nodes)
      nd.rbox = {{(R)nd.box.bmin.x, (R)nd.box.bmin.y, (R)nd.box.bmin.z}, {(R)nd.box.bmax.x, (R)nd.box.bmax.y, (R)nd.box.bmax.z}};
  }

  // ---- queries --------------------------------------------------------------------------------------
  struct SStats {
    uint64_t nodes_visited = 0, box_tests = 0, object_tests = 0;
  };
  // bvh.rs:153-230 with scene.rs:186-220's visitor.
  bool intersect(const Ray<R>& ray, R near, R far, SceneHit<R>* out, SStats* st) const {
    if (nodes.empty()) return false;
    R min_t = far;
    bool found = false;
    int64_t best_obj = -1, best_prim = -1;
    R d;
    st->box_tests++;
    if (!nodes[0].rbox.hit(ray, near, min_t, &d)) return false;
    struct E {
      uint64_t n;
      R d;
    };
    E stack[256];
    int sp = 0;
    stack[sp++] = {0, d};
    while (sp > 0) {
      E e = stack[--sp];
      if (canonical ? e.d > min_t : e.d >= min_t) continue;
      st->nodes_visited++;
      const SNode& nd = nodes[e.n];
      if (nd.count > 0) {
        for (uint64_t i = nd.left_first; i < nd.left_first + nd.count; ++i) {
          st->object_tests++;
          uint32_t k = prims[i];
          Isect<R> is;
          int64_t prim;
          if (!object_hit(k, ray, near, min_t, &is, &prim)) continue;
          bool take = !found || is.d < out->is.d;
          if (canonical && found && is.d == out->is.d)
            take = (int64_t)k < best_obj || ((int64_t)k == best_obj && prim < best_prim);
          if (!take) continue;
          out->is = is;
          out->object = k;
          out->prim = prim;
          best_obj = k;
          best_prim = prim;
          found = true;
          min_t = is.d;
        }
      } else {
        R ld, rd;
        st->box_tests += 2;
        bool lh = nodes[nd.left_first].rbox.hit(ray, near, min_t, &ld);
        bool rh = nodes[nd.left_first + 1].rbox.hit(ray, near, min_t, &rd);
        bool swap_order = lh && rh && rd < ld;
        E nearer = swap_order ? E{nd.left_first + 1, rd} : E{nd.left_first, ld};
        E farther = swap_order ? E{nd.left_first, ld} : E{nd.left_first + 1, rd};
        bool nh = swap_order ? rh : lh, fh = swap_order ? lh : rh;
        if (fh) stack[sp++] = farther;
        if (nh) stack[sp++] = nearer;
      }
    }
    return found;
  }
  // bvh.rs:239-275 — box test at pop, right pushed first so the left pops first.
  bool intersect_p(const Ray<R>& ray, R near, R far, SStats* st) const {
    if (nodes.empty()) return false;
    uint64_t stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const SNode& nd = nodes[stack[--sp]];
      R d;
      st->box_tests++;
      if (!nd.rbox.hit(ray, near, far, &d)) continue;
      st->nodes_visited++;
      if (nd.count > 0) {
        for (uint64_t i = nd.left_first; i < nd.left_first + nd.count; ++i) {
          st->object_tests++;
          if (object_any(prims[i], ray, near, far)) return true;
        }
      } else {
        stack[sp++] = nd.left_first + 1;
        stack[sp++] = nd.left_first;
      }
    }
    return false;
  }

  // ---- shading ------------------------------------------------------------------------------------
  // intersectable.rs:92-144
  static V3<R> to_local(const Isect<R>& i, V3<R> v) {
    V3<R> ss = normalized(i.dpdu);
    V3<R> ts = cross(i.n, ss);
    return {dot(ss, v), dot(ts, v), dot(i.n, v)};
  }
  static V3<R> to_world(const Isect<R>& i, V3<R> v) {
    V3<R> ns = i.n, ss = normalized(i.dpdu);
    V3<R> ts = cross(ns, ss);
    return {ss.x * v.x + ts.x * v.y + ns.x * v.z, ss.y * v.x + ts.y * v.y + ns.y * v.z,
            ss.z * v.x + ts.z * v.y + ns.z * v.z};
  }
  Spec<R> albedo(const MaterialDesc& m) const { return {{(R)m.rgb[0], (R)m.rgb[1], (R)m.rgb[2]}}; }

  // CheckerBoard::shade (textures/checker_board.rs:17-29) at the surface coordinates (u, v):
  // Rust's f64 % (fmod) and signum (+-1 by the sign bit).
  static bool checker_first(R u, R v, R scale) {
    R su = std::fmod(std::fabs(u) * scale, R(1));
    R sv = std::fmod(std::fabs(v) * scale, R(1));
    bool use_color = (su < R(0.5) && sv < R(0.5)) || (su >= R(0.5) && sv >= R(0.5));
    bool positive = std::copysign(R(1), u) * std::copysign(R(1), v) >= R(0);
    return use_color != positive;
  }
  // The material with its texture evaluated at the hit: rgb is the colour there.
  static MaterialDesc shaded(const MaterialDesc& m, const Isect<R>& is) {
    MaterialDesc r = m;
    if (m.tex == kCheckerBoard && !checker_first(is.u, is.v, (R)m.scale))
      for (int c = 0; c < 3; ++c) r.rgb[c] = m.rgb2[c];
    return r;
  }
  // BackgroundInfiniteLight::le (lights/background_infinite_light.rs:24-28), summed over the
  // background lights.
  Spec<R> background_le(V3<R> d) const {
    Spec<R> sum{{0, 0, 0}};
    V3<R> u = normalized(d);
    R factor = R(0.5) * (u.y + R(1));
    for (const LightDesc& L : lights) {
      if (L.kind != kBackgroundLight) continue;
      for (int c = 0; c < 3; ++c) sum.c[c] = sum.c[c] + ((R)L.pos[c] * (R(1) - factor) + (R)L.rgb[c] * factor);
    }
    return sum;
  }
  uint32_t n_background() const {
    uint32_t n = 0;
    for (const LightDesc& L : lights) n += L.kind == kBackgroundLight ? 1u : 0u;
    return n;
  }
  bool has_background() const { return n_background() > 0; }

  // Material::f (lambertian.rs:41-52); 0 for the others (materials.rs:32-34).
  Spec<R> mat_f(const MaterialDesc& m, const Isect<R>& is, V3<R> wo_w, V3<R> wi_w) const {
    if (m.kind == kLambertian) {
      V3<R> wi = to_local(is, wi_w), wo = to_local(is, wo_w);
      if (wo.z * wi.z > R(0)) return sdiv(albedo(m), R(M_PI));
      return {{0, 0, 0}};
    }
    if (m.kind == kGGX) return ggx_f(m, is, wo_w, wi_w);
    return {{0, 0, 0}};
  }

  // GGX (Trowbridge–Reitz) reflection with Smith masking and Schlick Fresnel on the albedo as F0;
  // sampled by the distribution of visible normals' simpler form: the half vector from D(h) cos.
  static R ggx_d(R cos_h, R a) {
    R a2 = a * a;
    R c2 = cos_h * cos_h;
    R den = c2 * (a2 - R(1)) + R(1);
    return a2 / (R(M_PI) * den * den);
  }
  static R ggx_g1(R cos_v, R a) {
    R c = std::fabs(cos_v);
    R a2 = a * a;
    return R(2) * c / (c + std::sqrt(a2 + (R(1) - a2) * c * c));
  }
  Spec<R> ggx_f(const MaterialDesc& m, const Isect<R>& is, V3<R> wo_w, V3<R> wi_w) const {
    V3<R> wo = to_local(is, wo_w), wi = to_local(is, wi_w);
    if (wo.z <= R(0) || wi.z <= R(0)) return {{0, 0, 0}};
    V3<R> h = normalized(wo + wi);
    R a = (R)m.roughness;
    R D = ggx_d(h.z, a);
    R G = ggx_g1(wo.z, a) * ggx_g1(wi.z, a);
    R c = R(1) - dot(wo, h);
    R c5 = c * c * c * c * c;
    Spec<R> F0 = albedo(m);
    Spec<R> F{{F0.c[0] + (R(1) - F0.c[0]) * c5, F0.c[1] + (R(1) - F0.c[1]) * c5, F0.c[2] + (R(1) - F0.c[2]) * c5}};
    R s = D * G / (R(4) * wo.z * wi.z);
    return sscale(F, s);
  }

  struct Scatter {
    Spec<R> att;
    Ray<R> ray;
    R pdf;
    bool specular;
  };
  bool scatter(const MaterialDesc& m, const Isect<R>& is, uint64_t seed, uint64_t path, uint32_t bounce, Scatter* s) const {
    if (m.kind == kLambertian) {
      // lambertian.rs:24-38, pdfs/cosine.rs:14-24
      R two_pi = R(2) * R(M_PI);
      R phi = two_pi * draw(seed, path, bounce, kScatter0);
      R r2 = draw(seed, path, bounce, kScatter1);
      R r2s = std::sqrt(r2);
      V3<R> lwi{std::cos(phi) * r2s, std::sin(phi) * r2s, std::sqrt(R(1) - r2)};
      V3<R> wi = to_world(is, lwi);
      s->ray = {is.p + is.n * R(0.001), normalized(wi)};
      s->att = sdiv(albedo(m), R(M_PI));
      s->pdf = std::fabs(lwi.z / R(M_PI));
      s->specular = false;
      return true;
    }
    if (m.kind == kDielectric) return dielectric(m, is, seed, path, bounce, s);
    if (m.kind == kMetal) {
      // metal.rs:22-54: the mirror direction plus fuzz times a point of the unit ball
      // (utils.rs:19-27, by rejection: attempt k draws three uniforms of its own)
      V3<R> lwo = normalized(to_local(is, is.wo));
      V3<R> lrefl{-lwo.x, -lwo.y, lwo.z};
      V3<R> q{0, 0, 0};
      for (uint32_t k = 0; k < kMetalAttempts; ++k) {
        V3<R> c{draw(seed, path, bounce, kMetal0 + 3 * k) * R(2) - R(1),
                draw(seed, path, bounce, kMetal0 + 3 * k + 1) * R(2) - R(1),
                draw(seed, path, bounce, kMetal0 + 3 * k + 2) * R(2) - R(1)};
        if (dot(c, c) < R(1)) {
          q = c;
          break;
        }
      }
      V3<R> lt = normalized(lrefl + q * (R)m.roughness);
      if (!(lt.z > R(0))) return false;
      s->ray = {is.p + is.n * R(0.001), normalized(to_world(is, lt))};
      s->att = sdiv(albedo(m), std::fabs(lt.z));
      s->pdf = R(1);
      s->specular = true;
      return true;
    }
    if (m.kind == kGGX) {
      V3<R> wo = to_local(is, is.wo);
      if (wo.z <= R(0)) return false;
      R a = (R)m.roughness;
      R u1 = draw(seed, path, bounce, kScatter0), u2 = draw(seed, path, bounce, kScatter1);
      // half vector from D(h) cos(theta_h): tan^2 = a^2 u / (1 - u)
      R phi = R(2) * R(M_PI) * u1;
      R cos_t = std::sqrt((R(1) - u2) / (R(1) + (a * a - R(1)) * u2));
      R sin_t = std::sqrt(std::fmax(R(0), R(1) - cos_t * cos_t));
      V3<R> h{sin_t * std::cos(phi), sin_t * std::sin(phi), cos_t};
      R woh = dot(wo, h);
      V3<R> wi = h * (R(2) * woh) - wo;
      if (wi.z <= R(0)) return false;
      R pdf = ggx_d(h.z, a) * h.z / (R(4) * woh);
      s->att = ggx_f(m, is, is.wo, to_world(is, wi));
      s->ray = {is.p + is.n * R(0.001), normalized(to_world(is, wi))};
      s->pdf = pdf;
      s->specular = false;
      return true;
    }
    return false;  // DiffuseLight does not scatter (materials.rs:20-22)
  }
  // dielectric.rs:129-157
  static bool refract(V3<R> wo, V3<R> n, R eta, V3<R>* t) {
    R ci = dot(wo, n);
    R s2i = R(1) - ci * ci;
    R s2t = eta * eta * s2i;
    R disc = R(1) - s2t;
    if (!(disc > R(0))) return false;
    R ct = std::sqrt(disc);
    V3<R> wi = wo * R(-1);
    *t = normalized(wi * eta + n * (eta * ci - ct));
    return true;
  }
  // dielectric.rs:168-182
  static R fresnel(V3<R> wo, V3<R> n, R n1, R n2) {
    R e = n1 / n2;
    R ci = wo.z * n.z;
    R ct = std::sqrt(R(1) - e * e * (R(1) - ci * ci));
    R ro = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct);
    R rp = (n2 * ci - n1 * ct) / (n2 * ci + n1 * ct);
    return (ro * ro + rp * rp) * R(0.5);
  }
  // dielectric.rs:31-127
  bool dielectric(const MaterialDesc& m, const Isect<R>& is, uint64_t seed, uint64_t path, uint32_t bounce, Scatter* s) const {
    V3<R> lwo = normalized(to_local(is, is.wo));
    V3<R> lrefl{-lwo.x, -lwo.y, lwo.z};
    Spec<R> att = albedo(m);
    V3<R> lnorm, wnorm;
    R ni, nt;
    if (lwo.z <= R(0)) {
      lnorm = {0, 0, -1};
      wnorm = is.n * R(-1);
      ni = (R)m.ior;
      nt = R(1);
    } else {
      lnorm = {0, 0, 1};
      wnorm = is.n;
      ni = R(1);
      nt = (R)m.ior;
    }
    V3<R> lrefr;
    V3<R> ldir, org;
    if (refract(lwo, lnorm, ni / nt, &lrefr)) {
      R refl = fresnel(lwo, lnorm, ni, nt);
      if (draw(seed, path, bounce, kCoin) < refl) {
        ldir = lrefl;
        org = is.p + wnorm * R(0.001);
      } else {
        ldir = lrefr;
        org = is.p + wnorm * R(-0.001);
      }
    } else {
      ldir = lrefl;
      org = is.p + wnorm * R(0.001);
    }
    s->ray = {org, normalized(to_world(is, ldir))};
    s->att = sdiv(att, std::fabs(ldir.z));
    s->pdf = R(1);
    s->specular = true;
    return true;
  }

  // Light::sample_li. Point light: point_light.rs:23-35 + VisibilityTester::between (lights.rs:38-46).
  // Area light (rectangle, uniform in area, one-sided along its normal): ideas/area_light.md §3.
  struct LiSample {
    Spec<R> li;
    V3<R> wi;
    R pdf;
    Ray<R> shadow;
    R dist;
  };
  bool sample_li(const LightDesc& L, const Isect<R>& is, uint64_t seed, uint64_t path, uint32_t bounce, LiSample* s) const {
    if (L.kind == kPointLight) {
      V3<R> pos{(R)L.pos[0], (R)L.pos[1], (R)L.pos[2]};
      V3<R> wi = pos - is.p;
      s->li = sdiv(Spec<R>{{(R)L.rgb[0], (R)L.rgb[1], (R)L.rgb[2]}}, dot(wi, wi));
      s->wi = normalized(wi);
      s->pdf = R(1);
      s->shadow = {is.p, normalized(normalized(pos - is.p))};  // Ray::spawn_from_through
      s->dist = length(pos - is.p);
      return true;
    }
    if (L.kind == kBackgroundLight) {
      // background_infinite_light.rs:31-44: a direction uniform on the sphere (pdfs/sphere.rs:
      // 14-23), VisibilityTester::towards_infinity (lights.rs:57-62)
      R phi = R(2) * R(M_PI) * draw(seed, path, bounce, kLightU);
      R cos_t = R(1) - R(2) * draw(seed, path, bounce, kLightV);
      R sin_t = std::sqrt(R(1) - cos_t * cos_t);
      V3<R> wi{std::cos(phi) * sin_t, std::sin(phi) * sin_t, cos_t};
      R factor = R(0.5) * (normalized(wi).y + R(1));
      s->li = {{(R)L.pos[0] * (R(1) - factor) + (R)L.rgb[0] * factor, (R)L.pos[1] * (R(1) - factor) + (R)L.rgb[1] * factor,
                (R)L.pos[2] * (R(1) - factor) + (R)L.rgb[2] * factor}};
      s->wi = wi;
      s->pdf = R(1) / (R(4) * R(M_PI));
      s->shadow = {is.p, normalized(wi)};
      s->dist = std::numeric_limits<R>::max();
      return true;
    }
    const Obj& o = objs[L.object];
    R w = (R)o.d.param[0], h = (R)o.d.param[1];
    R u1 = draw(seed, path, bounce, kLightU), u2 = draw(seed, path, bounce, kLightV);
    V3<R> pl_local{(u1 - R(0.5)) * w, R(0), (u2 - R(0.5)) * h};
    V3<R> pl = o.d.has_xform ? o.x.point(o.x.m, pl_local) : pl_local;
    V3<R> nl = o.d.has_xform ? normalized(o.x.normal(o.x.inv, V3<R>{0, 1, 0})) : V3<R>{0, 1, 0};
    V3<R> d = pl - is.p;
    R dist2 = dot(d, d);
    R dist = std::sqrt(dist2);
    V3<R> wi = d * (R(1) / dist);
    R cos_l = -dot(nl, wi);
    if (!(cos_l > R(0))) return false;
    R sx = o.d.has_xform ? length(o.x.dir(o.x.m, V3<R>{1, 0, 0})) : R(1);
    R sz = o.d.has_xform ? length(o.x.dir(o.x.m, V3<R>{0, 0, 1})) : R(1);
    R area = w * sx * h * sz;
    s->li = {{(R)L.rgb[0], (R)L.rgb[1], (R)L.rgb[2]}};
    s->wi = wi;
    s->pdf = dist2 / (area * cos_l);
    s->shadow = {is.p, wi};
    s->dist = dist - R(1e-4);  // stop short of the emitter's own surface
    return true;
  }

  // Density, in solid angle, with which scatter() produces direction wi_w at this vertex; 0 for
  // the specular and non-scattering materials.
  R bsdf_pdf(const MaterialDesc& m, const Isect<R>& is, V3<R> wo_w, V3<R> wi_w) const {
    if (m.kind == kLambertian) {
      R c = dot(is.n, wi_w);
      return c > R(0) ? c / R(M_PI) : R(0);
    }
    if (m.kind == kGGX) {
      V3<R> wo = to_local(is, wo_w), wi = to_local(is, wi_w);
      if (wo.z <= R(0) || wi.z <= R(0)) return R(0);
      V3<R> h = normalized(wo + wi);
      return ggx_d(h.z, (R)m.roughness) * h.z / (R(4) * dot(wo, h));
    }
    return R(0);
  }
  // Density, in solid angle and with the choice of light, with which next-event estimation picks
  // direction d from o towards the area light on object k, met at distance t; 0 if no area light
  // sits on k or d reaches its back.
  R light_pdf(uint32_t k, V3<R> d, R t) const {
    for (const LightDesc& L : lights) {
      if (L.kind == kPointLight || (uint32_t)L.object != k) continue;
      const Obj& o = objs[k];
      V3<R> nl = o.d.has_xform ? normalized(o.x.normal(o.x.inv, V3<R>{0, 1, 0})) : V3<R>{0, 1, 0};
      R cos_l = -dot(nl, d);
      if (!(cos_l > R(0))) return R(0);
      R sx = o.d.has_xform ? length(o.x.dir(o.x.m, V3<R>{1, 0, 0})) : R(1);
      R sz = o.d.has_xform ? length(o.x.dir(o.x.m, V3<R>{0, 0, 1})) : R(1);
      R area = (R)o.d.param[0] * sx * (R)o.d.param[1] * sz;
      return (R(1) / (R)lights.size()) * (t * t / (area * cos_l));
    }
    return R(0);
  }
  // power heuristic weight of the strategy with density a against b, a^2 / (a^2 + b^2), written as
  // 1 / (1 + (b / a)^2) so that an infinite density (a light sample at a grazing angle) gives 1 or
  // 0 rather than inf / inf; 0 when a is 0, 1/2 when both are infinite
  static R power(R a, R b) {
    if (!(a > R(0))) return R(0);
    R r = b / a;
    if (r != r) return R(0.5);
    return R(1) / (R(1) + r * r);
  }

  // integrators/path.rs:48-115 for one camera ray. With mis: next-event estimation weighted by
  // the power heuristic against the BSDF's density of the light's direction, and the emission a
  // BSDF-sampled ray meets after a non-specular vertex added too, weighted the other way (an
  // emitter no light samples gets weight 1).
  // NormalIntegrator::li (integrators/normal.rs:17-30): (n + 1) * 0.5 of the hit, black on a miss
  // (the trait's background_radiance, integrators.rs:9-11).
  Spec<R> li_normal(const Ray<R>& cam, R near, R far, uint64_t* rays_cast) const {
    SceneHit<R> hit;
    SStats st;
    (*rays_cast)++;
    if (!intersect(cam, near, far, &hit, &st)) return {{0, 0, 0}};
    const V3<R> n = hit.is.n;
    return {{(n.x + R(1)) * R(0.5), (n.y + R(1)) * R(0.5), (n.z + R(1)) * R(0.5)}};
  }
  // NaiveIntegrator::li (integrators/naive.rs:30-54), its recursion unrolled: at every hit the
  // emission times the throughput, then the scattered ray with the throughput times
  // attenuation * |cos| / pdf, until the depth is spent (li at depth 0 is black) or the material does
  // not scatter; on a miss the integrator's own sky (naive.rs:21-26) times the throughput. The
  // draws of vertex k are the path integrator's (bounce k).
  Spec<R> li_naive(const Ray<R>& cam, R near, R far, uint32_t max_depth, uint64_t seed, uint64_t path,
                   uint64_t* rays_cast) const {
    Spec<R> beta{{1, 1, 1}}, L{{0, 0, 0}};
    Ray<R> ray = cam;
    SStats st;
    for (uint32_t bounce = 1; bounce <= max_depth; ++bounce) {
      SceneHit<R> hit;
      (*rays_cast)++;
      if (!intersect(ray, near, far, &hit, &st)) {
        V3<R> u = normalized(ray.d);
        R t = R(0.5) * (u.y + R(1));
        Spec<R> sky{{R(1) * (R(1) - t) + R(0.5) * t, R(1) * (R(1) - t) + R(0.7) * t, R(1) * (R(1) - t) + R(1) * t}};
        L = sadd(L, smul(sky, beta));
        break;
      }
      const MaterialDesc m = shaded(mats[objs[hit.object].d.material], hit.is);
      if (m.kind == kDiffuseLight) L = sadd(L, smul(albedo(m), beta));
      if (bounce == max_depth) break;
      Scatter s;
      if (!scatter(m, hit.is, seed, path, bounce, &s)) break;
      R c = std::fabs(dot(s.ray.d, hit.is.n));
      beta = smul(beta, sdiv(sscale(s.att, c), s.pdf));
      if (!(std::isfinite(beta.c[0]) && std::isfinite(beta.c[1]) && std::isfinite(beta.c[2]))) break;
      ray = s.ray;
    }
    return L;
  }

  Spec<R> li(const Ray<R>& cam, R near, R far, uint32_t max_depth, uint64_t seed, uint64_t path, uint64_t* rays_cast) const {
    if (integrator == 2) return li_normal(cam, near, far, rays_cast);
    if (integrator == 1) return li_naive(cam, near, far, max_depth, seed, path, rays_cast);
    Spec<R> beta{{1, 1, 1}}, L{{0, 0, 0}};
    Ray<R> ray = cam;
    bool specular = true;
    R prev_pdf = R(0);  // BSDF density of the direction that led here (non-specular vertices)
    uint32_t bounce = 0;
    SStats st;
    for (;;) {
      SceneHit<R> hit;
      (*rays_cast)++;
      if (!intersect(ray, near, far, &hit, &st)) {
        // path.rs:101-108: the background's radiance after a specular bounce; with mis also after a
        // non-specular one, weighted against next-event estimation's density
        if (has_background()) {
          if (specular) {
            L = sadd(L, smul(background_le(ray.d), beta));
          } else if (mis) {
            R w = power(prev_pdf, (R(1) / (R)lights.size()) * (R(1) / (R(4) * R(M_PI))));
            L = sadd(L, sscale(smul(background_le(ray.d), beta), w));
          }
        }
        break;
      }
      const MaterialDesc m = shaded(mats[objs[hit.object].d.material], hit.is);
      if (specular && m.kind == kDiffuseLight) L = sadd(L, smul(albedo(m), beta));
      if (mis && !specular && m.kind == kDiffuseLight) {
        R w = power(prev_pdf, light_pdf((uint32_t)hit.object, ray.d, hit.is.d));
        L = sadd(L, sscale(smul(albedo(m), beta), w));
      }
      bounce += 1;
      if (bounce == max_depth) break;
      if (!lights.empty()) {
        uint32_t n = (uint32_t)lights.size();
        R u = draw(seed, path, bounce, kLightPick);
        R fi = std::fmin(u * (R)n, (R)n - R(1));
        const LightDesc& Ld = lights[(uint32_t)fi];
        R pl = R(1) / (R)n;
        LiSample ls;
        if (sample_li(Ld, hit.is, seed, path, bounce, &ls)) {
          Spec<R> f = sscale(mat_f(m, hit.is, ray.d * R(-1), ls.wi), std::fabs(dot(ls.wi, hit.is.n)));
          Spec<R> c = sdiv(smul(smul(ls.li, f), beta), pl * ls.pdf);
          if (mis && Ld.kind != kPointLight) c = sscale(c, power(pl * ls.pdf, bsdf_pdf(m, hit.is, ray.d * R(-1), ls.wi)));
          (*rays_cast)++;
          if (!intersect_p(ls.shadow, R(1e-5), ls.dist, &st)) L = sadd(L, c);
        }
      }
      Scatter s;
      if (!scatter(m, hit.is, seed, path, bounce, &s)) break;
      R c = std::fabs(dot(s.ray.d, hit.is.n));
      beta = smul(beta, sdiv(sscale(s.att, c), s.pdf));
      // a non-finite throughput (grazing specular event, 1/|cos| = inf) ends the path, as on the
      // device; the reference would carry a NaN
      if (!(std::isfinite(beta.c[0]) && std::isfinite(beta.c[1]) && std::isfinite(beta.c[2]))) break;
      if (roulette) {
        // path.rs:93-100: continue with probability q, the throughput's largest component (capped
        // at 1 here, so that the estimator stays unbiased), and divide the throughput by q; the path
        // survives only when the draw is below q, so a zero throughput (q = 0) always ends it
        const R q = std::fmin(R(1), std::fmax(std::fmax(beta.c[0], beta.c[1]), beta.c[2]));
        if (!(draw(seed, path, bounce, kRoulette) < q)) break;
        beta = sdiv(beta, q);
      }
      ray = s.ray;
      specular = s.specular;
      prev_pdf = s.pdf;
    }
    return L;
  }
};

}  // namespace orc
