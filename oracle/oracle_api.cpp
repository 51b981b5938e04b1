// ORACLE — test infrastructure, not product code.
//
// C entry points over the CPU restatements in mesh_oracle.hpp / camera_oracle.hpp, loaded by the
// tests and by the CPU-baseline leg of bench.py through ctypes. Precision 64 is "Oracle A" (the
// reference's f64 arithmetic), precision 32 is "Oracle B" (the same statements in FP32).
#include <atomic>
#include <chrono>
#include <cstdio>
#include <thread>
#include <vector>

#include "camera_oracle.hpp"
#include "mesh_oracle.hpp"

using namespace orc;

namespace {

struct MeshHandle {
  int precision;
  MeshBVH<double>* a = nullptr;
  MeshBVH<float>* b = nullptr;
  ~MeshHandle() {
    delete a;
    delete b;
  }
};

template <class R>
const typename MeshBVH<R>::Node* node_ptr(const MeshBVH<R>* m) {
  return m->nodes.data();
}

// Runs fn(i) for i in [0, n) over `threads` threads in static contiguous chunks.
template <class F>
void parallel_for(uint64_t n, int threads, F fn) {
  if (threads <= 1 || n < 1024) {
    for (uint64_t i = 0; i < n; ++i) fn(i, 0);
    return;
  }
  std::vector<std::thread> pool;
  uint64_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    uint64_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    pool.emplace_back([=]() {
      for (uint64_t i = lo; i < hi; ++i) fn(i, t);
    });
  }
  for (auto& th : pool) th.join();
}

// mode: 0 reference-ordered closest hit, 1 canonical closest hit, 2 brute force, 3 any-hit.
template <class R, class RayT>
void trace_impl(const MeshBVH<R>* m, const RayT* rays, uint64_t n, int mode, int64_t* prim, double* t, double* u,
                double* v, uint32_t* stats_per_ray, uint64_t* stats_sum, int threads) {
  std::vector<Stats> per_thread(std::max(threads, 1));
  parallel_for(n, threads, [&](uint64_t i, int tid) {
    const RayT* r = rays + 8 * i;
    Ray<R> ray{{(R)r[0], (R)r[1], (R)r[2]}, {(R)r[4], (R)r[5], (R)r[6]}};
    R near = (R)r[3], far = (R)r[7];
    Stats st;
    TriHit<R> h{};
    uint64_t id = 0;
    bool found;
    if (mode == 3) {
      found = m->intersect_p(ray, near, far);
    } else if (mode == 2) {
      found = m->brute_force(ray, near, far, &h, &id);
    } else {
      found = m->traverse(ray, near, far, mode == 1, &st, &h, &id);
    }
    if (prim) prim[i] = found ? (mode == 3 ? 1 : (int64_t)id) : -1;
    if (found && mode != 3) {
      if (t) t[i] = (double)h.t;
      if (u) u[i] = (double)h.u;
      if (v) v[i] = (double)h.v;
    } else {
      if (t) t[i] = 0.0;
      if (u) u[i] = 0.0;
      if (v) v[i] = 0.0;
    }
    if (stats_per_ray) {
      stats_per_ray[3 * i] = (uint32_t)st.nodes_visited;
      stats_per_ray[3 * i + 1] = (uint32_t)st.box_tests;
      stats_per_ray[3 * i + 2] = (uint32_t)st.triangle_tests;
    }
    Stats& acc = per_thread[tid];
    acc.nodes_visited += st.nodes_visited;
    acc.box_tests += st.box_tests;
    acc.triangle_tests += st.triangle_tests;
  });
  if (stats_sum) {
    stats_sum[0] = stats_sum[1] = stats_sum[2] = 0;
    for (auto& s : per_thread) {
      stats_sum[0] += s.nodes_visited;
      stats_sum[1] += s.box_tests;
      stats_sum[2] += s.triangle_tests;
    }
  }
}

}  // namespace

extern "C" {

void* orc_mesh_create(const float* xyz, uint64_t nv, const uint32_t* idx, uint64_t nt, int precision, int bins,
                      double t_trav, int max_leaf) {
  if (nt == 0 || bins < 2) return nullptr;  // triangle_mesh.rs:28 refuses an empty mesh
  BuildParams p;
  p.bins = bins;
  p.t_trav = t_trav;
  p.max_leaf = max_leaf;
  auto* h = new MeshHandle();
  h->precision = precision;
  if (precision == 64)
    h->a = new MeshBVH<double>(xyz, nv, idx, nt, p);
  else
    h->b = new MeshBVH<float>(xyz, nv, idx, nt, p);
  return h;
}

void orc_mesh_destroy(void* h) { delete (MeshHandle*)h; }

void orc_mesh_build_stats(void* hv, uint64_t* out) {
  auto* h = (MeshHandle*)hv;
  BuildStats s = h->a ? h->a->build_stats() : h->b->build_stats();
  out[0] = s.node_count;
  out[1] = s.leaf_count;
  out[2] = s.max_depth;
  out[3] = s.max_leaf_tri_count;
  out[4] = s.total_leaf_tri_count;
}

uint64_t orc_mesh_node_count(void* hv) {
  auto* h = (MeshHandle*)hv;
  return h->a ? h->a->nodes.size() : h->b->nodes.size();
}

// Tree in the reference's own layout: node boxes (f64, min xyz then max xyz), left_first,
// tri_count (0 = interior), and the triangle permutation tri_idx.
void orc_mesh_export(void* hv, double* boxes, uint64_t* left_first, uint64_t* tri_count, uint64_t* tri_idx) {
  auto* h = (MeshHandle*)hv;
  auto dump = [&](auto* m) {
    for (size_t i = 0; i < m->nodes.size(); ++i) {
      const auto& n = m->nodes[i];
      double* b = boxes + 6 * i;
      b[0] = n.bbox.bmin.x; b[1] = n.bbox.bmin.y; b[2] = n.bbox.bmin.z;
      b[3] = n.bbox.bmax.x; b[4] = n.bbox.bmax.y; b[5] = n.bbox.bmax.z;
      left_first[i] = n.left_first;
      tri_count[i] = n.tri_count;
    }
    for (size_t i = 0; i < m->tri_idx.size(); ++i) tri_idx[i] = m->tri_idx[i];
  };
  if (h->a) dump(h->a); else dump(h->b);
}

// Rays as 8 floats (ox, oy, oz, tmin, dx, dy, dz, tmax); Oracle A widens them exactly to f64.
int orc_mesh_trace_f32(void* hv, const float* rays, uint64_t n, int mode, int64_t* prim, double* t, double* u,
                       double* v, uint32_t* stats_per_ray, uint64_t* stats_sum, int threads) {
  auto* h = (MeshHandle*)hv;
  if (h->a)
    trace_impl(h->a, rays, n, mode, prim, t, u, v, stats_per_ray, stats_sum, threads);
  else
    trace_impl(h->b, rays, n, mode, prim, t, u, v, stats_per_ray, stats_sum, threads);
  return 0;
}

// Rays as 8 doubles; Oracle B rounds them to FP32.
int orc_mesh_trace_f64(void* hv, const double* rays, uint64_t n, int mode, int64_t* prim, double* t, double* u,
                       double* v, uint32_t* stats_per_ray, uint64_t* stats_sum, int threads) {
  auto* h = (MeshHandle*)hv;
  if (h->a)
    trace_impl(h->a, rays, n, mode, prim, t, u, v, stats_per_ray, stats_sum, threads);
  else
    trace_impl(h->b, rays, n, mode, prim, t, u, v, stats_per_ray, stats_sum, threads);
  return 0;
}

// Split cost of the best binned plane for the node spanning tri_idx[left_first, +count) of an
// unbuilt tree, and the exhaustive recount of the same plane (triangle_mesh/bvh.rs:867-894).
// Returns 0 when no plane separates the range.
int orc_split_costs(const float* xyz, uint64_t nv, const uint32_t* idx, uint64_t nt, int bins, uint64_t left_first,
                    uint64_t count, double* binned, double* exhaustive) {
  BuildParams p;
  p.bins = bins;
  p.max_leaf = 0;
  // Build with a huge t_trav so the root stays a leaf: tri_idx keeps its identity order.
  p.t_trav = 1e300;
  MeshBVH<double> m(xyz, nv, idx, nt, p);
  MeshBVH<double>::Node node{AABB<double>::empty(), left_first, count};
  m.update_bounds(node);
  auto s = m.find_best_split_plane(node);
  if (!s.valid) return 0;
  *binned = s.cost;
  *exhaustive = m.exhaustive_split_cost(node, s);
  return 1;
}

// The bvh_stats ray protocol (bvh_stats.rs:345-401) as 8 doubles per ray.
void orc_orbit_rays(const float* xyz, uint64_t nv, unsigned side, double fov, int views, double* out) {
  std::vector<double> v(nv * 3);
  for (uint64_t i = 0; i < nv * 3; ++i) v[i] = (double)xyz[i];
  std::vector<Ray<double>> rays((uint64_t)views * side * side);
  orbit_rays(v.data(), nv, side, fov, views, rays.data());
  for (size_t i = 0; i < rays.size(); ++i) {
    double* o = out + 8 * i;
    o[0] = rays[i].o.x; o[1] = rays[i].o.y; o[2] = rays[i].o.z; o[3] = 1e-4;
    o[4] = rays[i].d.x; o[5] = rays[i].d.y; o[6] = rays[i].d.z; o[7] = 1000.0;
  }
}

// One slab test (aabound.rs:323-355) at either precision, for the reference's box cases.
int orc_box_hit(int precision, const double* box6, const double* ray6, double tmin, double tmax, double* entry) {
  if (precision == 64) {
    AABB<double> b{{box6[0], box6[1], box6[2]}, {box6[3], box6[4], box6[5]}};
    Ray<double> r{{ray6[0], ray6[1], ray6[2]}, {ray6[3], ray6[4], ray6[5]}};
    return b.hit(r, tmin, tmax, entry) ? 1 : 0;
  }
  AABB<float> b{{(float)box6[0], (float)box6[1], (float)box6[2]}, {(float)box6[3], (float)box6[4], (float)box6[5]}};
  Ray<float> r{{(float)ray6[0], (float)ray6[1], (float)ray6[2]}, {(float)ray6[3], (float)ray6[4], (float)ray6[5]}};
  float e = 0;
  int ok = b.hit(r, (float)tmin, (float)tmax, &e) ? 1 : 0;
  *entry = e;
  return ok;
}

double orc_slab_widening(int precision) { return precision == 64 ? slab_widening<double>() : (double)slab_widening<float>(); }

}  // extern "C"

// ---- scene level -------------------------------------------------------------------------------
#include "scene_oracle.hpp"

namespace {
struct SceneHandle {
  SceneOracle<double>* a = nullptr;
  SceneOracle<float>* b = nullptr;
  ~SceneHandle() {
    delete a;
    delete b;
  }
};

template <class R, class RayT>
void scene_trace_impl(const SceneOracle<R>* s, const RayT* rays, uint64_t n, int any, int64_t* obj, int64_t* prim,
                      double* t, double* u, double* v, double* p, uint64_t* stats_sum, int threads) {
  std::vector<typename SceneOracle<R>::SStats> per(std::max(threads, 1));
  parallel_for(n, threads, [&](uint64_t i, int tid) {
    const RayT* r = rays + 8 * i;
    Ray<R> ray{{(R)r[0], (R)r[1], (R)r[2]}, {(R)r[4], (R)r[5], (R)r[6]}};
    R near = (R)r[3], far = (R)r[7];
    auto& st = per[tid];
    if (any) {
      obj[i] = s->intersect_p(ray, near, far, &st) ? 1 : -1;
      return;
    }
    SceneHit<R> h;
    bool f = s->intersect(ray, near, far, &h, &st);
    obj[i] = f ? h.object : -1;
    if (prim) prim[i] = f ? h.prim : -1;
    if (t) t[i] = f ? (double)h.is.d : 0.0;
    if (u) u[i] = f ? (double)h.is.u : 0.0;
    if (v) v[i] = f ? (double)h.is.v : 0.0;
    if (p) {
      p[3 * i] = f ? (double)h.is.p.x : 0.0;
      p[3 * i + 1] = f ? (double)h.is.p.y : 0.0;
      p[3 * i + 2] = f ? (double)h.is.p.z : 0.0;
    }
  });
  if (stats_sum) {
    stats_sum[0] = stats_sum[1] = stats_sum[2] = 0;
    for (auto& q : per) {
      stats_sum[0] += q.nodes_visited;
      stats_sum[1] += q.box_tests;
      stats_sum[2] += q.object_tests;
    }
  }
}

template <class R>
void render_impl(const SceneOracle<R>* s, uint32_t w, uint32_t h, double fov, const double* eye, const double* look,
                 uint32_t spp, uint32_t max_depth, uint64_t seed, double near, double far, uint32_t y0, uint32_t y1,
                 double* film, uint64_t* rays_cast, int threads) {
  PinHole cam(w, h, fov, near, far, m4_look_at({eye[0], eye[1], eye[2]}, {look[0], look[1], look[2]}, {0, 1, 0}));
  s->draw_spp = spp;
  std::vector<uint64_t> per(std::max(threads, 1), 0);
  const uint64_t npix = (uint64_t)w * (y1 - y0);
  parallel_for(npix, threads, [&](uint64_t k, int tid) {
    const uint64_t px = k % w, py = y0 + k / w;
    const uint64_t pixel = py * w + px;
    Spec<R> sum{{0, 0, 0}};
    double acc[3] = {0, 0, 0};
    for (uint32_t sidx = 0; sidx < spp; ++sidx) {
      const uint64_t path = pixel * spp + sidx;
      // renderers/mt.rs:123-127: pixel + jitter, camera ray, li
      // the f64 camera (the reference's arithmetic) for Oracle A, the renderer's FP32 camera for B
      Ray<R> ray;
      if (sizeof(R) == 8) {
        double jx = s->jitter(seed, pixel, sidx, spp, kJitterX), jy = s->jitter(seed, pixel, sidx, spp, kJitterY);
        Ray<double> cr = cam.get_ray((double)px + jx, (double)py + jy);
        ray = {{(R)cr.o.x, (R)cr.o.y, (R)cr.o.z}, {(R)cr.d.x, (R)cr.d.y, (R)cr.d.z}};
      } else {
        float jx = s->jitter_f32(seed, pixel, sidx, spp, kJitterX), jy = s->jitter_f32(seed, pixel, sidx, spp, kJitterY);
        Ray<float> cr = cam.get_ray_f32((float)px + jx, (float)py + jy);
        ray = {{(R)cr.o.x, (R)cr.o.y, (R)cr.o.z}, {(R)cr.d.x, (R)cr.d.y, (R)cr.d.z}};
      }
      Spec<R> L = s->li(ray, (R)near, (R)far, max_depth, seed, path, &per[tid]);
      for (int c = 0; c < 3; ++c) acc[c] += (double)L.c[c];
      (void)sum;
    }
    for (int c = 0; c < 3; ++c) film[3 * k + c] = acc[c] / (double)spp;
  });
  if (rays_cast) {
    *rays_cast = 0;
    for (auto x : per) *rays_cast += x;
  }
}
}  // namespace

extern "C" {

// precision 64: reference arithmetic and tie order; 32: FP32 with the order-independent tie rule.
void* orc_scene_create(int precision, const float* verts, const uint64_t* vert_off, const uint32_t* idx,
                       const uint64_t* idx_off, uint32_t n_mesh, const void* objects, uint32_t n_obj,
                       const void* materials, uint32_t n_mat, const void* lights, uint32_t n_light, const void* csg,
                       uint32_t n_csg, int bins, double t_trav, int max_leaf) {
  BuildParams bp;
  bp.bins = bins;
  bp.t_trav = t_trav;
  bp.max_leaf = max_leaf;
  auto* h = new SceneHandle();
  if (precision == 64) {
    h->a = new SceneOracle<double>();
    h->a->build(verts, vert_off, idx, idx_off, n_mesh, (const ObjectDesc*)objects, n_obj, (const MaterialDesc*)materials,
                n_mat, (const LightDesc*)lights, n_light, (const CsgElemDesc*)csg, n_csg, bp, false);
  } else {
    h->b = new SceneOracle<float>();
    h->b->build(verts, vert_off, idx, idx_off, n_mesh, (const ObjectDesc*)objects, n_obj, (const MaterialDesc*)materials,
                n_mat, (const LightDesc*)lights, n_light, (const CsgElemDesc*)csg, n_csg, bp, true);
  }
  return h;
}

void orc_scene_destroy(void* h) { delete (SceneHandle*)h; }

uint32_t orc_desc_sizes(uint32_t* obj, uint32_t* mat, uint32_t* light) {
  *obj = sizeof(ObjectDesc);
  *mat = sizeof(MaterialDesc);
  *light = sizeof(LightDesc);
  return sizeof(CsgElemDesc);
}

int orc_scene_trace_f32(void* hv, const float* rays, uint64_t n, int any, int64_t* obj, int64_t* prim, double* t,
                        double* u, double* v, double* p, uint64_t* stats, int threads) {
  auto* h = (SceneHandle*)hv;
  if (h->a) scene_trace_impl(h->a, rays, n, any, obj, prim, t, u, v, p, stats, threads);
  else scene_trace_impl(h->b, rays, n, any, obj, prim, t, u, v, p, stats, threads);
  return 0;
}

int orc_scene_trace_f64(void* hv, const double* rays, uint64_t n, int any, int64_t* obj, int64_t* prim, double* t,
                        double* u, double* v, double* p, uint64_t* stats, int threads) {
  auto* h = (SceneHandle*)hv;
  if (h->a) scene_trace_impl(h->a, rays, n, any, obj, prim, t, u, v, p, stats, threads);
  else scene_trace_impl(h->b, rays, n, any, obj, prim, t, u, v, p, stats, threads);
  return 0;
}

// Pinhole rays through pixel centres (pin_hole.rs:17-50), 8 doubles each, row-major.
void orc_pinhole_rays(uint32_t w, uint32_t hgt, double fov, const double* eye, const double* look, const double* up,
                      double near, double far, double* out) {
  PinHole cam(w, hgt, fov, near, far, m4_look_at({eye[0], eye[1], eye[2]}, {look[0], look[1], look[2]}, {up[0], up[1], up[2]}));
  for (uint32_t y = 0; y < hgt; ++y)
    for (uint32_t x = 0; x < w; ++x) {
      Ray<double> r = cam.get_ray((double)x + 0.5, (double)y + 0.5);
      double* o = out + 8 * ((uint64_t)y * w + x);
      o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z; o[3] = near;
      o[4] = r.d.x; o[5] = r.d.y; o[6] = r.d.z; o[7] = far;
    }
}

// The scene-mode measurement of bvh_stats.rs:254-286 in f64: primary rays through every pixel
// centre, and from each hit a shadow ray towards `target` over [shadow_near, distance]. out:
// primary rays, primary hits, shadow hits; stats: primary {nodes, boxes, objects}, shadow {...}.
int orc_scene_bvh_stats(void* hv, uint32_t w, uint32_t hgt, double fov, const double* eye, const double* look,
                        const double* up, double near, double far, const double* target, double shadow_near,
                        uint64_t* out, uint64_t* stats) {
  auto* h = (SceneHandle*)hv;
  if (!h->a) return -1;
  const SceneOracle<double>& s = *h->a;
  PinHole cam(w, hgt, fov, near, far, m4_look_at({eye[0], eye[1], eye[2]}, {look[0], look[1], look[2]}, {up[0], up[1], up[2]}));
  SceneOracle<double>::SStats ps, ss;
  uint64_t ph = 0, sh = 0;
  const V3<double> tgt{target[0], target[1], target[2]};
  for (uint32_t y = 0; y < hgt; ++y)
    for (uint32_t x = 0; x < w; ++x) {
      Ray<double> r = cam.get_ray((double)x + 0.5, (double)y + 0.5);
      SceneHit<double> hit;
      if (!s.intersect(r, near, far, &hit, &ps)) continue;
      ph++;
      Ray<double> sr{hit.is.p, normalized(normalized(tgt - hit.is.p))};  // Ray::spawn_from_through
      double dist = length(tgt - hit.is.p);
      if (s.intersect_p(sr, shadow_near, dist, &ss)) sh++;
    }
  out[0] = (uint64_t)w * hgt;
  out[1] = ph;
  out[2] = sh;
  stats[0] = ps.nodes_visited; stats[1] = ps.box_tests; stats[2] = ps.object_tests;
  stats[3] = ss.nodes_visited; stats[4] = ss.box_tests; stats[5] = ss.object_tests;
  return 0;
}

// 32 bits of coordinate `axis` of the Sobol sampler's point `sample` of pixel `pixel` in stratified
// pair `pair`, for n consecutive samples from `first`.
void orc_sobol_bits(uint64_t seed, uint64_t pixel, uint32_t first, uint32_t n, uint32_t pair, uint32_t axis,
                    uint32_t* out) {
  for (uint32_t i = 0; i < n; ++i) out[i] = sobol_bits(seed, pixel, first + i, pair, axis);
}

// The sampler of the renders (0: counter-based draws; 1: stratified pairs from scrambled Sobol points).
int orc_scene_set_sampler(void* hv, int kind) {
  auto* h = (SceneHandle*)hv;
  if (kind < 0 || kind > 1) return -1;
  if (h->a) h->a->sampler = kind;
  if (h->b) h->b->sampler = kind;
  return 0;
}

// Multiple importance sampling in the path integrator (0: the reference's estimator).
int orc_scene_set_mis(void* hv, int mis) {
  auto* h = (SceneHandle*)hv;
  if (h->a) h->a->mis = mis != 0;
  if (h->b) h->b->mis = mis != 0;
  return 0;
}

// The integrator of li (integrators.rs:23-28): 0 path, 1 naive, 2 normal.
int orc_scene_set_integrator(void* hv, int kind) {
  auto* h = (SceneHandle*)hv;
  if (kind < 0 || kind > 2) return -1;
  if (h->a) h->a->integrator = kind;
  if (h->b) h->b->integrator = kind;
  return 0;
}

int orc_scene_set_roulette(void* hv, int on) {
  auto* h = (SceneHandle*)hv;
  if (h->a) h->a->roulette = on != 0;
  if (h->b) h->b->roulette = on != 0;
  return 0;
}

// ---- shading primitives, for the closed-form checks of the oracle ------------------------------
}  // extern "C"

namespace {
// A canonical shading frame: p at the origin, n = +z, dpdu = +x (so the local frame is the world's).
template <class R>
Isect<R> canonical_isect(const double* wo) {
  Isect<R> is{};
  is.n = {0, 0, 1};
  is.dpdu = {1, 0, 0};
  is.dpdv = {0, 1, 0};
  is.wo = {(R)wo[0], (R)wo[1], (R)wo[2]};
  return is;
}
template <class R>
double fresnel_impl(const double* wo, const double* n, double n1, double n2) {
  return (double)SceneOracle<R>::fresnel({(R)wo[0], (R)wo[1], (R)wo[2]}, {(R)n[0], (R)n[1], (R)n[2]}, (R)n1, (R)n2);
}
template <class R>
int refract_impl(const double* wo, const double* n, double eta, double* t) {
  V3<R> r{};
  const bool ok = SceneOracle<R>::refract({(R)wo[0], (R)wo[1], (R)wo[2]}, {(R)n[0], (R)n[1], (R)n[2]}, (R)eta, &r);
  t[0] = r.x;
  t[1] = r.y;
  t[2] = r.z;
  return ok ? 1 : 0;
}
template <class R>
void material_eval_impl(const MaterialDesc* m, const double* wo, const double* wi, double* f, double* pdf) {
  SceneOracle<R> s;
  Isect<R> is = canonical_isect<R>(wo);
  V3<R> wiv{(R)wi[0], (R)wi[1], (R)wi[2]};
  Spec<R> v = s.mat_f(*m, is, is.wo, wiv);
  for (int c = 0; c < 3; ++c) f[c] = v.c[c];
  *pdf = s.bsdf_pdf(*m, is, is.wo, wiv);
}
template <class R>
int material_scatter_impl(const MaterialDesc* m, const double* wo, uint64_t seed, uint64_t path, uint32_t bounce,
                          double* att, double* dir, double* pdf, int* specular) {
  SceneOracle<R> s;
  Isect<R> is = canonical_isect<R>(wo);
  typename SceneOracle<R>::Scatter sc{};
  if (!s.scatter(*m, is, seed, path, bounce, &sc)) return 0;
  for (int c = 0; c < 3; ++c) att[c] = sc.att.c[c];
  dir[0] = sc.ray.d.x;
  dir[1] = sc.ray.d.y;
  dir[2] = sc.ray.d.z;
  *pdf = sc.pdf;
  *specular = sc.specular ? 1 : 0;
  return 1;
}
template <class R>
int sample_light_impl(const SceneOracle<R>* s, uint32_t light, const double* p, const double* n, uint64_t seed,
                      uint64_t path, uint32_t bounce, double* li, double* wi, double* pdf, double* dist) {
  if (light >= s->lights.size()) return -1;
  Isect<R> is{};
  is.p = {(R)p[0], (R)p[1], (R)p[2]};
  is.n = {(R)n[0], (R)n[1], (R)n[2]};
  typename SceneOracle<R>::LiSample ls{};
  if (!s->sample_li(s->lights[light], is, seed, path, bounce, &ls)) return 0;
  for (int c = 0; c < 3; ++c) li[c] = ls.li.c[c];
  wi[0] = ls.wi.x;
  wi[1] = ls.wi.y;
  wi[2] = ls.wi.z;
  *pdf = ls.pdf;
  *dist = ls.dist;
  return 1;
}
}  // namespace

extern "C" {


// dielectric.rs:168-182 (wo and n in the local frame: only their z components enter).
double orc_fresnel(int precision, const double* wo, const double* n, double n1, double n2) {
  return precision == 64 ? fresnel_impl<double>(wo, n, n1, n2) : fresnel_impl<float>(wo, n, n1, n2);
}

// dielectric.rs:129-157: 1 and the refracted direction, or 0 (total internal reflection).
int orc_refract(int precision, const double* wo, const double* n, double eta, double* t) {
  return precision == 64 ? refract_impl<double>(wo, n, eta, t) : refract_impl<float>(wo, n, eta, t);
}

double orc_ggx_d(int precision, double cos_h, double a) {
  return precision == 64 ? SceneOracle<double>::ggx_d(cos_h, a) : (double)SceneOracle<float>::ggx_d((float)cos_h, (float)a);
}
double orc_ggx_g1(int precision, double cos_v, double a) {
  return precision == 64 ? SceneOracle<double>::ggx_g1(cos_v, a)
                         : (double)SceneOracle<float>::ggx_g1((float)cos_v, (float)a);
}

// Material::f and the density of its own sampling, for outgoing wo and incoming wi in the
// canonical frame (n = +z).
void orc_material_eval(int precision, const void* material, const double* wo, const double* wi, double* f, double* pdf) {
  const auto* m = (const MaterialDesc*)material;
  if (precision == 64) material_eval_impl<double>(m, wo, wi, f, pdf);
  else material_eval_impl<float>(m, wo, wi, f, pdf);
}

// Material::scatter in the canonical frame under the counter sampler's draws for (seed, path,
// bounce): 1 and the attenuation, direction, density and specular flag, or 0 (no scattering).
int orc_material_scatter(int precision, const void* material, const double* wo, uint64_t seed, uint64_t path,
                         uint32_t bounce, double* att, double* dir, double* pdf, int* specular) {
  const auto* m = (const MaterialDesc*)material;
  return precision == 64 ? material_scatter_impl<double>(m, wo, seed, path, bounce, att, dir, pdf, specular)
                         : material_scatter_impl<float>(m, wo, seed, path, bounce, att, dir, pdf, specular);
}

// Light::sample_li of the scene's light `light` from point p (normal n) under the counter sampler's
// draws: 1 and the radiance, direction, density (solid angle) and shadow-ray length, 0 when the
// sample is not usable (the back of a one-sided emitter), -1 for no such light.
int orc_scene_sample_light(void* hv, uint32_t light, const double* p, const double* n, uint64_t seed, uint64_t path,
                           uint32_t bounce, double* li, double* wi, double* pdf, double* dist) {
  auto* h = (SceneHandle*)hv;
  return h->a ? sample_light_impl(h->a, light, p, n, seed, path, bounce, li, wi, pdf, dist)
              : sample_light_impl(h->b, light, p, n, seed, path, bounce, li, wi, pdf, dist);
}

// Rows [y0, y1) of a width x height render: film receives the per-pixel mean radiance (f64 RGB).
int orc_scene_render(void* hv, uint32_t w, uint32_t hgt, double fov, const double* eye, const double* look, uint32_t spp,
                     uint32_t max_depth, uint64_t seed, double near, double far, uint32_t y0, uint32_t y1, double* film,
                     uint64_t* rays_cast, int threads) {
  auto* h = (SceneHandle*)hv;
  if (h->a) render_impl(h->a, w, hgt, fov, eye, look, spp, max_depth, seed, near, far, y0, y1, film, rays_cast, threads);
  else render_impl(h->b, w, hgt, fov, eye, look, spp, max_depth, seed, near, far, y0, y1, film, rays_cast, threads);
  return 0;
}

}  // extern "C"

// ---- neural radiance cache -----------------------------------------------------------------------
#include "nrc_oracle.hpp"

extern "C" {

int orc_nrc_nparams(void) { return nrc::n_params(); }

void orc_nrc_encode(const float* raw, uint64_t n, const float* lo, const float* hi, float* out) {
  for (uint64_t i = 0; i < n; ++i) nrc::encode(raw + 12 * i, lo, hi, out + (size_t)nrc::kIn * i);
}

void orc_nrc_forward(const float* params, const float* x, uint64_t n, int bf16, double* out, int threads) {
  parallel_for(n, threads, [&](uint64_t i, int) {
    nrc::Forward f;
    nrc::forward(params, x + (size_t)nrc::kIn * i, bf16 != 0, &f);
    for (int c = 0; c < nrc::kOut; ++c) out[3 * i + c] = f.act[nrc::kLayers][c];
  });
}

double orc_nrc_loss_grad(const float* params, const float* x, const float* targets, uint64_t n, int bf16, double* grad) {
  std::vector<double> g;
  double loss = nrc::loss_grad(params, x, targets, (int)n, bf16 != 0, &g);
  std::memcpy(grad, g.data(), g.size() * sizeof(double));
  return loss;
}

void orc_nrc_adam(float* params, float* m, float* v, const float* grad, int n, int t, float lr, float b1, float b2,
                  float eps) {
  nrc::adam(params, m, v, grad, n, t, lr, b1, b2, eps);
}

}  // extern "C"
