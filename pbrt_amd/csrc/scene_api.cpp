// C ABI of the scene seam and the renderer: Scene::{new, add_object, add_light, commit, intersect,
// intersect_p} (src/scene.rs:80-150) and renderers::render (src/renderers/mt.rs:23) over the device
// scene of scene.hip. See include/pbrt_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bvh_gpu.hpp"
#include "camera.hpp"
#include "common.hpp"
#include "context.hpp"
#include "nrc.hpp"
#include "nrc_render.hpp"
#include "rng.hpp"
#include "scene.hpp"

namespace pbrt {
namespace {

struct HostObject {
  int32_t shape, mesh, material, has_xform;  // for a CSG object, mesh = its first element
  double param[3];                           // CSG: operation, element count
  double m[16], minv[16];
};

struct HostCsgElem {
  int32_t shape;
  double param[3];
  double m[16], minv[16];
};

struct Box64 {
  double lo[3], hi[3];
  static Box64 empty() { return {{DBL_MAX, DBL_MAX, DBL_MAX}, {-DBL_MAX, -DBL_MAX, -DBL_MAX}}; }
  void grow(const Box64& b) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::fmin(lo[a], b.lo[a]);
      hi[a] = std::fmax(hi[a], b.hi[a]);
    }
  }
  double centroid(int a) const { return (hi[a] + lo[a]) * 0.5; }
};

template <class T>
void dev_upload(T** dst, const std::vector<T>& src) {
  const size_t n = std::max<size_t>(1, src.size());
  PBRT_HIP(hipMalloc(dst, n * sizeof(T)));
  if (!src.empty()) PBRT_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
}

}  // namespace
}  // namespace pbrt

using namespace pbrt;

struct pbrt_scene {
  pbrt_ctx* ctx = nullptr;
  pbrt_bvh_params mesh_params{12, 4, 1.0};
  bool device_build = false;  // mesh BVHs built on the device at commit (one batched build)
  pbrt_scene_launch launch{16, 0, 4, 0};  // scene walks: LDS stack, workgroups per CU, tree width, min waves
  uint32_t blocks_fit[3] = {0, 0, 0};  // workgroups per CU that fit, per stack depth (8, 16, 24)
  struct MeshRange {
    uint32_t node_base;
    uint64_t n_nodes;
    int32_t root_ref;
  };
  std::vector<MeshRange> mesh_ranges;  // each mesh's binary nodes in blas (for a later collapse)
  std::vector<DevObject> dobjs_host;   // the device objects as uploaded
  double build_seconds = 0.0;  // wall time of the mesh BVH builds and the 4-wide collapse at commit
  std::vector<std::vector<float>> mesh_xyz;
  std::vector<std::vector<uint32_t>> mesh_idx;
  std::vector<HostObject> objects;
  std::vector<HostCsgElem> csg;
  std::vector<DevMaterial> materials;
  struct HL {
    int32_t kind, object;
    double pos[3], rgb[3];
  };
  std::vector<HL> lights;
  bool committed = false;
  DevNode *tlas = nullptr, *blas = nullptr;
  DevNode4* blas4 = nullptr;
  DevTri* tris = nullptr;
  DevObject* objs = nullptr;
  float* verts = nullptr;
  uint32_t* idx = nullptr;
  DevMaterial* mats = nullptr;
  DevLight* dlights = nullptr;
  DevCsgElem* dcsg = nullptr;
  RootDesc root{};
  bool has_tlas = false;
  uint32_t tlas_depth = 0, blas_depth = 0;
  pbrt_build_stats tlas_stats{};
  int* status = nullptr;
  uint2* spill = nullptr;
  uint64_t spill_entries = 0;
  // Every walk on the render stream or a caller's stream uses the one spill area: work on another
  // stream than the last one waits for the last one's end (event spill_done).
  hipEvent_t spill_done = nullptr;
  hipStream_t spill_stream = nullptr;
  PathBuffers pb{};
  std::mutex mu;
};

namespace {

void check_scene(pbrt_scene* s, bool need_commit) {
  if (!s) throw Error(PBRT_ERR_INVALID_ARG, "null scene");
  if (need_commit && !s->committed) throw Error(PBRT_ERR_INVALID_ARG, "scene not committed");
  if (!need_commit && s->committed) throw Error(PBRT_ERR_INVALID_ARG, "scene already committed");
  PBRT_HIP(hipSetDevice(s->ctx->device));
}

SceneView view_of(const pbrt_scene* s) {
  SceneView v{};
  v.tlas = s->tlas;
  v.root = s->root;
  v.objs = s->objs;
  v.blas = s->blas;
  v.blas4 = s->launch.width == 4 ? s->blas4 : nullptr;
  v.stack = s->launch.lds_stack;
  v.occ5 = s->launch.min_waves == 5 ? 1u : 0u;
  v.tris = s->tris;
  v.verts = s->verts;
  v.idx = s->idx;
  v.mats = s->mats;
  v.lights = s->dlights;
  v.csg = s->dcsg;
  v.n_objects = (uint32_t)s->objects.size();
  v.n_lights = (uint32_t)s->lights.size();
  v.has_tlas = s->has_tlas ? 1 : 0;
  for (const auto& l : s->lights) v.n_background += l.kind == kLightBackground ? 1u : 0u;
  return v;
}

// Samples per pixel in flight per pass of a plain render when the caller leaves it open: at most
// about 8M paths, and half the samples, so that a render of two passes or more runs them side by
// side (small scene at 256^2 x 16 spp: 2.17 ms in one pass, 2.04 in two; room at 512^2 x 64: 51.6
// ms in four passes of 16, 51.0 in two of 32).
uint32_t default_samples_per_pass(uint32_t npix, uint32_t spp) {
  const uint32_t budget = std::max(1u, (1u << 23) / std::max(1u, npix));
  return std::max(1u, std::min(budget, (spp + 1) / 2));
}

uint32_t persistent_grid(pbrt_scene* s, uint32_t n) {
  // the launch's workgroups (of 2 waves) per CU, or by default twice as many as fit on a CU at once
  // (8 at most fit): the second half queues and starts as workgroups of the first finish, so CUs
  // whose first workgroups end early take on more rays instead of idling while the slowest walks of
  // a grid that fits exactly drain (small scene 899 -> 988 Mrays/s, room 2,932 -> 3,021, jobs
  // 702 -> 714; the cache render unchanged)
  const uint32_t k = s->launch.lds_stack == 8 ? 0 : s->launch.lds_stack == 24 ? 2 : 1;
  if (!s->blocks_fit[k]) {
    const uint32_t b = scene_blocks_per_cu(s->launch.lds_stack);
    s->blocks_fit[k] = b ? std::min(b, 8u) : 8u;
  }
  const uint32_t per_cu = s->launch.blocks_per_cu ? s->launch.blocks_per_cu : 2u * s->blocks_fit[k];
  const uint32_t cap = (uint32_t)std::max(1, s->ctx->cu_count) * per_cu;
  const uint32_t need = (n + kSceneBlock - 1) / kSceneBlock;
  return std::max(1u, std::min(cap, need));
}

// Stream order of the work that shares the scene's spill area (the caller holds s->mu): before work
// on stream st, wait for the last work if it ran on another stream; after it, mark its end.
void spill_order(pbrt_scene* s, hipStream_t st) {
  if (!s->spill_done) PBRT_HIP(hipEventCreateWithFlags(&s->spill_done, hipEventDisableTiming));
  if (s->spill_stream && s->spill_stream != st) PBRT_HIP(hipStreamWaitEvent(st, s->spill_done, 0));
}
void spill_mark(pbrt_scene* s, hipStream_t st) {
  PBRT_HIP(hipEventRecord(s->spill_done, st));
  s->spill_stream = st;
}

// The integrator of a render (integrators.rs:23-28): the path integrator takes every option; the
// naive and normal integrators of the reference have none of them.
void check_integrator(const pbrt_render_config& c, bool with_cache) {
  if (c.integrator > PBRT_INTEGRATOR_NORMAL)
    throw Error(PBRT_ERR_INVALID_ARG, "integrator must be PBRT_INTEGRATOR_PATH, _NAIVE or _NORMAL");
  if (c.integrator != PBRT_INTEGRATOR_PATH && (c.mis || c.roulette || with_cache))
    throw Error(PBRT_ERR_INVALID_ARG, "mis, roulette and the radiance cache are options of the path integrator");
}

// The HBM spill area for a grid: a lane's stack holds at most (top-level + mesh depth) entries.
uint32_t ensure_spill(pbrt_scene* s, uint32_t grid) {
  // pending entries of the 4-wide walk: at most three per level
  const uint32_t depth = s->tlas_depth + (s->launch.width == 4 ? 3 : 1) * s->blas_depth + 2;
  const uint32_t stack = s->launch.lds_stack;
  const uint32_t per = depth > stack ? depth - stack : 1u;
  const uint64_t need = (uint64_t)per * grid * kSceneBlock;
  if (need > s->spill_entries) {
    if (s->spill) PBRT_HIP(hipFree(s->spill));
    s->spill = nullptr;
    s->spill_entries = 0;
    PBRT_HIP(hipMalloc(&s->spill, need * sizeof(uint2)));
    s->spill_entries = need;
  }
  return per;
}

// AABoundingBox::transform (src/geom/aabound.rs): the eight corners through the transform.
Box64 world_box(const Box64& b, const HostObject& o) {
  if (!o.has_xform) return b;
  const double* M = o.m;
  Box64 r = Box64::empty();
  const double xs[2] = {b.lo[0], b.hi[0]}, ys[2] = {b.lo[1], b.hi[1]}, zs[2] = {b.lo[2], b.hi[2]};
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) {
        const double x = xs[i], y = ys[j], z = zs[k];
        double px = M[0] * x + M[1] * y + M[2] * z + M[3];
        double py = M[4] * x + M[5] * y + M[6] * z + M[7];
        double pz = M[8] * x + M[9] * y + M[10] * z + M[11];
        const double w = M[12] * x + M[13] * y + M[14] * z + M[15];
        if (w != 1.0) {
          px = px * (1.0 / w);
          py = py * (1.0 / w);
          pz = pz * (1.0 / w);
        }
        r.lo[0] = std::fmin(r.lo[0], px); r.hi[0] = std::fmax(r.hi[0], px);
        r.lo[1] = std::fmin(r.lo[1], py); r.hi[1] = std::fmax(r.hi[1], py);
        r.lo[2] = std::fmin(r.lo[2], pz); r.hi[2] = std::fmax(r.hi[2], pz);
      }
  return r;
}

// CSG bounds in the object's frame (csg/union.rs:58-72: union of the elements' boxes,
// csg/intersection.rs:72-85: their intersection, csg/substraction.rs:94-104: the base's box).
Box64 csg_bounds(const pbrt_scene* s, const HostObject& o) {
  const int op = (int)o.param[0];
  const uint32_t cnt = (uint32_t)o.param[1];
  auto elem_box = [&](uint32_t k) {
    const HostCsgElem& e = s->csg[(size_t)o.mesh + k];
    HostObject t{};
    t.has_xform = 1;
    std::memcpy(t.m, e.m, sizeof(t.m));
    const Box64 lb = e.shape == kShapeSphere
                         ? Box64{{-e.param[0], -e.param[0], -e.param[0]}, {e.param[0], e.param[0], e.param[0]}}
                         : Box64{{0.0 - e.param[0] * 0.5, 0.0 - e.param[1] * 0.5, 0.0 - e.param[2] * 0.5},
                                 {e.param[0] * 0.5, e.param[1] * 0.5, e.param[2] * 0.5}};
    return world_box(lb, t);
  };
  Box64 b = elem_box(0);
  if (op == kCsgSubtraction) return b;
  for (uint32_t k = 1; k < cnt; ++k) {
    const Box64 e = elem_box(k);
    for (int a = 0; a < 3; ++a) {
      if (op == kCsgUnion) {
        b.lo[a] = std::fmin(b.lo[a], e.lo[a]);
        b.hi[a] = std::fmax(b.hi[a], e.hi[a]);
      } else {
        b.lo[a] = std::fmax(b.lo[a], e.lo[a]);
        b.hi[a] = std::fmin(b.hi[a], e.hi[a]);
      }
    }
  }
  return b;
}

// BVH::subdivide (src/bvh.rs:291-335): median split on the widest centroid axis, stable sort,
// one object per leaf; returns the node's box. Nodes are emitted in preorder.
struct TlasBuilder {
  const std::vector<Box64>& boxes;
  std::vector<uint32_t>& order;
  std::vector<BuildNode>& nodes;
  uint32_t max_depth = 0;
  Box64 build(uint32_t first, uint32_t count, uint32_t depth, uint32_t* me_out) {
    const uint32_t me = (uint32_t)nodes.size();
    nodes.push_back(BuildNode{});
    *me_out = me;
    max_depth = std::max(max_depth, depth);
    if (count == 1) {
      const Box64 b = boxes[order[first]];
      BuildNode& n = nodes[me];
      for (int a = 0; a < 3; ++a) {
        n.lo[a] = b.lo[a];
        n.hi[a] = b.hi[a];
      }
      n.leaf = true;
      n.first = first;
      n.count = 1;
      return b;
    }
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (uint32_t i = first; i < first + count; ++i)
      for (int a = 0; a < 3; ++a) {
        const double c = boxes[order[i]].centroid(a);
        mn[a] = std::fmin(mn[a], c);
        mx[a] = std::fmax(mx[a], c);
      }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
      if (mx[a] - mn[a] > mx[axis] - mn[axis]) axis = a;
    std::stable_sort(order.begin() + first, order.begin() + first + count,
                     [&](uint32_t p, uint32_t q) { return boxes[p].centroid(axis) < boxes[q].centroid(axis); });
    const uint32_t lc = count / 2;
    uint32_t l, r;
    const Box64 lb = build(first, lc, depth + 1, &l);
    const Box64 rb = build(first + lc, count - lc, depth + 1, &r);
    Box64 b = Box64::empty();
    b.grow(lb);
    b.grow(rb);
    BuildNode& n = nodes[me];
    for (int a = 0; a < 3; ++a) {
      n.lo[a] = b.lo[a];
      n.hi[a] = b.hi[a];
    }
    n.leaf = false;
    n.left = l;
    n.right = r;
    return b;
  }
};

void flatten_tlas(const std::vector<BuildNode>& b, std::vector<DevNode>* out, RootDesc* root) {
  std::vector<int32_t> dev_index(b.size(), -1);
  int32_t n_int = 0;
  for (size_t i = 0; i < b.size(); ++i)
    if (!b[i].leaf) dev_index[i] = n_int++;
  auto ref_of = [&](uint32_t i) { return b[i].leaf ? leaf_ref(b[i].first, 1) : dev_index[i]; };
  out->assign(n_int, DevNode{});
  for (size_t i = 0; i < b.size(); ++i) {
    if (b[i].leaf) continue;
    DevNode& d = (*out)[dev_index[i]];
    const BuildNode &L = b[b[i].left], &R = b[b[i].right];
    d.xy0[0] = (float)L.lo[0]; d.xy0[1] = (float)L.hi[0]; d.xy0[2] = (float)L.lo[1]; d.xy0[3] = (float)L.hi[1];
    d.xy1[0] = (float)R.lo[0]; d.xy1[1] = (float)R.hi[0]; d.xy1[2] = (float)R.lo[1]; d.xy1[3] = (float)R.hi[1];
    d.z01[0] = (float)L.lo[2]; d.z01[1] = (float)L.hi[2]; d.z01[2] = (float)R.lo[2]; d.z01[3] = (float)R.hi[2];
    d.ref[0] = ref_of(b[i].left);
    d.ref[1] = ref_of(b[i].right);
  }
  for (int a = 0; a < 3; ++a) {
    root->lo[a] = (float)b[0].lo[a];
    root->hi[a] = (float)b[0].hi[a];
  }
  root->ref = ref_of(0);
}

void free_paths(PathBuffers& p) {
  for (void* q : {(void*)p.ray_o, (void*)p.ray_d, (void*)p.beta, (void*)p.L, (void*)p.info, (void*)p.queue[0],
                  (void*)p.queue[1], (void*)p.counts, (void*)p.hit, (void*)p.sh_o, (void*)p.sh_d, (void*)p.sh_c})
    if (q) (void)hipFree(q);
  p = PathBuffers{};
}

void ensure_paths(pbrt_scene* s, uint32_t capacity) {
  PathBuffers& p = s->pb;
  if (p.capacity >= capacity) return;
  free_paths(p);
  const size_t n = capacity;
  PBRT_HIP(hipMalloc(&p.ray_o, n * sizeof(float4)));
  PBRT_HIP(hipMalloc(&p.ray_d, n * sizeof(float4)));
  PBRT_HIP(hipMalloc(&p.beta, n * sizeof(float4)));
  PBRT_HIP(hipMalloc(&p.L, n * sizeof(float4)));
  PBRT_HIP(hipMalloc(&p.info, n * sizeof(uint4)));
  PBRT_HIP(hipMalloc(&p.queue[0], n * sizeof(uint32_t)));
  PBRT_HIP(hipMalloc(&p.queue[1], n * sizeof(uint32_t)));
  PBRT_HIP(hipMalloc(&p.counts, kCounterWords * sizeof(uint32_t)));
  PBRT_HIP(hipMalloc(&p.hit, n * sizeof(SceneHitRec)));
  PBRT_HIP(hipMalloc(&p.sh_o, n * sizeof(float4)));
  PBRT_HIP(hipMalloc(&p.sh_d, n * sizeof(float4)));
  PBRT_HIP(hipMalloc(&p.sh_c, n * sizeof(float4)));
  p.capacity = capacity;
}

}  // namespace

extern "C" {

int pbrt_scene_create(pbrt_ctx* ctx, const pbrt_bvh_params* mesh_params, pbrt_scene** out) {
  return guarded([&] {
    if (!ctx || !out) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    auto s = std::make_unique<pbrt_scene>();
    s->ctx = ctx;
    if (mesh_params) s->mesh_params = *mesh_params;
    if (s->mesh_params.max_leaf > kMaxLeafTris) throw Error(PBRT_ERR_INVALID_ARG, "max_leaf above 32");
    *out = s.release();
  });
}

int pbrt_scene_destroy(pbrt_scene* s) {
  return guarded([&] {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    for (void* p : {(void*)s->tlas, (void*)s->blas, (void*)s->blas4, (void*)s->tris, (void*)s->objs, (void*)s->verts, (void*)s->idx,
                    (void*)s->mats, (void*)s->dlights, (void*)s->dcsg, (void*)s->status, (void*)s->spill})
      if (p) (void)hipFree(p);
    free_paths(s->pb);
    if (s->spill_done) (void)hipEventDestroy(s->spill_done);
    delete s;
  });
}

int pbrt_scene_add_mesh(pbrt_scene* s, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt, uint32_t* id) {
  return guarded([&] {
    check_scene(s, false);
    if (nt == 0) throw Error(PBRT_ERR_EMPTY_MESH, "a TriangleMesh needs at least one triangle");
    if (!xyz || !idx) throw Error(PBRT_ERR_INVALID_ARG, "null vertex or index array");
    for (uint64_t i = 0; i < 3ull * nt; ++i)
      if (idx[i] >= nv) throw Error(PBRT_ERR_INVALID_ARG, "vertex index out of range");
    s->mesh_xyz.emplace_back(xyz, xyz + 3 * (size_t)nv);
    s->mesh_idx.emplace_back(idx, idx + 3 * (size_t)nt);
    if (id) *id = (uint32_t)s->mesh_xyz.size() - 1;
  });
}

int pbrt_scene_add_material(pbrt_scene* s, int kind, const double rgb[3], double ior, double roughness, uint32_t* id) {
  return guarded([&] {
    check_scene(s, false);
    if (kind < kMatLambertian || kind > kMatMetal) throw Error(PBRT_ERR_INVALID_ARG, "unknown material kind");
    DevMaterial m{};
    m.kind = kind;
    for (int c = 0; c < 3; ++c) m.rgb[c] = (float)rgb[c];
    m.ior = (float)ior;
    m.roughness = (float)roughness;
    s->materials.push_back(m);
    if (id) *id = (uint32_t)s->materials.size() - 1;
  });
}

int pbrt_scene_set_texture(pbrt_scene* s, uint32_t material, int kind, const double rgb2[3], double scale) {
  return guarded([&] {
    check_scene(s, false);
    if (material >= s->materials.size()) throw Error(PBRT_ERR_INVALID_ARG, "no such material");
    if (kind != kTexColor && kind != kTexChecker) throw Error(PBRT_ERR_INVALID_ARG, "unknown texture kind");
    if (kind == kTexChecker && !rgb2) throw Error(PBRT_ERR_INVALID_ARG, "a checkerboard needs its second colour");
    DevMaterial& m = s->materials[material];
    m.tex = kind;
    m.scale = (float)scale;
    for (int c = 0; c < 3; ++c) m.rgb2[c] = kind == kTexChecker ? (float)rgb2[c] : 0.0f;
  });
}

int pbrt_scene_add_object(pbrt_scene* s, int shape, int mesh, const double param[3], uint32_t material,
                          const double* m16, const double* minv16, uint32_t* id) {
  return guarded([&] {
    check_scene(s, false);
    if (shape < kShapeMesh || shape > kShapeRect) throw Error(PBRT_ERR_INVALID_ARG, "unknown shape kind");
    if (shape == kShapeMesh && (mesh < 0 || (size_t)mesh >= s->mesh_xyz.size()))
      throw Error(PBRT_ERR_INVALID_ARG, "no such mesh");
    if (material >= s->materials.size()) throw Error(PBRT_ERR_INVALID_ARG, "no such material");
    if ((m16 == nullptr) != (minv16 == nullptr)) throw Error(PBRT_ERR_INVALID_ARG, "a transform needs its inverse");
    HostObject o{};
    o.shape = shape;
    o.mesh = mesh;
    o.material = (int32_t)material;
    for (int k = 0; k < 3; ++k) o.param[k] = param ? param[k] : 0.0;
    if (m16) {
      for (int k = 12; k < 16; ++k)
        if (m16[k] != (k == 15 ? 1.0 : 0.0) || minv16[k] != (k == 15 ? 1.0 : 0.0))
          throw Error(PBRT_ERR_INVALID_ARG, "transforms must be affine (last row 0 0 0 1)");
      o.has_xform = 1;
      std::memcpy(o.m, m16, sizeof(o.m));
      std::memcpy(o.minv, minv16, sizeof(o.minv));
    }
    s->objects.push_back(o);
    if (id) *id = (uint32_t)s->objects.size() - 1;
  });
}

int pbrt_scene_add_csg(pbrt_scene* s, int op, uint32_t n, const int* shapes, const double* params, const double* m,
                       const double* minv, uint32_t material, const double* m16, const double* minv16, uint32_t* id) {
  return guarded([&] {
    check_scene(s, false);
    if (op < kCsgUnion || op > kCsgSubtraction) throw Error(PBRT_ERR_INVALID_ARG, "unknown CSG operation");
    if (n == 0 || n > 64 || !shapes || !params || !m || !minv)
      throw Error(PBRT_ERR_INVALID_ARG, "a CSG object needs 1 to 64 elements with transforms");
    if (material >= s->materials.size()) throw Error(PBRT_ERR_INVALID_ARG, "no such material");
    if ((m16 == nullptr) != (minv16 == nullptr)) throw Error(PBRT_ERR_INVALID_ARG, "a transform needs its inverse");
    auto affine = [](const double* a) {
      for (int k = 12; k < 16; ++k)
        if (a[k] != (k == 15 ? 1.0 : 0.0)) return false;
      return true;
    };
    HostObject o{};
    o.shape = kShapeCsg;
    o.mesh = (int32_t)s->csg.size();
    o.material = (int32_t)material;
    o.param[0] = op;
    o.param[1] = n;
    if (m16) {
      if (!affine(m16) || !affine(minv16)) throw Error(PBRT_ERR_INVALID_ARG, "transforms must be affine");
      o.has_xform = 1;
      std::memcpy(o.m, m16, sizeof(o.m));
      std::memcpy(o.minv, minv16, sizeof(o.minv));
    }
    std::vector<HostCsgElem> els(n);
    for (uint32_t k = 0; k < n; ++k) {
      if (shapes[k] != kShapeSphere && shapes[k] != kShapeBox)
        throw Error(PBRT_ERR_INVALID_ARG, "CSG elements are spheres and boxes");
      if (!affine(m + 16 * k) || !affine(minv + 16 * k)) throw Error(PBRT_ERR_INVALID_ARG, "transforms must be affine");
      els[k].shape = shapes[k];
      for (int a = 0; a < 3; ++a) els[k].param[a] = params[3 * k + a];
      std::memcpy(els[k].m, m + 16 * k, sizeof(els[k].m));
      std::memcpy(els[k].minv, minv + 16 * k, sizeof(els[k].minv));
    }
    s->csg.insert(s->csg.end(), els.begin(), els.end());
    s->objects.push_back(o);
    if (id) *id = (uint32_t)s->objects.size() - 1;
  });
}

int pbrt_scene_add_point_light(pbrt_scene* s, const double pos[3], const double rgb[3]) {
  return guarded([&] {
    check_scene(s, false);
    pbrt_scene::HL l{kLightPoint, -1, {pos[0], pos[1], pos[2]}, {rgb[0], rgb[1], rgb[2]}};
    s->lights.push_back(l);
  });
}

int pbrt_scene_add_background_light(pbrt_scene* s, const double down[3], const double up[3]) {
  return guarded([&] {
    check_scene(s, false);
    if (!down || !up) throw Error(PBRT_ERR_INVALID_ARG, "null colour");
    pbrt_scene::HL l{kLightBackground, -1, {down[0], down[1], down[2]}, {up[0], up[1], up[2]}};
    s->lights.push_back(l);
  });
}

int pbrt_scene_add_area_light(pbrt_scene* s, uint32_t object, const double rgb[3]) {
  return guarded([&] {
    check_scene(s, false);
    if (object >= s->objects.size() || s->objects[object].shape != kShapeRect)
      throw Error(PBRT_ERR_INVALID_ARG, "an area light needs a rectangle object");
    pbrt_scene::HL l{kLightArea, (int32_t)object, {0, 0, 0}, {rgb[0], rgb[1], rgb[2]}};
    s->lights.push_back(l);
  });
}

// The 4-wide trees of every mesh, collapsed on the device from the binary nodes in s->blas
// (gpu_collapse4: the host collapse's rule, laid out level by level over all meshes, references
// global, so every mesh's node base is 0); root4 receives each mesh's root reference.
static void collapse_scene_meshes(pbrt_scene* s, std::vector<int32_t>* base4, std::vector<int32_t>* root4) {
  const size_t nm = s->mesh_ranges.size();
  std::vector<uint64_t> node_base(nm);
  std::vector<int32_t> root_ref(nm);
  uint64_t n_binary = 0;
  for (size_t k = 0; k < nm; ++k) {
    const auto& r = s->mesh_ranges[k];
    node_base[k] = r.node_base;
    root_ref[k] = r.root_ref;
    n_binary = std::max<uint64_t>(n_binary, (uint64_t)r.node_base + r.n_nodes);
  }
  DevNode4* d = nullptr;
  uint64_t n4 = 0;
  gpu_collapse4(s->ctx->stream, s->blas, node_base, root_ref, n_binary, &d, &n4, root4);
  // the kernels address 4-wide nodes by 32-bit byte offsets
  if (n4 * sizeof(DevNode4) > 0xffffffffull) {
    (void)hipFree(d);
    throw Error(PBRT_ERR_LAYOUT, "4-wide trees above 4 GiB (use the binary trees, mesh width 2)");
  }
  base4->assign(nm, 0);
  s->blas4 = d;
}

// After commit: the 4-wide trees on first use, and the device objects' references to them.
static void ensure_scene_blas4(pbrt_scene* s) {
  if (s->blas4 || s->mesh_ranges.empty()) return;
  std::vector<int32_t> base4, root4;
  collapse_scene_meshes(s, &base4, &root4);
  for (DevObject& d : s->dobjs_host)
    if (d.shape == kShapeMesh) {
      d.node_base4 = base4[d.mesh];
      d.root_ref4 = root4[d.mesh];
    }
  if (!s->dobjs_host.empty())
    PBRT_HIP(hipMemcpy(s->objs, s->dobjs_host.data(), s->dobjs_host.size() * sizeof(DevObject),
                       hipMemcpyHostToDevice));
}

int pbrt_scene_set_launch(pbrt_scene* s, const pbrt_scene_launch* l) {
  return guarded([&] {
    if (!s || !l) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    PBRT_HIP(hipSetDevice(s->ctx->device));
    pbrt_scene_launch v = *l;
    if (!v.lds_stack) v.lds_stack = 16;
    if (!v.width) v.width = 4;
    if (v.lds_stack != 8 && v.lds_stack != 16 && v.lds_stack != 24)
      throw Error(PBRT_ERR_INVALID_ARG, "lds_stack must be 8, 16 or 24");
    if (v.blocks_per_cu > 32) throw Error(PBRT_ERR_INVALID_ARG, "blocks_per_cu above 32");
    if (v.width != 2 && v.width != 4) throw Error(PBRT_ERR_INVALID_ARG, "width must be 2 or 4");
    if (v.min_waves != 0 && v.min_waves != 5) throw Error(PBRT_ERR_INVALID_ARG, "min_waves must be 0 or 5");
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->committed && v.width == 4) ensure_scene_blas4(s);
    s->launch = v;
  });
}

int pbrt_scene_get_launch(pbrt_scene* s, pbrt_scene_launch* out) {
  return guarded([&] {
    if (!s || !out) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    *out = s->launch;
  });
}

int pbrt_scene_set_mesh_build(pbrt_scene* s, int where) {
  return guarded([&] {
    check_scene(s, false);
    if (where != PBRT_BUILD_HOST && where != PBRT_BUILD_DEVICE)
      throw Error(PBRT_ERR_INVALID_ARG, "mesh build must be PBRT_BUILD_HOST or PBRT_BUILD_DEVICE");
    s->device_build = where == PBRT_BUILD_DEVICE;
  });
}

int pbrt_scene_set_mesh_width(pbrt_scene* s, uint32_t width) {
  return guarded([&] {
    check_scene(s, false);
    if (width != 2 && width != 4) throw Error(PBRT_ERR_INVALID_ARG, "mesh tree width must be 2 or 4");
    s->launch.width = width;
  });
}

int pbrt_scene_build_seconds(pbrt_scene* s, double* seconds) {
  return guarded([&] {
    check_scene(s, true);
    if (!seconds) throw Error(PBRT_ERR_INVALID_ARG, "null output");
    *seconds = s->build_seconds;
  });
}

// Scene::commit (scene.rs:101-115): per-mesh BVHs, the object BVH, upload.
int pbrt_scene_commit(pbrt_scene* s) {
  return guarded([&] {
    check_scene(s, false);
    // mesh BVHs, flattened into shared node / triangle arrays
    std::vector<DevNode> blas;
    std::vector<DevTri> tris;
    std::vector<float> verts;
    std::vector<uint32_t> idx;
    struct MeshInfo {
      int32_t node_base, tri_base, root_ref;
      int32_t node_base4 = 0, root_ref4 = 0;
      uint64_t n_nodes = 0;
      float lo[3], hi[3];
      uint32_t vert_base, idx_base;
      Box64 bounds;
    };
    std::vector<MeshInfo> info;
    // vertices and indices of every mesh, concatenated; TriangleMesh::get_bounding_box
    // (triangle_mesh.rs:150-174): every vertex
    const size_t nm = s->mesh_xyz.size();
    info.resize(nm);
    for (size_t k = 0; k < nm; ++k) {
      const auto& x = s->mesh_xyz[k];
      const auto& ix = s->mesh_idx[k];
      const uint32_t nv = (uint32_t)(x.size() / 3);
      MeshInfo& mi = info[k];
      mi.vert_base = (uint32_t)(verts.size() / 3);
      mi.idx_base = (uint32_t)(idx.size() / 3);
      mi.bounds = Box64::empty();
      for (uint32_t v = 0; v < nv; ++v)
        for (int a = 0; a < 3; ++a) {
          mi.bounds.lo[a] = std::fmin((double)x[3 * v + a], mi.bounds.lo[a]);
          mi.bounds.hi[a] = std::fmax((double)x[3 * v + a], mi.bounds.hi[a]);
        }
      verts.insert(verts.end(), x.begin(), x.end());
      idx.insert(idx.end(), ix.begin(), ix.end());
    }
    const auto build_t0 = std::chrono::steady_clock::now();
    double upload_s = 0.0;
    if (s->device_build && nm > 0) {
      // one batched device build (bvh_gpu.hip): every mesh's nodes and triangles in place
      dev_upload(&s->verts, verts);
      dev_upload(&s->idx, idx);
      std::vector<GpuMeshSpan> spans(nm);
      for (size_t k = 0; k < nm; ++k)
        spans[k] = GpuMeshSpan{info[k].vert_base, (uint32_t)(s->mesh_xyz[k].size() / 3),
                               (uint32_t)(s->mesh_idx[k].size() / 3)};
      GpuBuiltMany g;
      gpu_build_meshes(s->ctx->stream, s->verts, s->idx, spans, s->mesh_params, &g);
      s->blas = g.nodes;
      s->tris = g.tris;
      s->blas_depth = (uint32_t)g.stats.max_depth;
      for (size_t k = 0; k < nm; ++k) {
        MeshInfo& mi = info[k];
        mi.node_base = (int32_t)g.meshes[k].node_base;
        mi.n_nodes = g.meshes[k].n_nodes;
        mi.tri_base = (int32_t)g.meshes[k].tri_base;
        mi.root_ref = g.meshes[k].root.ref;
        for (int a = 0; a < 3; ++a) {
          mi.lo[a] = g.meshes[k].root.lo[a];
          mi.hi[a] = g.meshes[k].root.hi[a];
        }
      }
    } else {
      // the meshes' trees are independent: build them on the host's cores, then flatten in order
      std::vector<BuiltBVH> built(nm);
      {
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::atomic<size_t> next{0};
        std::vector<std::string> errors(hw);
        auto work = [&](unsigned w) {
          try {
            for (size_t k = next++; k < nm; k = next++) {
              const auto& x = s->mesh_xyz[k];
              const auto& ix = s->mesh_idx[k];
              built[k] = build_mesh_bvh(x.data(), (uint32_t)(x.size() / 3), ix.data(), (uint32_t)(ix.size() / 3),
                                        s->mesh_params);
            }
          } catch (const std::exception& e) {
            errors[w] = e.what();
          }
        };
        std::vector<std::thread> pool;
        for (unsigned w = 1; w < std::min<size_t>(hw, nm); ++w) pool.emplace_back(work, w);
        work(0);
        for (auto& t : pool) t.join();
        for (const auto& e : errors)
          if (!e.empty()) throw Error(PBRT_ERR_INTERNAL, "mesh BVH build: " + e);
      }
      for (size_t k = 0; k < nm; ++k) {
        const auto& x = s->mesh_xyz[k];
        const auto& ix = s->mesh_idx[k];
        BuiltBVH b = std::move(built[k]);
        s->blas_depth = std::max<uint32_t>(s->blas_depth, (uint32_t)b.stats.max_depth);
        std::vector<DevNode> n;
        std::vector<DevTri> t;
        RootDesc r;
        flatten_mesh_bvh(b, x.data(), ix.data(), &n, &t, &r);
        MeshInfo& mi = info[k];
        mi.node_base = (int32_t)blas.size();
        mi.n_nodes = n.size();
        mi.tri_base = (int32_t)tris.size();
        mi.root_ref = r.ref;
        for (int a = 0; a < 3; ++a) {
          mi.lo[a] = r.lo[a];
          mi.hi[a] = r.hi[a];
        }
        blas.insert(blas.end(), n.begin(), n.end());
        tris.insert(tris.end(), t.begin(), t.end());
      }
      // the arrays go to the device now (the 4-wide collapse runs there); the upload is not counted
      // as build time
      const auto up_t0 = std::chrono::steady_clock::now();
      dev_upload(&s->blas, blas);
      dev_upload(&s->tris, tris);
      dev_upload(&s->verts, verts);
      dev_upload(&s->idx, idx);
      upload_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - up_t0).count();
    }
    s->mesh_ranges.resize(nm);
    for (size_t k = 0; k < nm; ++k) s->mesh_ranges[k] = {(uint32_t)info[k].node_base, info[k].n_nodes, info[k].root_ref};
    if (s->launch.width == 4 && nm > 0) {
      std::vector<int32_t> base4, root4;
      collapse_scene_meshes(s, &base4, &root4);
      for (size_t k = 0; k < nm; ++k) {
        info[k].node_base4 = base4[k];
        info[k].root_ref4 = root4[k];
      }
    }
    s->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - build_t0).count() - upload_s;
    // object boxes (local bounds through the transform) and the median-split object BVH
    const uint32_t no = (uint32_t)s->objects.size();
    std::vector<Box64> boxes(no);
    for (uint32_t k = 0; k < no; ++k) {
      const HostObject& o = s->objects[k];
      Box64 b;
      switch (o.shape) {
        case kShapeSphere: b = {{-o.param[0], -o.param[0], -o.param[0]}, {o.param[0], o.param[0], o.param[0]}}; break;
        case kShapeBox:
          b = {{0.0 - o.param[0] * 0.5, 0.0 - o.param[1] * 0.5, 0.0 - o.param[2] * 0.5},
               {o.param[0] * 0.5, o.param[1] * 0.5, o.param[2] * 0.5}};
          break;
        case kShapeRect:
          b = {{-o.param[0] / 2.0, 0.0, -o.param[1] / 2.0}, {o.param[0] / 2.0, 0.0, o.param[1] / 2.0}};
          break;
        case kShapeCsg: b = csg_bounds(s, o); break;
        default: b = info[o.mesh].bounds; break;
      }
      boxes[k] = world_box(b, o);
    }
    std::vector<uint32_t> order(no);
    for (uint32_t k = 0; k < no; ++k) order[k] = k;
    std::vector<DevNode> tlas;
    if (no > 0) {
      std::vector<BuildNode> nodes;
      TlasBuilder tb{boxes, order, nodes};
      uint32_t root;
      tb.build(0, no, 1, &root);
      s->tlas_depth = tb.max_depth;
      flatten_tlas(nodes, &tlas, &s->root);
      s->has_tlas = true;
      s->tlas_stats.node_count = nodes.size();
    }
    std::vector<uint32_t> pos_of(no);
    std::vector<DevObject> dobjs(no);
    for (uint32_t p = 0; p < no; ++p) {
      const uint32_t k = order[p];
      pos_of[k] = p;
      const HostObject& o = s->objects[k];
      DevObject& d = dobjs[p];
      d.shape = o.shape;
      d.material = o.material;
      d.has_xform = o.has_xform;
      d.orig = (int32_t)k;
      for (int a = 0; a < 3; ++a) d.param[a] = (float)o.param[a];
      for (int a = 0; a < 12; ++a) {
        d.m[a] = (float)o.m[a];
        d.minv[a] = (float)o.minv[a];
      }
      d.mesh = o.mesh;
      if (o.shape == kShapeMesh) {
        const MeshInfo& mi = info[o.mesh];
        d.node_base = mi.node_base;
        d.tri_base = mi.tri_base;
        d.root_ref = mi.root_ref;
        for (int a = 0; a < 3; ++a) {
          d.root_lo[a] = mi.lo[a];
          d.root_hi[a] = mi.hi[a];
        }
        d.vert_base = mi.vert_base;
        d.idx_base = mi.idx_base;
        d.node_base4 = mi.node_base4;
        d.root_ref4 = mi.root_ref4;
      }
    }
    std::vector<DevLight> dl;
    for (const auto& l : s->lights) {
      DevLight d{};
      d.kind = l.kind;
      d.object = l.kind == kLightArea ? (int32_t)pos_of[l.object] : -1;
      for (int a = 0; a < 3; ++a) {
        d.pos[a] = (float)l.pos[a];
        d.rgb[a] = (float)l.rgb[a];
      }
      dl.push_back(d);
    }
    std::vector<DevCsgElem> dc(std::max<size_t>(1, s->csg.size()));
    for (size_t k = 0; k < s->csg.size(); ++k) {
      const HostCsgElem& e = s->csg[k];
      DevCsgElem& d = dc[k];
      d = DevCsgElem{};
      d.shape = e.shape;
      for (int a = 0; a < 3; ++a) d.param[a] = (float)e.param[a];
      for (int a = 0; a < 12; ++a) {
        d.m[a] = (float)e.m[a];
        d.minv[a] = (float)e.minv[a];
      }
    }
    dev_upload(&s->dcsg, dc);
    dev_upload(&s->tlas, tlas);
    dev_upload(&s->objs, dobjs);
    s->dobjs_host = dobjs;
    dev_upload(&s->mats, s->materials);
    dev_upload(&s->dlights, dl);
    PBRT_HIP(hipMalloc(&s->status, sizeof(int)));
    PBRT_HIP(hipMemset(s->status, 0, sizeof(int)));
    s->committed = true;
  });
}

int pbrt_scene_info(pbrt_scene* s, uint32_t* objects, uint32_t* tlas_nodes, uint32_t* tlas_depth, uint32_t* blas_depth) {
  return guarded([&] {
    check_scene(s, true);
    if (objects) *objects = (uint32_t)s->objects.size();
    if (tlas_nodes) *tlas_nodes = (uint32_t)s->tlas_stats.node_count;
    if (tlas_depth) *tlas_depth = s->tlas_depth;
    if (blas_depth) *blas_depth = s->blas_depth;
  });
}

static int scene_trace(pbrt_scene* s, const pbrt_ray* rays, uint32_t n, pbrt_scene_hit* hits, uint8_t* occ,
                       void* stream, int any) {
  return guarded([&] {
    check_scene(s, true);
    if (n == 0) return;
    if (!rays || (any ? !occ : !hits)) throw Error(PBRT_ERR_INVALID_ARG, "null buffer");
    std::lock_guard<std::mutex> lock(s->mu);
    SceneLaunch a{};
    a.view = view_of(s);
    a.rays = rays;
    a.hits = reinterpret_cast<SceneHitRec*>(hits);
    a.occluded = occ;
    a.n = n;
    a.grid = persistent_grid(s, n);
    a.spill_per_thread = ensure_spill(s, a.grid);
    a.spill = s->spill;
    a.status = s->status;
    a.stream = stream ? stream : s->ctx->stream;
    spill_order(s, (hipStream_t)a.stream);
    PBRT_HIP(launch_scene_trace(a, any));
    spill_mark(s, (hipStream_t)a.stream);
  });
}

int pbrt_scene_trace_closest(pbrt_scene* s, const pbrt_ray* rays, uint32_t n, pbrt_scene_hit* hits, void* stream) {
  return scene_trace(s, rays, n, hits, nullptr, stream, 0);
}

int pbrt_scene_trace_any(pbrt_scene* s, const pbrt_ray* rays, uint32_t n, uint8_t* occ, void* stream) {
  return scene_trace(s, rays, n, nullptr, occ, stream, 1);
}

int pbrt_scene_autotune(pbrt_scene* s, const pbrt_ray* rays, uint32_t n, pbrt_scene_hit* hits, uint32_t reps,
                        pbrt_scene_launch* chosen, float* best_ms) {
  return guarded([&] {
    check_scene(s, true);
    if (!rays || !hits || !n) throw Error(PBRT_ERR_INVALID_ARG, "null argument or empty batch");
    if (reps == 0) reps = 2;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      ensure_scene_blas4(s);
    }
    hipEvent_t e0, e1;
    PBRT_HIP(hipEventCreate(&e0));
    PBRT_HIP(hipEventCreate(&e1));
    const pbrt_scene_launch prior = s->launch;
    pbrt_scene_launch best = prior;
    float best_t = 1e30f;
    try {
      hipStream_t st = s->ctx->stream;
      for (uint32_t stack : {8u, 16u, 24u})
        for (uint32_t blocks : {2u, 4u, 6u, 8u, 16u, 32u})
          for (uint32_t width : {2u, 4u}) {
            std::lock_guard<std::mutex> lock(s->mu);
            s->launch = pbrt_scene_launch{stack, blocks, width, prior.min_waves};
            SceneLaunch a{};
            a.view = view_of(s);
            a.rays = rays;
            a.hits = reinterpret_cast<SceneHitRec*>(hits);
            a.n = n;
            a.grid = persistent_grid(s, n);
            a.spill_per_thread = ensure_spill(s, a.grid);
            a.spill = s->spill;
            a.status = s->status;
            a.stream = st;
            spill_order(s, st);
            PBRT_HIP(launch_scene_trace(a, 0));  // warm-up
            float cand = 1e30f;
            for (uint32_t r = 0; r < reps; ++r) {
              PBRT_HIP(hipEventRecord(e0, st));
              PBRT_HIP(launch_scene_trace(a, 0));
              PBRT_HIP(hipEventRecord(e1, st));
              spill_mark(s, st);
              PBRT_HIP(hipEventSynchronize(e1));
              float ms = 0.f;
              PBRT_HIP(hipEventElapsedTime(&ms, e0, e1));
              cand = std::min(cand, ms);
            }
            if (cand < best_t) {
              best_t = cand;
              best = s->launch;
            }
          }
    } catch (...) {
      s->launch = prior;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    s->launch = best;
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = best_t;
  });
}

int pbrt_scene_status(pbrt_scene* s, int* out) {
  return guarded([&] {
    check_scene(s, true);
    PBRT_HIP(hipDeviceSynchronize());
    PBRT_HIP(hipMemcpy(out, s->status, sizeof(int), hipMemcpyDeviceToHost));
  });
}

// renderers::mt::render (src/renderers/mt.rs:23, 112-130) on the device: every pixel's `spp`
// samples traced by the wavefront of scene.hip; film_out receives the per-pixel mean radiance
// (f64 RGB, row-major), before the reference's gamma and 8-bit quantisation (spectrum.rs:13-30).
int pbrt_render(pbrt_scene* s, const pbrt_render_config* c, double* film_out, pbrt_render_stats* stats) {
  return pbrt_render_ex(s, c, nullptr, nullptr, nullptr, film_out, stats);
}

int pbrt_scene_autotune_render(pbrt_scene* s, const pbrt_render_config* c, uint32_t reps, pbrt_scene_launch* chosen,
                               float* best_ms) {
  return guarded([&] {
    check_scene(s, true);
    if (!c) throw Error(PBRT_ERR_INVALID_ARG, "null config");
    if (reps == 0) reps = 2;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      ensure_scene_blas4(s);
    }
    const pbrt_scene_launch prior = s->launch;
    pbrt_scene_launch best = prior;
    double best_s = 1e30;
    for (uint32_t stack : {8u, 16u, 24u})
      for (uint32_t blocks : {2u, 4u, 6u, 8u, 16u, 32u})
        for (uint32_t width : {2u, 4u})
          for (uint32_t waves : {0u, 5u}) {
            {
              std::lock_guard<std::mutex> lock(s->mu);
              s->launch = pbrt_scene_launch{stack, blocks, width, waves};
            }
            double cand = 1e30;
            for (uint32_t r = 0; r <= reps; ++r) {  // the first run warms up
              pbrt_render_stats st{};
              const int rc = pbrt_render(s, c, nullptr, &st);
              if (rc != PBRT_OK) {
                s->launch = prior;
                throw Error(rc, pbrt_last_error());
              }
              if (r) cand = std::min(cand, st.seconds);
            }
            if (cand < best_s) {
              best_s = cand;
              best = s->launch;
            }
          }
    s->launch = best;
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = (float)(best_s * 1e3);
  });
}

int pbrt_render_ex(pbrt_scene* s, const pbrt_render_config* c, pbrt_nrc* nrc, const pbrt_nrc_render* nc,
                   pbrt_progress* progress, double* film_out, pbrt_render_stats* stats) {
  return guarded([&] {
    check_scene(s, true);
    if (!c || !c->width || !c->height || !c->spp || !c->max_depth) throw Error(PBRT_ERR_INVALID_ARG, "bad render config");
    if ((uint64_t)c->width * c->height > (1u << 26)) throw Error(PBRT_ERR_INVALID_ARG, "image larger than 2^26 pixels");
    pbrt_nrc_render ncfg{3, 33, 65536, 1, nullptr, 0, 0};
    if (nrc) {
      if (nrc->ctx != s->ctx) throw Error(PBRT_ERR_INVALID_ARG, "cache and scene on different contexts");
      if (nc) ncfg = *nc;
      if (ncfg.query_vertex < 1 || ncfg.query_vertex > c->max_depth)
        throw Error(PBRT_ERR_INVALID_ARG, "query_vertex must be in [1, max_depth]");
      if (ncfg.train_batch == 0 || ncfg.train_batch > 65536) throw Error(PBRT_ERR_INVALID_ARG, "train_batch in [1, 65536]");
    }
    std::lock_guard<std::mutex> lock(s->mu);
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t npix = c->width * c->height;
    uint32_t per_pass = c->samples_per_pass ? c->samples_per_pass
                                            : (nrc ? 1u : default_samples_per_pass(npix, c->spp));
    per_pass = std::min(per_pass, c->spp);
    if ((uint64_t)npix * per_pass > (1u << 28)) throw Error(PBRT_ERR_INVALID_ARG, "pass larger than 2^28 paths");
    ensure_paths(s, npix * per_pass);
    hipStream_t st = s->ctx->stream;
    spill_order(s, st);
    const uint32_t npass_paths = npix * per_pass;
    const uint32_t passes_total = (c->spp + per_pass - 1) / per_pass;

    // every device allocation of the call, released on every exit path
    std::vector<void*> owned;
    struct Release {
      std::vector<void*>& v;
      ~Release() {
        for (void* q : v) (void)hipFree(q);
      }
    } release{owned};
    auto alloc = [&](size_t bytes) {
      void* q = nullptr;
      PBRT_HIP(hipMalloc(&q, bytes ? bytes : 16));
      owned.push_back(q);
      return q;
    };
    std::vector<hipEvent_t> events;
    struct ReleaseEvents {
      std::vector<hipEvent_t>& v;
      ~ReleaseEvents() {
        for (hipEvent_t e : v) (void)hipEventDestroy(e);
      }
    } release_events{events};

    double* film = (double*)alloc((size_t)npix * 3 * sizeof(double));
    PBRT_HIP(hipMemsetAsync(film, 0, (size_t)npix * 3 * sizeof(double), st));
    PBRT_HIP(hipMemsetAsync(s->pb.counts, 0, kCounterWords * sizeof(uint32_t), st));
    PathBuffers pb = s->pb;
    pb.film = film;
    const Pinhole cam(c->width, c->height, c->fov, c->near, c->far,
                      look_at({c->eye[0], c->eye[1], c->eye[2]}, {c->look[0], c->look[1], c->look[2]},
                              {c->up[0], c->up[1], c->up[2]}));
    RenderParams p{};
    p.width = c->width;
    p.height = c->height;
    p.spp = c->spp;
    p.max_depth = c->max_depth;
    p.seed = c->seed;
    for (int i = 0; i < 16; ++i) {
      p.r2s[i] = (float)cam.r2s.m[i / 4][i % 4];
      p.c2w[i] = (float)cam.c2w.m[i / 4][i % 4];
    }
    p.near = (float)c->near;
    p.far = (float)c->far;
    if (c->mis > 1) throw Error(PBRT_ERR_INVALID_ARG, "mis must be 0 or 1");
    p.mis = c->mis;
    if (c->roulette > 1) throw Error(PBRT_ERR_INVALID_ARG, "roulette must be 0 or 1");
    if (c->roulette && nrc) throw Error(PBRT_ERR_INVALID_ARG, "Russian roulette is not supported with the radiance cache");
    p.roulette = c->roulette;
    check_integrator(*c, nrc != nullptr);
    p.integrator = c->integrator;
    if (c->sampler > PBRT_SAMPLER_SOBOL) throw Error(PBRT_ERR_INVALID_ARG, "sampler must be PBRT_SAMPLER_INDEPENDENT or _SOBOL");
    p.sampler = c->sampler;

    // radiance cache buffers
    NrcTrainBuffers tb{};
    uint16_t* batch_x = nullptr;
    uint32_t ranks_cap = 0;
    if (nrc) {
      p.nrc = 1;
      p.query_vertex = ncfg.query_vertex;
      p.train_stride = ncfg.train_stride;
      pb.q_raw = (float4*)alloc((size_t)npass_paths * 3 * sizeof(float4));
      pb.q_beta = (float4*)alloc((size_t)npass_paths * sizeof(float4));
      if (ncfg.train_stride) {
        ranks_cap = npass_paths / ncfg.train_stride + 2;
        pb.train_cap = ranks_cap;
        pb.rec = (float4*)alloc((size_t)kRecFields * c->max_depth * ranks_cap * sizeof(float4));
        const size_t slots = (size_t)ranks_cap * ncfg.query_vertex;
        tb.slot_raw = (float4*)alloc(slots * 3 * sizeof(float4));
        tb.slot_t = (float4*)alloc(slots * sizeof(float4));
        tb.idx = (uint32_t*)alloc(slots * sizeof(uint32_t));
        tb.block_cap = (uint32_t)((slots + 255) / 256);
        tb.block_count = (uint32_t*)alloc((size_t)tb.block_cap * sizeof(uint32_t));
        const size_t rows = (size_t)ncfg.train_batch * per_pass;  // one batch per frame of a pass
        tb.batch_raw = (float4*)alloc(rows * 3 * sizeof(float4));
        tb.batch_t = (float*)alloc(rows * 3 * sizeof(float));
        batch_x = (uint16_t*)alloc(rows * kNrcInPad * 2);
      }
      // the cache's encoding box: the scene's bounds, widened by 1 % (+1e-3) on every side
      float lo[3] = {-1.f, -1.f, -1.f}, hi[3] = {1.f, 1.f, 1.f};
      if (s->has_tlas)
        for (int a = 0; a < 3; ++a) {
          const float e = s->root.hi[a] - s->root.lo[a];
          lo[a] = s->root.lo[a] - 0.01f * e - 1e-3f;
          hi[a] = s->root.hi[a] + 0.01f * e + 1e-3f;
        }
      std::memcpy(nrc->lo, lo, sizeof lo);
      std::memcpy(nrc->hi, hi, sizeof hi);
    }

    // convergence record
    double* ref = nullptr;
    double* rmse_dev = nullptr;
    double* rmse_partial = nullptr;
    constexpr uint32_t kRmsePartials = 1024;
    const uint32_t n_record = progress ? std::min(progress->capacity, passes_total) : 0;
    if (progress) {
      if (n_record && (!progress->ms || (progress->reference && !progress->rmse)))
        throw Error(PBRT_ERR_INVALID_ARG, "progress needs ms (and rmse with a reference)");
      events.resize(n_record + 1);
      for (auto& e : events) PBRT_HIP(hipEventCreate(&e));
      if (progress->reference) {
        ref = (double*)alloc((size_t)npix * 3 * sizeof(double));
        PBRT_HIP(hipMemcpyAsync(ref, progress->reference, (size_t)npix * 3 * sizeof(double), hipMemcpyHostToDevice,
                                st));
        rmse_dev = (double*)alloc((size_t)std::max(1u, n_record) * sizeof(double));
        rmse_partial = (double*)alloc(kRmsePartials * sizeof(double));
      }
      PBRT_HIP(hipStreamSynchronize(st));
    }

    const SceneView v = view_of(s);
    const uint32_t grid = persistent_grid(s, npass_paths);
    const uint32_t spill_per = ensure_spill(s, grid);
    uint64_t passes = 0;
    // without the cache, several passes: two path sets and a second stream, pass k on set k % 2, so
    // one pass's bounces run beside the other's and the last, slowest walks of each traversal launch
    // overlap the other pass's work instead of leaving CUs idle; the film accumulations (and the
    // convergence records) wait for each other's events and run in pass order, so the film is that
    // of running the passes one after the other.
    const bool two = !nrc && passes_total > 1;
    // with the cache: three path sets and three streams. Pass k (set k % 3) traces bounces 1..
    // query_vertex on one of two front streams (the render stream for even k, a second one for odd
    // k), so two passes' fronts run side by side and fill each other's traversal tails; the
    // remaining bounces (training paths, paths past a specular vertex), the cache queries, the film
    // accumulation and the cache training run on a third, back stream, under later fronts. The back
    // stream runs the passes' back halves in order, so pass k's queries see the training of every
    // earlier pass, and a set is reused only once its last pass's back half has finished: the
    // results are those of running every step in order on one stream. The fronts do not read the
    // cache.
    const int nsets = nrc ? (passes_total > 1 ? 3 : 2) : (two ? 2 : 1);
    PathBuffers sets[3] = {pb, pb, pb};
    hipStream_t aux = nullptr, side = nullptr;
    uint2 *spill_aux = nullptr, *spill_side = nullptr;
    hipEvent_t ev_front[3] = {nullptr, nullptr, nullptr}, ev_back[3] = {nullptr, nullptr, nullptr};
    struct ReleaseAux {
      hipStream_t& s;
      hipStream_t& s2;
      hipEvent_t* a;
      hipEvent_t* b;
      ~ReleaseAux() {
        for (int i = 0; i < 3; ++i) {
          if (a[i]) (void)hipEventDestroy(a[i]);
          if (b[i]) (void)hipEventDestroy(b[i]);
        }
        if (s) (void)hipStreamDestroy(s);
        if (s2) (void)hipStreamDestroy(s2);
      }
    } release_aux{aux, side, ev_front, ev_back};
    for (int k = 1; k < nsets; ++k) {
      PathBuffers& o = sets[k];
      const size_t n = npass_paths;
      o.ray_o = (float4*)alloc(n * sizeof(float4));
      o.ray_d = (float4*)alloc(n * sizeof(float4));
      o.beta = (float4*)alloc(n * sizeof(float4));
      o.L = (float4*)alloc(n * sizeof(float4));
      o.info = (uint4*)alloc(n * sizeof(uint4));
      o.queue[0] = (uint32_t*)alloc(n * sizeof(uint32_t));
      o.queue[1] = (uint32_t*)alloc(n * sizeof(uint32_t));
      o.counts = (uint32_t*)alloc(kCounterWords * sizeof(uint32_t));
      o.hit = (SceneHitRec*)alloc(n * sizeof(SceneHitRec));
      o.sh_o = (float4*)alloc(n * sizeof(float4));
      o.sh_d = (float4*)alloc(n * sizeof(float4));
      o.sh_c = (float4*)alloc(n * sizeof(float4));
      o.capacity = (uint32_t)n;
      if (nrc && ncfg.train_stride)
        o.rec = (float4*)alloc((size_t)kRecFields * c->max_depth * ranks_cap * sizeof(float4));
      // each set queues its own cache queries: the back half of one pass answers them while the next
      // pass's front (the other set) queues new ones
      if (nrc) {
        o.q_raw = (float4*)alloc(n * 3 * sizeof(float4));
        o.q_beta = (float4*)alloc(n * sizeof(float4));
      }
      PBRT_HIP(hipMemsetAsync(o.counts, 0, kCounterWords * sizeof(uint32_t), st));
    }
    if (nsets > 1) {
      spill_aux = (uint2*)alloc((size_t)spill_per * grid * kSceneBlock * sizeof(uint2));
      PBRT_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
      for (int i = 0; i < 3; ++i) {
        PBRT_HIP(hipEventCreateWithFlags(&ev_front[i], hipEventDisableTiming));
        PBRT_HIP(hipEventCreateWithFlags(&ev_back[i], hipEventDisableTiming));
      }
      if (nsets == 3) {
        spill_side = (uint2*)alloc((size_t)spill_per * grid * kSceneBlock * sizeof(uint2));
        PBRT_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
      }
      PBRT_HIP(hipStreamSynchronize(st));
    }
    if (!events.empty()) PBRT_HIP(hipEventRecord(events[0], st));
    for (uint32_t first = 0; first < c->spp; first += per_pass) {
      p.pass_first_sample = first;
      p.pass_samples = std::min(per_pass, c->spp - first);
      const uint32_t npass = npix * p.pass_samples;
      if (!nrc) {
        const int S = two ? (int)(passes & 1) : 0;
        const PathBuffers& B = sets[S];
        hipStream_t on = S ? aux : st;
        uint2* sp = S ? spill_aux : s->spill;
        PBRT_HIP(launch_generate(v, p, B, on));
        int q = 0;
        for (uint32_t b = 0; b < c->max_depth; ++b) {
          PBRT_HIP(launch_extend(v, B, q, sp, spill_per, grid, s->status, on));
          PBRT_HIP(launch_reset_queue(B, q, on));
          PBRT_HIP(launch_shade(v, p, B, q, on));
          PBRT_HIP(launch_shadow(v, B, sp, spill_per, grid, s->status, on));
          q ^= 1;
        }
        // the film takes the passes in order
        if (two && passes > 0) PBRT_HIP(hipStreamWaitEvent(on, ev_back[S ^ 1], 0));
        PBRT_HIP(launch_accumulate(p, B, on));
        if (passes < n_record) {
          if (ref)
            PBRT_HIP(launch_rmse(film, 1.0 / (double)(first + p.pass_samples), ref, npix * 3, rmse_partial,
                                 kRmsePartials, rmse_dev + passes, on));
          PBRT_HIP(hipEventRecord(events[passes + 1], on));
        }
        if (two) PBRT_HIP(hipEventRecord(ev_back[S], on));
        passes++;
        continue;
      }
      const int S = (int)(passes % (uint64_t)nsets);
      PathBuffers& B = sets[S];
      const bool odd = nsets == 3 && (passes & 1);
      hipStream_t F = odd ? side : st;
      uint2* fsp = odd ? spill_side : s->spill;
      if (p.train_stride) p.train_offset = (uint32_t)(hash4(c->seed, passes, 0x747261696eull, 0) % p.train_stride);
      // front: the set is free once the last pass that used it has finished its back half
      if (passes >= (uint64_t)nsets) PBRT_HIP(hipStreamWaitEvent(F, ev_back[S], 0));
      // front (a front stream): the camera rays and every bounce up to the query vertex, where all
      // but the training paths end; the last of those bounces' shadow rays (training paths and
      // paths past a specular vertex only) are left to the back
      PBRT_HIP(launch_generate_nrc(p, B, F));
      int q = 0;
      const uint32_t front = std::min(p.query_vertex, c->max_depth);
      for (uint32_t b = 0; b < front; ++b) {
        PBRT_HIP(launch_extend(v, B, q, fsp, spill_per, grid, s->status, F));
        PBRT_HIP(launch_reset_queue(B, q, F));
        PBRT_HIP(launch_shade_nrc(v, p, B, q, F));
        if (b + 1 < front) PBRT_HIP(launch_shadow_nrc(v, p, B, fsp, spill_per, grid, s->status, F));
        q ^= 1;
      }
      PBRT_HIP(hipEventRecord(ev_front[S], F));
      // back (second stream, overlapping the next pass's front): the last front shadow rays, the
      // remaining bounces, the training on this pass's training paths (this stream trains the cache
      // in pass order), then the cache queries of the whole pass, answered by the cache trained on
      // this pass and every earlier one, and the film
      PBRT_HIP(hipStreamWaitEvent(aux, ev_front[S], 0));
      PBRT_HIP(launch_shadow_nrc(v, p, B, spill_aux, spill_per, grid, s->status, aux));
      for (uint32_t b = front; b < c->max_depth; ++b) {
        PBRT_HIP(launch_extend(v, B, q, spill_aux, spill_per, grid, s->status, aux));
        PBRT_HIP(launch_reset_queue(B, q, aux));
        PBRT_HIP(launch_shade_nrc(v, p, B, q, aux, 4u * (uint32_t)s->ctx->cu_count));
        PBRT_HIP(launch_shadow_nrc(v, p, B, spill_aux, spill_per, grid, s->status, aux));
        q ^= 1;
      }
      if (p.train_stride) {
        const uint32_t ranks = (npass - 1 + p.train_offset) / p.train_stride + 1;
        PBRT_HIP(launch_nrc_targets(p, B, tb, ranks, aux));
        // one batch per frame (sample) of the pass, each trained train_steps times, in frame order
        const uint32_t T = ncfg.train_batch, frames = p.pass_samples;
        PBRT_HIP(launch_nrc_select(B, tb, ranks * p.query_vertex, T, frames, aux));
        PBRT_HIP(nrc_launch_encode((const float*)tb.batch_raw, T * frames, nullptr, nrc->lo, nrc->hi, batch_x, aux));
        for (uint32_t f = 0; f < frames; ++f)
          for (uint32_t k = 0; k < std::max(1u, ncfg.train_steps); ++k)
            PBRT_HIP(nrc_train_device(nrc, batch_x + (size_t)f * T * kNrcInPad, tb.batch_t + (size_t)f * T * 3, T,
                                      counter(B.counts, kCntBatch), aux));
        if (ncfg.debug && ncfg.debug_slots && passes + 1 == passes_total) {
          const size_t ns = std::min<size_t>(ncfg.debug_slots, (size_t)ranks * p.query_vertex);
          PBRT_HIP(hipMemcpyAsync(ncfg.debug, tb.slot_raw, ns * 3 * sizeof(float4), hipMemcpyDeviceToHost, aux));
          PBRT_HIP(hipMemcpyAsync(ncfg.debug + (size_t)ncfg.debug_slots * 12, tb.slot_t, ns * sizeof(float4),
                                  hipMemcpyDeviceToHost, aux));
        }
      }
      PBRT_HIP(nrc_launch_query_apply(nrc->dev.infer(), (const float*)B.q_raw, B.q_beta, B.L,
                                      reinterpret_cast<unsigned long long*>(counter(B.counts, kCntQueries)),
                                      nrc->lo, nrc->hi, npass, counter(B.counts, kCntQuery),
                                      (uint32_t)s->ctx->cu_count, aux));
      PBRT_HIP(launch_accumulate(p, B, aux));
      if (passes < n_record) {
        if (ref)
          PBRT_HIP(launch_rmse(film, 1.0 / (double)(first + p.pass_samples), ref, npix * 3, rmse_partial, kRmsePartials,
                               rmse_dev + passes, aux));
        PBRT_HIP(hipEventRecord(events[passes + 1], aux));
      }
      PBRT_HIP(hipEventRecord(ev_back[S], aux));
      passes++;
    }
    if (side) PBRT_HIP(hipStreamSynchronize(side));
    if (aux) PBRT_HIP(hipStreamSynchronize(aux));
    uint64_t rays = 0, queries = 0;
    PBRT_HIP(hipStreamSynchronize(st));
    spill_mark(s, st);
    for (int k = 0; k < nsets; ++k) {
      uint64_t r = 0;
      PBRT_HIP(hipMemcpy(&r, counter(sets[k].counts, kCntRays), sizeof(r), hipMemcpyDeviceToHost));
      rays += r;
      if (nrc) {
        PBRT_HIP(hipMemcpy(&r, counter(sets[k].counts, kCntQueries), sizeof(r), hipMemcpyDeviceToHost));
        queries += r;
      }
    }
    if (film_out) {
      std::vector<double> sum((size_t)npix * 3);
      PBRT_HIP(hipMemcpy(sum.data(), film, sum.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < sum.size(); ++i) film_out[i] = sum[i] / (double)c->spp;
    }
    if (progress) {
      for (uint32_t k = 0; k < n_record; ++k) {
        float ms = 0.f;
        PBRT_HIP(hipEventElapsedTime(&ms, events[0], events[k + 1]));
        progress->ms[k] = ms;
      }
      if (ref && n_record) PBRT_HIP(hipMemcpy(progress->rmse, rmse_dev, n_record * sizeof(double), hipMemcpyDeviceToHost));
      progress->count = n_record;
    }
    int status = 0;
    PBRT_HIP(hipMemcpy(&status, s->status, sizeof(int), hipMemcpyDeviceToHost));
    if (status) throw Error(PBRT_ERR_INTERNAL, "a traversal stack overflowed");
    if (stats) {
      stats->rays = rays;
      stats->paths = (uint64_t)npix * c->spp;
      stats->passes = passes;
      stats->cache_queries = queries;
      stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  });
}

int pbrt_render_jobs(pbrt_scene* s, const pbrt_render_config* cfg, uint32_t n, double* const* films_out,
                     int use_graphs, pbrt_render_stats* stats) {
  return guarded([&] {
    check_scene(s, true);
    if (!cfg || n == 0 || n > 16) throw Error(PBRT_ERR_INVALID_ARG, "1 to 16 jobs");
    uint64_t total_pix = 0;
    uint32_t max_depth = 0;
    for (uint32_t k = 0; k < n; ++k) {
      const pbrt_render_config& c = cfg[k];
      if (!c.width || !c.height || !c.spp || !c.max_depth) throw Error(PBRT_ERR_INVALID_ARG, "bad render config");
      total_pix += (uint64_t)c.width * c.height;
      max_depth = std::max(max_depth, c.max_depth);
    }
    if (total_pix > (1u << 26)) throw Error(PBRT_ERR_INVALID_ARG, "more than 2^26 pixels in the batch");
    std::lock_guard<std::mutex> lock(s->mu);
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = s->ctx->stream;
    spill_order(s, st);
    // the pass schedule, fixed in advance: each pass, every unfinished job takes its share of the
    // path budget, at least one sample per pixel
    const uint64_t budget = 1u << 23;  // about 8M paths (4M: 833 against 814 ms on the jobs workload)
    std::vector<JobDesc> base(n);
    uint64_t film_words = 0, pix = 0;
    for (uint32_t k = 0; k < n; ++k) {
      const pbrt_render_config& c = cfg[k];
      const Pinhole cam(c.width, c.height, c.fov, c.near, c.far,
                        look_at({c.eye[0], c.eye[1], c.eye[2]}, {c.look[0], c.look[1], c.look[2]},
                                {c.up[0], c.up[1], c.up[2]}));
      JobDesc& J = base[k];
      J = JobDesc{};
      for (int i = 0; i < 16; ++i) {
        J.r2s[i] = (float)cam.r2s.m[i / 4][i % 4];
        J.c2w[i] = (float)cam.c2w.m[i / 4][i % 4];
      }
      J.seed = c.seed;
      J.film_base = film_words;
      J.width = c.width;
      J.height = c.height;
      J.spp = c.spp;
      J.max_depth = c.max_depth;
      J.pix_base = (uint32_t)pix;
      J.near = (float)c.near;
      J.far = (float)c.far;
      film_words += (uint64_t)c.width * c.height * 3;
      pix += (uint64_t)c.width * c.height;
    }
    std::vector<JobDesc> sched;
    std::vector<uint32_t> pass_paths;
    std::vector<uint32_t> done(n, 0);
    uint64_t capacity = 0;
    for (;;) {
      uint32_t active = 0;
      for (uint32_t k = 0; k < n; ++k) active += done[k] < cfg[k].spp ? 1 : 0;
      if (!active) break;
      uint64_t paths = 0;
      for (uint32_t k = 0; k < n; ++k) {
        JobDesc J = base[k];
        const uint64_t npix = (uint64_t)J.width * J.height;
        const uint32_t left = J.spp - done[k];
        const uint64_t share = std::max<uint64_t>(1, budget / ((uint64_t)active * npix));
        J.pass_first = done[k];
        J.pass_samples = left ? (uint32_t)std::min<uint64_t>(left, share) : 0;
        J.path_base = (uint32_t)paths;
        paths += npix * J.pass_samples;
        done[k] += J.pass_samples;
        sched.push_back(J);
      }
      pass_paths.push_back((uint32_t)paths);
      capacity = std::max(capacity, paths);
    }
    if (capacity > (1u << 28)) throw Error(PBRT_ERR_INVALID_ARG, "a pass larger than 2^28 paths");
    const uint32_t passes = (uint32_t)pass_paths.size();
    ensure_paths(s, (uint32_t)capacity);

    std::vector<void*> owned;
    struct Release {
      std::vector<void*>& v;
      ~Release() {
        for (void* q : v) (void)hipFree(q);
      }
    } release{owned};
    auto alloc = [&](size_t bytes) {
      void* q = nullptr;
      PBRT_HIP(hipMalloc(&q, bytes ? bytes : 16));
      owned.push_back(q);
      return q;
    };
    hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};  // [set * 2 + queue]
    hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    struct ReleaseGraphs {
      hipGraphExec_t* e;
      hipGraph_t* g;
      ~ReleaseGraphs() {
        for (int i = 0; i < 4; ++i) {
          if (e[i]) (void)hipGraphExecDestroy(e[i]);
          if (g[i]) (void)hipGraphDestroy(g[i]);
        }
      }
    } release_graphs{exec, graph};
    // two path sets on two streams: pass k runs on set k % 2, so one pass's bounces run beside the
    // other's and the last, slowest walks of each traversal launch overlap the other pass's work
    // instead of leaving CUs idle. Each set has its own queues, counters, job table and spill area;
    // the film accumulations wait for each other's events and run in pass order, so the films are
    // those of running the passes one after the other.
    hipStream_t side = nullptr;
    hipEvent_t acc_ev[2] = {nullptr, nullptr};
    struct ReleaseSide {
      hipStream_t& s;
      hipEvent_t* e;
      ~ReleaseSide() {
        for (int i = 0; i < 2; ++i)
          if (e[i]) (void)hipEventDestroy(e[i]);
        if (s) (void)hipStreamDestroy(s);
      }
    } release_side{side, acc_ev};

    double* film = (double*)alloc(film_words * sizeof(double));
    JobDesc* all = (JobDesc*)alloc(sched.size() * sizeof(JobDesc));
    JobDesc* cur = (JobDesc*)alloc(n * sizeof(JobDesc));
    PBRT_HIP(hipMemsetAsync(film, 0, film_words * sizeof(double), st));
    PBRT_HIP(hipMemsetAsync(s->pb.counts, 0, kCounterWords * sizeof(uint32_t), st));
    PBRT_HIP(hipMemcpyAsync(all, sched.data(), sched.size() * sizeof(JobDesc), hipMemcpyHostToDevice, st));
    PBRT_HIP(hipStreamSynchronize(st));
    PathBuffers pb = s->pb;
    pb.film = film;
    const int nsets = passes > 1 ? 2 : 1;
    PathBuffers sets[2] = {pb, pb};
    JobDesc* curs[2] = {cur, cur};
    hipStream_t streams[2] = {st, st};
    uint2* spills[2] = {nullptr, nullptr};
    RenderParams p{};
    p.jobs = cur;
    p.n_jobs = n;
    for (uint32_t k = 0; k < n; ++k)
      if (cfg[k].mis != cfg[0].mis || cfg[k].mis > 1) throw Error(PBRT_ERR_INVALID_ARG, "mis must be 0 or 1, the same for every job");
    p.mis = cfg[0].mis;
    for (uint32_t k = 0; k < n; ++k)
      if (cfg[k].roulette != cfg[0].roulette || cfg[k].roulette > 1)
        throw Error(PBRT_ERR_INVALID_ARG, "roulette must be 0 or 1, the same for every job");
    p.roulette = cfg[0].roulette;
    for (uint32_t k = 0; k < n; ++k) {
      check_integrator(cfg[k], false);
      if (cfg[k].integrator != cfg[0].integrator)
        throw Error(PBRT_ERR_INVALID_ARG, "the integrator must be the same for every job");
    }
    p.integrator = cfg[0].integrator;
    for (uint32_t k = 0; k < n; ++k)
      if (cfg[k].sampler != cfg[0].sampler || cfg[k].sampler > PBRT_SAMPLER_SOBOL)
        throw Error(PBRT_ERR_INVALID_ARG, "sampler must be PBRT_SAMPLER_INDEPENDENT or _SOBOL, the same for every job");
    p.sampler = cfg[0].sampler;
    p.batch_pixels = (uint32_t)pix;
    const SceneView v = view_of(s);
    const uint32_t grid = persistent_grid(s, (uint32_t)capacity);
    const uint32_t spill_per = ensure_spill(s, grid);
    spills[0] = s->spill;
    if (nsets == 2) {
      PathBuffers& o = sets[1];
      const size_t m = capacity;
      o.ray_o = (float4*)alloc(m * sizeof(float4));
      o.ray_d = (float4*)alloc(m * sizeof(float4));
      o.beta = (float4*)alloc(m * sizeof(float4));
      o.L = (float4*)alloc(m * sizeof(float4));
      o.info = (uint4*)alloc(m * sizeof(uint4));
      o.queue[0] = (uint32_t*)alloc(m * sizeof(uint32_t));
      o.queue[1] = (uint32_t*)alloc(m * sizeof(uint32_t));
      o.counts = (uint32_t*)alloc(kCounterWords * sizeof(uint32_t));
      o.hit = (SceneHitRec*)alloc(m * sizeof(SceneHitRec));
      o.sh_o = (float4*)alloc(m * sizeof(float4));
      o.sh_d = (float4*)alloc(m * sizeof(float4));
      o.sh_c = (float4*)alloc(m * sizeof(float4));
      o.capacity = (uint32_t)m;
      curs[1] = (JobDesc*)alloc(n * sizeof(JobDesc));
      spills[1] = (uint2*)alloc((size_t)spill_per * grid * kSceneBlock * sizeof(uint2));
      PBRT_HIP(hipMemsetAsync(o.counts, 0, kCounterWords * sizeof(uint32_t), st));
      PBRT_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
      streams[1] = side;
      for (int i = 0; i < 2; ++i) PBRT_HIP(hipEventCreateWithFlags(&acc_ev[i], hipEventDisableTiming));
      PBRT_HIP(hipStreamSynchronize(st));
    }
    auto bounce = [&](int set, int q, hipStream_t on) -> hipError_t {
      RenderParams ps = p;
      ps.jobs = curs[set];
      const PathBuffers& B = sets[set];
      hipError_t e = launch_extend(v, B, q, spills[set], spill_per, grid, s->status, on);
      if (e == hipSuccess) e = launch_reset_queue(B, q, on);
      if (e == hipSuccess) e = launch_shade(v, ps, B, q, on);
      if (e == hipSuccess) e = launch_shadow(v, B, spills[set], spill_per, grid, s->status, on);
      return e;
    };
    if (use_graphs) {
      for (int set = 0; set < nsets; ++set)
        for (int q = 0; q < 2; ++q) {
          PBRT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
          const hipError_t e = bounce(set, q, st);
          const hipError_t e2 = hipStreamEndCapture(st, &graph[set * 2 + q]);
          PBRT_HIP(e);
          PBRT_HIP(e2);
          PBRT_HIP(hipGraphInstantiate(&exec[set * 2 + q], graph[set * 2 + q], nullptr, nullptr, 0));
        }
    }
    for (uint32_t k = 0; k < passes; ++k) {
      const int S = (int)(k % (uint32_t)nsets);
      hipStream_t on = streams[S];
      PBRT_HIP(hipMemcpyAsync(curs[S], all + (size_t)k * n, n * sizeof(JobDesc), hipMemcpyDeviceToDevice, on));
      RenderParams ps = p;
      ps.jobs = curs[S];
      ps.pass_paths = pass_paths[k];
      PBRT_HIP(launch_generate_jobs(ps, sets[S], on));
      int q = 0;
      for (uint32_t b = 0; b < max_depth; ++b) {
        if (use_graphs) PBRT_HIP(hipGraphLaunch(exec[S * 2 + q], on));
        else PBRT_HIP(bounce(S, q, on));
        q ^= 1;
      }
      // the film takes the passes in order
      if (nsets == 2 && k > 0) PBRT_HIP(hipStreamWaitEvent(on, acc_ev[S ^ 1], 0));
      PBRT_HIP(launch_accumulate_jobs(ps, sets[S], on));
      if (nsets == 2) PBRT_HIP(hipEventRecord(acc_ev[S], on));
    }
    if (nsets == 2) PBRT_HIP(hipStreamSynchronize(side));
    PBRT_HIP(hipStreamSynchronize(st));
    spill_mark(s, st);
    uint64_t rays = 0;
    for (int set = 0; set < nsets; ++set) {
      uint64_t r = 0;
      PBRT_HIP(hipMemcpy(&r, counter(sets[set].counts, kCntRays), sizeof(r), hipMemcpyDeviceToHost));
      rays += r;
    }
    if (films_out) {
      std::vector<double> sum(film_words);
      PBRT_HIP(hipMemcpy(sum.data(), film, film_words * sizeof(double), hipMemcpyDeviceToHost));
      for (uint32_t k = 0; k < n; ++k) {
        if (!films_out[k]) continue;
        const size_t m = (size_t)cfg[k].width * cfg[k].height * 3;
        for (size_t i = 0; i < m; ++i) films_out[k][i] = sum[base[k].film_base + i] / (double)cfg[k].spp;
      }
    }
    int status = 0;
    PBRT_HIP(hipMemcpy(&status, s->status, sizeof(int), hipMemcpyDeviceToHost));
    if (status) throw Error(PBRT_ERR_INTERNAL, "a traversal stack overflowed");
    if (stats) {
      uint64_t paths = 0;
      for (uint32_t k = 0; k < n; ++k) paths += (uint64_t)cfg[k].width * cfg[k].height * cfg[k].spp;
      stats->rays = rays;
      stats->paths = paths;
      stats->passes = passes;
      stats->cache_queries = 0;
      stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  });
}

}  // extern "C"
