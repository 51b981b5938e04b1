// Device scene layout and the launch interface of the scene and renderer kernels (scene.hip).
#pragma once
#include "common.hpp"

namespace pbrt {

enum ShapeKind : int32_t { kShapeMesh = 0, kShapeSphere = 1, kShapeBox = 2, kShapeRect = 3, kShapeCsg = 4 };
enum CsgOp : int32_t { kCsgUnion = 0, kCsgIntersection = 1, kCsgSubtraction = 2 };

// One element of a CSG object (shapes/csg.rs: Elem): a sphere (param[0] = r) or a box (param =
// extents) and its transform to the CSG object's frame (rows 0-2 of the 4x4 and of its inverse).
// A CSG DevObject has shape kShapeCsg, param[0] = the operation, param[1] = the element count and
// mesh = its first element in SceneView::csg.
struct alignas(16) DevCsgElem {
  int32_t shape, pad0, pad1, pad2;
  float param[4];
  float m[12], minv[12];
};
static_assert(sizeof(DevCsgElem) == 128, "CSG element is 128 bytes");
enum MatKind : int32_t { kMatLambertian = 0, kMatDielectric = 1, kMatDiffuseLight = 2, kMatGGX = 3, kMatMetal = 4 };
enum LightKind : int32_t { kLightPoint = 0, kLightArea = 1, kLightBackground = 2 };
enum TexKind : int32_t { kTexColor = 0, kTexChecker = 1 };

// One scene object (a shape, a material and an optional rigid transform), in the leaf order of
// the scene BVH. Matrices are the first three rows of the 4x4 local->world and world->local
// transforms (the fourth row of an affine transform is 0 0 0 1 and contributes w = 1).
struct alignas(16) DevObject {
  int32_t shape, material, has_xform, orig;   // orig: the object's index in the caller's scene
  float param[4];                             // sphere r; box extents; rectangle width, height
  float m[12], minv[12];
  int32_t node_base, tri_base, root_ref, mesh;
  float root_lo[4], root_hi[4];               // mesh BVH root box
  uint32_t vert_base, idx_base;               // mesh vertices / indices for shading
  int32_t node_base4, root_ref4;              // the mesh's 4-wide tree (scenes committed with width 4)
};

struct alignas(16) DevMaterial {
  int32_t kind, tex;     // tex: kTexColor (rgb) or kTexChecker (rgb, rgb2, scale)
  float rgb[3];
  float ior, roughness;  // roughness: GGX alpha, or Metal's fuzz
  float scale;
  float rgb2[3];
  float pad;
};

struct alignas(16) DevLight {
  int32_t kind, object;  // object: leaf-order index of an area light's rectangle
  float pos[3], rgb[3];  // background light: pos = the colour looking down, rgb = looking up
};

// Device view of a committed scene.
struct SceneView {
  const DevNode* tlas;
  RootDesc root;
  const DevObject* objs;
  const DevNode* blas;
  const DevNode4* blas4;  // the meshes' 4-wide trees, or null (the binary trees are walked)
  const DevTri* tris;
  const float* verts;     // xyz per vertex, all meshes
  const uint32_t* idx;    // 3 per triangle, relative to the mesh's vert_base
  const DevMaterial* mats;
  const DevLight* lights;
  const DevCsgElem* csg;
  uint32_t n_objects, n_lights;
  int has_tlas;
  uint32_t n_background;  // background lights among the lights
  uint32_t stack;         // LDS stack entries per lane of the scene walks: 8, 16 or 24 (kernel variant)
  uint32_t occ5;          // 1: the extension walk's variant compiled for five waves per SIMD
};

// Closest hit at scene level: object (caller's index), primitive (mesh triangle or -1), distance
// and the shape's (u, v) — barycentrics for a triangle.
struct alignas(16) SceneHitRec {
  float t, u, v;
  int32_t obj;   // leaf-order object index, -1 on a miss
  int32_t prim;  // original triangle index for meshes, element index for CSG, -1 otherwise
  int32_t orig;  // caller's object index
  int32_t face;  // box face (x-, x+, y-, y+, z-, z+) of a box hit; for a CSG hit | 8 when the normal
                 // is flipped (a subtracted element's surface)
  int32_t pad;
};

struct SceneLaunch {
  SceneView view;
  const pbrt_ray* rays;
  SceneHitRec* hits;
  uint8_t* occluded;
  uint32_t n;
  uint2* spill;
  uint32_t spill_per_thread;
  int* status;
  void* stream;
  uint32_t grid, block;
};

hipError_t launch_scene_trace(const SceneLaunch& a, int any_hit);

// ---- wavefront renderer ----
// One render job of a batch rendered together (pbrt_render_jobs): its camera, sampler seed and
// sizes, the samples of the current pass, and where its paths, pixels and film start.
struct JobDesc {
  float r2s[16], c2w[16];    // the job's pinhole matrices, rounded to FP32
  uint64_t seed, film_base;  // film_base: first double of the job's film
  uint32_t width, height, spp, max_depth;
  uint32_t pass_first, pass_samples;  // samples [pass_first, pass_first + pass_samples) this pass
  uint32_t path_base, pix_base;       // first path of the job in the pass, first pixel in the batch
  float near, far;
  uint32_t pad[2];
};

struct RenderParams {
  uint32_t width, height, spp, max_depth;
  uint32_t pass_first_sample, pass_samples;  // samples [first, first + count) of every pixel
  uint64_t seed;
  float r2s[16], c2w[16];  // pinhole raster->camera and camera->world, built in f64, rounded to FP32 (row-major)
  float near, far;
  // radiance cache (nrc != 0): a path that is not a training path ends into the cache at its first
  // non-specular, non-emissive vertex at or past `query_vertex`; path j of the pass is a training
  // path when (j + train_offset) % train_stride == 0 (never if train_stride is 0), traced to
  // max_depth with its vertices recorded.
  uint32_t nrc, query_vertex, train_stride, train_offset;
  // a batch of jobs (n_jobs > 0): camera, seed, spp and depth come from jobs[info.w] of each path
  const JobDesc* jobs;
  uint32_t n_jobs, pass_paths, batch_pixels;
  uint32_t mis;  // multiple importance sampling (power heuristic) instead of the specular guard
  uint32_t roulette;  // Russian roulette after each scattering
  uint32_t integrator;  // PBRT_INTEGRATOR_PATH, _NAIVE or _NORMAL
  uint32_t sampler;     // PBRT_SAMPLER_INDEPENDENT or PBRT_SAMPLER_SOBOL
};

// Per-vertex record of a training path, field-major: rec[(f * max_depth + m - 1) * train_cap + rank]
// for vertex m (1-based). Fields: throughput on arrival (w: 1 = recorded, 3 = eligible as a cache
// query point), emission added at m, next-event contribution added at m, and the 12 raw cache
// inputs (position, outgoing direction, normal, albedo) in three float4.
enum RecField : uint32_t { kRecBeta = 0, kRecEmit = 1, kRecNee = 2, kRecRaw = 3, kRecFields = 6 };

// Queue and batch counters of the renderer, each on its own 256-byte line (atomics to one L2
// line serialise): counter(base, k) for k in Counter.
constexpr uint32_t kCounterStride = 64;
enum Counter : uint32_t {
  kCntQueue0 = 0, kCntQueue1 = 1, kCntShadow = 2, kCntRays = 4 /* uint64 */, kCntQuery = 6, kCntValid = 7,
  kCntBatch = 8, kCntQueries = 10 /* uint64: cache queries answered, every pass */, kCounters = 16
};
constexpr uint32_t kCounterWords = kCounters * kCounterStride;
__host__ __device__ inline uint32_t* counter(uint32_t* base, uint32_t k) { return base + (size_t)k * kCounterStride; }

struct PathBuffers {
  float4* ray_o;   // origin xyz, near
  float4* ray_d;   // direction xyz, far
  float4* beta;    // throughput rgb, specular flag (1/0)
  float4* L;       // radiance rgb
  uint4* info;     // pixel, sample, bounce, alive
  uint32_t* queue[2];
  uint32_t* counts;  // kCounterWords words: see Counter
  SceneHitRec* hit;
  float4* sh_o;    // shadow ray origin, near
  float4* sh_d;    // shadow ray direction, far
  float4* sh_c;    // contribution rgb, path index bits
  double* film;    // width * height * 3, running sum
  uint32_t capacity;
  // radiance cache (the pass's query count in counter kCntQuery)
  float4* q_raw;   // [capacity][3] raw cache inputs of each query
  float4* q_beta;  // throughput rgb, path index bits in w
  float4* rec;     // training-path records (see RecField)
  uint32_t train_cap;
};

__host__ __device__ inline bool is_training_path(const RenderParams& p, uint32_t j, uint32_t* rank) {
  if (!p.nrc || p.train_stride == 0) return false;
  const uint32_t k = j + p.train_offset;
  if (k % p.train_stride != 0) return false;
  *rank = k / p.train_stride;
  return true;
}

hipError_t launch_generate(const SceneView& v, const RenderParams& p, const PathBuffers& b, void* stream);
// Workgroups of the scene traversal kernels resident on one CU at once, for an LDS stack of `stack`
// entries (0 if the runtime cannot tell).
uint32_t scene_blocks_per_cu(uint32_t stack);
hipError_t launch_extend(const SceneView& v, const PathBuffers& b, int q, uint2* spill, uint32_t spill_per_thread,
                         uint32_t grid, int* status, void* stream);
hipError_t launch_shade(const SceneView& v, const RenderParams& p, const PathBuffers& b, int q, void* stream);
hipError_t launch_shadow(const SceneView& v, const PathBuffers& b, uint2* spill, uint32_t spill_per_thread,
                         uint32_t grid, int* status, void* stream);
hipError_t launch_accumulate(const RenderParams& p, const PathBuffers& b, void* stream);
hipError_t launch_reset_queue(const PathBuffers& b, int q, void* stream);
// grid_cap > 0: at most that many workgroups, striding over the queue (for queues known to be short,
// such as the training paths' bounces past the query vertex, whose full-size grids would take CU
// slots from the traversal launches running beside them); 0: one lane per path of the set.
hipError_t launch_shade_nrc(const SceneView& v, const RenderParams& p, const PathBuffers& b, int q, void* stream,
                            uint32_t grid_cap = 0);
hipError_t launch_shadow_nrc(const SceneView& v, const RenderParams& p, const PathBuffers& b, uint2* spill,
                             uint32_t spill_per_thread, uint32_t grid, int* status, void* stream);
hipError_t launch_generate_nrc(const RenderParams& p, const PathBuffers& b, void* stream);
// a pass of a job batch: p.n_jobs jobs, p.pass_paths paths, p.batch_pixels pixels in all
hipError_t launch_generate_jobs(const RenderParams& p, const PathBuffers& b, void* stream);
hipError_t launch_accumulate_jobs(const RenderParams& p, const PathBuffers& b, void* stream);

constexpr uint32_t kSceneBlock = 128;
constexpr int kSceneStack = 16;

}  // namespace pbrt
