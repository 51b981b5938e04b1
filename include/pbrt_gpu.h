/*
 * pbrt_gpu.h — C ABI of the MI355X ray hot path (gfx950, HIP).
 *
 * The drop-in boundary for nmichel/pbrt's ray path. Every entry point returns an int status
 * (PBRT_OK = 0, negative = error; pbrt_last_error() gives a thread-local message); no C++ exception
 * crosses the boundary. Input arrays are copied at call time, so the caller keeps ownership.
 * Parameters ending in `_dev` are device pointers (HBM); `stream` is a hipStream_t (NULL = the
 * context's own stream). Committed meshes and scenes are immutable and may be shared by threads.
 *
 * Reference interfaces replaced are cited as file:line into the reference source tree.
 */
#ifndef PBRT_GPU_H
#define PBRT_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference panics instead: triangle_mesh.rs:28, config.rs:155-166) ---- */
#define PBRT_OK 0
#define PBRT_ERR_INVALID_ARG (-1)
#define PBRT_ERR_EMPTY_MESH (-2)
#define PBRT_ERR_DEVICE (-3)
#define PBRT_ERR_OUT_OF_MEMORY (-4)
#define PBRT_ERR_LAYOUT (-5)
#define PBRT_ERR_INTERNAL (-6)

typedef struct pbrt_ctx pbrt_ctx;
typedef struct pbrt_mesh pbrt_mesh;

/* A ray as stored in HBM: 32 bytes, two 16-byte halves. The direction is used as given (the
 * reference normalises in Ray::new, src/geom/ray.rs:10-18; callers pass normalised directions).
 * tmin/tmax are the reference's `near`/`far` arguments (triangle_mesh/bvh.rs:350). A tmax above
 * FLT_MAX (i.e. +inf) is treated as FLT_MAX. */
typedef struct pbrt_ray {
  float ox, oy, oz, tmin;
  float dx, dy, dz, tmax;
} pbrt_ray;

/* Closest hit: 16 bytes. prim = index of the triangle in the caller's index array (the reference's
 * `current_tri_idx`, triangle_mesh/bvh.rs:489), -1 on a miss; t, u, v as in
 * TriangleIntersection (triangle_mesh.rs:176-181). */
typedef struct pbrt_hit {
  int32_t prim;
  float t, u, v;
} pbrt_hit;

/* Mesh BVH build parameters. The reference is bins = 8, t_trav = 0, max_leaf = 0 (no bound)
 * (triangle_mesh/bvh.rs:129, 725-735). A node is split when t_trav*A + (A_L*N_L + A_R*N_R) < A*N,
 * and always when N > max_leaf (max_leaf > 0). Device leaves hold at most 32 triangles. */
typedef struct pbrt_bvh_params {
  uint32_t bins;
  uint32_t max_leaf;
  double t_trav;
} pbrt_bvh_params;

/* triangle_mesh/bvh.rs:65-80 (BuildStats). */
typedef struct pbrt_build_stats {
  uint64_t node_count, leaf_count, max_depth, max_leaf_tri_count, total_leaf_tri_count;
} pbrt_build_stats;

/* Launch shape of the traversal kernels. Zero fields take the tuned defaults. */
typedef struct pbrt_launch {
  uint32_t block;        /* threads per workgroup (64, 128, 256) */
  uint32_t waves_per_cu; /* waves per CU in the grid (up to 128): the persistent grid, or more than fit
                            at once, in which case workgroups queue and start as others finish */
  uint32_t lds_stack;    /* LDS short-stack entries per lane (8, 12, 16, 24, 32); deeper entries spill to HBM */
  uint32_t refill;       /* 0: a lane takes its next ray as soon as its ray is done; N in 1..64: done lanes
                            wait and take their next rays together, once N lanes of the wave wait or no
                            lane of it is still tracing */
  uint32_t width;        /* tree walked by closest- and any-hit: 0 or 2 the binary tree, 4 a 4-wide tree
                            collapsed from it (128-byte nodes; built on first use). Same hits either way;
                            pbrt_mesh_trace_closest_stats always counts on the binary tree */
  uint32_t packet;       /* 1: closest hit by wave packets (64 consecutive rays walk the binary tree on one
                            wave-wide stack; for coherent rays; block, lds_stack, refill unused); needs
                            width 2 and a tree at most 33 levels deep. Same hits */
} pbrt_launch;

/* ---- context, errors, device plumbing ---- */
const char* pbrt_last_error(void);
int pbrt_version(void);
int pbrt_ctx_create(int device, pbrt_ctx** out);
int pbrt_ctx_destroy(pbrt_ctx* ctx);
int pbrt_ctx_stream(pbrt_ctx* ctx, void** stream_out);
int pbrt_ctx_device_info(pbrt_ctx* ctx, int* cu_count, char* name, size_t name_len);
int pbrt_dev_alloc(pbrt_ctx* ctx, size_t bytes, void** dptr_out);
int pbrt_dev_free(pbrt_ctx* ctx, void* dptr);
int pbrt_memcpy_h2d(pbrt_ctx* ctx, void* dst_dev, const void* src, size_t bytes);
int pbrt_memcpy_d2h(pbrt_ctx* ctx, void* dst, const void* src_dev, size_t bytes);
int pbrt_memset_dev(pbrt_ctx* ctx, void* dst_dev, int value, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (bytes a multiple of 16); also the known-size
 * reference for calibrating hardware byte counters. */
int pbrt_dev_copy(pbrt_ctx* ctx, void* dst_dev, const void* src_dev, size_t bytes, void* stream);
/* A non-blocking stream of default priority on the context's device (NULL stream arguments elsewhere
 * mean the context's own stream). */
int pbrt_stream_create(pbrt_ctx* ctx, void** stream_out);
int pbrt_stream_destroy(pbrt_ctx* ctx, void* stream);
int pbrt_stream_sync(pbrt_ctx* ctx, void* stream);
int pbrt_event_create(pbrt_ctx* ctx, void** ev_out);
int pbrt_event_record(void* ev, void* stream);
/* Work queued on `stream` (NULL: the context's) after this call waits, on the device, for the work
 * before the last record of `ev`. */
int pbrt_stream_wait_event(pbrt_ctx* ctx, void* stream, void* ev);
int pbrt_event_elapsed_ms(void* ev_start, void* ev_end, float* ms_out);
int pbrt_event_destroy(void* ev);

/* ---- mesh: TriangleMesh::new (triangle_mesh.rs:27-41) = BVHTree::new + build (bvh.rs:179-230) ----
 * xyz: nv*3 FP32 vertex coordinates; idx: nt*3 vertex indices. An empty mesh is
 * PBRT_ERR_EMPTY_MESH (the reference asserts, triangle_mesh.rs:28). The BVH is built on the host
 * and flattened into 64-byte device nodes (two child boxes each) and 48-byte triangles. With
 * ctx == NULL the mesh is built and flattened on the host only (for inspection; tracing it fails). */
int pbrt_mesh_create(pbrt_ctx* ctx, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                     const pbrt_bvh_params* params, pbrt_mesh** out);
/* The same mesh with its BVH built on the device (BVHTree::build, bvh.rs:240-309, 609-735, as a
 * level-synchronous GPU build): the tree, build statistics and device arrays are those
 * pbrt_mesh_create gives, byte for byte. Needs a context. Beyond pbrt_mesh_create's errors, a node
 * of more than 256 triangles sharing one centroid (no plane separates them) is refused with
 * PBRT_ERR_INVALID_ARG; pbrt_mesh_export_tree is not available on such a mesh. */
int pbrt_mesh_create_device(pbrt_ctx* ctx, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                            const pbrt_bvh_params* params, pbrt_mesh** out);
/* Host copies of the device arrays (node_bytes / tri_bytes of pbrt_mesh_device_bytes; NULL to
 * skip), the root box (min xyz, max xyz) and the root reference (>= 0 node, < 0 leaf). */
int pbrt_mesh_copy_layout(const pbrt_mesh* mesh, void* nodes_out, void* tris_out, float* root_box6, int32_t* root_ref);
int pbrt_mesh_destroy(pbrt_mesh* mesh);
/* BVHTree::build_stats (triangle_mesh/bvh.rs:539-546); triangle_count (bvh.rs:529-531). */
int pbrt_mesh_build_stats(const pbrt_mesh* mesh, pbrt_build_stats* out);
int pbrt_mesh_triangle_count(const pbrt_mesh* mesh, uint32_t* out);
/* Bytes of the device BVH: nodes and triangles. */
int pbrt_mesh_device_bytes(const pbrt_mesh* mesh, uint64_t* node_bytes, uint64_t* tri_bytes);
/* Host copy of the built tree, in preorder: per node its box (min xyz, max xyz, f64), whether it is
 * a leaf, child node indices (interior), and for leaves [first, first+count) into leaf_tris, which
 * lists original triangle indices. Arrays sized by build_stats.node_count and the triangle count. */
int pbrt_mesh_export_tree(const pbrt_mesh* mesh, double* boxes, uint8_t* is_leaf, uint32_t* left, uint32_t* right,
                          uint32_t* first, uint32_t* count, uint32_t* leaf_tris);

/* Batched closest hit: BVHTree::query (triangle_mesh/bvh.rs:350-352, ordered traversal 450-526),
 * with the order-independent tie rule (smallest t, then smallest triangle index). */
int pbrt_mesh_trace_closest(pbrt_mesh* mesh, const pbrt_ray* rays_dev, uint32_t n, pbrt_hit* hits_dev, void* stream);
/* Same, counting the work per ray like BVHTree::query_instrumented (bvh.rs:359-361): stats_dev
 * receives n x {nodes_visited, box_tests, triangle_tests} (uint32). */
int pbrt_mesh_trace_closest_stats(pbrt_mesh* mesh, const pbrt_ray* rays_dev, uint32_t n, pbrt_hit* hits_dev,
                                  uint32_t* stats_dev, void* stream);
/* Batched any hit: BVHTree::intersect_p (triangle_mesh/bvh.rs:381-420); occluded_dev[i] = 0 or 1. */
int pbrt_mesh_trace_any(pbrt_mesh* mesh, const pbrt_ray* rays_dev, uint32_t n, uint8_t* occluded_dev, void* stream);
/* Closest hit as pbrt_mesh_trace_closest, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid allocation); no counterpart
 * in the reference. Calls no hipMalloc or hipFree: the traversal stack's spill past the LDS stack
 * goes to a 64 MiB array the library holds on every device it runs on until the process ends, not
 * to the mesh's own spill area. Same hits as pbrt_mesh_trace_closest with the mesh's launch, whose
 * persistent grid is capped at the threads whose spill the array holds (fewer resident waves for
 * deep trees with short LDS stacks). A width-4 launch must have been set by pbrt_mesh_set_launch
 * (which builds the 4-wide tree). A tree whose spill for one workgroup exceeds the array is refused
 * with PBRT_ERR_INVALID_ARG. Every no-alloc call on a device shares that device's array: calls that
 * overlap on one device (two streams, two threads) race and must be ordered by the caller. */
int pbrt_mesh_trace_closest_noalloc(pbrt_mesh* mesh, const pbrt_ray* rays_dev, uint32_t n, pbrt_hit* hits_dev,
                                    void* stream);
/* The records a traversal fetches, per ray, on the tree the mesh's launch walks (width 2 or 4), as
 * BVHTree::query_instrumented counts work (triangle_mesh/bvh.rs:18-47,354-361) but for the device
 * layout: work_dev receives n x {node records (64 bytes each on the binary tree, 128 bytes on the
 * 4-wide tree), triangle records (48 bytes)} (uint32). out_dev receives what pbrt_mesh_trace_closest
 * (any_hit = 0, pbrt_hit) or pbrt_mesh_trace_any (any_hit != 0, bytes) gives. The ray order of
 * pbrt_mesh_set_ray_tiles is followed; the counts do not depend on the launch shape. Not available
 * with a packet launch (PBRT_ERR_INVALID_ARG). */
int pbrt_mesh_trace_work(pbrt_mesh* mesh, const pbrt_ray* rays_dev, uint32_t n, void* out_dev, int any_hit,
                         uint32_t* work_dev, void* stream);
/* Launch shape used by the mesh's traversal kernels (all fields 0 = defaults). */
int pbrt_mesh_set_launch(pbrt_mesh* mesh, const pbrt_launch* launch);
/* Ray order of the mesh's traversal kernels (no counterpart in the reference, which queries one ray
 * at a time): image_width 0 (the default) takes each batch's rays in batch order; image_width W > 0
 * declares every later batch an image W rays wide stored row after row (W a multiple of 16, the
 * batch's ray count a multiple of 16 W or the call fails with PBRT_ERR_INVALID_ARG), which the
 * closest-hit, instrumented and any-hit kernels take in tiles of 8 x 16 rays so that a wave starts
 * on an 8 x 8 patch of the image. Every result is stored at its ray's own index and is the same as
 * in batch order. Packet traversal ignores the setting. */
int pbrt_mesh_set_ray_tiles(pbrt_mesh* mesh, uint32_t image_width);
int pbrt_mesh_get_launch(const pbrt_mesh* mesh, pbrt_launch* out);
/* Autotuner (no counterpart in the reference, which has one CPU code path): times every supported
 * launch shape — block size x LDS stack depth (8 to 32) x waves per CU (8 to 128) x refill policy (0, 1, 4) x tree
 * width (2, 4), and for closest hit the wave-packet walk at each waves per CU — of the closest-hit
 * (any_hit = 0) or any-hit kernel on the caller's rays, `reps` timed runs each, and keeps the
 * fastest for this mesh. out_dev receives hits (pbrt_hit) or occlusion bytes and is overwritten. */
int pbrt_mesh_autotune(pbrt_mesh* mesh, const pbrt_ray* rays_dev, uint32_t n, void* out_dev, int any_hit,
                       uint32_t reps, pbrt_launch* chosen, float* best_ms);

/* Sticky error word of the mesh's kernels (waits for the device): 0 = fine, bit 0 = a traversal
 * stack exceeded its bound (cannot happen for a tree of the reported depth).
 * A mesh keeps an HBM stack spill area for each stream it traces on (up to four): traces on
 * different streams run at once. Past four streams an area changes hands, and a trace on its new
 * stream waits on the device for the old stream's last trace on it. Traces on one stream run in
 * stream order. */
int pbrt_mesh_status(pbrt_mesh* mesh, int* status_out);

/* ---- scene: Scene::{new, add_object, add_light, commit} (src/scene.rs:80-115) ----------------
 * Construction mirrors the reference as data: meshes, materials, objects (a shape + material,
 * optionally under a rigid affine transform given with its inverse, as Transformed does,
 * src/objects/transformed.rs) and lights. commit() builds one BVH per mesh and the reference's
 * median-split object BVH (src/bvh.rs:291-335) and uploads the scene. */
typedef struct pbrt_scene pbrt_scene;

#define PBRT_SHAPE_MESH 0      /* TriangleMesh (triangle_mesh.rs) */
#define PBRT_SHAPE_SPHERE 1    /* Sphere(r) at the origin (sphere.rs), param[0] = r */
#define PBRT_SHAPE_BOX 2       /* AABox(extent) centred at the origin (aabox.rs), param = extent */
#define PBRT_SHAPE_RECT 3      /* Rectangle(w, h) in y = 0, normal +y (rectangle.rs), param = w, h */
#define PBRT_SHAPE_CSG 4       /* CSG of spheres and boxes (shapes/csg/*.rs), see pbrt_scene_add_csg */
#define PBRT_CSG_UNION 0        /* csg/union.rs */
#define PBRT_CSG_INTERSECTION 1 /* csg/intersection.rs */
#define PBRT_CSG_SUBTRACTION 2  /* csg/substraction.rs: elements 1.. subtracted from element 0 */
#define PBRT_MAT_LAMBERTIAN 0  /* materials/lambertian.rs */
#define PBRT_MAT_DIELECTRIC 1  /* materials/dielectric.rs */
#define PBRT_MAT_DIFFUSE_LIGHT 2 /* materials/diffuse_light.rs (rgb = emitted radiance) */
#define PBRT_MAT_GGX 3         /* glossy microfacet reflector (no counterpart in the reference) */
#define PBRT_MAT_METAL 4       /* materials/metal.rs: roughness = fuzz */
#define PBRT_TEX_COLOR 0       /* textures/plain_color.rs: the material's rgb */
#define PBRT_TEX_CHECKER 1     /* textures/checker_board.rs: rgb, rgb2, scale at the surface (u, v) */

/* Scene-level closest hit: 32 bytes. object = the caller's object index (-1 on a miss); prim = the
 * mesh triangle (-1 for other shapes); t = distance along the ray (local distance for a transformed
 * object, as the reference); u, v = triangle barycentrics (0 for other shapes). */
typedef struct pbrt_scene_hit {
  float t, u, v;
  int32_t leaf_slot;
  int32_t prim, object;
  int32_t reserved0, reserved1;
} pbrt_scene_hit;

int pbrt_scene_create(pbrt_ctx* ctx, const pbrt_bvh_params* mesh_params, pbrt_scene** out);
int pbrt_scene_destroy(pbrt_scene* scene);
int pbrt_scene_add_mesh(pbrt_scene* scene, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                        uint32_t* mesh_id);
int pbrt_scene_add_material(pbrt_scene* scene, int kind, const double rgb[3], double ior, double roughness,
                            uint32_t* material_id);
/* m16 / minv16: row-major 4x4 local->world and world->local (both NULL for no transform). */
int pbrt_scene_add_object(pbrt_scene* scene, int shape, int mesh, const double param[3], uint32_t material,
                          const double* m16, const double* minv16, uint32_t* object_id);
/* PointLight (lights/point_light.rs): radiant intensity rgb at pos. */
/* A CSG object: csg::{Union, Intersection, Substraction}::new(elements) (shapes/csg/*.rs) wrapped in
 * Simple (+ Transformed when m16 != NULL). Element k is shapes[k] (PBRT_SHAPE_SPHERE with params[3k]
 * = r, or PBRT_SHAPE_BOX with params[3k..3k+2] = extents) under the affine transform m[16k..] (row-
 * major 4x4) with inverse minv[16k..], in the order the constructor receives them. 1 to 64
 * elements. A hit reports the element it lies on as its primitive. */
int pbrt_scene_add_csg(pbrt_scene* scene, int op, uint32_t n, const int* shapes, const double* params,
                       const double* m, const double* minv, uint32_t material, const double* m16,
                       const double* minv16, uint32_t* id);
int pbrt_scene_add_point_light(pbrt_scene* scene, const double pos[3], const double rgb[3]);
/* The material's texture (Lambertian / Metal / Dielectric / DiffuseLight take theirs from a texture
 * in the reference, textures/*.rs): PBRT_TEX_COLOR (the material's rgb) or PBRT_TEX_CHECKER
 * (CheckerBoard::new(rgb, rgb2, scale)). */
int pbrt_scene_set_texture(pbrt_scene* scene, uint32_t material, int kind, const double rgb2[3], double scale);
/* BackgroundInfiniteLight::new(down, up) (lights/background_infinite_light.rs): radiance blended
 * from `down` (direction y = -1) to `up` (y = +1); seen by rays that leave the scene after a
 * specular bounce, and sampled by next-event estimation over the sphere of directions. */
int pbrt_scene_add_background_light(pbrt_scene* scene, const double down[3], const double up[3]);
/* Area light on a rectangle object (no counterpart in the reference; follows ideas/area_light.md):
 * uniform area sampling, one-sided emission along the rectangle's normal, radiance rgb. */
int pbrt_scene_add_area_light(pbrt_scene* scene, uint32_t object, const double rgb[3]);
/* Where pbrt_scene_commit builds the meshes' BVHs: PBRT_BUILD_HOST (default; the host's cores, one
 * mesh per thread) or PBRT_BUILD_DEVICE (every mesh in one batched device build, the arrays of
 * pbrt_mesh_create_device: the same trees, so the same hits). pbrt_scene_build_seconds: wall time
 * of that step in the last commit, with the 4-wide collapse (on the device) when the scene walks
 * 4-wide trees, without the upload of host-built arrays. */
#define PBRT_BUILD_HOST 0
#define PBRT_BUILD_DEVICE 1
int pbrt_scene_set_mesh_build(pbrt_scene* scene, int where);
/* Tree the scene walks inside its meshes: 4 (default) trees collapsed to 4-wide nodes on the device
 * at commit (the rule of pbrt_launch.width 4 for a mesh), 2 the binary trees. Same hits and films
 * either way. */
int pbrt_scene_set_mesh_width(pbrt_scene* scene, uint32_t width);
/* Launch settings of the scene walks (pbrt_scene_trace_*, the renderer's extension and shadow rays):
 * LDS stack entries per lane (8, 16, 24; deeper entries spill to HBM), workgroups of 128 threads
 * per CU in the grid (1-32; beyond what fits on a CU at once, workgroups queue and start as others
 * finish; 0 = twice as many as fit), mesh tree width (2, 4; 4 collapses the trees on first use),
 * min_waves (0: the renderer's extension walk at the compiler's register count; 5: its variant
 * compiled for five waves per SIMD). Same hits and films for every setting. Defaults 16, 0, 4, 0
 * (width 0 also means 4). */
typedef struct pbrt_scene_launch {
  uint32_t lds_stack, blocks_per_cu, width, min_waves;
} pbrt_scene_launch;
int pbrt_scene_set_launch(pbrt_scene* scene, const pbrt_scene_launch* launch);
int pbrt_scene_get_launch(pbrt_scene* scene, pbrt_scene_launch* out);
/* Scene autotuner (no counterpart in the reference): times the committed scene's closest-hit walk
 * on the caller's rays for every LDS stack depth x workgroups per CU (2, 4, 6, 8, 16, 32) x mesh tree
 * width, `reps` timed runs each, and keeps the fastest (min_waves is left as set: the walk timed
 * here has one variant). hits_dev (n entries) is overwritten. */
int pbrt_scene_autotune(pbrt_scene* scene, const pbrt_ray* rays_dev, uint32_t n, pbrt_scene_hit* hits_dev,
                        uint32_t reps, pbrt_scene_launch* chosen, float* best_ms);
int pbrt_scene_build_seconds(pbrt_scene* scene, double* seconds);
int pbrt_scene_commit(pbrt_scene* scene);
int pbrt_scene_info(pbrt_scene* scene, uint32_t* objects, uint32_t* tlas_nodes, uint32_t* tlas_depth,
                    uint32_t* mesh_depth);
/* Scene::intersect (scene.rs:118-120) batched: closest hit per ray. Ties at equal distance go to the
 * smaller (object, triangle) index. */
int pbrt_scene_trace_closest(pbrt_scene* scene, const pbrt_ray* rays_dev, uint32_t n, pbrt_scene_hit* hits_dev,
                             void* stream);
/* Scene::intersect_p (scene.rs:130-132) batched. */
int pbrt_scene_trace_any(pbrt_scene* scene, const pbrt_ray* rays_dev, uint32_t n, uint8_t* occluded_dev, void* stream);
int pbrt_scene_status(pbrt_scene* scene, int* status_out);

/* ---- renderer: renderers::mt::render (src/renderers/mt.rs:23) with PathIntegrator::li
 * (src/integrators/path.rs:48-115) as a device wavefront. Pinhole camera looking from eye at look
 * (pin_hole.rs), fov in radians; each sample jitters its pixel by two uniforms; every random number
 * is drawn from a counter keyed by (seed, pixel * spp + sample, bounce, slot). */
typedef struct pbrt_render_config {
  uint32_t width, height, spp, max_depth;
  uint64_t seed;
  double fov;
  double eye[3], look[3], up[3];
  double near, far;
  uint32_t samples_per_pass; /* samples of every pixel in flight at once (0 = half the samples, at
                              * most about 8M paths) */
  uint32_t mis;              /* 0: the reference's estimator (integrators/path.rs:48-115: next-event
                                estimation, emission only after a specular bounce); 1: multiple
                                importance sampling of next-event estimation and the BSDF sample,
                                power heuristic (the reference's plan, ideas/area_light.md:158-165) */
  uint32_t roulette;         /* 1: Russian roulette after each scattering, as integrators/path.rs:93-100
                                has it commented out (continue with probability q = the throughput's
                                largest component, capped at 1 here, the throughput divided by q);
                                not with the radiance cache. 0: paths end only at max_depth */
  uint32_t integrator;       /* the reference's integrators::Type (src/integrators.rs:23-28, chosen in
                                src/main.rs:16-20): PBRT_INTEGRATOR_PATH (0), PBRT_INTEGRATOR_NAIVE or
                                PBRT_INTEGRATOR_NORMAL; any other value is PBRT_ERR_INVALID_ARG, and so
                                are mis, roulette and the radiance cache with anything but PATH */
  uint32_t sampler;          /* PBRT_SAMPLER_INDEPENDENT (0): every uniform from the counter keyed by
                                (seed, pixel * spp + sample, bounce, slot); PBRT_SAMPLER_SOBOL (1): the
                                pixel jitter and, at vertices 1 and 2, the light and BSDF samples' pairs
                                from the pixel's Owen-scrambled Sobol (0,2)-sequence at its sample index
                                (stratified across a pixel's samples; the other draws as for 0). The
                                reference samples with an unseeded thread RNG (utils.rs:5-7); both are
                                unbiased, the second converges faster at equal spp */
  uint32_t reserved;
} pbrt_render_config;
#define PBRT_SAMPLER_INDEPENDENT 0
#define PBRT_SAMPLER_SOBOL 1
#define PBRT_INTEGRATOR_PATH 0   /* PathIntegrator::li, integrators/path.rs:48-115 */
#define PBRT_INTEGRATOR_NAIVE 1  /* NaiveIntegrator::li, integrators/naive.rs:30-54: emission at every hit,
                                    the scattered ray followed to max_depth, no next-event estimation,
                                    the sky gradient of naive.rs:21-26 on a miss */
#define PBRT_INTEGRATOR_NORMAL 2 /* NormalIntegrator::li, integrators/normal.rs:17-30: (n + 1) / 2 of the
                                    camera ray's hit, black on a miss */

typedef struct pbrt_render_stats {
  uint64_t rays;   /* every ray traced: camera, extension and shadow */
  uint64_t paths;  /* width * height * spp */
  uint64_t passes;
  double seconds;  /* host wall time of the call */
  uint64_t cache_queries; /* paths ended into the radiance cache (pbrt_render_ex with a cache), else 0 */
} pbrt_render_stats;

/* film_out (host, may be NULL): width*height*3 doubles, the per-pixel mean radiance, accumulated
 * in FP64 in a fixed sample order. */
int pbrt_render(pbrt_scene* scene, const pbrt_render_config* config, double* film_out, pbrt_render_stats* stats);
/* The scene autotuner on renders: the candidates of pbrt_scene_autotune x min_waves (0, 5), each timed by a
 * pbrt_render of `config` (no film; one warm-up and `reps` timed renders, host wall time), and the
 * fastest kept. A short configuration (a few milliseconds) is enough. */
int pbrt_scene_autotune_render(pbrt_scene* scene, const pbrt_render_config* config, uint32_t reps,
                               pbrt_scene_launch* chosen, float* best_ms);

/* Film output (host). pbrt_film_to_rgba8: width*height RGBA bytes of a film as the reference
 * writes its image — Spectrum::gamma_correct then to_rgb (spectrum.rs:13-30, renderers/mt.rs:83-90):
 * sqrt, clamp to [0, 1], x255, truncated; alpha 255. tonemap 1 applies Reinhard x / (1 + x) per
 * channel first (an option the reference lacks), 0 none. pbrt_film_write_png: the same pixels as an
 * 8-bit RGBA PNG file (image_write, renderers/mt.rs:152-165), stored deflate blocks: the file is a
 * pure function of the pixels. */
int pbrt_film_to_rgba8(const double* film, uint32_t width, uint32_t height, uint32_t tonemap, uint8_t* rgba_out);
int pbrt_film_write_png(const double* film, uint32_t width, uint32_t height, uint32_t tonemap, const char* path);

/* ---- scene files (host) ----
 * pbrt_ply_read: an ASCII PLY mesh as the reference loads it (load_ply_mesh, loader/mesh_loader.rs;
 * loader/ply/ply_loader.rs:251-255 reads ASCII only): the x, y, z vertex properties as FP32
 * (correctly rounded), triangle faces only, each face's indices reversed when `reverse` != 0.
 * Call with xyz_out or idx_out NULL to get the counts in *nv / *nt; then with arrays of at least
 * 3 * *nv floats and 3 * *nt indices (the counts in *nv / *nt are their capacities). */
int pbrt_ply_read(const char* path, int reverse, float* xyz_out, uint32_t* nv, uint32_t* idx_out, uint32_t* nt);
/* pbrt_scene_load_stage: Loader::load_scene (src/loader.rs:29-46) — a `.stage` file parsed
 * (loader/parser.rs, loader/parser/lexer.rs) and built as SceneBuilderVisitor does
 * (loader/visitors/scene_builder.rs): objects in reverse file order, transform steps applied in
 * listed order, mesh files read relative to base_dir (NULL: the working directory, as the
 * reference), then a point light of intensity 15 at (0, 2, 1) and BackgroundInfiniteLight(white,
 * (0.5, 0.7, 1.0)). A file may instead declare its lights after `scene`, beside its objects (the
 * `light` production the reference plans, IDEAS.md:25-27, 180-199): `light point color r g b
 * [transform { ... }]` (the transform applied to the origin), `light uniform_infinite color r g b`,
 * `light background_infinite color r g b color r g b` (y = -1, then y = +1). Such a file gets only
 * its declared lights, in file order, after area lights for its emissive (diffuse_light, colour
 * texture) rectangles, registered as the objects are built. Returns a committed scene; view_out (may be NULL) receives the pin-hole camera
 * (eye, look, up) and the renderer's defaults fov 90 degrees, near 1e-4, far 1000, other fields
 * untouched. A compound object becomes its children; cylinders, planes and the thin-lens camera
 * (no device counterpart) are refused with PBRT_ERR_INVALID_ARG, as is any syntax error (the
 * message names the line). */
int pbrt_scene_load_stage(pbrt_ctx* ctx, const char* path, const char* base_dir, const pbrt_bvh_params* mesh_params,
                          pbrt_scene** scene_out, pbrt_render_config* view_out);

/* A batch of n render jobs (1 to 16) on one scene traced together (no counterpart in the
 * reference, which renders one image per run): every pass gives each unfinished job a share of
 * about 8M paths (at least one sample per pixel), and the jobs' paths share the wavefront's
 * device-resident queues and its persistent traversal kernels. Each job's film is bit for bit what
 * pbrt_render returns for that job alone. films_out: n host pointers (NULL entries skipped), each
 * width*height*3 doubles. use_graphs != 0: each bounce's kernels (closest hit, queue reset, shade,
 * shadow rays) are captured once per queue parity and path set as a hipGraph and replayed for every
 * bounce of every pass. Consecutive passes run side by side on the render stream and a second
 * stream the call creates (two path sets; the film takes the passes in order), and the call
 * returns once both are done. stats: totals over the batch. */
int pbrt_render_jobs(pbrt_scene* scene, const pbrt_render_config* configs, uint32_t n, double* const* films_out,
                     int use_graphs, pbrt_render_stats* stats);

/* ---- neural radiance cache ---------------------------------------------------------------------
 * No counterpart in the reference, whose only integrator is plain path tracing
 * (src/integrators/path.rs:13-97). An MLP 81 -> 64 -> 64 -> 64 -> 64 -> 3 (ReLU hidden layers,
 * linear output) maps an encoded path vertex to outgoing radiance: inputs are the position under a
 * 12-frequency sin/cos encoding relative to the scene box, the outgoing direction, the normal and
 * the albedo (12 raw floats per vertex: p, wo, n, albedo). Weights and activations are BF16 on the
 * matrix cores with FP32 accumulation; FP32 master parameters and Adam moments. Flat parameter
 * order: per layer W [out][in] row-major, then b [out]; PBRT_NRC_PARAMS floats.
 * Every `_dev` pointer is device memory, every other pointer host memory. */
#define PBRT_NRC_PARAMS 17923
#define PBRT_NRC_FEATURES 96 /* BF16 features per encoded vertex (81 used, zero padded) */
typedef struct pbrt_nrc pbrt_nrc;

typedef struct pbrt_nrc_optim {
  float lr, beta1, beta2, eps;
  /* Inference reads an exponential moving average of the parameters, avg = ema * avg + (1 - ema) *
   * params after every step, starting from the parameters when set; 0 = the parameters. */
  float ema;
} pbrt_nrc_optim;

/* Seeded He-uniform hidden weights, zero biases and output layer (an untrained cache answers 0),
 * Adam lr 1e-3, betas 0.9 / 0.99, eps 1e-8 (override with pbrt_nrc_set_optim); scene box [0, 1]^3
 * until pbrt_nrc_set_bounds. */
int pbrt_nrc_create(pbrt_ctx* ctx, uint64_t seed, pbrt_nrc** out);
int pbrt_nrc_destroy(pbrt_nrc* nrc);
int pbrt_nrc_set_bounds(pbrt_nrc* nrc, const float lo[3], const float hi[3]);
int pbrt_nrc_set_optim(pbrt_nrc* nrc, const pbrt_nrc_optim* optim);
/* FP32 master parameters in flat order (n must be PBRT_NRC_PARAMS); setting them also resets the
 * Adam moments and the step count. */
int pbrt_nrc_set_params(pbrt_nrc* nrc, const float* params, uint32_t n);
int pbrt_nrc_get_params(pbrt_nrc* nrc, float* params, uint32_t n);
/* raw_dev: n x 12 floats -> x_dev: n x PBRT_NRC_FEATURES BF16 (uint16 bit patterns). */
int pbrt_nrc_encode(pbrt_nrc* nrc, const float* raw_dev, uint32_t n, uint16_t* x_dev, void* stream);
/* x_dev (encoded) -> out_dev: n x 3 FP32 radiance. */
int pbrt_nrc_infer(pbrt_nrc* nrc, const uint16_t* x_dev, uint32_t n, float* out_dev, void* stream);
/* raw_dev (n x 12 floats, rows 16-byte aligned) -> out_dev (n x 3 FP32): encode and infer in one
 * kernel, the features computed in registers (equal to pbrt_nrc_encode followed by pbrt_nrc_infer). */
int pbrt_nrc_query(pbrt_nrc* nrc, const float* raw_dev, uint32_t n, float* out_dev, void* stream);
/* Loss (relative L2: mean over samples and channels of (y - t)^2 / (y^2 + 0.01), y held constant
 * in the denominator) and its gradient in flat order, without an update. grads may be NULL. */
int pbrt_nrc_gradients(pbrt_nrc* nrc, const uint16_t* x_dev, const float* target_dev, uint32_t n, float* grads,
                       float* loss);
/* One optimisation step on a batch: gradient, Adam update, refreshed BF16 weights; all on the
 * stream, no host synchronisation. The batch loss lands in loss_dev (one float) if not NULL. */
int pbrt_nrc_train_step(pbrt_nrc* nrc, const uint16_t* x_dev, const float* target_dev, uint32_t n, float* loss_dev,
                        void* stream);
/* Adam steps taken since creation or the last set_params. */
int pbrt_nrc_step_count(pbrt_nrc* nrc, uint64_t* steps);
/* One 16 x 16 x 32 BF16 matrix-core product D = A B (A 16 x 32 row-major, B 32 x 16 row-major,
 * host BF16 bit patterns in, FP32 out): the operand layout the network's kernels assume. */
int pbrt_nrc_mfma_check(pbrt_ctx* ctx, const uint16_t* a, const uint16_t* b, float* d);

/* ---- rendering with the radiance cache, and progressive convergence --------------------------- */
typedef struct pbrt_nrc_render {
  /* A path that is not a training path ends into the cache at its first non-specular,
   * non-emissive vertex at or past query_vertex (3: after two bounces); the cache's answer times
   * the path throughput replaces everything the path would have gathered from there on. */
  uint32_t query_vertex;
  /* One path in train_stride of every pass is a training path (33: 3 %), traced to max_depth
   * without the cache; 0 = no training. Its vertices 1 .. query_vertex give training samples whose
   * target is the radiance gathered over the next max_depth - query_vertex vertices divided by the
   * path throughput at the vertex (the same window the cache stands for at query_vertex). */
  uint32_t train_stride;
  uint32_t train_batch;  /* samples per optimisation step, at most (65536) */
  uint32_t train_steps;  /* optimisation steps per frame (1); a pass of F samples per pixel holds F
                          * frames and trains F batches, each drawn evenly from its frame's slots */
  /* Optional host output of the last pass's training slots, for inspection: slot s = rank *
   * query_vertex + (v - 1); the first debug_slots * 12 floats are the raw cache inputs, the next
   * debug_slots * 4 the target rgb and a validity flag. debug_slots = 0 disables it. */
  float* debug;
  uint32_t debug_slots;
  uint32_t reserved;
} pbrt_nrc_render;

typedef struct pbrt_progress {
  const double* reference; /* host width*height*3 film to measure against (NULL: no measure) */
  double* rmse;            /* host, per pass: RMSE over pixels and channels of the running mean
                            * against the reference, reduced in FP64 on the device */
  double* ms;              /* host, per pass: milliseconds from the first launch to the pass's end
                            * (HIP events on the render stream) */
  uint32_t capacity;       /* entries available in rmse / ms */
  uint32_t count;          /* out: passes recorded */
} pbrt_progress;

/* pbrt_render with the radiance cache (nrc != NULL) and/or a per-pass convergence record
 * (progress != NULL). With the cache, every sample of every pixel is a frame; a pass holds
 * config->samples_per_pass frames (0: 1) traced together, then trains the cache once per frame
 * (frames later in a pass query a cache up to samples_per_pass - 1 steps older). The cache's scene
 * box is set to the scene's bounds. With nrc == NULL this is pbrt_render. Passes run side by side
 * on the render stream and streams the call creates (plain: two, consecutive passes on either; with
 * the cache: a second stream for every other pass's first bounces and a third for the passes' last
 * bounces, queries, training and film), with as many path sets; the film, the convergence record
 * and the cache's training take the passes in order, so results are those of running the passes
 * one after the other, and the call returns once every stream is done. */
int pbrt_render_ex(pbrt_scene* scene, const pbrt_render_config* config, pbrt_nrc* nrc, const pbrt_nrc_render* nrc_cfg,
                   pbrt_progress* progress, double* film_out, pbrt_render_stats* stats);

/* ---- synthetic inputs and ray generation (host, f64 as the reference, rays rounded to FP32) ----
 * pinhole camera of src/cameras/pin_hole.rs:17-50 looking from `eye` at `look` (up = +y, fov in
 * radians); writes width*height rays through pixel centres (px + 0.5, py + 0.5), row-major. */
int pbrt_pinhole_rays(uint32_t width, uint32_t height, double fov, const double eye[3], const double look[3],
                      float tmin, float tmax, pbrt_ray* out);
/* Camera `view` (0..5) of the bvh_stats orbit framing a vertex set (src/bin/bvh_stats.rs:345-401). */
int pbrt_orbit_rays(const float* xyz, uint32_t nv, uint32_t view, uint32_t side, double fov, float tmin, float tmax,
                    pbrt_ray* out);
/* n seeded rays, origins uniform in a ball, directions uniform on the sphere (src/pdfs/sphere.rs:14-23);
 * ray k uses counter first + k. */
int pbrt_random_rays(const double center[3], double radius, uint64_t seed, uint64_t first, uint32_t n, float tmin,
                     float tmax, pbrt_ray* out);
/* Seeded icosphere of `subdiv` subdivisions (10*4^s+2 vertices, 20*4^s triangles) wound so the reference's
 * mesh normal (triangle_mesh.rs:119) points outward; radial displacement: noise_kind 0 none,
 * 1 per-vertex uniform +-amp*radius, 2 fBm value noise (octaves, base frequency freq) * amp*radius.
 * With null outputs only the sizes are written. */
int pbrt_gen_icosphere(uint32_t subdiv, double radius, const double center[3], int noise_kind, double amp, int octaves,
                       double freq, uint64_t seed, float* xyz_out, uint32_t* idx_out, uint32_t* nv_out,
                       uint32_t* nt_out);

#ifdef __cplusplus
}
#endif
#endif /* PBRT_GPU_H */
