// C ABI of the ray path: context and device plumbing, mesh create/trace. See include/pbrt_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "bvh_gpu.hpp"
#include "common.hpp"
#include "context.hpp"
#include "traverse.hpp"

namespace pbrt {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

}  // namespace pbrt

using namespace pbrt;

// The spill area of pbrt_mesh_trace_closest_noalloc: traversal stack entries past the LDS stack,
// 64 MiB per device, held by every device the library runs on until the process ends. Host code
// never uses the array's name as an address: each device's copy is found once, at context creation
// with the device current, by hipGetSymbolAddress, and that address is what the kernels receive.
// Nothing in it needs to start at zero (every entry is written before it is read). All no-alloc
// calls on one device share it: calls that overlap on one device race.
constexpr uint64_t kNoallocSpillBytes = 64ull << 20;
constexpr uint64_t kNoallocSpillEntries = kNoallocSpillBytes / sizeof(uint2);
static __device__ __align__(256) uint2 scratch[kNoallocSpillEntries];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};


struct pbrt_mesh {
  pbrt_ctx* ctx = nullptr;
  uint32_t nt = 0;
  pbrt_build_stats stats{};
  BuiltBVH host;            // kept for export_tree
  std::vector<uint32_t> idx;
  DevNode* nodes = nullptr;
  DevTri* tris = nullptr;
  uint64_t n_nodes = 0;
  DevNode4* nodes4 = nullptr;  // the 4-wide tree, collapsed on first use
  uint64_t n_nodes4 = 0;
  int32_t root4 = 0;
  uint32_t depth4 = 0;         // interior levels of the 4-wide tree
  RootDesc root{};
  // HBM overflow areas of the traversal stacks, one per stream the mesh has traced on (up to
  // kSpillAreas): launches on different streams run at once, each on its own area. Past that many
  // streams an area changes hands, and its new stream waits for the end of the old one's last
  // launch on it (event `done`); launches on one stream are ordered by the stream.
  struct SpillArea {
    hipStream_t stream = nullptr;
    uint2* p = nullptr;
    uint64_t entries = 0;
    hipEvent_t done = nullptr;
  };
  static constexpr int kSpillAreas = 4;
  SpillArea spill[kSpillAreas];
  int spill_next = 0;  // the area to hand over next once all are taken
  int* status = nullptr;
  pbrt_launch launch{128, 16, 16, 0, 2, 0};
  uint32_t tile_w = 0;  // pbrt_mesh_set_ray_tiles
  // Guards every field above that a call reads or writes after creation (launch settings, the
  // 4-wide tree, the spill areas and their streams).
  mutable std::mutex mu;
};

namespace {

void check_ctx(pbrt_ctx* c) {
  if (!c) throw Error(PBRT_ERR_INVALID_ARG, "null context");
  PBRT_HIP(hipSetDevice(c->device));
}

hipStream_t stream_of(pbrt_ctx* c, void* s) { return s ? (hipStream_t)s : c->stream; }

LaunchArgs base_args(pbrt_mesh* m, uint32_t n, void* stream) {
  LaunchArgs a{};
  a.nodes = m->nodes;
  a.tris = m->tris;
  a.root = m->root;
  a.n = n;
  a.status = m->status;
  a.stream = stream_of(m->ctx, stream);
  a.block = m->launch.block;
  a.lds_stack = m->launch.lds_stack;
  a.waves_per_cu = m->launch.waves_per_cu;
  a.refill = m->launch.refill;
  a.width = m->launch.width == 4 ? 4 : 2;
  a.nodes4 = m->nodes4;
  a.root4 = m->root4;
  a.packet = m->launch.packet;
  a.tile_w = m->tile_w;
  if (m->tile_w && n % (16ull * m->tile_w))
    throw Error(PBRT_ERR_INVALID_ARG, "ray count not a multiple of sixteen image rows (pbrt_mesh_set_ray_tiles)");
  a.cu_count = (uint32_t)m->ctx->cu_count;
  return a;
}

// The 4-wide tree, collapsed on the host from the device's binary nodes (either builder's).
void ensure_bvh4(pbrt_mesh* m) {
  if (m->nodes4 || !m->ctx) return;
  std::vector<DevNode> n2(m->n_nodes);
  if (m->n_nodes) PBRT_HIP(hipMemcpy(n2.data(), m->nodes, m->n_nodes * sizeof(DevNode), hipMemcpyDeviceToHost));
  std::vector<DevNode4> n4;
  int32_t root4 = 0;
  collapse_bvh4(n2, m->root.ref, &n4, &root4);
  // the kernels address 4-wide nodes by 32-bit byte offsets
  if ((uint64_t)n4.size() * sizeof(DevNode4) > 0xffffffffull)
    throw Error(PBRT_ERR_LAYOUT, "4-wide tree above 4 GiB (use the binary tree, launch width 2)");
  DevNode4* d = nullptr;
  PBRT_HIP(hipMalloc(&d, std::max<size_t>(1, n4.size()) * sizeof(DevNode4)));
  if (!n4.empty()) {
    const hipError_t e = hipMemcpy(d, n4.data(), n4.size() * sizeof(DevNode4), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(d);
      PBRT_HIP(e);
    }
  }
  m->nodes4 = d;
  m->n_nodes4 = n4.size();
  m->root4 = root4;
  // interior levels of the 4-wide tree (iterative walk; a leaf root has none)
  uint32_t depth = 0;
  std::vector<std::pair<int32_t, uint32_t>> todo;
  if (root4 >= 0) todo.push_back({root4, 1});
  while (!todo.empty()) {
    const auto [ref, level] = todo.back();
    todo.pop_back();
    depth = std::max(depth, level);
    const DevNode4& nd = n4[ref];
    for (int c = 0; c < nd.count[0]; ++c)
      if (nd.ref[c] >= 0) todo.push_back({nd.ref[c], level + 1});
  }
  m->depth4 = depth;
}

// The spill area of stream s, sized for the launch shape (the caller holds m->mu): the stream's own,
// or a free one, or the next one handed over (s then waits for its old stream's last launch on it).
pbrt_mesh::SpillArea& prepare_spill(pbrt_mesh* m, LaunchArgs* a) {
  const uint32_t threads = traversal_grid_threads(*a);
  if (threads == 0) throw Error(PBRT_ERR_INVALID_ARG, "unsupported block size");
  // pending entries: at most depth - 1 on the binary tree, 3 per interior level on the 4-wide one
  const uint64_t depth = a->width == 4 ? 3ull * m->depth4 : m->stats.max_depth;
  const uint32_t per_thread = depth + 1 > a->lds_stack ? (uint32_t)(depth + 1 - a->lds_stack) : 1;
  const uint64_t need = (uint64_t)per_thread * threads;
  const hipStream_t s = (hipStream_t)a->stream;
  int k = -1;
  for (int i = 0; i < pbrt_mesh::kSpillAreas && k < 0; ++i)
    if (m->spill[i].done && m->spill[i].stream == s) k = i;
  for (int i = 0; i < pbrt_mesh::kSpillAreas && k < 0; ++i)
    if (!m->spill[i].done) k = i;
  if (k < 0) {
    k = m->spill_next;
    m->spill_next = (m->spill_next + 1) % pbrt_mesh::kSpillAreas;
    PBRT_HIP(hipStreamWaitEvent(s, m->spill[k].done, 0));
  }
  pbrt_mesh::SpillArea& area = m->spill[k];
  if (!area.done) PBRT_HIP(hipEventCreateWithFlags(&area.done, hipEventDisableTiming));
  area.stream = s;
  if (need > area.entries) {
    // hipFree waits for the device, so no launch still reads the old area
    if (area.p) PBRT_HIP(hipFree(area.p));
    area.p = nullptr;
    area.entries = 0;
    PBRT_HIP(hipMalloc(&area.p, need * sizeof(uint2)));
    area.entries = need;
  }
  a->spill = area.p;
  a->spill_entries_per_thread = per_thread;
  a->spill_capacity_entries = area.entries;
  return area;
}

// One traversal launch (the caller holds m->mu).
void run(pbrt_mesh* m, LaunchArgs& a, int kind) {
  if (a.n == 0) return;
  pbrt_mesh::SpillArea& area = prepare_spill(m, &a);
  PBRT_HIP(launch_traversal(a, kind));
  PBRT_HIP(hipEventRecord(area.done, (hipStream_t)a.stream));
}

}  // namespace

extern "C" {

const char* pbrt_last_error(void) { return g_last_error.c_str(); }
int pbrt_version(void) { return 1; }

int pbrt_ctx_create(int device, pbrt_ctx** out) {
  return guarded([&] {
    if (!out) throw Error(PBRT_ERR_INVALID_ARG, "null output");
    int n = 0;
    PBRT_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw Error(PBRT_ERR_INVALID_ARG, "no such device");
    if (device >= kMaxDevices) throw Error(PBRT_ERR_INVALID_ARG, "device index above 63");
    auto c = std::make_unique<pbrt_ctx>();
    c->device = device;
    PBRT_HIP(hipSetDevice(device));
    {
      std::lock_guard<std::mutex> lock(g_scratch_mu);
      if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        PBRT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(PBRT_ERR_DEVICE, "no device address for the no-alloc spill array");
        g_scratch_ptr[device] = p;
      }
      c->scratch = g_scratch_ptr[device];
    }
    hipDeviceProp_t prop;
    PBRT_HIP(hipGetDeviceProperties(&prop, device));
    c->cu_count = prop.multiProcessorCount;
    c->name = prop.gcnArchName;
    PBRT_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    *out = c.release();
  });
}

int pbrt_ctx_destroy(pbrt_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int pbrt_ctx_stream(pbrt_ctx* ctx, void** s) {
  return guarded([&] {
    if (!ctx || !s) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    *s = ctx->stream;
  });
}

int pbrt_ctx_device_info(pbrt_ctx* ctx, int* cu, char* name, size_t len) {
  return guarded([&] {
    if (!ctx) throw Error(PBRT_ERR_INVALID_ARG, "null context");
    if (cu) *cu = ctx->cu_count;
    if (name && len) {
      std::strncpy(name, ctx->name.c_str(), len - 1);
      name[len - 1] = 0;
    }
  });
}

int pbrt_dev_alloc(pbrt_ctx* ctx, size_t bytes, void** p) {
  return guarded([&] {
    check_ctx(ctx);
    if (!p) throw Error(PBRT_ERR_INVALID_ARG, "null output");
    PBRT_HIP(hipMalloc(p, bytes ? bytes : 1));
  });
}

int pbrt_dev_free(pbrt_ctx* ctx, void* p) {
  return guarded([&] {
    check_ctx(ctx);
    if (p) PBRT_HIP(hipFree(p));
  });
}

int pbrt_memcpy_h2d(pbrt_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    check_ctx(ctx);
    if (bytes) PBRT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  });
}

int pbrt_memcpy_d2h(pbrt_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    check_ctx(ctx);
    PBRT_HIP(hipStreamSynchronize(ctx->stream));
    if (bytes) PBRT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  });
}

int pbrt_memset_dev(pbrt_ctx* ctx, void* dst, int value, size_t bytes) {
  return guarded([&] {
    check_ctx(ctx);
    if (bytes) PBRT_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
  });
}

int pbrt_dev_copy(pbrt_ctx* ctx, void* dst, const void* src, size_t bytes, void* s) {
  return guarded([&] {
    check_ctx(ctx);
    if (bytes % 16) throw Error(PBRT_ERR_INVALID_ARG, "copy size must be a multiple of 16 bytes");
    PBRT_HIP(launch_copy(src, dst, bytes, stream_of(ctx, s)));
  });
}

int pbrt_stream_create(pbrt_ctx* ctx, void** s) {
  return guarded([&] {
    check_ctx(ctx);
    if (!s) throw Error(PBRT_ERR_INVALID_ARG, "null output");
    hipStream_t st = nullptr;
    PBRT_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *s = st;
  });
}

int pbrt_stream_destroy(pbrt_ctx* ctx, void* s) {
  return guarded([&] {
    check_ctx(ctx);
    if (s) PBRT_HIP(hipStreamDestroy((hipStream_t)s));
  });
}

int pbrt_stream_sync(pbrt_ctx* ctx, void* s) {
  return guarded([&] {
    check_ctx(ctx);
    PBRT_HIP(hipStreamSynchronize(stream_of(ctx, s)));
  });
}

int pbrt_event_create(pbrt_ctx* ctx, void** ev) {
  return guarded([&] {
    check_ctx(ctx);
    hipEvent_t e;
    PBRT_HIP(hipEventCreate(&e));
    *ev = e;
  });
}

int pbrt_event_record(void* ev, void* s) {
  return guarded([&] { PBRT_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)s)); });
}

int pbrt_stream_wait_event(pbrt_ctx* ctx, void* s, void* ev) {
  return guarded([&] {
    check_ctx(ctx);
    if (!ev) throw Error(PBRT_ERR_INVALID_ARG, "null event");
    PBRT_HIP(hipStreamWaitEvent(stream_of(ctx, s), (hipEvent_t)ev, 0));
  });
}

int pbrt_event_elapsed_ms(void* a, void* b, float* ms) {
  return guarded([&] {
    PBRT_HIP(hipEventSynchronize((hipEvent_t)b));
    PBRT_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
  });
}

int pbrt_event_destroy(void* ev) {
  return guarded([&] { PBRT_HIP(hipEventDestroy((hipEvent_t)ev)); });
}

int pbrt_mesh_create(pbrt_ctx* ctx, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                     const pbrt_bvh_params* params, pbrt_mesh** out) {
  return guarded([&] {
    if (ctx) check_ctx(ctx);
    if (!out) throw Error(PBRT_ERR_INVALID_ARG, "null output");
    if (nt == 0) throw Error(PBRT_ERR_EMPTY_MESH, "a TriangleMesh needs at least one triangle");
    if (!xyz || !idx) throw Error(PBRT_ERR_INVALID_ARG, "null vertex or index array");
    pbrt_bvh_params p = params ? *params : pbrt_bvh_params{12, 4, 1.0};
    if (p.max_leaf > kMaxLeafTris) throw Error(PBRT_ERR_INVALID_ARG, "max_leaf above the device bound of 32");
    auto m = std::make_unique<pbrt_mesh>();
    m->ctx = ctx;
    m->nt = nt;
    m->host = build_mesh_bvh(xyz, nv, idx, nt, p);
    m->stats = m->host.stats;
    m->idx.assign(idx, idx + 3 * (size_t)nt);
    std::vector<DevNode> nodes;
    std::vector<DevTri> tris;
    flatten_mesh_bvh(m->host, xyz, idx, &nodes, &tris, &m->root);
    m->n_nodes = nodes.size();
    if (!ctx) {  // host-only mesh: built and flattened for inspection, never traced
      *out = m.release();
      return;
    }
    PBRT_HIP(hipMalloc(&m->nodes, std::max<size_t>(1, nodes.size()) * sizeof(DevNode)));
    PBRT_HIP(hipMalloc(&m->tris, tris.size() * sizeof(DevTri)));
    PBRT_HIP(hipMalloc(&m->status, sizeof(int)));
    if (!nodes.empty()) PBRT_HIP(hipMemcpy(m->nodes, nodes.data(), nodes.size() * sizeof(DevNode), hipMemcpyHostToDevice));
    PBRT_HIP(hipMemcpy(m->tris, tris.data(), tris.size() * sizeof(DevTri), hipMemcpyHostToDevice));
    PBRT_HIP(hipMemset(m->status, 0, sizeof(int)));
    *out = m.release();
  });
}

int pbrt_mesh_create_device(pbrt_ctx* ctx, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                            const pbrt_bvh_params* params, pbrt_mesh** out) {
  return guarded([&] {
    check_ctx(ctx);
    if (!out) throw Error(PBRT_ERR_INVALID_ARG, "null output");
    if (nt == 0) throw Error(PBRT_ERR_EMPTY_MESH, "a TriangleMesh needs at least one triangle");
    if (!xyz || !idx) throw Error(PBRT_ERR_INVALID_ARG, "null vertex or index array");
    pbrt_bvh_params p = params ? *params : pbrt_bvh_params{12, 4, 1.0};
    if (p.max_leaf > kMaxLeafTris) throw Error(PBRT_ERR_INVALID_ARG, "max_leaf above the device bound of 32");
    auto m = std::make_unique<pbrt_mesh>();
    m->ctx = ctx;
    m->nt = nt;
    m->idx.assign(idx, idx + 3 * (size_t)nt);
    float* xyz_d = nullptr;
    uint32_t* idx_d = nullptr;
    GpuBuilt g;
    try {
      PBRT_HIP(hipMalloc(&xyz_d, std::max<size_t>(1, 3 * (size_t)nv) * sizeof(float)));
      PBRT_HIP(hipMalloc(&idx_d, 3 * (size_t)nt * sizeof(uint32_t)));
      PBRT_HIP(hipMemcpyAsync(xyz_d, xyz, 3 * (size_t)nv * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
      PBRT_HIP(hipMemcpyAsync(idx_d, idx, 3 * (size_t)nt * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
      gpu_build_mesh_bvh(ctx->stream, xyz_d, nv, idx_d, nt, p, &g);
    } catch (...) {
      (void)hipFree(xyz_d);
      (void)hipFree(idx_d);
      throw;
    }
    (void)hipFree(xyz_d);
    (void)hipFree(idx_d);
    m->nodes = g.nodes;
    m->tris = g.tris;
    m->n_nodes = g.n_nodes;
    m->root = g.root;
    m->stats = g.stats;
    PBRT_HIP(hipMalloc(&m->status, sizeof(int)));
    PBRT_HIP(hipMemset(m->status, 0, sizeof(int)));
    *out = m.release();
  });
}

int pbrt_mesh_copy_layout(const pbrt_mesh* m, void* nodes_out, void* tris_out, float* root_box6, int32_t* root_ref) {
  return guarded([&] {
    if (!m) throw Error(PBRT_ERR_INVALID_ARG, "null mesh");
    if (!m->ctx) throw Error(PBRT_ERR_INVALID_ARG, "a host-only mesh has no device layout");
    PBRT_HIP(hipSetDevice(m->ctx->device));
    PBRT_HIP(hipStreamSynchronize(m->ctx->stream));
    if (nodes_out && m->n_nodes)
      PBRT_HIP(hipMemcpy(nodes_out, m->nodes, m->n_nodes * sizeof(DevNode), hipMemcpyDeviceToHost));
    if (tris_out) PBRT_HIP(hipMemcpy(tris_out, m->tris, (size_t)m->nt * sizeof(DevTri), hipMemcpyDeviceToHost));
    if (root_box6)
      for (int k = 0; k < 3; ++k) {
        root_box6[k] = m->root.lo[k];
        root_box6[3 + k] = m->root.hi[k];
      }
    if (root_ref) *root_ref = m->root.ref;
  });
}

int pbrt_mesh_destroy(pbrt_mesh* m) {
  return guarded([&] {
    if (!m) return;
    if (m->ctx) (void)hipSetDevice(m->ctx->device);
    (void)hipFree(m->nodes);
    (void)hipFree(m->nodes4);
    (void)hipFree(m->tris);
    for (auto& area : m->spill) {
      (void)hipFree(area.p);
      if (area.done) (void)hipEventDestroy(area.done);
    }
    (void)hipFree(m->status);
    delete m;
  });
}

int pbrt_mesh_build_stats(const pbrt_mesh* m, pbrt_build_stats* out) {
  return guarded([&] {
    if (!m || !out) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    *out = m->stats;
  });
}

int pbrt_mesh_triangle_count(const pbrt_mesh* m, uint32_t* out) {
  return guarded([&] {
    if (!m || !out) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    *out = m->nt;
  });
}

int pbrt_mesh_device_bytes(const pbrt_mesh* m, uint64_t* nb, uint64_t* tb) {
  return guarded([&] {
    if (!m) throw Error(PBRT_ERR_INVALID_ARG, "null mesh");
    if (nb) *nb = m->n_nodes * sizeof(DevNode);
    if (tb) *tb = (uint64_t)m->nt * sizeof(DevTri);
  });
}

int pbrt_mesh_export_tree(const pbrt_mesh* m, double* boxes, uint8_t* is_leaf, uint32_t* left, uint32_t* right,
                          uint32_t* first, uint32_t* count, uint32_t* leaf_tris) {
  return guarded([&] {
    if (!m) throw Error(PBRT_ERR_INVALID_ARG, "null mesh");
    const auto& b = m->host;
    if (b.nodes.empty()) throw Error(PBRT_ERR_INVALID_ARG, "the tree was built on the device: no host tree to export");
    for (size_t i = 0; i < b.nodes.size(); ++i) {
      const BuildNode& n = b.nodes[i];
      for (int k = 0; k < 3; ++k) {
        boxes[6 * i + k] = n.lo[k];
        boxes[6 * i + 3 + k] = n.hi[k];
      }
      is_leaf[i] = n.leaf ? 1 : 0;
      left[i] = n.leaf ? 0 : n.left;
      right[i] = n.leaf ? 0 : n.right;
      first[i] = n.leaf ? n.first : 0;
      count[i] = n.leaf ? n.count : 0;
    }
    for (size_t i = 0; i < b.order.size(); ++i) leaf_tris[i] = b.order[i];
  });
}

int pbrt_mesh_set_ray_tiles(pbrt_mesh* m, uint32_t image_width) {
  return guarded([&] {
    if (!m) throw Error(PBRT_ERR_INVALID_ARG, "null mesh");
    if (image_width % 16) throw Error(PBRT_ERR_INVALID_ARG, "image width must be a multiple of 16");
    // sixteen rows of the image must fit a batch of at most 2^32 - 1 rays
    if (image_width >= (1u << 28)) throw Error(PBRT_ERR_INVALID_ARG, "image width too large (at most 2^28 - 16)");
    std::lock_guard<std::mutex> lock(m->mu);
    m->tile_w = image_width;
  });
}

int pbrt_mesh_set_launch(pbrt_mesh* m, const pbrt_launch* l) {
  return guarded([&] {
    if (!m || !l) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    pbrt_launch v = *l;
    if (!v.block) v.block = 128;
    if (!v.waves_per_cu) v.waves_per_cu = 16;
    if (!v.lds_stack) v.lds_stack = 16;
    if (v.block != 64 && v.block != 128 && v.block != 256) throw Error(PBRT_ERR_INVALID_ARG, "block must be 64, 128 or 256");
    if (v.lds_stack != 8 && v.lds_stack != 12 && v.lds_stack != 16 && v.lds_stack != 24 && v.lds_stack != 32)
      throw Error(PBRT_ERR_INVALID_ARG, "lds_stack must be 8, 12, 16, 24 or 32");
    if (v.waves_per_cu > 128) throw Error(PBRT_ERR_INVALID_ARG, "waves_per_cu above 128");
    if (v.refill > 64) throw Error(PBRT_ERR_INVALID_ARG, "refill above 64 lanes");
    if (!v.width) v.width = 2;
    if (v.width != 2 && v.width != 4) throw Error(PBRT_ERR_INVALID_ARG, "width must be 2 or 4");
    if (v.packet > 1) throw Error(PBRT_ERR_INVALID_ARG, "packet must be 0 or 1");
    if (v.packet && v.width != 2) throw Error(PBRT_ERR_INVALID_ARG, "packet traversal walks the binary tree (width 2)");
    if (v.packet && m->stats.max_depth > kPacketStackDepth + 1)
      throw Error(PBRT_ERR_INVALID_ARG, "tree too deep for packet traversal");
    std::lock_guard<std::mutex> lock(m->mu);
    if (v.width == 4) {
      if (!m->ctx) throw Error(PBRT_ERR_INVALID_ARG, "mesh was built without a device context");
      check_ctx(m->ctx);
      ensure_bvh4(m);
    }
    m->launch = v;
  });
}

static int trace_common(pbrt_mesh* m, const pbrt_ray* rays, uint32_t n, pbrt_hit* hits, uint8_t* occ, uint32_t* stats,
                        void* stream, int kind) {
  return guarded([&] {
    if (!m) throw Error(PBRT_ERR_INVALID_ARG, "null mesh");
    if (!m->ctx) throw Error(PBRT_ERR_INVALID_ARG, "mesh was built without a device context");
    check_ctx(m->ctx);
    if (n && !rays) throw Error(PBRT_ERR_INVALID_ARG, "null rays");
    if (n && kind != 2 && !hits) throw Error(PBRT_ERR_INVALID_ARG, "null hits");
    if (n && kind == 2 && !occ) throw Error(PBRT_ERR_INVALID_ARG, "null occlusion output");
    if (n && kind == 1 && !stats) throw Error(PBRT_ERR_INVALID_ARG, "null stats output");
    std::lock_guard<std::mutex> lock(m->mu);
    LaunchArgs a = base_args(m, n, stream);
    a.rays = rays;
    a.hits = hits;
    a.occluded = occ;
    a.stats = stats;
    run(m, a, kind);
  });
}

int pbrt_mesh_trace_closest(pbrt_mesh* m, const pbrt_ray* rays, uint32_t n, pbrt_hit* hits, void* stream) {
  return trace_common(m, rays, n, hits, nullptr, nullptr, stream, 0);
}

int pbrt_mesh_trace_closest_stats(pbrt_mesh* m, const pbrt_ray* rays, uint32_t n, pbrt_hit* hits, uint32_t* stats,
                                  void* stream) {
  return trace_common(m, rays, n, hits, nullptr, stats, stream, 1);
}

int pbrt_mesh_trace_any(pbrt_mesh* m, const pbrt_ray* rays, uint32_t n, uint8_t* occ, void* stream) {
  return trace_common(m, rays, n, nullptr, occ, nullptr, stream, 2);
}

int pbrt_mesh_trace_closest_noalloc(pbrt_mesh* m, const pbrt_ray* rays, uint32_t n, pbrt_hit* hits, void* stream) {
  return guarded([&] {
    if (!m) throw Error(PBRT_ERR_INVALID_ARG, "null mesh");
    if (!m->ctx) throw Error(PBRT_ERR_INVALID_ARG, "mesh was built without a device context");
    check_ctx(m->ctx);
    if (n && (!rays || !hits)) throw Error(PBRT_ERR_INVALID_ARG, "null rays or hits");
    if (!m->ctx->scratch) throw Error(PBRT_ERR_INTERNAL, "the context has no spill array address");
    std::lock_guard<std::mutex> lock(m->mu);
    LaunchArgs a = base_args(m, n, stream);
    if (n == 0) return;
    if (a.width == 4 && !m->nodes4)
      throw Error(PBRT_ERR_INVALID_ARG, "the 4-wide tree is built by pbrt_mesh_set_launch; set the launch first");
    a.rays = rays;
    a.hits = hits;
    if (!a.packet) {
      const uint64_t depth = a.width == 4 ? 3ull * m->depth4 : m->stats.max_depth;
      const uint64_t per_thread = depth + 1 > a.lds_stack ? depth + 1 - a.lds_stack : 1;
      if (per_thread * a.block > kNoallocSpillEntries)
        throw Error(PBRT_ERR_INVALID_ARG,
                    "tree too deep for the no-alloc spill array: " + std::to_string(per_thread * a.block * sizeof(uint2)) +
                        " bytes per workgroup needed, " + std::to_string(kNoallocSpillBytes) + " available");
      // the persistent grid shrinks to the threads whose spill the array holds
      a.max_threads = (uint32_t)std::min<uint64_t>(kNoallocSpillEntries / per_thread, 1u << 30);
      a.spill = (uint2*)m->ctx->scratch;
      a.spill_entries_per_thread = (uint32_t)per_thread;
      a.spill_capacity_entries = kNoallocSpillEntries;
    }
    PBRT_HIP(launch_traversal(a, 0));
  });
}

int pbrt_mesh_trace_work(pbrt_mesh* m, const pbrt_ray* rays, uint32_t n, void* out_dev, int any_hit, uint32_t* work,
                         void* stream) {
  return guarded([&] {
    if (!m) throw Error(PBRT_ERR_INVALID_ARG, "null mesh");
    if (!m->ctx) throw Error(PBRT_ERR_INVALID_ARG, "mesh was built without a device context");
    check_ctx(m->ctx);
    if (n && (!rays || !out_dev || !work)) throw Error(PBRT_ERR_INVALID_ARG, "null rays, output or work array");
    std::lock_guard<std::mutex> lock(m->mu);
    if (m->launch.packet) throw Error(PBRT_ERR_INVALID_ARG, "the packet walk has no per-ray fetch counters");
    LaunchArgs a = base_args(m, n, stream);
    if (n == 0) return;
    a.rays = rays;
    a.hits = any_hit ? nullptr : (pbrt_hit*)out_dev;
    a.occluded = any_hit ? (uint8_t*)out_dev : nullptr;
    a.block = kWorkBlock;
    a.lds_stack = kWorkLdsStack;
    a.refill = 0;
    pbrt_mesh::SpillArea& area = prepare_spill(m, &a);
    PBRT_HIP(launch_traversal_work(a, any_hit ? 1 : 0, work));
    PBRT_HIP(hipEventRecord(area.done, (hipStream_t)a.stream));
  });
}

// Times every supported launch shape of one traversal kernel on the caller's rays and keeps the
// fastest: block {64, 128, 256} x LDS stack {8, 12, 16, 24, 32} x waves per CU in the grid
// {8, 12, 16, 24, 32, 64, 128} x refill {0, 1, 4} x tree width {2, 4}; each candidate is run
// `reps` times after one warm-up launch, scored by its fastest run. The hits/occlusion output
// buffer is overwritten.
int pbrt_mesh_autotune(pbrt_mesh* m, const pbrt_ray* rays, uint32_t n, void* out_dev, int any_hit, uint32_t reps,
                       pbrt_launch* chosen, float* best_ms) {
  return guarded([&] {
    if (!m || !rays || !out_dev || !n) throw Error(PBRT_ERR_INVALID_ARG, "null argument or empty batch");
    if (!m->ctx) throw Error(PBRT_ERR_INVALID_ARG, "mesh was built without a device context");
    check_ctx(m->ctx);
    if (reps == 0) reps = 2;
    std::lock_guard<std::mutex> lock(m->mu);
    hipEvent_t e0, e1;
    PBRT_HIP(hipEventCreate(&e0));
    PBRT_HIP(hipEventCreate(&e1));
    const uint32_t blocks[] = {64, 128, 256};
    const uint32_t depths[] = {8, 12, 16, 24, 32};
    const uint32_t waves[] = {8, 12, 16, 24, 32, 64, 128};
    const uint32_t refills[] = {0, 1, 4};
    const uint32_t widths[] = {2, 4};
    ensure_bvh4(m);
    pbrt_launch best{};
    float best_t = 1e30f;
    const int kind = any_hit ? 2 : 0;
    auto time_launch = [&](const pbrt_launch& L) {
      m->launch = L;
      LaunchArgs a = base_args(m, n, nullptr);
      a.rays = rays;
      a.hits = any_hit ? nullptr : (pbrt_hit*)out_dev;
      a.occluded = any_hit ? (uint8_t*)out_dev : nullptr;
      run(m, a, kind);  // warm-up
      float cand = 1e30f;
      for (uint32_t r = 0; r < reps; ++r) {
        PBRT_HIP(hipEventRecord(e0, a.stream ? (hipStream_t)a.stream : m->ctx->stream));
        run(m, a, kind);
        PBRT_HIP(hipEventRecord(e1, (hipStream_t)a.stream));
        PBRT_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PBRT_HIP(hipEventElapsedTime(&ms, e0, e1));
        cand = std::min(cand, ms);
      }
      if (cand < best_t) {
        best_t = cand;
        best = m->launch;
      }
    };
    try {
      for (uint32_t b : blocks)
        for (uint32_t d : depths)
          for (uint32_t w : waves)
            for (uint32_t f : refills)
              for (uint32_t wd : widths)
                if (w >= b / 64) time_launch(pbrt_launch{b, w, d, f, wd, 0});
      if (!any_hit && m->stats.max_depth <= kPacketStackDepth + 1)
        for (uint32_t w : waves) time_launch(pbrt_launch{64, w, 8, 0, 2, 1});
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    m->launch = best;
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = best_t;
  });
}

int pbrt_mesh_get_launch(const pbrt_mesh* m, pbrt_launch* out) {
  return guarded([&] {
    if (!m || !out) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lock(m->mu);
    *out = m->launch;
  });
}

int pbrt_mesh_status(pbrt_mesh* m, int* status_out) {
  return guarded([&] {
    if (!m || !status_out) throw Error(PBRT_ERR_INVALID_ARG, "null argument");
    if (!m->ctx) throw Error(PBRT_ERR_INVALID_ARG, "mesh was built without a device context");
    check_ctx(m->ctx);
    PBRT_HIP(hipDeviceSynchronize());
    PBRT_HIP(hipMemcpy(status_out, m->status, sizeof(int), hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
