// Mesh BVH built on the device: the binned surface-area heuristic of the reference's BVHTree
// (src/shapes/triangle_mesh/bvh.rs:240-309, 609-735), giving the tree the host builder gives
// (bvh_build.cpp) and its flattened arrays byte for byte.
//
// Top levels (nodes of more than kSubtreeMax triangles), one pass each over the level's triangles
// in chunks of kChunk per 256-thread workgroup: node and centroid bounds, then per-axis bins
// (LDS atomics on order-preserving integer keys, merged into the node's bins with global atomics),
// then one thread per node scans the bins (the host's scan order and arithmetic: f64 centroids,
// bin index floor((c - cmin) / w), half areas and costs in f64), then a stable partition
// (per-chunk counts, offsets, ballot ranks). The host reads each level's decisions to form the
// next level. Below kSubtreeMax, one wave builds a whole subtree with its triangle indices in LDS:
// the same three steps per node, wave-synchronous, smaller child first so the task stack stays
// within log2(kSubtreeMax) entries.
//
// Layout without a serial flattening: an interior node that splits [f, f + c) at m is recorded in
// slot m - 1 (split positions are distinct across a tree). Its preorder index is the number of
// interior nodes whose split position is at most f, plus the number of its ancestors whose left
// subtree holds it — a prefix sum over the used slots and a count carried down the build. Leaves
// are references (first, count) as in the host layout, their triangles sorted by index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bvh_gpu.hpp"

namespace pbrt {
namespace {

constexpr uint32_t kSubtreeMax = 256;
constexpr uint32_t kChunk = 4096;
constexpr int kTopBlock = 256;
constexpr int kStackMax = 24;

enum : int { kErrIndex = 1, kErrLeaf = 2, kErrStack = 4 };
enum : int32_t { kLeaf = 0, kSplit = 1, kHalve = 2 };

struct Task {
  uint32_t f, c, al, depth;  // segment of `order`, ancestors holding it on their left, depth (root 1)
  int32_t parent;            // slot of the parent (-1 - m: the root of mesh m)
  uint32_t side;             // 0: left child, 1: right child
};
struct Chunk {
  uint32_t node, start, len;
};
struct Decision {
  int32_t mode, axis;
  uint32_t boundary, lc;
  double cmin, w;
};
struct NodeBounds {
  uint32_t blo[3], bhi[3];
  unsigned long long clo[3], chi[3];
};
struct BuildSlot {  // an interior node, at slot (split position - 1)
  float lbox[6], rbox[6];
  int32_t ref[2];   // >= 0: slot of an interior child; < 0: leaf reference
  uint32_t f, al;
};
static_assert(sizeof(BuildSlot) == 64, "slot is 64 bytes");

struct Tris {  // per-triangle build inputs, structure of arrays
  const float* lo[3];
  const float* hi[3];
  const double* c[3];
};
struct TrisW {
  float* lo[3];
  float* hi[3];
  double* c[3];
};
struct Params {
  uint32_t K, max_leaf;
  double t_trav;
};
struct MeshDev {  // a mesh of the batch: its triangles at [tri_start, tri_start + nt) of the build order
  uint32_t tri_start, nt, vert_base, nv;
};

__device__ inline uint32_t mesh_of(const MeshDev* M, uint32_t nm, uint32_t pos) {
  uint32_t lo = 0, hi = nm - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) / 2;
    if (M[mid].tri_start <= pos)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// ---- order-preserving keys (no NaN inputs; -0 orders below +0) ------------------------------------
__device__ inline uint32_t fkey(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float kfloat(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
__device__ inline unsigned long long dkey(double d) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(d);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}
__device__ inline double kdouble(unsigned long long k) {
  return __longlong_as_double((long long)((k >> 63) ? (k & ~(1ull << 63)) : ~k));
}
constexpr uint32_t kKeyPosInf = 0xFF800000u;  // fkey(+inf): identity of min
constexpr uint32_t kKeyNegInf = 0x007FFFFFu;  // fkey(-inf): identity of max
constexpr unsigned long long kDKeyPosInf = 0xFFF0000000000000ull;
constexpr unsigned long long kDKeyNegInf = 0x000FFFFFFFFFFFFFull;

__device__ inline uint32_t bin_init(uint32_t field) {
  return field < 3 ? kKeyPosInf : field < 6 ? kKeyNegInf : 0u;
}

// bvh.rs:173-176 with Rust's saturating float -> usize cast (bvh_build.cpp bin_of)
__device__ inline uint32_t bin_of(double c, double cmin, double w, uint32_t K) {
  const double q = floor((c - cmin) / w);
  if (!(q > 0.0)) return 0;
  if (q >= (double)(K - 1)) return K - 1;
  return (uint32_t)q;
}

// aabound.rs:123-130: half the surface area, 0 for an empty box
__device__ inline double half_area(const double lo[3], const double hi[3]) {
  if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) return 0.0;
  const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  return dx * dy + dy * dz + dz * dx;
}

__device__ inline void grow_bin(double lo[3], double hi[3], const uint32_t* bin) {
  for (int a = 0; a < 3; ++a) {
    lo[a] = fmin(lo[a], (double)kfloat(bin[a]));
    hi[a] = fmax(hi[a], (double)kfloat(bin[3 + a]));
  }
}

// One axis of find_best_split_plane (bvh.rs:609-685): the first strictly cheapest boundary.
// B: the axis's K bins of 7 words (3 min keys, 3 max keys, count); ra / rc: K - 1 entries of scratch.
__device__ bool axis_best(const uint32_t* B, uint32_t K, double* ra, uint32_t* rc, double* cost_out, uint32_t* b_out) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t rs = 0;
  for (int m = (int)K - 2; m >= 0; --m) {
    grow_bin(lo, hi, B + (size_t)(m + 1) * 7);
    rs += B[(size_t)(m + 1) * 7 + 6];
    ra[m] = half_area(lo, hi);
    rc[m] = rs;
  }
  for (int a = 0; a < 3; ++a) {
    lo[a] = INFINITY;
    hi[a] = -INFINITY;
  }
  uint32_t ls = 0;
  bool found = false;
  double best = 0.0;
  for (uint32_t b = 0; b + 1 < K; ++b) {
    grow_bin(lo, hi, B + (size_t)b * 7);
    ls += B[(size_t)b * 7 + 6];
    if (ls == 0 || rc[b] == 0) continue;
    const double cost = half_area(lo, hi) * (double)ls + ra[b] * (double)rc[b];
    if (!found || cost < best) {
      best = cost;
      *b_out = b;
      found = true;
    }
  }
  *cost_out = best;
  return found;
}

// The node's choice (bvh.rs:240-309 as bvh_build.cpp Builder::build): split on the best plane when
// it is cheaper than a leaf or the leaf bound forces it; halve by index when no plane separates
// and the bound forces a split; else a leaf.
__device__ inline int32_t choose(bool found, double split_cost, double area, uint32_t c, const Params& P) {
  const double leaf_cost = area * (double)c;
  const bool over = P.max_leaf > 0 && c > P.max_leaf;
  if (found && (split_cost + P.t_trav * area < leaf_cost || over)) return kSplit;
  return over ? kHalve : kLeaf;
}

__device__ inline void union_bins(const uint32_t* B, uint32_t b0, uint32_t b1, float box[6]) {
  uint32_t lo[3] = {kKeyPosInf, kKeyPosInf, kKeyPosInf}, hi[3] = {kKeyNegInf, kKeyNegInf, kKeyNegInf};
  for (uint32_t b = b0; b < b1; ++b)
    for (int a = 0; a < 3; ++a) {
      lo[a] = min(lo[a], B[(size_t)b * 7 + a]);
      hi[a] = max(hi[a], B[(size_t)b * 7 + 3 + a]);
    }
  for (int a = 0; a < 3; ++a) {
    box[a] = kfloat(lo[a]);
    box[3 + a] = kfloat(hi[a]);
  }
}

__device__ inline void link(BuildSlot* slots, int32_t* root_ref, int32_t parent, uint32_t side, int32_t ref) {
  if (parent < 0)
    root_ref[-1 - parent] = ref;
  else
    slots[parent].ref[side] = ref;
}

__device__ inline int32_t dev_leaf_ref(uint32_t first, uint32_t count) {
  return ~(int32_t)((first << kLeafCountBits) | (count - 1));
}

// ---- wave reductions ------------------------------------------------------------------------------
__device__ inline float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ inline float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}
__device__ inline uint32_t wave_umin(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
  return v;
}
__device__ inline uint32_t wave_umax(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
  return v;
}
__device__ inline unsigned long long shfl_xor64(unsigned long long v, int o) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, o);
  const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), o);
  return ((unsigned long long)hi << 32) | lo;
}
__device__ inline unsigned long long wave_umin64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, shfl_xor64(v, o));
  return v;
}
__device__ inline unsigned long long wave_umax64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, shfl_xor64(v, o));
  return v;
}
__device__ inline uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o);
  return v;
}
__device__ inline uint64_t lanes_below() { return (1ull << __lane_id()) - 1ull; }

// ---- kernels ----------------------------------------------------------------------------------------
__global__ void k_prep(const float* xyz, const uint32_t* idx, uint32_t nt, const MeshDev* M, uint32_t nm, TrisW T,
                       uint32_t* order, int* err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const MeshDev md = M[mesh_of(M, nm, i)];
  uint32_t a = idx[3 * (size_t)i], b = idx[3 * (size_t)i + 1], c = idx[3 * (size_t)i + 2];
  if (a >= md.nv || b >= md.nv || c >= md.nv) {
    atomicOr(err, kErrIndex);
    a = b = c = 0;
  }
  a += md.vert_base;
  b += md.vert_base;
  c += md.vert_base;
  for (int k = 0; k < 3; ++k) {
    const float v0 = xyz[3 * (size_t)a + k], v1 = xyz[3 * (size_t)b + k], v2 = xyz[3 * (size_t)c + k];
    float lo = fminf(fminf(v0, v1), v2), hi = fmaxf(fmaxf(v0, v1), v2);
    if (lo == 0.0f) lo = 0.0f;
    if (hi == 0.0f) hi = 0.0f;
    double s = (double)v0;  // centroid: ((v0 + v1) + v2) * (1/3) in f64, bvh.rs:579-585
    s = s + (double)v1;
    s = s + (double)v2;
    T.lo[k][i] = lo;
    T.hi[k][i] = hi;
    T.c[k][i] = s * (1.0 / 3.0);
  }
  order[i] = i;
}

__global__ void k_level_init(uint32_t nl, uint32_t K, NodeBounds* nb, uint32_t* bins) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t per = 3 * K * 7;
  if (i < nl) {
    NodeBounds& b = nb[i];
    for (int a = 0; a < 3; ++a) {
      b.blo[a] = kKeyPosInf;
      b.bhi[a] = kKeyNegInf;
      b.clo[a] = kDKeyPosInf;
      b.chi[a] = kDKeyNegInf;
    }
  }
  for (size_t j = i; j < (size_t)nl * per; j += (size_t)gridDim.x * blockDim.x) bins[j] = bin_init((uint32_t)(j % 7));
}

__global__ __launch_bounds__(kTopBlock) void k_top_bounds(const Chunk* chunks, const uint32_t* order, Tris T,
                                                          NodeBounds* nb) {
  __shared__ uint32_t s_b[6];
  __shared__ unsigned long long s_c[6];
  const Chunk ch = chunks[blockIdx.x];
  const int tid = threadIdx.x;
  if (tid < 6) {
    s_b[tid] = tid < 3 ? kKeyPosInf : kKeyNegInf;
    s_c[tid] = tid < 3 ? kDKeyPosInf : kDKeyNegInf;
  }
  uint32_t blo[3] = {kKeyPosInf, kKeyPosInf, kKeyPosInf}, bhi[3] = {kKeyNegInf, kKeyNegInf, kKeyNegInf};
  unsigned long long clo[3] = {kDKeyPosInf, kDKeyPosInf, kDKeyPosInf}, chi[3] = {kDKeyNegInf, kDKeyNegInf, kDKeyNegInf};
  for (uint32_t j = tid; j < ch.len; j += kTopBlock) {
    const uint32_t t = order[ch.start + j];
    for (int a = 0; a < 3; ++a) {
      blo[a] = min(blo[a], fkey(T.lo[a][t]));
      bhi[a] = max(bhi[a], fkey(T.hi[a][t]));
      const unsigned long long k = dkey(T.c[a][t]);
      clo[a] = min(clo[a], k);
      chi[a] = max(chi[a], k);
    }
  }
  for (int a = 0; a < 3; ++a) {
    blo[a] = wave_umin(blo[a]);
    bhi[a] = wave_umax(bhi[a]);
    clo[a] = wave_umin64(clo[a]);
    chi[a] = wave_umax64(chi[a]);
  }
  __syncthreads();
  if (__lane_id() == 0)
    for (int a = 0; a < 3; ++a) {
      atomicMin(&s_b[a], blo[a]);
      atomicMax(&s_b[3 + a], bhi[a]);
      atomicMin(&s_c[a], clo[a]);
      atomicMax(&s_c[3 + a], chi[a]);
    }
  __syncthreads();
  NodeBounds& g = nb[ch.node];
  if (tid < 3) {
    atomicMin(&g.blo[tid], s_b[tid]);
    atomicMax(&g.bhi[tid], s_b[3 + tid]);
    atomicMin(&g.clo[tid], s_c[tid]);
    atomicMax(&g.chi[tid], s_c[3 + tid]);
  }
}

__global__ __launch_bounds__(kTopBlock) void k_top_bins(const Chunk* chunks, const uint32_t* order, Tris T,
                                                        const NodeBounds* nb, uint32_t K, uint32_t* bins) {
  extern __shared__ uint32_t s_bins[];  // 3 * K * 7
  const Chunk ch = chunks[blockIdx.x];
  const int tid = threadIdx.x;
  const uint32_t per = 3 * K * 7;
  for (uint32_t i = tid; i < per; i += kTopBlock) s_bins[i] = bin_init(i % 7);
  const NodeBounds b = nb[ch.node];
  double cmin[3], w[3];
  bool act[3];
  for (int a = 0; a < 3; ++a) {
    cmin[a] = kdouble(b.clo[a]);
    const double cmax = kdouble(b.chi[a]);
    act[a] = !(cmin[a] == cmax);
    w[a] = (cmax - cmin[a]) / (double)K;
  }
  __syncthreads();
  for (uint32_t j = tid; j < ch.len; j += kTopBlock) {
    const uint32_t t = order[ch.start + j];
    uint32_t lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = fkey(T.lo[a][t]);
      hi[a] = fkey(T.hi[a][t]);
    }
    for (int a = 0; a < 3; ++a) {
      if (!act[a]) continue;
      uint32_t* B = s_bins + (a * K + bin_of(T.c[a][t], cmin[a], w[a], K)) * 7;
      for (int k = 0; k < 3; ++k) {
        atomicMin(&B[k], lo[k]);
        atomicMax(&B[3 + k], hi[k]);
      }
      atomicAdd(&B[6], 1u);
    }
  }
  __syncthreads();
  uint32_t* G = bins + (size_t)ch.node * per;
  for (uint32_t i = tid; i < 3 * K; i += kTopBlock) {
    const uint32_t* B = s_bins + i * 7;
    if (B[6] == 0) continue;
    for (int k = 0; k < 3; ++k) {
      atomicMin(&G[i * 7 + k], B[k]);
      atomicMax(&G[i * 7 + 3 + k], B[3 + k]);
    }
    atomicAdd(&G[i * 7 + 6], B[6]);
  }
}

__global__ void k_top_decide(uint32_t nl, const Task* tasks, const NodeBounds* nb, const uint32_t* bins, Params P,
                             double* ra, uint32_t* rc, Decision* dec, BuildSlot* slots, uint32_t* used,
                             int32_t* root_ref) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nl) return;
  const uint32_t K = P.K;
  const Task t = tasks[j];
  const NodeBounds b = nb[j];
  const uint32_t* NB = bins + (size_t)j * 3 * K * 7;
  double lo[3], hi[3], cmin[3], w[3];
  int axis = -1;
  uint32_t bnd = 0;
  double best = 0.0;
  for (int a = 0; a < 3; ++a) {
    lo[a] = (double)kfloat(b.blo[a]);
    hi[a] = (double)kfloat(b.bhi[a]);
    cmin[a] = kdouble(b.clo[a]);
    const double cmax = kdouble(b.chi[a]);
    w[a] = (cmax - cmin[a]) / (double)K;
    if (cmin[a] == cmax) continue;
    double cost;
    uint32_t bb;
    if (axis_best(NB + (size_t)a * K * 7, K, ra + (size_t)j * K, rc + (size_t)j * K, &cost, &bb) &&
        (axis < 0 || cost < best)) {
      axis = a;
      bnd = bb;
      best = cost;
    }
  }
  Decision d{};
  d.mode = choose(axis >= 0, best, half_area(lo, hi), t.c, P);
  if (d.mode == kSplit) {
    const uint32_t* B = NB + (size_t)axis * K * 7;
    d.axis = axis;
    d.boundary = bnd;
    d.cmin = cmin[axis];
    d.w = w[axis];
    uint32_t lc = 0;
    for (uint32_t k = 0; k <= bnd; ++k) lc += B[(size_t)k * 7 + 6];
    d.lc = lc;
    const uint32_t s = t.f + lc - 1;
    BuildSlot& S = slots[s];
    union_bins(B, 0, bnd + 1, S.lbox);
    union_bins(B, bnd + 1, K, S.rbox);
    S.ref[0] = S.ref[1] = 0;
    S.f = t.f;
    S.al = t.al;
    used[s] = 1;
    link(slots, root_ref, t.parent, t.side, (int32_t)s);
  }
  dec[j] = d;
}

__global__ __launch_bounds__(kTopBlock) void k_top_count(const Chunk* chunks, const uint32_t* order, Tris T,
                                                         const Decision* dec, uint32_t K, uint32_t* chunk_left) {
  __shared__ uint32_t s_sum;
  const Chunk ch = chunks[blockIdx.x];
  const Decision d = dec[ch.node];
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  uint32_t n = 0;
  for (uint32_t j = threadIdx.x; j < ch.len; j += kTopBlock) {
    const uint32_t t = order[ch.start + j];
    n += bin_of(T.c[d.axis][t], d.cmin, d.w, K) <= d.boundary ? 1u : 0u;
  }
  n = wave_sum(n);
  if (__lane_id() == 0) atomicAdd(&s_sum, n);
  __syncthreads();
  if (threadIdx.x == 0) chunk_left[blockIdx.x] = s_sum;
}

// Stable partition of each chunk into `out`: left-goers from f + lbase, right-goers from
// f + lc + rbase, in chunk order.
__global__ __launch_bounds__(kTopBlock) void k_top_scatter(const Chunk* chunks, const uint32_t* order, uint32_t* out,
                                                           Tris T, const Decision* dec, const Task* tasks, uint32_t K,
                                                           const uint32_t* lbase, const uint32_t* rbase) {
  __shared__ uint32_t s_l[kTopBlock / 64], s_r[kTopBlock / 64];
  const Chunk ch = chunks[blockIdx.x];
  const Decision d = dec[ch.node];
  const uint32_t f = tasks[ch.node].f;
  uint32_t lpos = f + lbase[blockIdx.x], rpos = f + d.lc + rbase[blockIdx.x];
  const int wv = threadIdx.x / 64;
  for (uint32_t r = 0; r < ch.len; r += kTopBlock) {
    const uint32_t j = r + threadIdx.x;
    const bool valid = j < ch.len;
    const uint32_t t = valid ? order[ch.start + j] : 0u;
    const bool left = valid && bin_of(T.c[d.axis][t], d.cmin, d.w, K) <= d.boundary;
    const uint64_t ml = __ballot(left), mv = __ballot(valid);
    const uint64_t mr = mv & ~ml;
    if (__lane_id() == 0) {
      s_l[wv] = (uint32_t)__popcll(ml);
      s_r[wv] = (uint32_t)__popcll(mr);
    }
    __syncthreads();
    uint32_t lo = 0, ro = 0, lt = 0, rt = 0;
    for (int k = 0; k < kTopBlock / 64; ++k) {
      if (k < wv) {
        lo += s_l[k];
        ro += s_r[k];
      }
      lt += s_l[k];
      rt += s_r[k];
    }
    if (left)
      out[lpos + lo + (uint32_t)__popcll(ml & lanes_below())] = t;
    else if (valid)
      out[rpos + ro + (uint32_t)__popcll(mr & lanes_below())] = t;
    lpos += lt;
    rpos += rt;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kTopBlock) void k_top_copy(const Chunk* chunks, const uint32_t* from, uint32_t* to) {
  const Chunk ch = chunks[blockIdx.x];
  for (uint32_t j = threadIdx.x; j < ch.len; j += kTopBlock) to[ch.start + j] = from[ch.start + j];
}

// One wave builds the subtree of one task, its indices held in LDS.
__global__ __launch_bounds__(64) void k_subtree(const Task* tasks, uint32_t* order, Tris T, Params P,
                                                BuildSlot* slots, uint32_t* used, int32_t* root_ref,
                                                unsigned long long* stats, int* err) {
  extern __shared__ __align__(16) unsigned char lds[];
  const uint32_t K = P.K;
  double* s_ra = (double*)lds;                              // 3 * K
  uint32_t* s_ordA = (uint32_t*)(s_ra + 3 * K);             // kSubtreeMax
  uint32_t* s_ordB = s_ordA + kSubtreeMax;                  // kSubtreeMax
  uint32_t* s_bins = s_ordB + kSubtreeMax;                  // 3 * K * 7
  uint32_t* s_rc = s_bins + 3 * K * 7;                      // 3 * K
  Task* s_stack = (Task*)(s_rc + 3 * K);                    // kStackMax
  const int lane = threadIdx.x;
  const Task root = tasks[blockIdx.x];
  const uint32_t f0 = root.f;
  for (uint32_t i = lane; i < root.c; i += 64) s_ordA[i] = order[f0 + i];
  if (lane == 0) s_stack[0] = Task{0, root.c, root.al, root.depth, root.parent, root.side};
  int sp = 1;
  unsigned long long n_nodes = 0, n_leaves = 0, n_tris = 0, max_leaf = 0, max_depth = 0;
  __syncthreads();
  while (sp > 0) {
    const Task t = s_stack[--sp];
    __syncthreads();
    // node and centroid bounds; a node of at most 64 triangles keeps one per lane in registers
    const bool small = t.c <= 64;
    uint32_t su = 0;
    float slo[3] = {INFINITY, INFINITY, INFINITY}, shi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double sc[3] = {0.0, 0.0, 0.0};
    if (small && (uint32_t)lane < t.c) {
      su = s_ordA[t.f + lane];
      for (int a = 0; a < 3; ++a) {
        slo[a] = T.lo[a][su];
        shi[a] = T.hi[a][su];
        sc[a] = T.c[a][su];
      }
    }
    float blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    unsigned long long clo[3] = {kDKeyPosInf, kDKeyPosInf, kDKeyPosInf}, chi[3] = {kDKeyNegInf, kDKeyNegInf, kDKeyNegInf};
    if (small) {
      if ((uint32_t)lane < t.c)
        for (int a = 0; a < 3; ++a) {
          blo[a] = slo[a];
          bhi[a] = shi[a];
          clo[a] = chi[a] = dkey(sc[a]);
        }
    } else {
      for (uint32_t i = lane; i < t.c; i += 64) {
        const uint32_t u = s_ordA[t.f + i];
        for (int a = 0; a < 3; ++a) {
          blo[a] = fminf(blo[a], T.lo[a][u]);
          bhi[a] = fmaxf(bhi[a], T.hi[a][u]);
          const unsigned long long k = dkey(T.c[a][u]);
          clo[a] = min(clo[a], k);
          chi[a] = max(chi[a], k);
        }
      }
    }
    double lo[3], hi[3], cmin[3], w[3];
    bool act[3];
    bool any = false;
    for (int a = 0; a < 3; ++a) {
      lo[a] = (double)wave_min(blo[a]);
      hi[a] = (double)wave_max(bhi[a]);
      cmin[a] = kdouble(wave_umin64(clo[a]));
      const double cmax = kdouble(wave_umax64(chi[a]));
      act[a] = !(cmin[a] == cmax);
      w[a] = (cmax - cmin[a]) / (double)K;
      any = any || act[a];
    }
    // bins and the best plane: its axis, boundary, left count and the two child boxes
    int axis = -1;
    uint32_t bnd = 0, lc = 0;
    double best = 0.0;
    float lbox[6], rbox[6];
    if (any) {
      for (uint32_t i = lane; i < 3 * K * 7; i += 64) s_bins[i] = bin_init(i % 7);
      __syncthreads();
      auto bin_one = [&](float tl0, float tl1, float tl2, float th0, float th1, float th2, const double* tc) {
        const uint32_t klo[3] = {fkey(tl0), fkey(tl1), fkey(tl2)}, khi[3] = {fkey(th0), fkey(th1), fkey(th2)};
        for (int a = 0; a < 3; ++a) {
          if (!act[a]) continue;
          uint32_t* B = s_bins + (a * K + bin_of(tc[a], cmin[a], w[a], K)) * 7;
          for (int k = 0; k < 3; ++k) {
            atomicMin(&B[k], klo[k]);
            atomicMax(&B[3 + k], khi[k]);
          }
          atomicAdd(&B[6], 1u);
        }
      };
      if (small) {
        if ((uint32_t)lane < t.c) bin_one(slo[0], slo[1], slo[2], shi[0], shi[1], shi[2], sc);
      } else {
        for (uint32_t i = lane; i < t.c; i += 64) {
          const uint32_t u = s_ordA[t.f + i];
          const double tc[3] = {T.c[0][u], T.c[1][u], T.c[2][u]};
          bin_one(T.lo[0][u], T.lo[1][u], T.lo[2][u], T.hi[0][u], T.hi[1][u], T.hi[2][u], tc);
        }
      }
      __syncthreads();
      if (3 * K <= 64) {
        // one lane per (axis, bin): prefix and suffix unions by shuffles, then the first cheapest
        // (axis, boundary) in the host's scan order
        const uint32_t L = (uint32_t)lane, a = L / K, k = L % K;
        const bool in = L < 3 * K;
        uint32_t plo[3], phi[3], pc = 0, qlo[3], qhi[3], qc = 0;
        for (int e = 0; e < 3; ++e) {
          plo[e] = qlo[e] = in ? s_bins[L * 7 + e] : kKeyPosInf;
          phi[e] = qhi[e] = in ? s_bins[L * 7 + 3 + e] : kKeyNegInf;
        }
        if (in) pc = qc = s_bins[L * 7 + 6];
        for (uint32_t d = 1; d < K; d <<= 1) {
          const int from_lo = (int)(L >= d ? L - d : L), from_hi = (int)(L + d < 64 ? L + d : L);
          const bool take_p = in && k >= d, take_q = in && k + d < K;
          for (int e = 0; e < 3; ++e) {
            const uint32_t a0 = (uint32_t)__shfl((int)plo[e], from_lo), a1 = (uint32_t)__shfl((int)phi[e], from_lo);
            const uint32_t b0 = (uint32_t)__shfl((int)qlo[e], from_hi), b1 = (uint32_t)__shfl((int)qhi[e], from_hi);
            if (take_p) {
              plo[e] = min(plo[e], a0);
              phi[e] = max(phi[e], a1);
            }
            if (take_q) {
              qlo[e] = min(qlo[e], b0);
              qhi[e] = max(qhi[e], b1);
            }
          }
          const uint32_t ac = (uint32_t)__shfl((int)pc, from_lo), bc = (uint32_t)__shfl((int)qc, from_hi);
          if (take_p) pc += ac;
          if (take_q) qc += bc;
        }
        // right of boundary k: the suffix from bin k + 1
        const int nxt = (int)(L + 1 < 64 ? L + 1 : L);
        uint32_t rlo[3], rhi[3];
        for (int e = 0; e < 3; ++e) {
          rlo[e] = (uint32_t)__shfl((int)qlo[e], nxt);
          rhi[e] = (uint32_t)__shfl((int)qhi[e], nxt);
        }
        const uint32_t rcount = (uint32_t)__shfl((int)qc, nxt);
        const bool axis_on = a == 0 ? act[0] : a == 1 ? act[1] : a == 2 ? act[2] : false;
        bool valid = in && axis_on && k + 1 < K && pc > 0 && rcount > 0;
        double cost = 0.0;
        if (valid) {
          const double pl[3] = {kfloat(plo[0]), kfloat(plo[1]), kfloat(plo[2])};
          const double ph[3] = {kfloat(phi[0]), kfloat(phi[1]), kfloat(phi[2])};
          const double ql[3] = {kfloat(rlo[0]), kfloat(rlo[1]), kfloat(rlo[2])};
          const double qh[3] = {kfloat(rhi[0]), kfloat(rhi[1]), kfloat(rhi[2])};
          cost = half_area(pl, ph) * (double)pc + half_area(ql, qh) * (double)rcount;
        }
        uint32_t who = valid ? L : 64u;
        for (int o = 32; o > 0; o >>= 1) {
          const double oc = __shfl_xor(cost, o);
          const uint32_t ow = (uint32_t)__shfl_xor((int)who, o);
          const bool ov = ow < 64u;
          if (ov && (!valid || oc < cost || (oc == cost && ow < who))) {
            cost = oc;
            who = ow;
            valid = true;
          }
        }
        if (who < 64u) {
          axis = (int)(who / K);
          bnd = who % K;
          best = cost;
          lc = (uint32_t)__shfl((int)pc, (int)who);
          for (int e = 0; e < 3; ++e) {
            lbox[e] = kfloat((uint32_t)__shfl((int)plo[e], (int)who));
            lbox[3 + e] = kfloat((uint32_t)__shfl((int)phi[e], (int)who));
            rbox[e] = kfloat((uint32_t)__shfl((int)rlo[e], (int)who));
            rbox[3 + e] = kfloat((uint32_t)__shfl((int)rhi[e], (int)who));
          }
        }
      } else {
        double cost = 0.0;
        uint32_t bb = 0;
        bool found = false;
        if (lane < 3 && act[lane]) found = axis_best(s_bins + lane * K * 7, K, s_ra + lane * K, s_rc + lane * K, &cost, &bb);
        for (int a = 0; a < 3; ++a) {  // the axes in order, each replacing only a strictly higher cost
          const bool fa = __shfl((int)found, a) != 0;
          const double ca = __shfl(cost, a);
          const uint32_t ba = (uint32_t)__shfl((int)bb, a);
          if (fa && (axis < 0 || ca < best)) {
            axis = a;
            bnd = ba;
            best = ca;
          }
        }
        if (axis >= 0) {
          const uint32_t* B = s_bins + axis * K * 7;
          for (uint32_t k = 0; k <= bnd; ++k) lc += B[k * 7 + 6];
          union_bins(B, 0, bnd + 1, lbox);
          union_bins(B, bnd + 1, K, rbox);
        }
      }
    }
    const int32_t mode = choose(axis >= 0, best, half_area(lo, hi), t.c, P);
    if (mode == kLeaf) {
      if (t.c > kMaxLeafTris) {
        if (lane == 0) atomicOr(err, kErrLeaf);
      } else {
        uint32_t u = 0, r = 0;
        if ((uint32_t)lane < t.c) {
          u = s_ordA[t.f + lane];
          for (uint32_t j = 0; j < t.c; ++j) r += s_ordA[t.f + j] < u ? 1u : 0u;
        }
        __syncthreads();
        if ((uint32_t)lane < t.c) s_ordA[t.f + r] = u;
      }
      if (lane == 0) {
        link(slots, root_ref, t.parent, t.side, dev_leaf_ref(f0 + t.f, t.c));
        ++n_nodes;
        ++n_leaves;
        n_tris += t.c;
        max_leaf = max(max_leaf, (unsigned long long)t.c);
        max_depth = max(max_depth, (unsigned long long)t.depth);
      }
      __syncthreads();
      continue;
    }
    if (mode == kSplit) {
      const double ca = cmin[axis], wa = w[axis];
      const double sca = axis == 0 ? sc[0] : axis == 1 ? sc[1] : sc[2];
      uint32_t lpos = 0, rpos = 0;
      for (uint32_t base = 0; base < t.c; base += 64) {
        const uint32_t i = base + lane;
        const bool valid = i < t.c;
        const uint32_t u = !valid ? 0u : small ? su : s_ordA[t.f + i];
        const bool left = valid && bin_of(small ? sca : T.c[axis][u], ca, wa, K) <= bnd;
        const uint64_t ml = __ballot(left), mr = __ballot(valid) & ~ml;
        if (left)
          s_ordB[t.f + lpos + (uint32_t)__popcll(ml & lanes_below())] = u;
        else if (valid)
          s_ordB[t.f + lc + rpos + (uint32_t)__popcll(mr & lanes_below())] = u;
        lpos += (uint32_t)__popcll(ml);
        rpos += (uint32_t)__popcll(mr);
      }
    } else {  // every centroid coincides and the leaf bound forces a split: sorted by index, halved
      lc = t.c / 2;
      for (uint32_t i = lane; i < t.c; i += 64) {
        const uint32_t u = s_ordA[t.f + i];
        uint32_t r = 0;
        for (uint32_t j = 0; j < t.c; ++j) r += s_ordA[t.f + j] < u ? 1u : 0u;
        s_ordB[t.f + r] = u;
      }
      __syncthreads();
      float hl[3] = {INFINITY, INFINITY, INFINITY}, hh[3] = {-INFINITY, -INFINITY, -INFINITY};
      float gl[3] = {INFINITY, INFINITY, INFINITY}, gh[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (uint32_t i = lane; i < t.c; i += 64) {
        const uint32_t u = s_ordB[t.f + i];
        for (int a = 0; a < 3; ++a) {
          if (i < lc) {
            hl[a] = fminf(hl[a], T.lo[a][u]);
            hh[a] = fmaxf(hh[a], T.hi[a][u]);
          } else {
            gl[a] = fminf(gl[a], T.lo[a][u]);
            gh[a] = fmaxf(gh[a], T.hi[a][u]);
          }
        }
      }
      for (int a = 0; a < 3; ++a) {
        lbox[a] = wave_min(hl[a]);
        lbox[3 + a] = wave_max(hh[a]);
        rbox[a] = wave_min(gl[a]);
        rbox[3 + a] = wave_max(gh[a]);
      }
    }
    __syncthreads();
    for (uint32_t i = lane; i < t.c; i += 64) s_ordA[t.f + i] = s_ordB[t.f + i];
    const uint32_t s = f0 + t.f + lc - 1;
    if (lane == 0) {
      BuildSlot& S = slots[s];
      for (int k = 0; k < 6; ++k) {
        S.lbox[k] = lbox[k];
        S.rbox[k] = rbox[k];
      }
      S.ref[0] = S.ref[1] = 0;
      S.f = f0 + t.f;
      S.al = t.al;
      used[s] = 1;
      link(slots, root_ref, t.parent, t.side, (int32_t)s);
      ++n_nodes;
    }
    const Task L{t.f, lc, t.al + 1, t.depth + 1, (int32_t)s, 0};
    const Task R{t.f + lc, t.c - lc, t.al, t.depth + 1, (int32_t)s, 1};
    if (sp + 2 > kStackMax) {  // unreachable: the smaller child goes first, so the stack holds at
      if (lane == 0) atomicOr(err, kErrStack);  // most log2(kSubtreeMax) + 1 entries
      break;
    }
    if (lane == 0) {
      s_stack[sp] = lc >= t.c - lc ? L : R;  // the larger child waits
      s_stack[sp + 1] = lc >= t.c - lc ? R : L;
    }
    sp += 2;
    __syncthreads();
  }
  __syncthreads();
  for (uint32_t i = lane; i < root.c; i += 64) order[f0 + i] = s_ordA[i];
  if (lane == 0) {
    atomicAdd(&stats[0], n_nodes);
    atomicAdd(&stats[1], n_leaves);
    atomicMax(&stats[2], max_depth);
    atomicMax(&stats[3], max_leaf);
    atomicAdd(&stats[4], n_tris);
  }
}

// ---- preorder numbering and the final arrays ---------------------------------------------------------
constexpr uint32_t kScanBlock = 256, kScanPer = 16, kScanTile = kScanBlock * kScanPer;

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const uint32_t* used, uint32_t n, uint32_t* tile_sum) {
  __shared__ uint32_t s[kScanBlock / 64];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
  uint32_t v = 0;
  for (uint32_t k = 0; k < kScanPer; ++k)
    if (base + k < n) v += used[base + k];
  v = wave_sum(v);
  if (__lane_id() == 0) s[threadIdx.x / 64] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t k = 0; k < kScanBlock / 64; ++k) t += s[k];
    tile_sum[blockIdx.x] = t;
  }
}

// exclusive scan of the tile sums in place, one workgroup
__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* tile_sum, uint32_t n) {
  __shared__ uint32_t s[1024];
  const uint32_t per = (n + 1023) / 1024, b = threadIdx.x * per;
  uint32_t v = 0;
  for (uint32_t k = 0; k < per; ++k)
    if (b + k < n) v += tile_sum[b + k];
  s[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int k = 0; k < 1024; ++k) {
      const uint32_t x = s[k];
      s[k] = run;
      run += x;
    }
  }
  __syncthreads();
  uint32_t run = s[threadIdx.x];
  for (uint32_t k = 0; k < per; ++k)
    if (b + k < n) {
      const uint32_t x = tile_sum[b + k];
      tile_sum[b + k] = run;
      run += x;
    }
}

// inclusive prefix of `used`
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const uint32_t* used, uint32_t n, const uint32_t* tile_off,
                                                           uint32_t* prefix) {
  __shared__ uint32_t s[kScanBlock];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
  uint32_t v = 0;
  for (uint32_t k = 0; k < kScanPer; ++k)
    if (base + k < n) v += used[base + k];
  s[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = tile_off[blockIdx.x];
    for (uint32_t k = 0; k < kScanBlock; ++k) {
      const uint32_t x = s[k];
      s[k] = run;
      run += x;
    }
  }
  __syncthreads();
  uint32_t run = s[threadIdx.x];
  for (uint32_t k = 0; k < kScanPer; ++k)
    if (base + k < n) {
      run += used[base + k];
      prefix[base + k] = run;
    }
}

__global__ void k_number(const uint32_t* used, uint32_t n, const BuildSlot* slots, const uint32_t* prefix,
                         uint32_t* pre) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !used[s]) return;
  const uint32_t f = slots[s].f;
  pre[s] = (f ? prefix[f - 1] : 0u) + slots[s].al;
}

// first interior node of each mesh in the global preorder: the used slots before its triangles
__global__ void k_mesh_base(const MeshDev* M, uint32_t nm, const uint32_t* prefix, uint32_t* node_base) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  node_base[m] = M[m].tri_start ? prefix[M[m].tri_start - 1] : 0u;
}

// A mesh's nodes keep references local to the mesh (node and triangle indices from its own
// first), as the host flattening writes them.
__device__ inline int32_t local_ref(int32_t r, const uint32_t* pre, uint32_t node_base, uint32_t tri_start) {
  if (r >= 0) return (int32_t)(pre[r] - node_base);
  const uint32_t v = ~(uint32_t)r;
  return dev_leaf_ref((v >> kLeafCountBits) - tri_start, (v & (kMaxLeafTris - 1)) + 1);
}

__global__ void k_nodes(const uint32_t* used, uint32_t n, const BuildSlot* slots, const uint32_t* pre, const MeshDev* M,
                        uint32_t nm, const uint32_t* node_base, DevNode* nodes) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !used[s]) return;
  const BuildSlot S = slots[s];
  const uint32_t m = mesh_of(M, nm, S.f);
  const uint32_t nb = node_base[m], ts = M[m].tri_start;
  DevNode d;
  d.xy0[0] = S.lbox[0]; d.xy0[1] = S.lbox[3]; d.xy0[2] = S.lbox[1]; d.xy0[3] = S.lbox[4];
  d.xy1[0] = S.rbox[0]; d.xy1[1] = S.rbox[3]; d.xy1[2] = S.rbox[1]; d.xy1[3] = S.rbox[4];
  d.z01[0] = S.lbox[2]; d.z01[1] = S.lbox[5]; d.z01[2] = S.rbox[2]; d.z01[3] = S.rbox[5];
  d.ref[0] = local_ref(S.ref[0], pre, nb, ts);
  d.ref[1] = local_ref(S.ref[1], pre, nb, ts);
  d.ref[2] = 0;
  d.ref[3] = 0;
  nodes[pre[s]] = d;
}

__global__ void k_tris(const uint32_t* order, uint32_t nt, const float* xyz, const uint32_t* idx, const MeshDev* M,
                       uint32_t nm, DevTri* tris) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  const MeshDev md = M[mesh_of(M, nm, i)];
  const uint32_t t = order[i];
  uint32_t a = idx[3 * (size_t)t], b = idx[3 * (size_t)t + 1], c = idx[3 * (size_t)t + 2];
  if (a >= md.nv || b >= md.nv || c >= md.nv) a = b = c = 0;  // flagged by k_prep; the build fails
  const float* p0 = xyz + 3 * (size_t)(md.vert_base + a);
  const float* p1 = xyz + 3 * (size_t)(md.vert_base + b);
  const float* p2 = xyz + 3 * (size_t)(md.vert_base + c);
  DevTri d;
  for (int k = 0; k < 3; ++k) {
    d.p0[k] = p0[k];
    d.e1[k] = p1[k] - p0[k];
    d.e2[k] = p2[k] - p0[k];
  }
  d.p0[3] = __uint_as_float(t - md.tri_start);  // the triangle's index within its mesh
  d.e1[3] = 0.f;
  d.e2[3] = 0.f;
  tris[i] = d;
}


// ---- 4-wide collapse (the host's collapse_bvh4 rule, level by level over every mesh) ------------------
// A frontier entry: the binary node (global index) a 4-wide node at the level's slot base + i is made from.
struct C4Entry {
  uint32_t bin, mesh;
};

struct C4Child {
  int32_t ref;  // binary reference, local to the mesh (or a leaf code)
  float lo[3], hi[3];
};

__device__ inline void c4_children(const DevNode& d, C4Child* c) {
  c[0] = C4Child{d.ref[0], {d.xy0[0], d.xy0[2], d.z01[0]}, {d.xy0[1], d.xy0[3], d.z01[1]}};
  c[1] = C4Child{d.ref[1], {d.xy1[0], d.xy1[2], d.z01[2]}, {d.xy1[1], d.xy1[3], d.z01[3]}};
}

__device__ inline double c4_area(const C4Child& c) {
  const double dx = (double)c.hi[0] - c.lo[0], dy = (double)c.hi[1] - c.lo[1], dz = (double)c.hi[2] - c.lo[2];
  return dx * dy + dy * dz + dz * dx;
}

// One 4-wide node per frontier entry: its binary node's two children, the interior child of largest
// surface area opened (the first of equal areas) until there are four; boxes, child count and the
// children's binary references written to out[base + i]; cnt[i] = its interior children.
__global__ void k_c4_expand(const DevNode* __restrict__ blas, const uint64_t* __restrict__ node_base,
                            const C4Entry* __restrict__ F, uint32_t n, uint32_t base, DevNode4* __restrict__ out,
                            uint32_t* __restrict__ cnt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const C4Entry e = F[i];
  const DevNode* nodes = blas + node_base[e.mesh];
  C4Child ch[4];
  int nc = 2;
  c4_children(blas[e.bin], ch);
  while (nc < 4) {
    int best = -1;
    double best_area = -1.0;
    for (int k = 0; k < nc; ++k)
      if (ch[k].ref >= 0 && c4_area(ch[k]) > best_area) {
        best = k;
        best_area = c4_area(ch[k]);
      }
    if (best < 0) break;
    C4Child sub[2];
    c4_children(nodes[ch[best].ref], sub);
    for (int k = nc; k > best + 1; --k) ch[k] = ch[k - 1];
    ch[best] = sub[0];
    ch[best + 1] = sub[1];
    ++nc;
  }
  DevNode4 d;
  uint32_t interior = 0;
  for (int k = 0; k < 4; ++k) {
    const bool on = k < nc;
    d.lox[k] = on ? ch[k].lo[0] : 0.f;
    d.loy[k] = on ? ch[k].lo[1] : 0.f;
    d.loz[k] = on ? ch[k].lo[2] : 0.f;
    d.hix[k] = on ? ch[k].hi[0] : 0.f;
    d.hiy[k] = on ? ch[k].hi[1] : 0.f;
    d.hiz[k] = on ? ch[k].hi[2] : 0.f;
    d.ref[k] = on ? ch[k].ref : 0;
    d.count[k] = 0;
    interior += (on && ch[k].ref >= 0) ? 1u : 0u;
  }
  d.count[0] = nc;
  out[base + i] = d;
  cnt[i] = interior;
}

// The interior children of entry i take the next level's slots next + (its exclusive prefix) + j,
// in child order; their references are rewritten to those slots and they form the next frontier.
__global__ void k_c4_link(const uint64_t* __restrict__ node_base, const C4Entry* __restrict__ F, uint32_t n,
                          uint32_t base, uint32_t next, const uint32_t* __restrict__ cnt,
                          const uint32_t* __restrict__ prefix, DevNode4* __restrict__ out, C4Entry* __restrict__ F2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const C4Entry e = F[i];
  DevNode4& d = out[base + i];
  uint32_t j = prefix[i] - cnt[i];
  for (int k = 0; k < 4; ++k) {
    if (k >= d.count[0] || d.ref[k] < 0) continue;
    F2[j] = C4Entry{(uint32_t)(node_base[e.mesh] + (uint64_t)d.ref[k]), e.mesh};
    d.ref[k] = (int32_t)(next + j);
    ++j;
  }
}

// ---- host orchestration -------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { PBRT_HIP(hipMalloc(&p, bytes ? bytes : 1)); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

uint32_t blocks_for(size_t n, uint32_t per) { return (uint32_t)std::max<size_t>(1, (n + per - 1) / per); }

size_t subtree_lds(uint32_t K) {
  return 3 * K * sizeof(double) + 2 * kSubtreeMax * sizeof(uint32_t) + 3 * K * 7 * sizeof(uint32_t) +
         3 * K * sizeof(uint32_t) + kStackMax * sizeof(Task);
}

}  // namespace

void gpu_build_meshes(hipStream_t st, const float* xyz, const uint32_t* idx, const std::vector<GpuMeshSpan>& spans,
                      const pbrt_bvh_params& bp, GpuBuiltMany* out) {
  if (spans.empty()) throw Error(PBRT_ERR_INVALID_ARG, "no mesh to build");
  if (bp.bins < 2 || bp.bins > 256) throw Error(PBRT_ERR_INVALID_ARG, "bins must be in [2, 256]");
  const uint32_t nm = (uint32_t)spans.size();
  std::vector<MeshDev> md(nm);
  uint64_t total = 0;
  for (uint32_t m = 0; m < nm; ++m) {
    if (spans[m].nt == 0) throw Error(PBRT_ERR_EMPTY_MESH, "a TriangleMesh needs at least one triangle");
    if (spans[m].nv == 0) throw Error(PBRT_ERR_INVALID_ARG, "vertex index out of range");
    md[m] = MeshDev{(uint32_t)total, spans[m].nt, spans[m].vert_base, spans[m].nv};
    total += spans[m].nt;
  }
  if (total - 1 > kMaxLeafFirst) throw Error(PBRT_ERR_LAYOUT, "meshes too large for the 27-bit leaf offset");
  const uint32_t nt = (uint32_t)total;
  const Params P{bp.bins, bp.max_leaf, bp.t_trav};
  const uint32_t K = P.K;
  const uint32_t ns = nt - 1;  // slots: split positions 1 .. nt - 1

  DevBuf d_mesh(nm * sizeof(MeshDev)), d_root(nm * sizeof(int32_t)), d_base(nm * sizeof(uint32_t));
  DevBuf tri_f(6 * (size_t)nt * sizeof(float)), tri_d(3 * (size_t)nt * sizeof(double));
  DevBuf order(nt * sizeof(uint32_t)), order2(nt * sizeof(uint32_t));
  DevBuf slots((size_t)std::max(ns, 1u) * sizeof(BuildSlot)), used((size_t)std::max(ns, 1u) * sizeof(uint32_t));
  DevBuf misc(64);  // err (int), stats (5 x u64 from offset 16)
  TrisW TW;
  Tris T;
  for (int a = 0; a < 3; ++a) {
    TW.lo[a] = tri_f.as<float>() + (size_t)a * nt;
    TW.hi[a] = tri_f.as<float>() + (size_t)(3 + a) * nt;
    TW.c[a] = tri_d.as<double>() + (size_t)a * nt;
    T.lo[a] = TW.lo[a];
    T.hi[a] = TW.hi[a];
    T.c[a] = TW.c[a];
  }
  const MeshDev* M = d_mesh.as<MeshDev>();
  int* err = misc.as<int>();
  int32_t* root_ref = d_root.as<int32_t>();
  unsigned long long* stats = (unsigned long long*)(misc.as<char>() + 16);
  PBRT_HIP(hipMemcpyAsync(d_mesh.p, md.data(), nm * sizeof(MeshDev), hipMemcpyHostToDevice, st));
  PBRT_HIP(hipMemsetAsync(misc.p, 0, 64, st));
  PBRT_HIP(hipMemsetAsync(used.p, 0, (size_t)std::max(ns, 1u) * sizeof(uint32_t), st));
  k_prep<<<blocks_for(nt, 256), 256, 0, st>>>(xyz, idx, nt, M, nm, TW, order.as<uint32_t>(), err);
  PBRT_HIP(hipGetLastError());

  // level buffers, sized for the widest level: the roots, or nodes above kSubtreeMax triangles
  const size_t max_nodes = nt / (kSubtreeMax + 1) + nm + 1;
  const size_t max_chunks = nt / kChunk + max_nodes + 1;
  DevBuf d_tasks(max_nodes * sizeof(Task)), d_nb(max_nodes * sizeof(NodeBounds));
  DevBuf d_bins(max_nodes * 3 * K * 7 * sizeof(uint32_t)), d_dec(max_nodes * sizeof(Decision));
  DevBuf d_ra(max_nodes * K * sizeof(double)), d_rc(max_nodes * K * sizeof(uint32_t));
  DevBuf d_chunks(max_chunks * sizeof(Chunk)), d_cl(max_chunks * sizeof(uint32_t));
  DevBuf d_lb(max_chunks * sizeof(uint32_t)), d_rb(max_chunks * sizeof(uint32_t));

  std::vector<Task> level(nm), bottom;
  for (uint32_t m = 0; m < nm; ++m) level[m] = Task{md[m].tri_start, md[m].nt, 0, 1, -1 - (int32_t)m, 0};
  std::vector<Chunk> chunks;
  auto make_chunks = [&](const std::vector<Task>& tasks) {
    chunks.clear();
    for (uint32_t j = 0; j < (uint32_t)tasks.size(); ++j)
      for (uint32_t off = 0; off < tasks[j].c; off += kChunk)
        chunks.push_back(Chunk{j, tasks[j].f + off, std::min(kChunk, tasks[j].c - off)});
    PBRT_HIP(hipMemcpyAsync(d_tasks.p, tasks.data(), tasks.size() * sizeof(Task), hipMemcpyHostToDevice, st));
    PBRT_HIP(hipMemcpyAsync(d_chunks.p, chunks.data(), chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice, st));
    return (uint32_t)chunks.size();
  };

  // the roots' boxes
  std::vector<NodeBounds> root_nb(nm);
  {
    const uint32_t nc = make_chunks(level);
    k_level_init<<<blocks_for((size_t)nm * 3 * K * 7, 256), 256, 0, st>>>(nm, K, d_nb.as<NodeBounds>(),
                                                                          d_bins.as<uint32_t>());
    k_top_bounds<<<nc, kTopBlock, 0, st>>>(d_chunks.as<Chunk>(), order.as<uint32_t>(), T, d_nb.as<NodeBounds>());
    PBRT_HIP(hipGetLastError());
    PBRT_HIP(hipMemcpyAsync(root_nb.data(), d_nb.p, nm * sizeof(NodeBounds), hipMemcpyDeviceToHost, st));
    PBRT_HIP(hipStreamSynchronize(st));
  }

  uint64_t top_interior = 0;
  std::vector<Decision> dec;
  std::vector<uint32_t> cl, lb, rb;
  while (!level.empty()) {
    std::vector<Task> large;
    for (const Task& t : level) (t.c > kSubtreeMax ? large : bottom).push_back(t);
    if (large.empty()) break;
    const uint32_t nl = (uint32_t)large.size();
    const uint32_t nc = make_chunks(large);
    k_level_init<<<blocks_for((size_t)nl * 3 * K * 7, 256), 256, 0, st>>>(nl, K, d_nb.as<NodeBounds>(),
                                                                          d_bins.as<uint32_t>());
    k_top_bounds<<<nc, kTopBlock, 0, st>>>(d_chunks.as<Chunk>(), order.as<uint32_t>(), T, d_nb.as<NodeBounds>());
    k_top_bins<<<nc, kTopBlock, 3 * K * 7 * sizeof(uint32_t), st>>>(d_chunks.as<Chunk>(), order.as<uint32_t>(), T,
                                                                     d_nb.as<NodeBounds>(), K, d_bins.as<uint32_t>());
    k_top_decide<<<blocks_for(nl, 64), 64, 0, st>>>(nl, d_tasks.as<Task>(), d_nb.as<NodeBounds>(), d_bins.as<uint32_t>(),
                                                    P, d_ra.as<double>(), d_rc.as<uint32_t>(), d_dec.as<Decision>(),
                                                    slots.as<BuildSlot>(), used.as<uint32_t>(), root_ref);
    PBRT_HIP(hipGetLastError());
    dec.resize(nl);
    PBRT_HIP(hipMemcpyAsync(dec.data(), d_dec.p, nl * sizeof(Decision), hipMemcpyDeviceToHost, st));
    PBRT_HIP(hipStreamSynchronize(st));
    for (uint32_t j = 0; j < nl; ++j) {
      if (dec[j].mode == kLeaf)
        throw Error(PBRT_ERR_LAYOUT, "leaf of " + std::to_string(large[j].c) + " triangles exceeds the device bound of 32");
      if (dec[j].mode == kHalve)
        throw Error(PBRT_ERR_INVALID_ARG, std::to_string(large[j].c) +
                                              " triangles share one centroid: too many for the device builder");
    }
    k_top_count<<<nc, kTopBlock, 0, st>>>(d_chunks.as<Chunk>(), order.as<uint32_t>(), T, d_dec.as<Decision>(), K,
                                          d_cl.as<uint32_t>());
    cl.resize(nc);
    PBRT_HIP(hipMemcpyAsync(cl.data(), d_cl.p, nc * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PBRT_HIP(hipStreamSynchronize(st));
    lb.resize(nc);
    rb.resize(nc);
    {
      uint32_t node = UINT32_MAX, l = 0, r = 0;
      for (uint32_t i = 0; i < nc; ++i) {
        if (chunks[i].node != node) {
          node = chunks[i].node;
          l = r = 0;
        }
        lb[i] = l;
        rb[i] = r;
        l += cl[i];
        r += chunks[i].len - cl[i];
      }
    }
    PBRT_HIP(hipMemcpyAsync(d_lb.p, lb.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    PBRT_HIP(hipMemcpyAsync(d_rb.p, rb.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    k_top_scatter<<<nc, kTopBlock, 0, st>>>(d_chunks.as<Chunk>(), order.as<uint32_t>(), order2.as<uint32_t>(), T,
                                            d_dec.as<Decision>(), d_tasks.as<Task>(), K, d_lb.as<uint32_t>(),
                                            d_rb.as<uint32_t>());
    k_top_copy<<<nc, kTopBlock, 0, st>>>(d_chunks.as<Chunk>(), order2.as<uint32_t>(), order.as<uint32_t>());
    PBRT_HIP(hipGetLastError());
    top_interior += nl;
    std::vector<Task> next;
    next.reserve(2 * nl);
    for (uint32_t j = 0; j < nl; ++j) {
      const Task& t = large[j];
      const uint32_t lc = dec[j].lc;
      const int32_t s = (int32_t)(t.f + lc - 1);
      next.push_back(Task{t.f, lc, t.al + 1, t.depth + 1, s, 0});
      next.push_back(Task{t.f + lc, t.c - lc, t.al, t.depth + 1, s, 1});
    }
    level.swap(next);
  }

  if (!bottom.empty()) {
    std::sort(bottom.begin(), bottom.end(), [](const Task& a, const Task& b) { return a.c > b.c; });
    DevBuf d_bottom(bottom.size() * sizeof(Task));
    PBRT_HIP(hipMemcpyAsync(d_bottom.p, bottom.data(), bottom.size() * sizeof(Task), hipMemcpyHostToDevice, st));
    k_subtree<<<(uint32_t)bottom.size(), 64, subtree_lds(K), st>>>(d_bottom.as<Task>(), order.as<uint32_t>(), T, P,
                                                                   slots.as<BuildSlot>(), used.as<uint32_t>(),
                                                                   root_ref, stats, err);
    PBRT_HIP(hipGetLastError());
    PBRT_HIP(hipStreamSynchronize(st));  // d_bottom is released at the end of this scope
  }

  // preorder numbering (all meshes in order, each mesh's nodes contiguous), then the arrays
  uint32_t n_interior = 0;
  std::vector<uint32_t> node_base(nm, 0);
  DevBuf prefix((size_t)std::max(ns, 1u) * sizeof(uint32_t)), pre((size_t)std::max(ns, 1u) * sizeof(uint32_t));
  if (ns > 0) {
    const uint32_t tiles = blocks_for(ns, kScanTile);
    DevBuf tile(tiles * sizeof(uint32_t));
    k_scan_tiles<<<tiles, kScanBlock, 0, st>>>(used.as<uint32_t>(), ns, tile.as<uint32_t>());
    k_scan_sums<<<1, 1024, 0, st>>>(tile.as<uint32_t>(), tiles);
    k_scan_apply<<<tiles, kScanBlock, 0, st>>>(used.as<uint32_t>(), ns, tile.as<uint32_t>(), prefix.as<uint32_t>());
    k_mesh_base<<<blocks_for(nm, 256), 256, 0, st>>>(M, nm, prefix.as<uint32_t>(), d_base.as<uint32_t>());
    PBRT_HIP(hipGetLastError());
    PBRT_HIP(hipMemcpyAsync(&n_interior, prefix.as<uint32_t>() + (ns - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PBRT_HIP(hipMemcpyAsync(node_base.data(), d_base.p, nm * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    k_number<<<blocks_for(ns, 256), 256, 0, st>>>(used.as<uint32_t>(), ns, slots.as<BuildSlot>(), prefix.as<uint32_t>(),
                                                  pre.as<uint32_t>());
    PBRT_HIP(hipStreamSynchronize(st));
  } else {
    PBRT_HIP(hipMemsetAsync(d_base.p, 0, nm * sizeof(uint32_t), st));
  }
  DevNode* nodes = nullptr;
  DevTri* tris = nullptr;
  PBRT_HIP(hipMalloc(&nodes, std::max<size_t>(1, n_interior) * sizeof(DevNode)));
  PBRT_HIP(hipMalloc(&tris, (size_t)nt * sizeof(DevTri)));
  try {
    if (ns > 0)
      k_nodes<<<blocks_for(ns, 256), 256, 0, st>>>(used.as<uint32_t>(), ns, slots.as<BuildSlot>(), pre.as<uint32_t>(),
                                                   M, nm, d_base.as<uint32_t>(), nodes);
    k_tris<<<blocks_for(nt, 256), 256, 0, st>>>(order.as<uint32_t>(), nt, xyz, idx, M, nm, tris);
    PBRT_HIP(hipGetLastError());
    int host_err = 0;
    std::vector<int32_t> rref(nm);
    unsigned long long st5[5];
    PBRT_HIP(hipMemcpyAsync(&host_err, err, sizeof(int), hipMemcpyDeviceToHost, st));
    PBRT_HIP(hipMemcpyAsync(rref.data(), root_ref, nm * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PBRT_HIP(hipMemcpyAsync(st5, stats, sizeof(st5), hipMemcpyDeviceToHost, st));
    PBRT_HIP(hipStreamSynchronize(st));
    if (host_err & kErrIndex) throw Error(PBRT_ERR_INVALID_ARG, "vertex index out of range");
    if (host_err & kErrLeaf) throw Error(PBRT_ERR_LAYOUT, "leaf of more than 32 triangles exceeds the device bound of 32");
    if (host_err & kErrStack) throw Error(PBRT_ERR_INTERNAL, "device BVH build: task stack overflow");
    out->nodes = nodes;
    out->tris = tris;
    out->n_nodes = n_interior;
    out->meshes.assign(nm, GpuMeshOut{});
    auto kf = [](uint32_t k) {
      const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
      float f;
      std::memcpy(&f, &b, 4);
      return f;
    };
    for (uint32_t m = 0; m < nm; ++m) {
      GpuMeshOut& o = out->meshes[m];
      o.node_base = node_base[m];
      o.n_nodes = (m + 1 < nm ? node_base[m + 1] : n_interior) - node_base[m];
      o.tri_base = md[m].tri_start;
      // a mesh's root, when interior, is first in its preorder; a leaf root is local too
      if (rref[m] >= 0) {
        o.root.ref = 0;
      } else {
        const uint32_t v = ~(uint32_t)rref[m];
        o.root.ref = leaf_ref((v >> kLeafCountBits) - md[m].tri_start, (v & (kMaxLeafTris - 1)) + 1);
      }
      for (int a = 0; a < 3; ++a) {
        o.root.lo[a] = kf(root_nb[m].blo[a]);
        o.root.hi[a] = kf(root_nb[m].bhi[a]);
      }
    }
    out->stats.node_count = top_interior + st5[0];
    out->stats.leaf_count = st5[1];
    out->stats.max_depth = st5[2];
    out->stats.max_leaf_tri_count = st5[3];
    out->stats.total_leaf_tri_count = st5[4];
  } catch (...) {
    (void)hipFree(nodes);
    (void)hipFree(tris);
    throw;
  }
}

void gpu_build_mesh_bvh(hipStream_t st, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                        const pbrt_bvh_params& p, GpuBuilt* out) {
  if (nt == 0) throw Error(PBRT_ERR_EMPTY_MESH, "a TriangleMesh needs at least one triangle");
  GpuBuiltMany many;
  gpu_build_meshes(st, xyz, idx, {GpuMeshSpan{0, nv, nt}}, p, &many);
  out->nodes = many.nodes;
  out->tris = many.tris;
  out->n_nodes = many.n_nodes;
  out->root = many.meshes[0].root;
  out->stats = many.stats;
}

void gpu_collapse4(hipStream_t st, const DevNode* blas, const std::vector<uint64_t>& node_base,
                   const std::vector<int32_t>& root_ref, uint64_t n_binary, DevNode4** out, uint64_t* n_out,
                   std::vector<int32_t>* root4) {
  const size_t nm = node_base.size();
  if (root_ref.size() != nm) throw Error(PBRT_ERR_INVALID_ARG, "one root per mesh");
  if (n_binary >= (1ull << 31)) throw Error(PBRT_ERR_LAYOUT, "more than 2^31 nodes for the 4-wide layout");
  root4->assign(nm, 0);
  // the first level: every mesh with an interior root, in mesh order (a leaf root stays a leaf)
  std::vector<C4Entry> first;
  for (size_t m = 0; m < nm; ++m) {
    if (root_ref[m] < 0) {
      (*root4)[m] = root_ref[m];
      continue;
    }
    (*root4)[m] = (int32_t)first.size();
    first.push_back(C4Entry{(uint32_t)(node_base[m] + (uint64_t)root_ref[m]), (uint32_t)m});
  }
  *out = nullptr;
  *n_out = 0;
  // each 4-wide node consumes at least one binary node, so n_binary bounds the node and frontier counts
  const size_t cap = std::max<size_t>(1, (size_t)n_binary);
  DevBuf d_base(nm * sizeof(uint64_t)), fa(cap * sizeof(C4Entry)), fb(cap * sizeof(C4Entry));
  DevBuf cnt(cap * sizeof(uint32_t)), prefix(cap * sizeof(uint32_t)), tile(blocks_for(cap, kScanTile) * sizeof(uint32_t));
  DevBuf wide(cap * sizeof(DevNode4));
  PBRT_HIP(hipMemcpyAsync(d_base.p, node_base.data(), nm * sizeof(uint64_t), hipMemcpyHostToDevice, st));
  if (!first.empty())
    PBRT_HIP(hipMemcpyAsync(fa.p, first.data(), first.size() * sizeof(C4Entry), hipMemcpyHostToDevice, st));
  C4Entry* F = fa.as<C4Entry>();
  C4Entry* F2 = fb.as<C4Entry>();
  uint32_t n = (uint32_t)first.size(), base = 0;
  while (n > 0) {
    if ((uint64_t)base + n > cap) throw Error(PBRT_ERR_INTERNAL, "4-wide collapse exceeds its node bound");
    hipLaunchKernelGGL(k_c4_expand, dim3(blocks_for(n, 256)), dim3(256), 0, st, blas, d_base.as<uint64_t>(), F, n,
                       base, wide.as<DevNode4>(), cnt.as<uint32_t>());
    const uint32_t tiles = blocks_for(n, kScanTile);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanBlock), 0, st, cnt.as<uint32_t>(), n, tile.as<uint32_t>());
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, tile.as<uint32_t>(), tiles);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(kScanBlock), 0, st, cnt.as<uint32_t>(), n, tile.as<uint32_t>(),
                       prefix.as<uint32_t>());
    const uint32_t next = base + n;
    hipLaunchKernelGGL(k_c4_link, dim3(blocks_for(n, 256)), dim3(256), 0, st, d_base.as<uint64_t>(), F, n, base, next,
                       cnt.as<uint32_t>(), prefix.as<uint32_t>(), wide.as<DevNode4>(), F2);
    PBRT_HIP(hipGetLastError());
    uint32_t n2 = 0;
    PBRT_HIP(hipMemcpyAsync(&n2, prefix.as<uint32_t>() + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PBRT_HIP(hipStreamSynchronize(st));
    base = next;
    n = n2;
    std::swap(F, F2);
  }
  // the exact-size array
  DevNode4* d = nullptr;
  PBRT_HIP(hipMalloc(&d, std::max<size_t>(1, base) * sizeof(DevNode4)));
  if (base) {
    const hipError_t e = hipMemcpyAsync(d, wide.p, (size_t)base * sizeof(DevNode4), hipMemcpyDeviceToDevice, st);
    const hipError_t e2 = e == hipSuccess ? hipStreamSynchronize(st) : e;
    if (e2 != hipSuccess) {
      (void)hipFree(d);
      PBRT_HIP(e2);
    }
  }
  *out = d;
  *n_out = base;
}

}  // namespace pbrt
