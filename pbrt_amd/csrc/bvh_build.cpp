// Host BVH builder for triangle meshes: binned surface-area heuristic, the algorithm of the
// reference's BVHTree (src/shapes/triangle_mesh/bvh.rs:240-309, 609-735), with the bin count,
// traversal cost and leaf bound as parameters, then flattened into the 64-byte two-child device
// nodes the traversal kernels read.
//
// The tree is a function of the triangle set alone: per-node centroid extents, bin boxes (min/max
// unions, exact) and bin counts do not depend on the order triangles are visited in, and candidate
// planes are scanned axis by axis, boundary by boundary, keeping the first strictly cheaper one
// (bvh.rs:658-681). So this builder — which keeps precomputed per-triangle boxes and partitions
// with std::partition — produces the same tree as the reference's in-place swap partition.
//
// Two choices make the flattened layout independent of the partition order as well, so that the
// device builder (bvh_gpu.hip) reproduces it byte for byte: a zero box bound is stored as +0
// (fmin / fmax of -0 and +0 depend on argument order; the two compare equal, so no decision
// changes), and each leaf's triangles are sorted by index.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "common.hpp"

namespace pbrt {
namespace {

struct TriInfo {
  float lo[3], hi[3];
  double c[3];  // centroid in f64: ((v0 + v1) + v2) * (1/3), bvh.rs:579-585
};

struct Box {
  double lo[3], hi[3];
  static Box empty() { return {{DBL_MAX, DBL_MAX, DBL_MAX}, {-DBL_MAX, -DBL_MAX, -DBL_MAX}}; }
  void grow(const TriInfo& t) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::fmin(lo[a], (double)t.lo[a]);
      hi[a] = std::fmax(hi[a], (double)t.hi[a]);
    }
  }
  void grow(const Box& b) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::fmin(lo[a], b.lo[a]);
      hi[a] = std::fmax(hi[a], b.hi[a]);
    }
  }
  // half surface area, 0 for the empty box (aabound.rs:123-130)
  double half_area() const {
    if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) return 0.0;
    double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
  }
};

// bvh.rs:173-176, with Rust's saturating float->usize cast.
inline uint32_t bin_of(double c, double cmin, double w, uint32_t K) {
  double q = std::floor((c - cmin) / w);
  if (!(q > 0.0)) return 0;
  if (q >= (double)(K - 1)) return K - 1;
  return (uint32_t)q;
}

class Builder {
 public:
  Builder(const std::vector<TriInfo>& ti, const pbrt_bvh_params& p, BuiltBVH* out) : ti_(ti), p_(p), out_(out) {
    K_ = p.bins;
    bins_.resize(K_);
    cnt_.resize(K_);
    la_.resize(K_);
    ra_.resize(K_);
    lc_.resize(K_);
    rc_.resize(K_);
  }

  uint32_t build(uint32_t first, uint32_t count, uint32_t depth) {
    auto& order = out_->order;
    Box box = Box::empty();
    for (uint32_t i = first; i < first + count; ++i) box.grow(ti_[order[i]]);
    const uint32_t me = (uint32_t)out_->nodes.size();
    out_->nodes.push_back(BuildNode{});
    {
      BuildNode& n = out_->nodes[me];
      for (int a = 0; a < 3; ++a) {
        n.lo[a] = box.lo[a];
        n.hi[a] = box.hi[a];
      }
    }
    const double area = box.half_area();
    const double leaf_cost = area * (double)count;
    int axis = -1;
    uint32_t boundary = 0;
    double cmin_s = 0, w_s = 0;
    const double split_cost = best_split(first, count, &axis, &boundary, &cmin_s, &w_s);
    const bool over = p_.max_leaf > 0 && count > p_.max_leaf;
    uint32_t mid;
    if (axis >= 0 && (split_cost + p_.t_trav * area < leaf_cost || over)) {
      auto it = std::partition(order.begin() + first, order.begin() + first + count, [&](uint32_t t) {
        return bin_of(ti_[t].c[axis], cmin_s, w_s, K_) <= boundary;
      });
      mid = (uint32_t)(it - order.begin());
    } else if (over) {
      // every centroid coincides: halve the triangles in index order
      std::sort(order.begin() + first, order.begin() + first + count);
      mid = first + count / 2;
    } else {
      mid = first;  // leaf
    }
    if (mid == first || mid == first + count) {
      std::sort(order.begin() + first, order.begin() + first + count);
      BuildNode& n = out_->nodes[me];
      n.leaf = true;
      n.first = first;
      n.count = count;
      auto& s = out_->stats;
      s.node_count++;
      s.leaf_count++;
      s.total_leaf_tri_count += count;
      s.max_leaf_tri_count = std::max<uint64_t>(s.max_leaf_tri_count, count);
      s.max_depth = std::max<uint64_t>(s.max_depth, depth);
      return me;
    }
    out_->stats.node_count++;
    uint32_t l = build(first, mid - first, depth + 1);
    uint32_t r = build(mid, first + count - mid, depth + 1);
    BuildNode& n = out_->nodes[me];
    n.leaf = false;
    n.left = l;
    n.right = r;
    return me;
  }

 private:
  // find_best_split_plane, bvh.rs:609-685. Returns the cost; *axis = -1 when no plane separates.
  double best_split(uint32_t first, uint32_t count, int* best_axis, uint32_t* best_b, double* best_cmin,
                    double* best_w) {
    const auto& order = out_->order;
    double best = 0;
    *best_axis = -1;
    for (int axis = 0; axis < 3; ++axis) {
      double cmin = DBL_MAX, cmax = -DBL_MAX;
      for (uint32_t i = first; i < first + count; ++i) {
        double c = ti_[order[i]].c[axis];
        cmin = std::fmin(cmin, c);
        cmax = std::fmax(cmax, c);
      }
      if (cmin == cmax) continue;
      const double w = (cmax - cmin) / (double)K_;
      for (uint32_t k = 0; k < K_; ++k) {
        bins_[k] = Box::empty();
        cnt_[k] = 0;
      }
      for (uint32_t i = first; i < first + count; ++i) {
        const TriInfo& t = ti_[order[i]];
        uint32_t b = bin_of(t.c[axis], cmin, w, K_);
        bins_[b].grow(t);
        cnt_[b]++;
      }
      Box lb = Box::empty(), rb = Box::empty();
      uint64_t ls = 0, rs = 0;
      for (uint32_t i = 0; i + 1 < K_; ++i) {
        lb.grow(bins_[i]);
        ls += cnt_[i];
        la_[i] = lb.half_area();
        lc_[i] = ls;
        uint32_t m = K_ - 2 - i;
        rb.grow(bins_[m + 1]);
        rs += cnt_[m + 1];
        ra_[m] = rb.half_area();
        rc_[m] = rs;
      }
      for (uint32_t b = 0; b + 1 < K_; ++b) {
        if (lc_[b] == 0 || rc_[b] == 0) continue;
        double cost = la_[b] * (double)lc_[b] + ra_[b] * (double)rc_[b];
        if (*best_axis < 0 || cost < best) {
          best = cost;
          *best_axis = axis;
          *best_b = b;
          *best_cmin = cmin;
          *best_w = w;
        }
      }
    }
    return best;
  }

  const std::vector<TriInfo>& ti_;
  pbrt_bvh_params p_;
  BuiltBVH* out_;
  uint32_t K_;
  std::vector<Box> bins_;
  std::vector<uint64_t> cnt_;
  std::vector<double> la_, ra_;
  std::vector<uint64_t> lc_, rc_;
};

}  // namespace

BuiltBVH build_mesh_bvh(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt, const pbrt_bvh_params& p) {
  if (nt == 0) throw Error(PBRT_ERR_EMPTY_MESH, "a TriangleMesh needs at least one triangle");
  if (p.bins < 2 || p.bins > 256) throw Error(PBRT_ERR_INVALID_ARG, "bins must be in [2, 256]");
  std::vector<TriInfo> ti(nt);
  for (uint32_t i = 0; i < nt; ++i) {
    const uint32_t a = idx[3 * i], b = idx[3 * i + 1], c = idx[3 * i + 2];
    if (a >= nv || b >= nv || c >= nv) throw Error(PBRT_ERR_INVALID_ARG, "vertex index out of range");
    const float* v0 = xyz + 3 * (size_t)a;
    const float* v1 = xyz + 3 * (size_t)b;
    const float* v2 = xyz + 3 * (size_t)c;
    TriInfo& t = ti[i];
    for (int k = 0; k < 3; ++k) {
      t.lo[k] = std::fmin(std::fmin(v0[k], v1[k]), v2[k]);
      t.hi[k] = std::fmax(std::fmax(v0[k], v1[k]), v2[k]);
      if (t.lo[k] == 0.0f) t.lo[k] = 0.0f;
      if (t.hi[k] == 0.0f) t.hi[k] = 0.0f;
      double s = (double)v0[k];
      s = s + (double)v1[k];
      s = s + (double)v2[k];
      t.c[k] = s * (1.0 / 3.0);
    }
  }
  BuiltBVH out;
  out.order.resize(nt);
  for (uint32_t i = 0; i < nt; ++i) out.order[i] = i;
  out.nodes.reserve(2 * (size_t)nt);
  Builder(ti, p, &out).build(0, nt, 1);
  return out;
}

namespace {

struct Child4 {
  int32_t ref;
  float lo[3], hi[3];
};

void children_of(const DevNode& d, Child4 c[2]) {
  c[0] = Child4{d.ref[0], {d.xy0[0], d.xy0[2], d.z01[0]}, {d.xy0[1], d.xy0[3], d.z01[1]}};
  c[1] = Child4{d.ref[1], {d.xy1[0], d.xy1[2], d.z01[2]}, {d.xy1[1], d.xy1[3], d.z01[3]}};
}

double child_area(const Child4& c) {
  const double dx = (double)c.hi[0] - c.lo[0], dy = (double)c.hi[1] - c.lo[1], dz = (double)c.hi[2] - c.lo[2];
  return dx * dy + dy * dz + dz * dx;
}

int32_t collapse_node(const std::vector<DevNode>& n2, int32_t ref2, std::vector<DevNode4>* n4) {
  if (ref2 < 0) return ref2;
  Child4 ch[4];
  int nc = 2;
  children_of(n2[ref2], ch);
  while (nc < 4) {
    int best = -1;
    double best_area = -1.0;
    for (int k = 0; k < nc; ++k)
      if (ch[k].ref >= 0 && child_area(ch[k]) > best_area) {
        best = k;
        best_area = child_area(ch[k]);
      }
    if (best < 0) break;
    Child4 sub[2];
    children_of(n2[ch[best].ref], sub);
    for (int k = nc; k > best + 1; --k) ch[k] = ch[k - 1];  // the opened child's two take its place
    ch[best] = sub[0];
    ch[best + 1] = sub[1];
    ++nc;
  }
  const int32_t me = (int32_t)n4->size();
  n4->push_back(DevNode4{});
  int32_t refs[4] = {0, 0, 0, 0};
  for (int k = 0; k < nc; ++k) refs[k] = collapse_node(n2, ch[k].ref, n4);
  DevNode4& d = (*n4)[me];
  for (int k = 0; k < 4; ++k) {
    const bool on = k < nc;
    d.lox[k] = on ? ch[k].lo[0] : 0.f;
    d.loy[k] = on ? ch[k].lo[1] : 0.f;
    d.loz[k] = on ? ch[k].lo[2] : 0.f;
    d.hix[k] = on ? ch[k].hi[0] : 0.f;
    d.hiy[k] = on ? ch[k].hi[1] : 0.f;
    d.hiz[k] = on ? ch[k].hi[2] : 0.f;
    d.ref[k] = refs[k];
    d.count[k] = 0;
  }
  d.count[0] = nc;
  return me;
}

}  // namespace

void collapse_bvh4(const std::vector<DevNode>& n2, int32_t root_ref, std::vector<DevNode4>* n4, int32_t* root4) {
  n4->clear();
  n4->reserve(n2.size() / 2 + 1);
  *root4 = collapse_node(n2, root_ref, n4);
}

void flatten_mesh_bvh(const BuiltBVH& b, const float* xyz, const uint32_t* idx, std::vector<DevNode>* nodes,
                      std::vector<DevTri>* tris, RootDesc* root) {
  const uint32_t nt = (uint32_t)b.order.size();
  if (nt - 1 > kMaxLeafFirst) throw Error(PBRT_ERR_LAYOUT, "mesh too large for the 27-bit leaf offset");
  // Triangles in leaf order (the builder's `order` already is), pre-gathered with their edges.
  tris->resize(nt);
  for (uint32_t i = 0; i < nt; ++i) {
    const uint32_t t = b.order[i];
    const float* p0 = xyz + 3 * (size_t)idx[3 * t];
    const float* p1 = xyz + 3 * (size_t)idx[3 * t + 1];
    const float* p2 = xyz + 3 * (size_t)idx[3 * t + 2];
    DevTri& d = (*tris)[i];
    for (int k = 0; k < 3; ++k) {
      d.p0[k] = p0[k];
      d.e1[k] = p1[k] - p0[k];
      d.e2[k] = p2[k] - p0[k];
    }
    std::memcpy(&d.p0[3], &t, 4);
    d.e1[3] = 0.f;
    d.e2[3] = 0.f;
  }
  // Interior nodes renumbered in preorder; leaves become references.
  std::vector<int32_t> dev_index(b.nodes.size(), -1);
  uint32_t n_int = 0;
  for (size_t i = 0; i < b.nodes.size(); ++i)
    if (!b.nodes[i].leaf) dev_index[i] = (int32_t)n_int++;
  auto ref_of = [&](uint32_t i) -> int32_t {
    const BuildNode& n = b.nodes[i];
    if (n.leaf) {
      if (n.count > kMaxLeafTris)
        throw Error(PBRT_ERR_LAYOUT, "leaf of " + std::to_string(n.count) + " triangles exceeds the device bound of 32");
      return leaf_ref(n.first, n.count);
    }
    return dev_index[i];
  };
  nodes->assign(n_int, DevNode{});
  for (size_t i = 0; i < b.nodes.size(); ++i) {
    const BuildNode& n = b.nodes[i];
    if (n.leaf) continue;
    DevNode& d = (*nodes)[dev_index[i]];
    const BuildNode& L = b.nodes[n.left];
    const BuildNode& R = b.nodes[n.right];
    // The boxes are unions of FP32 coordinates, so the f64 -> f32 conversion is exact.
    d.xy0[0] = (float)L.lo[0]; d.xy0[1] = (float)L.hi[0]; d.xy0[2] = (float)L.lo[1]; d.xy0[3] = (float)L.hi[1];
    d.xy1[0] = (float)R.lo[0]; d.xy1[1] = (float)R.hi[0]; d.xy1[2] = (float)R.lo[1]; d.xy1[3] = (float)R.hi[1];
    d.z01[0] = (float)L.lo[2]; d.z01[1] = (float)L.hi[2]; d.z01[2] = (float)R.lo[2]; d.z01[3] = (float)R.hi[2];
    d.ref[0] = ref_of(n.left);
    d.ref[1] = ref_of(n.right);
    d.ref[2] = 0;
    d.ref[3] = 0;
  }
  const BuildNode& r = b.nodes[0];
  for (int k = 0; k < 3; ++k) {
    root->lo[k] = (float)r.lo[k];
    root->hi[k] = (float)r.hi[k];
  }
  root->ref = ref_of(0);
}

}  // namespace pbrt
