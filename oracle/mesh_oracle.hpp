// ORACLE — test infrastructure, not product code.
//
// CPU restatement of the reference's mesh hot path (nmichel/pbrt, Rust, computed in f64):
// vector arithmetic, the widened slab test, Möller–Trumbore, the binned-SAH mesh BVH and its
// closest-hit / any-hit traversals. Templated on the scalar type:
//   * Real = double  -> "Oracle A": the reference's own arithmetic, tie order and parameters.
//   * Real = float   -> "Oracle B": the same statements in FP32, the bit-exact comparator for the
//                       device kernels (canonical tie rule: smallest t, then smallest primitive id).
// Only tests/, the smoke check and the CPU-baseline leg of bench.py may load this code, and only as
// a checker. Every function cites the reference file:line it restates.
//
// Parity pins (see tests/test_oracle_*.py):
//   * published known answers of docs/mesures_bvh.md:706-713 — hit counts, node/leaf counts and
//     per-ray box/triangle test counts over the bvh_stats ray protocol on the five shipped PLYs;
//   * brute force over all triangles (triangle_mesh/bvh.rs:792-816) and any-hit/closest agreement
//     with the far sweep (triangle_mesh/bvh.rs:937-975);
//   * the binned cost against the exhaustive scan (triangle_mesh/bvh.rs:324-347, 867-894);
//   * slab cases of aabound.rs:476-562; closed forms (M-T against an FP64 re-evaluation).
#pragma once
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace orc {

template <class R>
struct V3 {
  R x, y, z;
  R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
// src/geom/vector3.rs:291-296 — evaluated left to right.
template <class R> inline R dot(V3<R> u, V3<R> v) { return u.x * v.x + u.y * v.y + u.z * v.z; }
// src/geom/vector3.rs:298-307
template <class R> inline V3<R> cross(V3<R> u, V3<R> v) {
  return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
// src/geom/vector3.rs:69-74 — multiply by the reciprocal of the root.
template <class R> inline V3<R> normalized(V3<R> v) {
  R n = dot(v, v);
  R inv = R(1) / std::sqrt(n);
  return v * inv;
}
template <class R> inline R length(V3<R> v) { return std::sqrt(dot(v, v)); }

// Rust f64::min / f64::max ignore a NaN operand; so do C's fmin/fmax.
template <class R> inline R rmin(R a, R b) { return std::fmin(a, b); }
template <class R> inline R rmax(R a, R b) { return std::fmax(a, b); }

// src/geom/aabound.rs:18-39 — u = eps/2, gamma(3) = 3u/(1-3u), widening = 2 gamma(3), evaluated in
// the type's own arithmetic.
template <class R> inline R slab_widening() {
  const R u = std::numeric_limits<R>::epsilon() * R(0.5);
  const R g3 = R(3) * u / (R(1) - R(3) * u);
  return R(2) * g3;
}

template <class R>
struct Ray {
  V3<R> o, d;
};

template <class R>
struct AABB {
  V3<R> bmin, bmax;
  // aabound.rs:52-57 — inverted bounds are the union identity.
  static AABB empty() {
    const R M = std::numeric_limits<R>::max();
    return {{M, M, M}, {-M, -M, -M}};
  }
  // aabound.rs:86-88
  bool is_empty() const { return bmin.x > bmax.x || bmin.y > bmax.y || bmin.z > bmax.z; }
  // aabound.rs:369-373
  void combine_with(const AABB& o) {
    bmin = {rmin(bmin.x, o.bmin.x), rmin(bmin.y, o.bmin.y), rmin(bmin.z, o.bmin.z)};
    bmax = {rmax(bmax.x, o.bmax.x), rmax(bmax.y, o.bmax.y), rmax(bmax.z, o.bmax.z)};
  }
  // aabound.rs:123-130
  R half_area() const {
    if (is_empty()) return R(0);
    V3<R> d = bmax - bmin;
    return d.x * d.y + d.y * d.z + d.z * d.x;
  }
  // aabound.rs:323-355 — returns false on a miss, else the entry distance clamped to tmin.
  bool hit(const Ray<R>& ray, R tmin, R tmax, R* entry) const {
    const R W = slab_widening<R>();
    for (int i = 0; i < 3; ++i) {
      const R di = ray.d[i], oi = ray.o[i];
      const R lo = bmin[i], hi = bmax[i];
      if (di == R(0)) {
        if (oi < lo || oi > hi) return false;
        continue;
      }
      const R inv = R(1) / di;
      R t0 = (lo - oi) * inv;
      R t1 = (hi - oi) * inv;
      if (inv < R(0)) std::swap(t0, t1);
      t0 = t0 - std::fabs(t0) * W;
      t1 = t1 + std::fabs(t1) * W;
      tmin = rmax(t0, tmin);
      tmax = rmin(t1, tmax);
      if (tmax < tmin) return false;
    }
    *entry = tmin;
    return true;
  }
};

// src/shapes/triangle_mesh/triangle_mesh.rs:183-217 (Möller–Trumbore).
template <class R>
struct TriHit {
  R t, u, v;
};
template <class R>
inline bool intersect_tri(const Ray<R>& ray, V3<R> p0, V3<R> p1, V3<R> p2, TriHit<R>* out) {
  V3<R> e1 = p1 - p0;
  V3<R> e2 = p2 - p0;
  V3<R> pv = cross(ray.d, e2);
  R det = dot(e1, pv);
  if (std::fabs(det) < R(1e-8)) return false;
  R inv_det = R(1) / det;
  V3<R> s = ray.o - p0;
  R u = inv_det * dot(s, pv);
  if (u < R(0) || u > R(1)) return false;
  V3<R> q = cross(s, e1);
  R v = inv_det * dot(ray.d, q);
  if (v < R(0) || u + v > R(1)) return false;
  R t = inv_det * dot(e2, q);
  *out = {t, u, v};
  return true;
}

// Traversal work counters: triangle_mesh/bvh.rs:18-47.
struct Stats {
  uint64_t nodes_visited = 0, box_tests = 0, triangle_tests = 0;
};

// Build parameters. The reference is bins=8, t_trav=0, max_leaf=0 (unbounded)
// (triangle_mesh/bvh.rs:129, 733-735); the device configuration adds a leaf bound.
struct BuildParams {
  int bins = 8;
  double t_trav = 0.0;
  int max_leaf = 0;  // 0: no bound (reference)
};

struct BuildStats {
  uint64_t node_count = 0, leaf_count = 0, max_depth = 0, max_leaf_tri_count = 0, total_leaf_tri_count = 0;
};

// The mesh BVH of triangle_mesh/bvh.rs:116-736. The build always runs in f64 — the reference's
// precision — over vertices that are FP32 values widened exactly; the traversal runs in R.
template <class R>
class MeshBVH {
 public:
  struct Node {
    AABB<double> bbox;
    uint64_t left_first, tri_count;
    bool is_leaf() const { return tri_count > 0; }
  };
  struct Tri {
    uint32_t v0, v1, v2;
    V3<double> centroid;
  };

  std::vector<double> verts64;  // f64 vertex coordinates (the reference's storage)
  std::vector<R> vertsR;        // the same vertices in the traversal type
  std::vector<uint32_t> indices;
  std::vector<Tri> tris;
  std::vector<uint64_t> tri_idx;
  std::vector<Node> nodes;
  // Traversal copy of the nodes: box in the traversal type (exact: min/max of inputs) next to the
  // child/leaf fields, one record per node as the reference's BVHNode (64 bytes in f64).
  struct TNode {
    AABB<R> box;
    uint64_t left_first, tri_count;
  };
  std::vector<TNode> tnodes;
  BuildParams params;

  // triangle_mesh/bvh.rs:179-219 (BVHTree::new) + 221-230 (build).
  MeshBVH(const float* xyz, uint64_t nv, const uint32_t* idx, uint64_t nt, BuildParams p) : params(p) {
    verts64.resize(nv * 3);
    vertsR.resize(nv * 3);
    for (uint64_t i = 0; i < nv * 3; ++i) {
      verts64[i] = (double)xyz[i];
      vertsR[i] = (R)verts64[i];
    }
    indices.assign(idx, idx + nt * 3);
    tris.resize(nt);
    tri_idx.resize(nt);
    for (uint64_t i = 0; i < nt; ++i) {
      Tri t;
      t.v0 = idx[3 * i];
      t.v1 = idx[3 * i + 1];
      t.v2 = idx[3 * i + 2];
      // bvh.rs:579-585 — v0 += v1; += v2; *= 1/3.
      V3<double> c = vert64(t.v0);
      c = c + vert64(t.v1);
      c = c + vert64(t.v2);
      c = c * (1.0 / 3.0);
      t.centroid = c;
      tris[i] = t;
      tri_idx[i] = i;
    }
    nodes.reserve(nt * 2);
    Node root{AABB<double>::empty(), 0, nt};
    update_bounds(root);
    nodes.push_back(root);
    subdivide(0, 1);
    tnodes.resize(nodes.size());
    for (size_t i = 0; i < nodes.size(); ++i) {
      const auto& b = nodes[i].bbox;
      tnodes[i].box = {{(R)b.bmin.x, (R)b.bmin.y, (R)b.bmin.z}, {(R)b.bmax.x, (R)b.bmax.y, (R)b.bmax.z}};
      tnodes[i].left_first = nodes[i].left_first;
      tnodes[i].tri_count = nodes[i].tri_count;
    }
  }

  V3<double> vert64(uint32_t i) const { return {verts64[3 * i], verts64[3 * i + 1], verts64[3 * i + 2]}; }
  V3<R> vertR(uint32_t i) const { return {vertsR[3 * i], vertsR[3 * i + 1], vertsR[3 * i + 2]}; }

  // bvh.rs:564-572
  AABB<double> tri_box(const Tri& t) const {
    V3<double> a = vert64(t.v0), b = vert64(t.v1), c = vert64(t.v2);
    return {{rmin(rmin(a.x, b.x), c.x), rmin(rmin(a.y, b.y), c.y), rmin(rmin(a.z, b.z), c.z)},
            {rmax(rmax(a.x, b.x), c.x), rmax(rmax(a.y, b.y), c.y), rmax(rmax(a.z, b.z), c.z)}};
  }
  // bvh.rs:232-238
  void update_bounds(Node& n) const {
    for (uint64_t i = n.left_first; i < n.left_first + n.tri_count; ++i) n.bbox.combine_with(tri_box(tris[tri_idx[i]]));
  }

  // bvh.rs:173-176 — Rust's `as usize` saturates (negative or NaN -> 0) before the clamp.
  static uint64_t bin_index(double c, double cmin, double w, int K) {
    double q = std::floor((c - cmin) / w);
    if (!(q > 0.0)) return 0;
    if (q >= (double)(K - 1)) return (uint64_t)(K - 1);
    return (uint64_t)q;
  }

  struct Split {
    bool valid = false;
    int axis = 0;
    uint64_t boundary = 0;
    double cmin = 0, w = 0, cost = 0;
  };

  // bvh.rs:688-699
  void centroid_extent(const Node& n, int axis, double* mn, double* mx) const {
    double lo = std::numeric_limits<double>::max(), hi = -std::numeric_limits<double>::max();
    for (uint64_t i = n.left_first; i < n.left_first + n.tri_count; ++i) {
      double c = tris[tri_idx[i]].centroid[axis];
      lo = rmin(lo, c);
      hi = rmax(hi, c);
    }
    *mn = lo;
    *mx = hi;
  }

  // bvh.rs:609-685 (with bins = params.bins instead of the constant 8).
  Split find_best_split_plane(const Node& n) const {
    const int K = params.bins;
    Split best;
    std::vector<AABB<double>> bb(K);
    std::vector<uint64_t> bc(K);
    std::vector<double> la(K - 1), ra(K - 1);
    std::vector<uint64_t> lc(K - 1), rc(K - 1);
    for (int axis = 0; axis < 3; ++axis) {
      double cmin, cmax;
      centroid_extent(n, axis, &cmin, &cmax);
      if (cmin == cmax) continue;
      double w = (cmax - cmin) / (double)K;
      // fill_bins, bvh.rs:708-723
      for (int k = 0; k < K; ++k) {
        bb[k] = AABB<double>::empty();
        bc[k] = 0;
      }
      for (uint64_t i = n.left_first; i < n.left_first + n.tri_count; ++i) {
        const Tri& t = tris[tri_idx[i]];
        uint64_t b = bin_index(t.centroid[axis], cmin, w, K);
        bb[b].combine_with(tri_box(t));
        bc[b] += 1;
      }
      AABB<double> lb = AABB<double>::empty(), rb = AABB<double>::empty();
      uint64_t ls = 0, rs = 0;
      for (int i = 0; i < K - 1; ++i) {
        lb.combine_with(bb[i]);
        ls += bc[i];
        la[i] = lb.half_area();
        lc[i] = ls;
        int m = K - 2 - i;
        rb.combine_with(bb[m + 1]);
        rs += bc[m + 1];
        ra[m] = rb.half_area();
        rc[m] = rs;
      }
      for (int b = 0; b < K - 1; ++b) {
        if (lc[b] == 0 || rc[b] == 0) continue;
        double cost = la[b] * (double)lc[b] + ra[b] * (double)rc[b];
        if (!best.valid || cost < best.cost) {
          best.valid = true;
          best.axis = axis;
          best.boundary = (uint64_t)b;
          best.cmin = cmin;
          best.w = w;
          best.cost = cost;
        }
      }
    }
    return best;
  }

  // bvh.rs:324-347 (test-only in the reference): the same cost from a full scan.
  double exhaustive_split_cost(const Node& n, const Split& s) const {
    AABB<double> lb = AABB<double>::empty(), rb = AABB<double>::empty();
    uint64_t lc = 0, rc = 0;
    for (uint64_t i = n.left_first; i < n.left_first + n.tri_count; ++i) {
      const Tri& t = tris[tri_idx[i]];
      if (bin_index(t.centroid[s.axis], s.cmin, s.w, params.bins) <= s.boundary) {
        lc++;
        lb.combine_with(tri_box(t));
      } else {
        rc++;
        rb.combine_with(tri_box(t));
      }
    }
    return lb.half_area() * (double)lc + rb.half_area() * (double)rc;
  }

  // bvh.rs:240-309. Generalised: a split is taken when t_trav*A + cost < A*N (the reference has
  // t_trav = 0); a node above max_leaf is split even when the SAH prefers a leaf, by the best
  // binned plane or, when every centroid coincides, by halving its triangles in index order.
  void subdivide(uint64_t node_idx, uint64_t depth) {
    const uint64_t left_first = nodes[node_idx].left_first, tri_count = nodes[node_idx].tri_count;
    const double area = nodes[node_idx].bbox.half_area();
    const double leaf_cost = area * (double)tri_count;
    Split split = find_best_split_plane(nodes[node_idx]);
    const bool over = params.max_leaf > 0 && tri_count > (uint64_t)params.max_leaf;
    uint64_t left_end = left_first;
    if (split.valid && (split.cost + params.t_trav * area < leaf_cost || over)) {
      for (uint64_t i = left_first; i < left_first + tri_count; ++i) {
        double c = tris[tri_idx[i]].centroid[split.axis];
        if (bin_index(c, split.cmin, split.w, params.bins) <= split.boundary) {
          std::swap(tri_idx[i], tri_idx[left_end]);
          left_end++;
        }
      }
    } else if (over) {
      std::sort(tri_idx.begin() + left_first, tri_idx.begin() + left_first + tri_count);
      left_end = left_first + tri_count / 2;
    } else {
      return;
    }
    const uint64_t left_count = left_end - left_first;
    if (left_count == 0 || left_count == tri_count) return;
    Node l{AABB<double>::empty(), left_first, left_count};
    update_bounds(l);
    Node r{AABB<double>::empty(), left_end, tri_count - left_count};
    update_bounds(r);
    uint64_t li = nodes.size();
    nodes.push_back(l);
    nodes.push_back(r);
    nodes[node_idx].left_first = li;
    nodes[node_idx].tri_count = 0;
    subdivide(li, depth + 1);
    subdivide(li + 1, depth + 1);
  }

  // bvh.rs:539-562
  void collect(uint64_t i, uint64_t depth, BuildStats& s) const {
    const Node& n = nodes[i];
    s.node_count++;
    s.max_depth = std::max(s.max_depth, depth);
    if (n.is_leaf()) {
      s.leaf_count++;
      s.total_leaf_tri_count += n.tri_count;
      s.max_leaf_tri_count = std::max(s.max_leaf_tri_count, n.tri_count);
    } else {
      collect(n.left_first, depth + 1, s);
      collect(n.left_first + 1, depth + 1, s);
    }
  }
  BuildStats build_stats() const {
    BuildStats s;
    collect(0, 1, s);
    return s;
  }

  bool tri_test(const Ray<R>& ray, uint64_t orig, TriHit<R>* h) const {
    const uint32_t* ix = &indices[3 * orig];
    return intersect_tri(ray, vertR(ix[0]), vertR(ix[1]), vertR(ix[2]), h);
  }

  // bvh.rs:450-526 — ordered closest-hit. canonical=false: the reference's ties (strict <, first
  // found wins, pop skips entry >= min_t). canonical=true: the order-independent rule the device
  // kernels share: accept (t, id) lexicographically smaller; skip only entries > min_t.
  bool traverse(const Ray<R>& ray, R near, R far, bool canonical, Stats* st, TriHit<R>* out, uint64_t* out_id) const {
    R min_t = std::numeric_limits<R>::max();
    bool found = false;
    uint64_t best = ~0ull;
    TriHit<R> bh{};
    R rd;
    st->box_tests++;
    if (!tnodes[0].box.hit(ray, near, far, &rd)) return false;
    struct E {
      uint64_t n;
      R d;
    };
    E stack[512];
    int sp = 0;
    stack[sp++] = {0, rd};
    while (sp > 0) {
      E e = stack[--sp];
      if (canonical ? (e.d > min_t) : (e.d >= min_t)) continue;
      st->nodes_visited++;
      const TNode& n = tnodes[e.n];
      if (n.tri_count > 0) {
        for (uint64_t i = 0; i < n.tri_count; ++i) {
          uint64_t orig = tri_idx[n.left_first + i];
          st->triangle_tests++;
          TriHit<R> h;
          if (tri_test(ray, orig, &h)) {
            bool take = canonical ? (h.t >= near && h.t <= far && (h.t < min_t || (h.t == min_t && orig < best)))
                                  : (h.t < min_t && h.t >= near && h.t <= far);
            if (take) {
              min_t = h.t;
              bh = h;
              best = orig;
              found = true;
            }
          }
        }
      } else {
        uint64_t li = n.left_first, ri = n.left_first + 1;
        R ld, rdist;
        st->box_tests += 2;
        bool lh = tnodes[li].box.hit(ray, near, far, &ld);
        bool rh = tnodes[ri].box.hit(ray, near, far, &rdist);
        // bvh.rs:503-506: the right child goes first only when both are hit and it is strictly nearer.
        const bool swap_order = lh && rh && rdist < ld;
        const E nearer = swap_order ? E{ri, rdist} : E{li, ld};
        const E farther = swap_order ? E{li, ld} : E{ri, rdist};
        const bool nh = swap_order ? rh : lh;
        const bool fh = swap_order ? lh : rh;
        if (fh && (canonical ? farther.d <= min_t : farther.d < min_t)) stack[sp++] = farther;
        if (nh && (canonical ? nearer.d <= min_t : nearer.d < min_t)) stack[sp++] = nearer;
      }
    }
    if (found) {
      *out = bh;
      *out_id = best;
    }
    return found;
  }

  // bvh.rs:381-420 — unordered any-hit over a fixed [near, far].
  bool intersect_p(const Ray<R>& ray, R near, R far) const {
    R d;
    if (!tnodes[0].box.hit(ray, near, far, &d)) return false;
    uint64_t stack[1024];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const TNode& n = tnodes[stack[--sp]];
      if (n.tri_count > 0) {
        for (uint64_t i = 0; i < n.tri_count; ++i) {
          TriHit<R> h;
          if (tri_test(ray, tri_idx[n.left_first + i], &h) && h.t >= near && h.t <= far) return true;
        }
      } else {
        for (uint64_t c : {n.left_first, n.left_first + 1}) {
          if (tnodes[c].box.hit(ray, near, far, &d)) stack[sp++] = c;
        }
      }
    }
    return false;
  }

  // triangle_mesh/bvh.rs:792-816 — every triangle in index order, strict < (lowest index wins).
  bool brute_force(const Ray<R>& ray, R near, R far, TriHit<R>* out, uint64_t* out_id) const {
    bool found = false;
    for (uint64_t i = 0; i < tris.size(); ++i) {
      TriHit<R> h;
      if (!tri_test(ray, i, &h)) continue;
      if (h.t < near || h.t > far) continue;
      if (!found || h.t < out->t) {
        *out = h;
        *out_id = i;
        found = true;
      }
    }
    return found;
  }
};

}  // namespace orc
