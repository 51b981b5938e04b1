// ORACLE — test infrastructure, not product code.
//
// f64 restatement of the reference's 4x4 matrices and pinhole camera, and of the bvh_stats ray
// protocol (six orbit cameras x 200^2 pixel centres). Pinned by the published hit counts of
// docs/mesures_bvh.md:706-713, which depend on nothing but these rays and Möller–Trumbore.
#pragma once
#include <cmath>

#include "mesh_oracle.hpp"

namespace orc {

struct M4 {
  double m[4][4];
};
// src/geom/matrix4.rs:24-44
inline M4 m4_zero() { return M4{{{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}}}; }
inline M4 m4_translation(double x, double y, double z) { return M4{{{1, 0, 0, x}, {0, 1, 0, y}, {0, 0, 1, z}, {0, 0, 0, 1}}}; }
inline M4 m4_scale(double x, double y, double z) { return M4{{{x, 0, 0, 0}, {0, y, 0, 0}, {0, 0, z, 0}, {0, 0, 0, 1}}}; }
inline M4 m4_rotation_y(double th) {
  double s = std::sin(th), c = std::cos(th);
  return M4{{{c, 0, s, 0}, {0, 1, 0, 0}, {-s, 0, c, 0}, {0, 0, 0, 1}}};
}
// matrix4.rs:70-79
inline M4 m4_perspective(double fov, double n, double f) {
  double s = 1.0 / std::tan(fov / 2.0);
  return M4{{{s, 0, 0, 0}, {0, s, 0, 0}, {0, 0, f / (f - n), -f * n / (f - n)}, {0, 0, 1, 0}}};
}
// matrix4.rs:81-98
inline M4 m4_look_at(V3<double> pos, V3<double> look, V3<double> up) {
  V3<double> dir = normalized(look - pos);
  V3<double> nup = normalized(up);
  V3<double> right = normalized(cross(nup, dir));
  V3<double> rup = cross(dir, right);
  return M4{{{right.x, rup.x, dir.x, pos.x}, {right.y, rup.y, dir.y, pos.y}, {right.z, rup.z, dir.z, pos.z}, {0, 0, 0, 1}}};
}
// matrix4.rs:100-174 — Gauss–Jordan with full pivoting (pbrt's MatrixInvert).
inline M4 m4_inverse(const M4& a) {
  int indxc[4] = {0, 0, 0, 0}, indxr[4] = {0, 0, 0, 0}, ipiv[4] = {0, 0, 0, 0};
  M4 r = a;
  double(*minv)[4] = r.m;
  for (int i = 0; i < 4; ++i) {
    int irow = 0, icol = 0;
    double big = 0.0;
    for (int j = 0; j < 4; ++j) {
      if (ipiv[j] != 1) {
        for (int k = 0; k < 4; ++k) {
          if (ipiv[k] == 0) {
            if (std::fabs(minv[j][k]) >= big) {
              big = std::fabs(minv[j][k]);
              irow = j;
              icol = k;
            }
          }
        }
      }
    }
    ipiv[icol] += 1;
    if (irow != icol) {
      for (int k = 0; k < 4; ++k) std::swap(minv[irow][k], minv[icol][k]);
    }
    indxr[i] = irow;
    indxc[i] = icol;
    double pivinv = 1.0 / minv[icol][icol];
    minv[icol][icol] = 1.0;
    for (int j = 0; j < 4; ++j) minv[icol][j] *= pivinv;
    for (int j = 0; j < 4; ++j) {
      if (j != icol) {
        double save = minv[j][icol];
        minv[j][icol] = 0.0;
        for (int k = 0; k < 4; ++k) minv[j][k] -= minv[icol][k] * save;
      }
    }
  }
  for (int j = 3; j >= 0; --j) {
    if (indxr[j] != indxc[j]) {
      for (int k = 0; k < 4; ++k) std::swap(minv[k][indxr[j]], minv[k][indxc[j]]);
    }
  }
  return r;
}
// matrix4.rs:216-232 — the sum starts from 0.0 and accumulates k = 0..3.
inline M4 m4_mul(const M4& a, const M4& b) {
  M4 r = m4_zero();
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) r.m[i][j] += a.m[i][k] * b.m[k][j];
  return r;
}
// matrix4.rs:187-199 (and the identical Mul<&Vector3f> at 242-257): w-divide only when w != 1.
inline V3<double> m4_point(const M4& M, V3<double> v) {
  const double(*m)[4] = M.m;
  double x = m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z + m[0][3];
  double y = m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z + m[1][3];
  double z = m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z + m[2][3];
  double w = m[3][0] * v.x + m[3][1] * v.y + m[3][2] * v.z + m[3][3];
  V3<double> r{x, y, z};
  if (w != 1.0) r = r * (1.0 / w);
  return r;
}
// matrix4.rs:201-206
inline V3<double> m4_dir(const M4& M, V3<double> v) {
  const double(*m)[4] = M.m;
  return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
          m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
}

// src/cameras/pin_hole.rs:17-50
struct PinHole {
  M4 raster_to_screen, cam_to_world;
  PinHole(unsigned w, unsigned h, double fov, double near, double far, const M4& c2w) : cam_to_world(c2w) {
    double ar = (double)w / (double)h;
    double sx0, sx1, sy0, sy1;
    if (ar > 1.0) {
      sx0 = -ar; sx1 = ar; sy0 = -1.0; sy1 = 1.0;
    } else {
      sx0 = -1.0; sx1 = 1.0; sy0 = -1.0 / ar; sy1 = 1.0 / ar;
    }
    M4 t = m4_mul(m4_inverse(m4_perspective(fov, near, far)), m4_translation(sx0, sy1, 0.0));
    t = m4_mul(t, m4_scale(sx1 - sx0, sy0 - sy1, 1.0));
    t = m4_mul(t, m4_scale(1.0 / (double)w, 1.0 / (double)h, 1.0));
    raster_to_screen = t;
  }
  // pin_hole.rs:42-50; Ray::new normalises again (src/geom/ray.rs:10-18).
  Ray<double> get_ray(double px, double py) const {
    V3<double> cv = normalized(m4_point(raster_to_screen, {px, py, 0.0}));
    V3<double> o = m4_point(cam_to_world, {0.0, 0.0, 0.0});
    V3<double> d = normalized(m4_dir(cam_to_world, cv));
    return {o, d};
  }
  // The same ray in FP32 (the renderer's camera in Oracle B): both matrices rounded to FP32 once,
  // then matrix4.rs:187-206's statements and the two normalisations in FP32.
  Ray<float> get_ray_f32(float px, float py) const {
    float r[4][4], c[4][4];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        r[i][j] = (float)raster_to_screen.m[i][j];
        c[i][j] = (float)cam_to_world.m[i][j];
      }
    auto point = [](const float (*m)[4], float x, float y, float z) {
      float a = m[0][0] * x + m[0][1] * y + m[0][2] * z + m[0][3];
      float b = m[1][0] * x + m[1][1] * y + m[1][2] * z + m[1][3];
      float e = m[2][0] * x + m[2][1] * y + m[2][2] * z + m[2][3];
      float w = m[3][0] * x + m[3][1] * y + m[3][2] * z + m[3][3];
      if (w != 1.0f) {
        float iw = 1.0f / w;
        a = a * iw;
        b = b * iw;
        e = e * iw;
      }
      return V3<float>{a, b, e};
    };
    V3<float> cv = normalized(point(r, px, py, 0.0f));
    V3<float> o = point(c, 0.0f, 0.0f, 0.0f);
    V3<float> d = normalized(V3<float>{c[0][0] * cv.x + c[0][1] * cv.y + c[0][2] * cv.z,
                                       c[1][0] * cv.x + c[1][1] * cv.y + c[1][2] * cv.z,
                                       c[2][0] * cv.x + c[2][1] * cv.y + c[2][2] * cv.z});
    return {o, d};
  }
};

// src/bin/bvh_stats.rs:95-116, 345-401: the six orbit cameras framing the mesh's bounding sphere.
constexpr int kViewCount = 6;
constexpr unsigned kImageSide = 200;
inline double bvh_stats_fov() { return 40.0 * M_PI / 180.0; }
inline double bvh_stats_elevation() { return 20.0 * M_PI / 180.0; }

// triangle_mesh.rs:150-174 — the mesh box from every vertex.
inline AABB<double> vertex_bounds(const double* v, uint64_t nv) {
  double M = std::numeric_limits<double>::max();
  double lx = M, ly = M, lz = M, hx = -M, hy = -M, hz = -M;
  for (uint64_t i = 0; i < nv; ++i) {
    double x = v[3 * i], y = v[3 * i + 1], z = v[3 * i + 2];
    lx = std::fmin(x, lx); ly = std::fmin(y, ly); lz = std::fmin(z, lz);
    hx = std::fmax(x, hx); hy = std::fmax(y, hy); hz = std::fmax(z, hz);
  }
  return {{lx, ly, lz}, {hx, hy, hz}};
}

// bvh_stats.rs:385-401 — camera number `view`, `side` x `side` pixels, looking at `center`.
inline PinHole orbit_camera(int view, V3<double> center, double distance, unsigned side, double fov) {
  double az = 2.0 * M_PI * (double)view / (double)kViewCount;
  double el = bvh_stats_elevation();
  V3<double> dir{std::cos(el) * std::cos(az), std::sin(el), std::cos(el) * std::sin(az)};
  V3<double> pos = center + dir * distance;
  return PinHole(side, side, fov, 1e-4, 1000.0, m4_look_at(pos, center, {0.0, 1.0, 0.0}));
}

// bvh_stats.rs:345-382 — every ray of the protocol, in cast order (view, row, column). `side` and
// `fov` default to the protocol's 200 and 40 degrees; other values give the same framing at another
// resolution.
inline void orbit_rays(const double* verts, uint64_t nv, unsigned side, double fov, int views, Ray<double>* out) {
  AABB<double> bb = vertex_bounds(verts, nv);
  V3<double> center = (bb.bmax + bb.bmin) * 0.5;
  double radius = length(bb.bmax - bb.bmin) * 0.5;
  double distance = radius / std::sin(fov / 2.0);
  uint64_t k = 0;
  for (int v = 0; v < views; ++v) {
    PinHole cam = orbit_camera(v, center, distance, side, fov);
    for (unsigned py = 0; py < side; ++py)
      for (unsigned px = 0; px < side; ++px) out[k++] = cam.get_ray((double)px + 0.5, (double)py + 0.5);
  }
}

}  // namespace orc
