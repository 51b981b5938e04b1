// Host pinhole camera in f64: src/cameras/pin_hole.rs:17-50 with Matrix4 (src/geom/matrix4.rs):
// inverse perspective composed with the raster -> screen map, w-divide when w != 1, and the two
// normalisations of get_ray + Ray::new.
#pragma once
#include <cmath>
#include <utility>

namespace pbrt {

struct D3 {
  double x, y, z;
};
inline D3 d3add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 d3sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 d3scl(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double d3dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 d3cross(D3 u, D3 v) { return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x}; }
inline D3 d3norm(D3 v) {
  const double inv = 1.0 / std::sqrt(d3dot(v, v));
  return d3scl(v, inv);
}

struct Mat4 {
  double m[4][4];
};
// matrix4.rs:216-232: entries summed from 0.0 over k
inline Mat4 mat_mul(const Mat4& a, const Mat4& b) {
  Mat4 r{};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += a.m[i][k] * b.m[k][j];
      r.m[i][j] = s;
    }
  return r;
}
// Gauss-Jordan with full pivoting (matrix4.rs:100-174)
inline Mat4 mat_inverse(const Mat4& a) {
  int ic[4] = {0, 0, 0, 0}, ir[4] = {0, 0, 0, 0}, piv[4] = {0, 0, 0, 0};
  Mat4 r = a;
  for (int i = 0; i < 4; ++i) {
    int row = 0, col = 0;
    double big = 0.0;
    for (int j = 0; j < 4; ++j)
      if (piv[j] != 1)
        for (int k = 0; k < 4; ++k)
          if (piv[k] == 0 && std::fabs(r.m[j][k]) >= big) {
            big = std::fabs(r.m[j][k]);
            row = j;
            col = k;
          }
    piv[col] += 1;
    if (row != col)
      for (int k = 0; k < 4; ++k) std::swap(r.m[row][k], r.m[col][k]);
    ir[i] = row;
    ic[i] = col;
    const double p = 1.0 / r.m[col][col];
    r.m[col][col] = 1.0;
    for (int j = 0; j < 4; ++j) r.m[col][j] *= p;
    for (int j = 0; j < 4; ++j)
      if (j != col) {
        const double s = r.m[j][col];
        r.m[j][col] = 0.0;
        for (int k = 0; k < 4; ++k) r.m[j][k] -= r.m[col][k] * s;
      }
  }
  for (int j = 3; j >= 0; --j)
    if (ir[j] != ic[j])
      for (int k = 0; k < 4; ++k) std::swap(r.m[k][ir[j]], r.m[k][ic[j]]);
  return r;
}
inline D3 mat_point(const Mat4& M, D3 v) {
  const auto& m = M.m;
  D3 r{m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z + m[0][3], m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z + m[1][3],
       m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z + m[2][3]};
  const double w = m[3][0] * v.x + m[3][1] * v.y + m[3][2] * v.z + m[3][3];
  return w != 1.0 ? d3scl(r, 1.0 / w) : r;
}
inline D3 mat_dir(const Mat4& M, D3 v) {
  const auto& m = M.m;
  return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
          m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
}
// matrix4.rs:81-98
inline Mat4 look_at(D3 pos, D3 look, D3 up) {
  const D3 dir = d3norm(d3sub(look, pos));
  const D3 right = d3norm(d3cross(d3norm(up), dir));
  const D3 u = d3cross(dir, right);
  return Mat4{{{right.x, u.x, dir.x, pos.x}, {right.y, u.y, dir.y, pos.y}, {right.z, u.z, dir.z, pos.z}, {0, 0, 0, 1}}};
}

struct Pinhole {
  Mat4 r2s, c2w;
  Pinhole(uint32_t w, uint32_t h, double fov, double near, double far, const Mat4& cam_to_world) : c2w(cam_to_world) {
    const double ar = (double)w / (double)h;
    double x0 = -1.0, x1 = 1.0, y0 = -1.0 / ar, y1 = 1.0 / ar;
    if (ar > 1.0) {
      x0 = -ar; x1 = ar; y0 = -1.0; y1 = 1.0;
    }
    const double s = 1.0 / std::tan(fov / 2.0);
    const Mat4 persp{{{s, 0, 0, 0}, {0, s, 0, 0}, {0, 0, far / (far - near), -far * near / (far - near)}, {0, 0, 1, 0}}};
    const Mat4 tr{{{1, 0, 0, x0}, {0, 1, 0, y1}, {0, 0, 1, 0}, {0, 0, 0, 1}}};
    const Mat4 s1{{{x1 - x0, 0, 0, 0}, {0, y0 - y1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}}};
    const Mat4 s2{{{1.0 / (double)w, 0, 0, 0}, {0, 1.0 / (double)h, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}}};
    r2s = mat_mul(mat_mul(mat_mul(mat_inverse(persp), tr), s1), s2);
  }
  void ray(double px, double py, D3* o, D3* d) const {
    const D3 cv = d3norm(mat_point(r2s, {px, py, 0.0}));
    *o = mat_point(c2w, {0, 0, 0});
    *d = d3norm(mat_dir(c2w, cv));
  }
};

}  // namespace pbrt
