// Seeded synthetic inputs for the ray path: displaced icospheres, pinhole / orbit / incoherent ray
// batches. Host code, f64 arithmetic; rays are rounded to FP32 once, at the end.
//
// The pinhole camera restates src/cameras/pin_hole.rs:17-50 (inverse perspective composed with the
// raster -> screen map, w-divide when w != 1, two normalisations) and Matrix4::look_at
// (src/geom/matrix4.rs:81-98); the orbit framing restates src/bin/bvh_stats.rs:345-401.
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "camera.hpp"
#include "common.hpp"
#include "rng.hpp"

namespace pbrt {
namespace {

inline D3 add(D3 a, D3 b) { return d3add(a, b); }
inline D3 sub(D3 a, D3 b) { return d3sub(a, b); }
inline D3 scl(D3 a, double s) { return d3scl(a, s); }
inline double dot3(D3 a, D3 b) { return d3dot(a, b); }
inline D3 nrm(D3 v) { return d3norm(v); }

inline void put_ray(pbrt_ray* r, D3 o, D3 d, float tmin, float tmax) {
  r->ox = (float)o.x; r->oy = (float)o.y; r->oz = (float)o.z; r->tmin = tmin;
  r->dx = (float)d.x; r->dy = (float)d.y; r->dz = (float)d.z; r->tmax = tmax;
}

// ---- value noise (seeded lattice values, quintic fade, trilinear blend) ------------------------
inline double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
  const uint64_t h = hash4(seed, (uint64_t)x, (uint64_t)y, (uint64_t)z);
  return to_unit(h) * 2.0 - 1.0;
}
inline double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
double value_noise(D3 p, uint64_t seed) {
  const double fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
  const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
  const double tx = fade(p.x - fx), ty = fade(p.y - fy), tz = fade(p.z - fz);
  double c[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int k = 0; k < 2; ++k) c[a][b][k] = lattice(ix + a, iy + b, iz + k, seed);
  auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
  const double x00 = lerp(c[0][0][0], c[1][0][0], tx), x10 = lerp(c[0][1][0], c[1][1][0], tx);
  const double x01 = lerp(c[0][0][1], c[1][0][1], tx), x11 = lerp(c[0][1][1], c[1][1][1], tx);
  return lerp(lerp(x00, x10, ty), lerp(x01, x11, ty), tz);
}
// fBm: sum of `octaves` layers, lacunarity 2, gain 0.5, normalised to [-1, 1].
double fbm(D3 p, int octaves, uint64_t seed) {
  double sum = 0.0, amp = 1.0, norm = 0.0, f = 1.0;
  for (int o = 0; o < octaves; ++o) {
    sum += amp * value_noise(scl(p, f), seed + (uint64_t)o * 0x9E37ull);
    norm += amp;
    amp *= 0.5;
    f *= 2.0;
  }
  return sum / norm;
}

}  // namespace
}  // namespace pbrt

using namespace pbrt;

extern "C" {

// Icosphere of `subdiv` midpoint subdivisions (10*4^s + 2 vertices, 20*4^s triangles), radius r,
// centred at `center`. Faces wind so that the reference's mesh normal -normalize(cross(e1, e2))
// (triangle_mesh.rs:119) points outward. Displacement along the radius:
//   noise_kind 0: none; 1: per-vertex uniform in [-amp, amp] * r (seeded by vertex index);
//   2: fBm value noise of `octaves` octaves at base frequency `freq` (per unit of radius), * amp * r.
int pbrt_gen_icosphere(uint32_t subdiv, double radius, const double center[3], int noise_kind, double amp,
                       int octaves, double freq, uint64_t seed, float* xyz_out, uint32_t* idx_out, uint32_t* nv_out,
                       uint32_t* nt_out) {
  return guarded([&] {
    if (subdiv > 10) throw Error(PBRT_ERR_INVALID_ARG, "subdiv above 10");
    const uint64_t nv = 10ull * (1ull << (2 * subdiv)) + 2, nt = 20ull * (1ull << (2 * subdiv));
    if (nv_out) *nv_out = (uint32_t)nv;
    if (nt_out) *nt_out = (uint32_t)nt;
    if (!xyz_out || !idx_out) return;  // size query
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<D3> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : v) p = nrm(p);
    // counter-clockwise seen from outside
    std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6,
                               7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7,
                               9, 8, 1};
    v.reserve(nv);
    for (uint32_t s = 0; s < subdiv; ++s) {
      std::unordered_map<uint64_t, uint32_t> mid;
      mid.reserve(f.size());
      auto midpoint = [&](uint32_t a, uint32_t b) {
        const uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
        auto it = mid.find(key);
        if (it != mid.end()) return it->second;
        const uint32_t id = (uint32_t)v.size();
        v.push_back(nrm(scl(add(v[a], v[b]), 0.5)));
        mid.emplace(key, id);
        return id;
      };
      std::vector<uint32_t> g;
      g.reserve(f.size() * 4);
      for (size_t i = 0; i < f.size(); i += 3) {
        const uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
        const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
        const uint32_t tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
        g.insert(g.end(), tri, tri + 12);
      }
      f.swap(g);
    }
    if (v.size() != nv || f.size() != 3 * nt) throw Error(PBRT_ERR_INTERNAL, "icosphere size mismatch");
    for (uint64_t i = 0; i < nv; ++i) {
      double s = 1.0;
      if (noise_kind == 1) s = 1.0 + amp * (to_unit(hash4(seed, i, 0x1c05, 0)) * 2.0 - 1.0);
      else if (noise_kind == 2) s = 1.0 + amp * fbm(scl(v[i], freq), octaves, seed);
      const D3 p = add(scl(v[i], radius * s), {center[0], center[1], center[2]});
      xyz_out[3 * i] = (float)p.x;
      xyz_out[3 * i + 1] = (float)p.y;
      xyz_out[3 * i + 2] = (float)p.z;
    }
    // clockwise from outside, so that -cross(e1, e2) is the outward normal
    for (uint64_t i = 0; i < nt; ++i) {
      idx_out[3 * i] = f[3 * i];
      idx_out[3 * i + 1] = f[3 * i + 2];
      idx_out[3 * i + 2] = f[3 * i + 1];
    }
  });
}

int pbrt_pinhole_rays(uint32_t width, uint32_t height, double fov, const double eye[3], const double look[3],
                      float tmin, float tmax, pbrt_ray* out) {
  return guarded([&] {
    if (!width || !height || !out) throw Error(PBRT_ERR_INVALID_ARG, "bad size or null output");
    const Pinhole cam(width, height, fov, (double)tmin, (double)tmax,
                      look_at({eye[0], eye[1], eye[2]}, {look[0], look[1], look[2]}, {0, 1, 0}));
    for (uint32_t y = 0; y < height; ++y)
      for (uint32_t x = 0; x < width; ++x) {
        D3 o, d;
        cam.ray((double)x + 0.5, (double)y + 0.5, &o, &d);
        put_ray(out + (size_t)y * width + x, o, d, tmin, tmax);
      }
  });
}

// The bvh_stats framing (bvh_stats.rs:345-401) of a vertex set: camera `view` of the 6-view orbit,
// side x side pixel centres, FOV `fov`, 20 degrees of elevation, tangent to the bounding sphere.
int pbrt_orbit_rays(const float* xyz, uint32_t nv, uint32_t view, uint32_t side, double fov, float tmin, float tmax,
                    pbrt_ray* out) {
  return guarded([&] {
    if (!xyz || !nv || !out || !side) throw Error(PBRT_ERR_INVALID_ARG, "bad argument");
    double lo[3] = {1e308, 1e308, 1e308}, hi[3] = {-1e308, -1e308, -1e308};
    for (uint32_t i = 0; i < nv; ++i)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::fmin((double)xyz[3 * i + k], lo[k]);
        hi[k] = std::fmax((double)xyz[3 * i + k], hi[k]);
      }
    const D3 bmin{lo[0], lo[1], lo[2]}, bmax{hi[0], hi[1], hi[2]};
    const D3 center = scl(add(bmax, bmin), 0.5);
    const D3 diag = sub(bmax, bmin);
    const double radius = std::sqrt(dot3(diag, diag)) * 0.5;
    const double distance = radius / std::sin(fov / 2.0);
    const double az = 2.0 * M_PI * (double)view / 6.0;
    const double el = 20.0 * M_PI / 180.0;
    const D3 dir{std::cos(el) * std::cos(az), std::sin(el), std::cos(el) * std::sin(az)};
    const D3 pos = add(center, scl(dir, distance));
    const Pinhole cam(side, side, fov, 1e-4, 1000.0, look_at(pos, center, {0, 1, 0}));
    for (uint32_t y = 0; y < side; ++y)
      for (uint32_t x = 0; x < side; ++x) {
        D3 o, d;
        cam.ray((double)x + 0.5, (double)y + 0.5, &o, &d);
        put_ray(out + (size_t)y * side + x, o, d, tmin, tmax);
      }
  });
}

// Incoherent rays: origins uniform in the ball (center, radius), directions uniform on the sphere
// by the map of src/pdfs/sphere.rs:14-23 (phi = 2 pi u1, cos theta = 1 - 2 u2), seeded per ray.
int pbrt_random_rays(const double center[3], double radius, uint64_t seed, uint64_t first, uint32_t n, float tmin,
                     float tmax, pbrt_ray* out) {
  return guarded([&] {
    if (!out) throw Error(PBRT_ERR_INVALID_ARG, "null output");
    for (uint32_t i = 0; i < n; ++i) {
      const uint64_t k = first + i;
      auto u = [&](uint64_t slot) { return to_unit(hash4(seed, k, slot, 0x7a11)); };
      // direction
      const double phi = 2.0 * M_PI * u(0);
      const double ct = 1.0 - 2.0 * u(1);
      const double st = std::sqrt(1.0 - ct * ct);
      const D3 d = nrm({std::cos(phi) * st, std::sin(phi) * st, ct});
      // origin: uniform in the ball (direction uniform, radius r * cbrt(u))
      const double phi2 = 2.0 * M_PI * u(2);
      const double ct2 = 1.0 - 2.0 * u(3);
      const double st2 = std::sqrt(1.0 - ct2 * ct2);
      const double rr = radius * std::cbrt(u(4));
      const D3 o{center[0] + rr * std::cos(phi2) * st2, center[1] + rr * std::sin(phi2) * st2, center[2] + rr * ct2};
      put_ray(out + i, o, d, tmin, tmax);
    }
  });
}

}  // extern "C"
