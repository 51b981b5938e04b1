// ORACLE — test infrastructure, not product code.
//
// CPU restatement of the neural radiance cache the device trains and queries. The reference has
// no such component (its integrator is plain path tracing, src/integrators/path.rs), so this is
// written from the definition it implements, and pinned by closed forms (finite differences of the
// loss against the analytic backward pass; Adam against its published update rule, Kingma & Ba
// 2015) rather than by reference outputs.
//
// Network: 81 inputs (position under a 12-frequency sin/cos encoding, direction, normal, albedo)
// -> 4 hidden layers of 64 ReLU units -> 3 linear outputs (RGB radiance). Parameters are FP32
// masters; a forward pass reads them rounded to BF16 (round to nearest even), rounds every layer's
// input activations to BF16, and accumulates in FP64 here (FP32 on the device). Loss: relative L2,
// mean over samples and channels of (y - t)^2 / (sg(y)^2 + 0.01) with sg() held constant
// (Müller et al., "Real-time Neural Radiance Caching for Path Tracing", 2021).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {
namespace nrc {

constexpr int kIn = 81, kHidden = 64, kOut = 3, kLayers = 5;
constexpr int kFreq = 12;
constexpr int kInDim[kLayers] = {kIn, kHidden, kHidden, kHidden, kHidden};
constexpr int kOutDim[kLayers] = {kHidden, kHidden, kHidden, kHidden, kOut};

// Offsets of W_l ([out][in], row-major) and b_l in the flat parameter vector.
inline int w_off(int l) {
  int o = 0;
  for (int i = 0; i < l; ++i) o += kInDim[i] * kOutDim[i] + kOutDim[i];
  return o;
}
inline int b_off(int l) { return w_off(l) + kInDim[l] * kOutDim[l]; }
inline int n_params() { return w_off(kLayers); }

inline float bf16_round(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / NaN unchanged
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}
inline float bf16_from_bits(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

// sin(pi x) for x = 2^k p with p an FP32 value: x is exact in FP64, the period and the symmetries
// sin(pi (x - 1)) = -sin(pi x), sin(pi (1 - x)) = sin(pi x) reduce it exactly to [0, 1/2], where
// one FP64 sin is taken (so the exact zeros and ones of the function come out exact).
inline double sinpi(double x) {
  x = std::fmod(x, 2.0);
  if (x < 0.0) x += 2.0;
  double sign = 1.0;
  if (x >= 1.0) {
    x -= 1.0;
    sign = -1.0;
  }
  if (x > 0.5) x = 1.0 - x;
  return sign * std::sin(M_PI * x);
}
inline double cospi(double x) { return sinpi(x + 0.5); }

// The 81 input features of one sample from position, outgoing direction, normal and albedo. The
// position is normalised by the scene box in FP32, p_a = (x_a - lo_a) * (1 / (hi_a - lo_a)), and
// that FP32 value is encoded exactly: for each axis a, for each frequency k, sin(2^k pi p_a) and
// cos(2^k pi p_a), evaluated here in FP64 and rounded once.
inline void encode(const float* raw12, const float* lo, const float* hi, float* out81) {
  int j = 0;
  for (int a = 0; a < 3; ++a) {
    const float inv = 1.0f / (hi[a] - lo[a]);
    const float p = (raw12[a] - lo[a]) * inv;
    for (int k = 0; k < kFreq; ++k) {
      const double x = std::ldexp((double)p, k);
      out81[j++] = (float)sinpi(x);
      out81[j++] = (float)cospi(x);
    }
  }
  for (int i = 3; i < 12; ++i) out81[j++] = raw12[i];
}

struct Forward {
  std::vector<double> act[kLayers + 1];  // act[0] = inputs, act[l+1] = output of layer l (post-ReLU)
  std::vector<double> pre[kLayers];      // pre-activations
};

// One sample's forward pass. `bf16` emulates the device's rounding points.
inline void forward(const float* params, const float* x81, bool bf16, Forward* f) {
  f->act[0].resize(kIn);
  for (int i = 0; i < kIn; ++i) f->act[0][i] = bf16 ? bf16_round(x81[i]) : x81[i];
  for (int l = 0; l < kLayers; ++l) {
    const int ni = kInDim[l], no = kOutDim[l];
    const float* W = params + w_off(l);
    const float* b = params + b_off(l);
    f->pre[l].assign(no, 0.0);
    f->act[l + 1].assign(no, 0.0);
    for (int o = 0; o < no; ++o) {
      double s = 0.0;
      for (int i = 0; i < ni; ++i) s += (double)(bf16 ? bf16_round(W[o * ni + i]) : W[o * ni + i]) * f->act[l][i];
      s += (double)b[o];
      f->pre[l][o] = s;
      double a = (l == kLayers - 1) ? s : (s > 0.0 ? s : 0.0);
      if (bf16 && l < kLayers - 1) a = bf16_round((float)a);
      f->act[l + 1][o] = a;
    }
  }
}

// Loss and gradient over a batch (inputs n x 81, targets n x 3). grad has n_params() entries.
inline double loss_grad(const float* params, const float* x, const float* targets, int n, bool bf16,
                        std::vector<double>* grad) {
  grad->assign(n_params(), 0.0);
  double loss = 0.0;
  Forward f;
  const double inv = 1.0 / ((double)n * kOut);
  for (int s = 0; s < n; ++s) {
    forward(params, x + (size_t)s * kIn, bf16, &f);
    // dz: the gradient at a layer's pre-activation; with `bf16` the bias gradient and the two
    // products (weights, previous activations) take it rounded to BF16, as the device's MFMAs do
    // (the bias gradient is dz^T 1 on the matrix cores).
    std::vector<double> dz(kOut), dzr(kOut);
    for (int c = 0; c < kOut; ++c) {
      const double y = f.act[kLayers][c], t = targets[(size_t)s * kOut + c];
      const double den = y * y + 0.01;
      loss += (y - t) * (y - t) / den * inv;
      dz[c] = 2.0 * (y - t) / den * inv;
      if (bf16) dz[c] = (float)dz[c];
      dzr[c] = bf16 ? bf16_round((float)dz[c]) : dz[c];
    }
    for (int l = kLayers - 1; l >= 0; --l) {
      const int ni = kInDim[l], no = kOutDim[l];
      const float* W = params + w_off(l);
      for (int o = 0; o < no; ++o) {
        (*grad)[b_off(l) + o] += dzr[o];
        for (int i = 0; i < ni; ++i) (*grad)[w_off(l) + o * ni + i] += dzr[o] * f.act[l][i];
      }
      if (l == 0) break;
      std::vector<double> da(ni, 0.0);
      for (int i = 0; i < ni; ++i) {
        double s2 = 0.0;
        for (int o = 0; o < no; ++o) s2 += (double)(bf16 ? bf16_round(W[o * ni + i]) : W[o * ni + i]) * dzr[o];
        da[i] = s2;
      }
      dz.assign(ni, 0.0);
      dzr.assign(ni, 0.0);
      for (int i = 0; i < ni; ++i) {
        double g = f.act[l][i] > 0.0 ? da[i] : 0.0;  // ReLU' from the stored (post-ReLU) activation
        if (bf16) g = (float)g;
        dz[i] = g;
        dzr[i] = bf16 ? bf16_round((float)g) : g;
      }
    }
  }
  return loss;
}

// Adam (Kingma & Ba 2015, Algorithm 1) at step t (1-based), FP32 state as on the device.
inline void adam(float* params, float* m, float* v, const float* grad, int n, int t, float lr, float b1, float b2,
                 float eps) {
  const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
  for (int i = 0; i < n; ++i) {
    m[i] = b1 * m[i] + (1.0f - b1) * grad[i];
    v[i] = b2 * v[i] + (1.0f - b2) * grad[i] * grad[i];
    const float mh = m[i] / c1, vh = v[i] / c2;
    params[i] = params[i] - lr * mh / (std::sqrt(vh) + eps);
  }
}

}  // namespace nrc
}  // namespace orc
