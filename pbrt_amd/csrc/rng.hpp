// Counter-based random numbers shared by host and device code.
//
// The reference draws every uniform from an unseeded thread RNG (src/utils.rs:5-7), so no image of
// it is reproducible. This path keys every draw by a counter instead — (seed, pixel, sample, bounce,
// slot) — so that the host oracle and the device kernels draw the same numbers whatever the order
// paths are processed in, as the reference's own plan for a seeded sampler asks
// (ideas/rng_graine.md:65-84: reseed per pixel sample from a hash of the seed and the pixel).
// The mixer is the SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014), applied once per key word.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define PBRT_HD __host__ __device__ __forceinline__
#else
#define PBRT_HD inline
#endif

namespace pbrt {

PBRT_HD uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

PBRT_HD uint64_t hash4(uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  return mix64(mix64(mix64(mix64(a) ^ b) ^ c) ^ d);
}

// 53 high bits -> [0, 1)
PBRT_HD double to_unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }
// 24 high bits -> [0, 1) in FP32, exactly (rounding the 53-bit value to FP32 could give 1.0)
PBRT_HD float to_unit_f32(uint64_t h) { return (float)(h >> 40) * 0x1p-24f; }

// ---- stratified draws: scrambled Sobol (0,2)-sequences ------------------------------------------
// With the Sobol sampler a few 2D draws of every path (the pixel jitter; at vertices 1 and 2 the
// light sample's two uniforms and the BSDF sample's two) come, for pixel p and its sample index s,
// from a 2D Sobol (0,2)-sequence with Owen scrambling by hashing (Burley, "Practical Hash-based Owen
// Scrambling", JCGT 9(4), 2020): the index is shuffled by a nested uniform scramble, each coordinate
// scrambled the same way, every (pixel, pair) with its own seeds. The first 2^m samples of a pixel
// then form a (0,m,2)-net in each pair (one point in every elementary interval of area 2^-m), which
// is what lowers the variance of edges and of direct lighting at a fixed number of samples; pairs
// are decorrelated by their independent shuffles. Every other draw stays the counter-based one.

PBRT_HD uint32_t reverse_bits32(uint32_t x) {
  x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
  x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
  x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
  x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
  return (x >> 16) | (x << 16);
}
// The Laine-Karras style hash of Burley's paper: every step leaves bit i a function of bits <= i
// (additions and products by even constants carry upward only), so on reversed bits it is an Owen
// scramble: a digit flips as a function of the coarser digits only.
PBRT_HD uint32_t lk_hash(uint32_t x, uint32_t seed) {
  x += seed;
  x ^= x * 0x6c50b47cu;
  x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u;
  x ^= x * 0x8d22f6e6u;
  return x;
}
PBRT_HD uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
  return reverse_bits32(lk_hash(reverse_bits32(x), seed));
}
// The second dimension of the Sobol sequence: generator columns v_0 = 2^31, v_k = v_{k-1} ^ (v_{k-1} >> 1)
// (the first is the van der Corput radical inverse, reverse_bits32).
PBRT_HD uint32_t sobol_dim1(uint32_t i) {
  uint32_t r = 0, v = 0x80000000u;
  for (; i; i >>= 1, v ^= v >> 1)
    if (i & 1u) r ^= v;
  return r;
}
// The stratified pair a draw belongs to (-1: a counter-based draw) and its coordinate: the jitter
// (bounce 0, slots 0 / 1) is pair 0; at bounce b = 1, 2 the light sample (slots 3 / 4) is pair
// 2b - 1 and the BSDF sample (slots 5 / 6) pair 2b.
PBRT_HD int sobol_pair(uint32_t bounce, uint32_t slot, uint32_t* axis) {
  if (bounce == 0 && slot <= 1) {
    *axis = slot;
    return 0;
  }
  if (bounce >= 1 && bounce <= 2 && slot >= 3 && slot <= 6) {
    *axis = (slot - 3) & 1u;
    return (int)(2 * bounce - 1 + (slot >= 5 ? 1 : 0));
  }
  return -1;
}
// 32 bits of the coordinate `axis` of sample `sample` of pixel `pixel` in stratified pair `pair`.
PBRT_HD uint32_t sobol_draw32(uint64_t seed, uint64_t pixel, uint32_t sample, uint32_t pair, uint32_t axis) {
  const uint64_t h = hash4(seed, pixel, 0x50B0Cu + pair, 0);
  const uint32_t shuffled = nested_uniform_scramble(sample, (uint32_t)h);
  const uint32_t x = axis == 0 ? reverse_bits32(shuffled) : sobol_dim1(shuffled);
  return nested_uniform_scramble(x, (uint32_t)(h >> 32) ^ (axis ? 0x9E3779B9u : 0u));
}

}  // namespace pbrt
