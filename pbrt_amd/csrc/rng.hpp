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

}  // namespace pbrt
