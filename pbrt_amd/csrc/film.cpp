// Film output: the FP64 film to 8-bit RGBA the reference's way, and a PNG file of it.
//
// Conversion (src/spectrum.rs:13-30, src/renderers/mt.rs:83-90): per channel sqrt (gamma 2),
// clamp to [0, 1] with NaN-ignoring max / min (a NaN becomes 0, as f64::max does), x 255, truncate
// toward zero; alpha 255. Optional tone mapping (not in the reference): Reinhard x / (1 + x) on each
// channel before the gamma.
//
// PNG (the reference writes 8-bit RGBA through the png crate, renderers/mt.rs:152-165): signature,
// IHDR, one IDAT holding a zlib stream of stored (uncompressed) deflate blocks over scanlines with
// filter 0, IEND. Stored blocks make the file a pure function of the pixels — the same bytes on
// every run and host — at the cost of size; any PNG reader decodes the same pixels as the
// reference's compressed file would give.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace pbrt {
namespace {

uint8_t channel8(double v, bool tonemap) {
  if (tonemap) v = v / (1.0 + v);
  const double g = std::sqrt(v);
  const double c = std::fmin(1.0, std::fmax(0.0, g));
  return (uint8_t)(c * 255.0);
}

void to_rgba8(const double* film, uint32_t w, uint32_t h, bool tonemap, uint8_t* out) {
  const size_t n = (size_t)w * h;
  for (size_t i = 0; i < n; ++i) {
    out[4 * i + 0] = channel8(film[3 * i + 0], tonemap);
    out[4 * i + 1] = channel8(film[3 * i + 1], tonemap);
    out[4 * i + 2] = channel8(film[3 * i + 2], tonemap);
    out[4 * i + 3] = 255;
  }
}

// CRC-32 (ISO 3309, polynomial 0xEDB88320, reflected), as PNG chunks require
struct CrcTable {
  uint32_t t[256];
  CrcTable() {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      t[i] = c;
    }
  }
};
uint32_t crc32_update(uint32_t crc, const uint8_t* p, size_t n) {
  static const CrcTable tab;  // initialised once, thread-safe
  const uint32_t* table = tab.t;
  crc = ~crc;
  for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
  return ~crc;
}

void put32(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back((uint8_t)(x >> 24));
  v.push_back((uint8_t)(x >> 16));
  v.push_back((uint8_t)(x >> 8));
  v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t>& png, const char type[4], const std::vector<uint8_t>& data) {
  put32(png, (uint32_t)data.size());
  const size_t start = png.size();
  png.insert(png.end(), type, type + 4);
  png.insert(png.end(), data.begin(), data.end());
  put32(png, crc32_update(0, png.data() + start, png.size() - start));
}

std::vector<uint8_t> encode_png(const uint8_t* rgba, uint32_t w, uint32_t h) {
  // raw scanlines: filter byte 0, then the row
  const size_t row = (size_t)w * 4;
  std::vector<uint8_t> raw;
  raw.reserve((row + 1) * h);
  for (uint32_t y = 0; y < h; ++y) {
    raw.push_back(0);
    raw.insert(raw.end(), rgba + y * row, rgba + (y + 1) * row);
  }
  // zlib stream (RFC 1950): CMF 0x78 (deflate, 32K window), FLG 0x01 (no dictionary, fastest;
  // 0x7801 is a multiple of 31), stored blocks of at most 65,535 bytes (RFC 1951 3.2.4), Adler-32
  std::vector<uint8_t> z = {0x78, 0x01};
  size_t off = 0;
  do {
    const size_t len = std::min<size_t>(65535, raw.size() - off);
    const bool last = off + len == raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back((uint8_t)(len & 0xFF));
    z.push_back((uint8_t)(len >> 8));
    z.push_back((uint8_t)(~len & 0xFF));
    z.push_back((uint8_t)((~len >> 8) & 0xFF));
    z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
    off += len;
  } while (off < raw.size());
  uint32_t a = 1, b = 0;
  for (uint8_t c : raw) {
    a = (a + c) % 65521u;
    b = (b + a) % 65521u;
  }
  put32(z, (b << 16) | a);

  std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  std::vector<uint8_t> ihdr;
  put32(ihdr, w);
  put32(ihdr, h);
  ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // bit depth 8, colour type 6 (RGBA), deflate, filter 0, no interlace
  chunk(png, "IHDR", ihdr);
  chunk(png, "IDAT", z);
  chunk(png, "IEND", {});
  return png;
}

void check_film(const double* film, uint32_t w, uint32_t h) {
  if (!film) throw Error(PBRT_ERR_INVALID_ARG, "null film");
  if (!w || !h || (uint64_t)w * h > (1ull << 28)) throw Error(PBRT_ERR_INVALID_ARG, "film size out of range");
}

}  // namespace
}  // namespace pbrt

using namespace pbrt;

int pbrt_film_to_rgba8(const double* film, uint32_t width, uint32_t height, uint32_t tonemap, uint8_t* rgba_out) {
  return guarded([&] {
    check_film(film, width, height);
    if (!rgba_out) throw Error(PBRT_ERR_INVALID_ARG, "null output");
    if (tonemap > 1) throw Error(PBRT_ERR_INVALID_ARG, "tonemap must be 0 (none) or 1 (Reinhard)");
    to_rgba8(film, width, height, tonemap == 1, rgba_out);
  });
}

int pbrt_film_write_png(const double* film, uint32_t width, uint32_t height, uint32_t tonemap, const char* path) {
  return guarded([&] {
    check_film(film, width, height);
    if (!path) throw Error(PBRT_ERR_INVALID_ARG, "null path");
    if (tonemap > 1) throw Error(PBRT_ERR_INVALID_ARG, "tonemap must be 0 (none) or 1 (Reinhard)");
    std::vector<uint8_t> rgba((size_t)width * height * 4);
    to_rgba8(film, width, height, tonemap == 1, rgba.data());
    const std::vector<uint8_t> png = encode_png(rgba.data(), width, height);
    FILE* f = std::fopen(path, "wb");
    if (!f) throw Error(PBRT_ERR_INVALID_ARG, std::string("cannot open for writing: ") + path);
    const size_t wrote = std::fwrite(png.data(), 1, png.size(), f);
    const int closed = std::fclose(f);
    if (wrote != png.size() || closed != 0) throw Error(PBRT_ERR_INVALID_ARG, std::string("write failed: ") + path);
  });
}
