// Output encoders: LinColor layers to 8-bit sRGB RGBA (the step after the
// fine stage), and the BMP / PNG containers the reference writes them in.
//   k_encode_rgba8 -- RGBA::from(LinColor) per pixel (src/color.rs:164-175)
//   bmp_header     -- ImageRead::write_bmp's BITMAPV4HEADER (src/image.rs:164-213)
//   png_encode     -- ImageRead::write_png (src/image.rs:215-235), stored deflate
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <vector>
#include "rz_paint.hpp"

namespace rze {

// Rust's `f as u8`: truncation, saturating at 0 and 255, NaN to 0.
RZ_HD uint8_t sat_u8(float f) {
  if (!(f > 0.0f)) return 0;
  if (f >= 255.0f) return 255;
  return (uint8_t)f;
}

// RGBA::from(LinColor): l2s of the unmultiplied colour on every lane, each
// channel (v * 255 + 0.5) as u8; alpha from the premultiplied alpha.
RZ_HD uint32_t rgba8_of(rzp::C4 lin) {
  const rzp::C4 s = rzp::l2s(rzp::unmultiply(lin));
  const uint32_t r = sat_u8(s.r * 255.0f + 0.5f), g = sat_u8(s.g * 255.0f + 0.5f);
  const uint32_t b = sat_u8(s.b * 255.0f + 0.5f), a = sat_u8(lin.a * 255.0f + 0.5f);
  return r | (g << 8) | (b << 16) | (a << 24);  // bytes R, G, B, A
}

// One thread per pixel, consecutive threads along a row (16-byte loads, 4-byte
// stores, both coalesced). `flip` writes row y to row h - 1 - y (BMP order).
__global__ __launch_bounds__(256) void k_encode_rgba8(const rzp::C4* img, size_t stride_px, int w,
                                                      int h, uint8_t* out, size_t out_stride,
                                                      int flip) {
  const size_t n = (size_t)w * h;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t y = i / (size_t)w, x = i - y * (size_t)w;
    const size_t yo = flip ? (size_t)h - 1 - y : y;
    *reinterpret_cast<uint32_t*>(out + yo * out_stride + 4 * x) = rgba8_of(img[y * stride_px + x]);
  }
}

// BMP file header (14 bytes) and BITMAPV4HEADER (108 bytes) of a w x h
// 32-bit image: BI_BITFIELDS with the masks of bytes R, G, B, A, 2835 px/m,
// the sRGB colour space; the pixel rows follow, last row first.
constexpr size_t kBmpHeader = 122;
inline void bmp_header(uint32_t w, uint32_t h, uint8_t out[kBmpHeader]) {
  std::memset(out, 0, kBmpHeader);
  auto u32 = [&](size_t o, uint32_t v) {
    for (int k = 0; k < 4; ++k) out[o + k] = (uint8_t)(v >> (8 * k));
  };
  const uint32_t data = w * h * 4, off = (uint32_t)kBmpHeader;
  out[0] = 'B';
  out[1] = 'M';
  u32(2, data + off);
  u32(10, off);
  u32(14, 108);
  u32(18, w);
  u32(22, h);
  out[26] = 1;     // planes
  out[28] = 32;    // bits per pixel
  out[30] = 3;     // BI_BITFIELDS
  u32(34, data);
  u32(38, 2835);
  u32(42, 2835);
  u32(54, 0x000000FFu);  // red mask
  u32(58, 0x0000FF00u);  // green mask
  u32(62, 0x00FF0000u);  // blue mask
  u32(66, 0xFF000000u);  // alpha mask
  out[70] = 0x42;  // LCS_sRGB, 'sRGB' little-endian
  out[71] = 0x47;
  out[72] = 0x52;
  out[73] = 0x73;
}

// PNG of 8-bit RGBA rows: signature, IHDR, one IDAT holding a zlib stream of
// stored (uncompressed) deflate blocks over the filter-0 rows, IEND. CRC-32
// (ISO 3309 / ITU-T V.42, polynomial 0xEDB88320) and Adler-32 (RFC 1950)
// computed here.
inline uint32_t crc32_update(uint32_t c, const uint8_t* p, size_t n) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t v = i;
      for (int k = 0; k < 8; ++k) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
      table[i] = v;
    }
    init = true;
  }
  for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
  return c;
}

inline size_t png_size(size_t w, size_t h) {
  const size_t raw = h * (1 + 4 * w);
  const size_t blocks = raw ? (raw + 65534) / 65535 : 1;
  return 8 + (12 + 13) + (12 + 2 + raw + 5 * blocks + 4) + 12;
}

inline void png_encode(const uint8_t* rgba, size_t w, size_t h, size_t stride, uint8_t* out) {
  uint8_t* p = out;
  auto be32 = [&](uint32_t v) {
    p[0] = (uint8_t)(v >> 24);
    p[1] = (uint8_t)(v >> 16);
    p[2] = (uint8_t)(v >> 8);
    p[3] = (uint8_t)v;
    p += 4;
  };
  static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
  std::memcpy(p, sig, 8);
  p += 8;
  auto chunk_end = [&](uint8_t* type_at) {
    const uint32_t c = crc32_update(0xFFFFFFFFu, type_at, (size_t)(p - type_at)) ^ 0xFFFFFFFFu;
    be32(c);
  };
  // IHDR
  be32(13);
  uint8_t* t = p;
  std::memcpy(p, "IHDR", 4);
  p += 4;
  be32((uint32_t)w);
  be32((uint32_t)h);
  *p++ = 8;  // bit depth
  *p++ = 6;  // colour type RGBA
  *p++ = 0;  // deflate
  *p++ = 0;  // adaptive filtering
  *p++ = 0;  // no interlace
  chunk_end(t);
  // IDAT: zlib header, stored blocks of the rows (filter byte 0 each)
  const size_t raw = h * (1 + 4 * w);
  const size_t blocks = raw ? (raw + 65534) / 65535 : 1;
  be32((uint32_t)(2 + raw + 5 * blocks + 4));
  t = p;
  std::memcpy(p, "IDAT", 4);
  p += 4;
  *p++ = 0x78;  // CMF: deflate, 32 KiB window
  *p++ = 0x01;  // FLG: (0x78 << 8 | 0x01) % 31 == 0, no dictionary
  uint32_t s1 = 1, s2 = 0;  // Adler-32
  size_t left = raw, y = 0, x = 0;  // x: byte within the row incl. its filter byte
  if (!raw) {
    *p++ = 1;  // a final empty stored block
    *p++ = 0;
    *p++ = 0;
    *p++ = 0xFF;
    *p++ = 0xFF;
  }
  while (left) {
    const size_t n = left < 65535 ? left : 65535;
    left -= n;
    *p++ = left ? 0 : 1;  // BFINAL, BTYPE = 00
    *p++ = (uint8_t)n;
    *p++ = (uint8_t)(n >> 8);
    *p++ = (uint8_t)~n;
    *p++ = (uint8_t)(~n >> 8);
    for (size_t k = 0; k < n; ++k) {
      const uint8_t v = x == 0 ? 0 : rgba[y * stride + (x - 1)];
      *p++ = v;
      s1 = (s1 + v) % 65521u;
      s2 = (s2 + s1) % 65521u;
      if (++x == 1 + 4 * w) {
        x = 0;
        ++y;
      }
    }
  }
  be32((s2 << 16) | s1);
  chunk_end(t);
  // IEND
  be32(0);
  t = p;
  std::memcpy(p, "IEND", 4);
  p += 4;
  chunk_end(t);
}

}  // namespace rze
