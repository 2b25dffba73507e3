// Morton (Z-order) keys of the target order (order.cpp).  Host code with no
// dependencies, so that a test driver can include it.
#pragma once

#include <cstdint>

namespace lpm_amd {

// Spreads the low 20 bits of v to every third bit (bit b -> bit 3 b).
inline uint64_t spread3(uint64_t v) {
  v &= 0xFFFFF;
  v = (v | (v << 32)) & 0x001F00000000FFFFull;
  v = (v | (v << 16)) & 0x001F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// Morton key of a point of [-1, 1]^3: 20 bits per coordinate, interleaved.
inline uint64_t morton_key(double x, double y, double z) {
  auto q = [](double c) -> uint64_t {
    if (!(c > -1.0)) return 0;  // also NaN
    if (c >= 1.0) return (1u << 20) - 1;
    return static_cast<uint64_t>((c + 1.0) * 0.5 * (1u << 20));
  };
  return spread3(q(x)) | (spread3(q(y)) << 1) | (spread3(q(z)) << 2);
}

}  // namespace lpm_amd
