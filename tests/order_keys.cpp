// Host driver for tests/test_target_order.py: the Morton keys of the
// library's target order (lpm_amd/csrc/order_keys.hpp, used by order.cpp)
// for points read from a file, and a bit-by-bit check of the interleaving.
//   order_keys <in: n x 3 doubles> <out: n uint64 keys>
#include <cstdio>
#include <random>
#include <vector>

#include "order_keys.hpp"

using lpm_amd::morton_key;
using lpm_amd::spread3;

static uint64_t naive_spread(uint64_t v) {
  uint64_t r = 0;
  for (int b = 0; b < 20; ++b)
    if ((v >> b) & 1) r |= uint64_t(1) << (3 * b);
  return r;
}

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  std::mt19937_64 g(11);
  for (int i = 0; i < 100000; ++i) {
    const uint64_t v = g() & 0xFFFFF;
    if (spread3(v) != naive_spread(v)) {
      std::fprintf(stderr, "spread3 mismatch at %llu\n",
                   static_cast<unsigned long long>(v));
      return 3;
    }
  }
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 4;
  std::vector<double> p;
  double b[3];
  while (std::fread(b, sizeof(double), 3, f) == 3) p.insert(p.end(), b, b + 3);
  std::fclose(f);
  const size_t n = p.size() / 3;
  std::vector<uint64_t> k(n);
  for (size_t i = 0; i < n; ++i) k[i] = morton_key(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 5;
  std::fwrite(k.data(), sizeof(uint64_t), n, o);
  std::fclose(o);
  return 0;
}
