// Evaluates the C++ gallery functors (lpm/galleries.hpp) at points read from
// a binary file of n x 3 doubles and writes NOUT doubles per point:
//   0      SolidBodyRotation vorticity          1-3   its velocity
//   4      GaussianVortexSphere(2.5, 3, 0.7, -0.4), gauss_const from 1.3
//   5      RossbyHaurwitz54(0.1, -0.1)           6     its laplacian
//   7-9    RossbyWave54Velocity of that wave
//   10     RossbyHaurwitz54(0, 1) after set_stationary_wave_speed()
//   11-13  SWETestCase2(2, 2 pi, pi / 6) u0     14, 15, 16  zeta0, sfc0, sigma0
//   17     SphereTestCase2InitialSurface
// Usage: galleries_eval in.bin out.bin
#include <array>
#include <cstdio>
#include <vector>

#include "lpm/galleries.hpp"

using namespace Lpm;

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  std::FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 3;
  std::vector<double> x;
  double buf[3];
  while (std::fread(buf, sizeof(double), 3, f) == 3)
    x.insert(x.end(), buf, buf + 3);
  std::fclose(f);
  const size_t n = x.size() / 3;

  const SolidBodyRotation sbr;
  GaussianVortexSphere gauss(2.5, 3.0, 0.7, -0.4);
  gauss.set_gauss_const(1.3);
  const RossbyHaurwitz54 rh(0.1, -0.1);
  const RossbyWave54Velocity rhv(rh);
  RossbyHaurwitz54 rhs;
  rhs.set_stationary_wave_speed();
  const SWETestCase2 tc2(2.0);
  const SphereTestCase2InitialSurface surf;

  constexpr int NOUT = 18;
  std::vector<double> out(n * NOUT);
  for (size_t i = 0; i < n; ++i) {
    const std::array<Real, 3> p{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    double* o = &out[i * NOUT];
    o[0] = sbr(p[0], p[1], p[2]);
    sbr.init_velocity(o[1], o[2], o[3], p[0], p[1], p[2]);
    o[4] = gauss(p);
    o[5] = rh(p[0], p[1], p[2]);
    o[6] = rh.laplacian(p);
    const auto u = rhv(p.data());
    o[7] = u[0];
    o[8] = u[1];
    o[9] = u[2];
    o[10] = rhs(p);
    std::array<Real, 3> v;
    tc2.u0(v, p);
    o[11] = v[0];
    o[12] = v[1];
    o[13] = v[2];
    o[14] = tc2.zeta0(p);
    o[15] = tc2.sfc0(p);
    o[16] = tc2.sigma0(p);
    o[17] = surf(p);
    // the point and scalar overloads agree
    if (sbr(p) != o[0] || gauss(p[0], p[1], p[2]) != o[4] || rh(p) != o[5] ||
        rh.laplacian(p[0], p[1], p[2]) != o[6])
      return 4;
  }
  f = std::fopen(argv[2], "wb");
  if (!f) return 3;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  return 0;
}
