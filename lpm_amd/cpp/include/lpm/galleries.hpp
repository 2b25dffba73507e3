// Initial conditions and analytic solutions with the names and arguments of
// LPM's galleries: SolidBodyRotation, GaussianVortexSphere, RossbyHaurwitz54
// (src/lpm_vorticity_gallery.hpp:31-147), RossbyWave54Velocity
// (src/lpm_velocity_gallery.hpp:250-288), SWETestCase2
// (src/lpm_swe_problem_gallery.hpp:283-330) and
// SphereTestCase2InitialSurface (src/lpm_surface_gallery.hpp:120-134).
// Host functors: BVESphere / SWE init_* take any (x, y, z) -> Real callable;
// the point overloads take anything indexable as xyz[0..2].
#pragma once

#include <array>
#include <cmath>
#include <string>

#include "lpm.hpp"

namespace Lpm {

namespace constants {
constexpr Real GRAVITY = 9.80616;
}

namespace gallery_detail {
inline Real longitude(Real x, Real y) {
  const Real a = std::atan2(y, x);
  return a < 0 ? a + 2 * constants::PI : a;
}
inline Real latitude(Real x, Real y, Real z) {
  return std::atan2(z, std::sqrt(x * x + y * y));
}
}  // namespace gallery_detail

// Rigid rotation about the z axis with angular speed OMEGA = 2 pi:
// relative vorticity 2 OMEGA z, velocity OMEGA (-y, x, 0).
struct SolidBodyRotation {
  static constexpr Real OMEGA = 2 * constants::PI;
  static constexpr bool IsVorticity = true;

  Real operator()(Real, Real, Real z) const { return 2 * OMEGA * z; }
  template <typename Pt>
  Real operator()(const Pt& xyz) const {
    return 2 * OMEGA * xyz[2];
  }
  std::string name() const { return "rotation"; }
  void init_velocity(Real& u, Real& v, Real& w, Real x, Real y, Real) const {
    u = -OMEGA * y;
    v = OMEGA * x;
    w = 0;
  }
};

// strength exp(-b^2 (1 - x . c)) - gauss_const, centred at (init_lon,
// init_lat); set_gauss_const(sum of zeta A) removes the mean so that the
// total circulation vanishes.
struct GaussianVortexSphere {
  static constexpr bool IsVorticity = true;
  Real gauss_const = 0;
  Real vortex_strength;
  Real shape_parameter;
  std::array<Real, 3> xyz_ctr;

  explicit GaussianVortexSphere(Real str = 4 * constants::PI, Real b = 4,
                                Real init_lon = 0,
                                Real init_lat = constants::PI / 20)
      : vortex_strength(str),
        shape_parameter(b),
        xyz_ctr{std::cos(init_lon) * std::cos(init_lat),
                std::sin(init_lon) * std::cos(init_lat), std::sin(init_lat)} {}

  void set_gauss_const(Real vorticity_sum) {
    gauss_const = vorticity_sum / (4 * constants::PI);
  }
  Real operator()(Real x, Real y, Real z) const {
    const Real d = 1.0 - x * xyz_ctr[0] - y * xyz_ctr[1] - z * xyz_ctr[2];
    return vortex_strength * std::exp(-shape_parameter * shape_parameter * d) -
           gauss_const;
  }
  template <typename Pt>
  Real operator()(const Pt& xyz) const {
    return (*this)(xyz[0], xyz[1], xyz[2]);
  }
  std::string name() const { return "SphericalGaussianVortex"; }
};

// 2 u0 z + 30 amp cos(4 lon) P(z), P(z) = z (z^2 - 1)^2 (the Y_5^4
// harmonic): the Rossby-Haurwitz wave zeta = 2 w sin(lat) - 30 K cos^4(lat)
// sin(lat) cos(4 lon) is u0 = w, amp = -K.
struct RossbyHaurwitz54 {
  static constexpr bool IsVorticity = true;
  Real u0;
  Real rh54_amplitude;

  explicit RossbyHaurwitz54(Real zonal_background_velocity = 0,
                            Real wave_amp = 1)
      : u0(zonal_background_velocity), rh54_amplitude(wave_amp) {}

  // the background rotation for which the wave is stationary, (3 + R) R
  // u0 = 2 Omega with R = 4
  void set_stationary_wave_speed(Real Omega = 2 * constants::PI) {
    u0 = Omega / 14;
  }
  static Real legendreP54(Real z) { return z * (z * z - 1) * (z * z - 1); }
  Real operator()(Real x, Real y, Real z) const {
    return 2 * u0 * z + 30 * rh54_amplitude *
                            std::cos(4 * gallery_detail::longitude(x, y)) *
                            legendreP54(z);
  }
  // Laplacian on the unit sphere: the degree-1 term has eigenvalue -2, the
  // degree-5 term -30.
  Real laplacian(Real x, Real y, Real z) const {
    return -4 * u0 * z - 900 * rh54_amplitude *
                             std::cos(4 * gallery_detail::longitude(x, y)) *
                             legendreP54(z);
  }
  template <typename Pt>
  Real operator()(const Pt& xyz) const {
    return (*this)(xyz[0], xyz[1], xyz[2]);
  }
  template <typename Pt>
  Real laplacian(const Pt& xyz) const {
    return laplacian(xyz[0], xyz[1], xyz[2]);
  }
  std::string name() const { return "RossbyHaurwitz54"; }
};

// Velocity of the RossbyHaurwitz54 vorticity at t = 0 (zonal and meridional
// components of the Y_5^4 stream function plus the background rotation).
struct RossbyWave54Velocity {
  static constexpr int ndim = 3;
  Real background_rotation;
  Real rh54_amplitude;

  explicit RossbyWave54Velocity(Real u0 = 0, Real amp = 1)
      : background_rotation(u0), rh54_amplitude(amp) {}
  explicit RossbyWave54Velocity(const RossbyHaurwitz54& rh)
      : background_rotation(rh.u0), rh54_amplitude(rh.rh54_amplitude) {}

  std::array<Real, 3> operator()(const Real* x, Real /*t*/ = 0) const {
    const Real lat = gallery_detail::latitude(x[0], x[1], x[2]);
    const Real lon = gallery_detail::longitude(x[0], x[1]);
    const Real cl = std::cos(lat), sl = std::sin(lat);
    const Real cl3 = cl * cl * cl;
    const Real u = rh54_amplitude * 0.5 * std::cos(4 * lon) * cl3 *
                   (5 * std::cos(2 * lat) - 3);
    const Real v = rh54_amplitude * 4 * cl3 * sl * std::sin(4 * lon);
    return {-background_rotation * x[1] - u * std::sin(lon) -
                v * sl * std::cos(lon),
            background_rotation * x[0] + u * std::cos(lon) -
                v * sl * std::sin(lon),
            v * cl};
  }
  std::string name() const { return "RossbyWave54Velocity"; }
};

// Steady zonal geostrophic flow u = u0max cos(lat) with surface
// h0max - (Omega u0max + u0max^2 / 2) z^2 / g (h0max = equatorial height).
struct SWETestCase2 {
  Real h0max;
  Real Omega;
  Real u0max;
  Real g = constants::GRAVITY;

  explicit SWETestCase2(Real hh = 1, Real Omg = 2 * constants::PI,
                        Real uu = constants::PI / 6)
      : h0max(hh), Omega(Omg), u0max(uu) {}

  template <typename Vec, typename Pt>
  void u0(Vec& uvw, const Pt& xyz) const {
    // u0max cos(lat) times the eastward unit vector (-sin lon, cos lon, 0);
    // cos(lat) (-sin lon, cos lon) = (-y, x) / |xyz|
    const Real r =
        std::sqrt(xyz[0] * xyz[0] + xyz[1] * xyz[1] + xyz[2] * xyz[2]);
    uvw[0] = -u0max * xyz[1] / r;
    uvw[1] = u0max * xyz[0] / r;
    uvw[2] = 0;
  }
  template <typename Pt>
  Real zeta0(const Pt& xyz) const {
    return 2 * u0max * xyz[2];
  }
  template <typename Pt>
  Real sfc0(const Pt& xyz) const {
    return h0max - (Omega * u0max + 0.5 * u0max * u0max) * xyz[2] * xyz[2] / g;
  }
  template <typename Pt>
  Real h0(const Pt& xyz) const {
    return sfc0(xyz);
  }
  template <typename Pt>
  Real sigma0(const Pt&) const {
    return 0;
  }
  template <typename Pt>
  Real bottom_height(const Pt&) const {
    return 0;
  }
  std::string name() const { return "SWETestCase2"; }
};

// h0 + Omega u0 cos^2(lat) / g with h0 = 10, g = 1, Omega = 2 pi,
// u0 = 2 pi / 12.
struct SphereTestCase2InitialSurface {
  static constexpr Real h0 = 10;
  static constexpr Real g = 1.0;
  static constexpr Real Omega = 2 * constants::PI;
  static constexpr Real u0 = 2 * constants::PI / 12;

  template <typename Pt>
  Real operator()(const Pt& xyz) const {
    return h0 + Omega * u0 * (1 - xyz[2] * xyz[2]) / g;
  }
  std::string name() const { return "SphereTestCase2InitialSurface"; }
};

}  // namespace Lpm
