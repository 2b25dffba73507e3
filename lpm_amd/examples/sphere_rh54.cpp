// Stationary Rossby-Haurwitz 5-4 wave with Incompressible2D /
// Incompressible2DRK2, the counterpart of the reference's
// examples/sphere_rh54.cpp (its uniform-mesh path: no AMR, remeshing, FTLE or
// tracers).  The background rotation u0 = Omega / 14 makes the wave
// stationary, so the exact velocity and relative vorticity at each particle's
// current position are RossbyWave54Velocity and RossbyHaurwitz54 there.
// Reports, at t = 0 and after the steps, the velocity and vorticity ErrNorms
// (leaf faces, area-weighted) and the leaf-face totals, as JSON.
//   sphere_rh54 [-d depth] [-n nsteps] [-tf tfinal] [-seed cube|icos]
//               [-eps velocity_eps] [-precision fp64|mixed]
#include <cstdio>
#include <string>

#include "lpm/error.hpp"
#include "lpm/galleries.hpp"
#include "lpm/incompressible2d.hpp"

using namespace Lpm;

namespace {

// Divided faces are neither evaluated nor advanced (the reference evaluates
// them too, and on the icosahedral mesh a divided face shares its central
// child's centre, where the kernel is singular); positions are projected
// back onto the sphere after each stage, which the mixed sums assume.
constexpr unsigned kFlags = LPM_FLAG_SKIP_MASKED_TARGETS | LPM_RK4_RENORMALIZE;

struct Args {
  int depth = 4;
  int nsteps = 1;
  Real tfinal = 0.025;
  std::string seed = "cube";
  Real eps = 0;
  int precision = LPM_FP64;
};

struct Report {
  ErrNorms vel{0, 0, 0}, zeta{0, 0, 0};
  Real total_vorticity = 0, total_enstrophy = 0, total_kinetic_energy = 0;
};

template <typename Seed>
Report measure(Incompressible2D<Seed>& sphere, const RossbyHaurwitz54& zeta,
               const RossbyWave54Velocity& velocity) {
  const Context& ctx = *Context::default_context();
  const int64_t nf = sphere.n_faces_host();
  sphere.update_host();
  VectorField uex(static_cast<size_t>(nf));
  ScalarField zex(static_cast<size_t>(nf));
  for (int64_t i = 0; i < nf; ++i) {
    const Real x[3] = {sphere.face_x.host(i, 0), sphere.face_x.host(i, 1),
                       sphere.face_x.host(i, 2)};
    const auto u = velocity(x);
    for (int k = 0; k < 3; ++k) uex.host(i, k) = u[k];
    zex.host()[i] = zeta(x[0], x[1], x[2]);
  }
  uex.update_device();
  zex.update_device();
  Report r;
  r.vel = ErrNorms(ctx, nullptr, sphere.velocity_active, uex,
                   sphere.face_area.data(), sphere.face_mask.data(), nf);
  r.zeta = ErrNorms(ctx, nullptr, sphere.rel_vort_active.data(), zex.data(),
                    sphere.face_area.data(), sphere.face_mask.data(), nf, 1);
  r.total_vorticity = sphere.total_vorticity();
  r.total_enstrophy = sphere.total_enstrophy();
  r.total_kinetic_energy = sphere.total_kinetic_energy();
  return r;
}

std::string json(const Report& r) {
  char buf[512];
  std::snprintf(buf, sizeof buf,
                "{\"velocity_err\": [%.6e, %.6e, %.6e], "
                "\"vorticity_err\": [%.6e, %.6e, %.6e], "
                "\"total_vorticity\": %.9e, \"total_enstrophy\": %.9e, "
                "\"total_kinetic_energy\": %.9e}",
                r.vel.l1, r.vel.l2, r.vel.linf, r.zeta.l1, r.zeta.l2,
                r.zeta.linf, r.total_vorticity, r.total_enstrophy,
                r.total_kinetic_energy);
  return buf;
}

template <typename Seed>
int run(const Args& a) {
  const Real Omega = 2 * constants::PI;
  RossbyHaurwitz54 zeta;
  zeta.set_stationary_wave_speed(Omega);
  const RossbyWave54Velocity velocity(zeta);

  MeshSeed<Seed> seed;
  Index nv, ne, nf;
  seed.set_max_allocations(nv, ne, nf, a.depth);
  Incompressible2D<Seed> sphere(nv, ne, nf, Omega, a.eps);
  sphere.tree_init(a.depth, seed);
  sphere.init_vorticity(zeta);
  sphere.init_direct_sums(kFlags, a.precision);
  const Report r0 = measure(sphere, zeta, velocity);

  const Real dt = a.tfinal / a.nsteps;
  Incompressible2DRK2<Seed> solver(dt, sphere, kFlags, a.precision);
  for (int s = 0; s < a.nsteps; ++s) sphere.advance_timestep(solver);
  const Report r1 = measure(sphere, zeta, velocity);
  std::printf(
      "{\"seed\": \"%s\", \"depth\": %d, \"n_faces\": %d, \"u0\": %.17g, "
      "\"steps\": %d, \"dt\": %.6e, \"t\": %.6e, \"initial\": %s, "
      "\"final\": %s}\n",
      Seed::id_string().c_str(), a.depth, sphere.n_faces_host(), zeta.u0,
      a.nsteps, dt, sphere.t, json(r0).c_str(), json(r1).c_str());
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    Args a;
    for (int i = 1; i + 1 < argc; i += 2) {
      const std::string k = argv[i], v = argv[i + 1];
      if (k == "-d") a.depth = std::stoi(v);
      else if (k == "-n") a.nsteps = std::stoi(v);
      else if (k == "-tf") a.tfinal = std::stod(v);
      else if (k == "-seed") a.seed = v;
      else if (k == "-eps") a.eps = std::stod(v);
      else if (k == "-precision") a.precision = v == "mixed" ? LPM_MIXED : LPM_FP64;
      else throw std::runtime_error("unknown option " + k);
    }
    if (a.nsteps < 1) throw std::runtime_error("nsteps must be >= 1");
    if (a.seed == "icos") return run<IcosTriSphereSeed>(a);
    if (a.seed == "cube") return run<CubedSphereSeed>(a);
    throw std::runtime_error("unknown seed " + a.seed);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "sphere_rh54: %s\n", e.what());
    return 1;
  }
}
