// Steady zonal flow (Williamson test case 2 form) with SWE / SWERK4, the
// counterpart of the reference's examples/sphere_swe_tc2.cpp: u = u0 cos(lat),
// zeta = 2 u0 z, sigma = 0, s = h0 + (Omega u0 + u0^2/2) cos^2(lat) / g.
// Reports the initial velocity, double-dot and Laplacian errors against the
// closed forms (:231-251) and the drift of divergence and vorticity after
// the RK4 steps, as JSON.
//   sphere_swe_tc2 [-d depth] [-dt dt] [-n steps] [-precision fp64|mixed]
#include <cmath>
#include <cstdio>
#include <string>

#include "lpm/swe.hpp"

using namespace Lpm;

int main(int argc, char** argv) {
  try {
    int depth = 4, nsteps = 3, precision = LPM_FP64;
    Real dt = 0.01;
    for (int i = 1; i + 1 < argc; i += 2) {
      const std::string k = argv[i], v = argv[i + 1];
      if (k == "-d") depth = std::stoi(v);
      else if (k == "-dt") dt = std::stod(v);
      else if (k == "-n") nsteps = std::stoi(v);
      else if (k == "-precision") precision = v == "mixed" ? LPM_MIXED : LPM_FP64;
      else throw std::runtime_error("unknown option " + k);
    }
    const Real Omega = 2 * constants::PI, u0 = 0.2, g = 1.0, h0 = 1.0;
    MeshSeed<IcosTriSphereSeed> seed;
    Index nv, ne, nf;
    seed.set_max_allocations(nv, ne, nf, depth);
    SWE<IcosTriSphereSeed> swe(nv, ne, nf, CoriolisSphere(Omega), g);
    swe.tree_init(depth, seed);
    const Real h = swe.appx_mesh_size();
    swe.set_kernel_parameters(h, 2 * h);
    swe.init_surface([&](Real, Real, Real z) {
      return h0 + (Omega * u0 + 0.5 * u0 * u0) * (1 - z * z) / g;
    });
    swe.init_vorticity([&](Real, Real, Real z) { return 2 * u0 * z; });
    swe.init_divergence([](Real, Real, Real) { return 0.0; });
    swe.init_direct_sums(precision);
    swe.update_host();
    double eu = 0, nu = 0, ed = 0, nd = 0, el = 0, nl = 0;
    for (Index i = 0; i < swe.n_vertices_host(); ++i) {
      const Real x = swe.vert_x.host(i, 0), y = swe.vert_x.host(i, 1),
                 z = swe.vert_x.host(i, 2);
      const Real ue[3] = {-u0 * y, u0 * x, 0};
      for (int k = 0; k < 3; ++k) {
        const Real d = swe.velocity_passive.host(i, k) - ue[k];
        eu += d * d;
        nu += ue[k] * ue[k];
      }
      const Real dde = -2 * u0 * u0 * z * z;
      const Real le = (u0 * u0 + 2 * Omega * u0) * (2 * z * z - (1 - z * z)) / g;
      ed += std::pow(swe.double_dot_passive.host()[i] - dde, 2);
      nd += dde * dde;
      el += std::pow(swe.surf_lap_passive.host()[i] - le, 2);
      nl += le * le;
    }
    SWERK4<IcosTriSphereSeed> solver(dt, swe, precision, 0);
    for (int s = 0; s < nsteps; ++s) swe.advance_timestep(solver);
    swe.update_host();
    Real max_sigma = 0;
    for (Index i = 0; i < swe.n_faces_host(); ++i)
      if (!swe.face_mask.host()[i])
        max_sigma = std::fmax(max_sigma, std::fabs(swe.div_active.host()[i]));
    std::printf(
        "{\"depth\": %d, \"steps\": %d, \"t\": %.4f, \"velocity_rel_l2\": %.6e, "
        "\"ddot_rel_l2\": %.6e, \"laps_rel_l2\": %.6e, "
        "\"max_abs_divergence\": %.6e}\n",
        depth, nsteps, swe.t, std::sqrt(eu / nu), std::sqrt(ed / nd),
        std::sqrt(el / nl), max_sigma);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "sphere_swe_tc2: %s\n", e.what());
    return 1;
  }
}
