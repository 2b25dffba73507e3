// Solid-body rotation driver for BVESphere / BVERK4, the counterpart of the
// reference's examples/bve_rotation.cpp: zeta = 2 Omega_r z with
// Omega_r = 2 pi, background rotation set to 0, RK4 to tf; reports velocity,
// stream-function and position errors against the exact rotation as JSON.
//   bve_rotation [-d depth] [-dt dt] [-tf tfinal] [-seed cube|icos]
//                [-precision fp64|mixed] [-renorm 0|1] [-o case_name]
//                [-f output_interval]
// -f n > 0 writes .vtp frames (vtk_interface) at t = 0, every n steps and at
// the end, named case_name_<seed><depth>_dt<dt>_<frame>.vtp, as the
// reference's driver does.
// -renorm 1 projects RK4 stage positions back onto the sphere.  Without it
// (the reference's stepper) stage positions sit off the sphere by
// O((dt |u|)^2), which shifts 1 - x.y of nearest neighbours by the same
// amount: once that is comparable to the leaf spacing squared the kernel is
// near-singular and the step loses accuracy.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "lpm/bve_sphere.hpp"
#include "lpm/error.hpp"
#include "lpm/galleries.hpp"
#include "lpm/vtk.hpp"

using namespace Lpm;

namespace {

struct Args {
  int depth = 3;
  Real dt = 0.01;
  Real tf = 0.03;
  std::string seed = "cube";
  int precision = LPM_FP64;
  bool renorm = false;
  std::string case_name = "bve_rotation";
  int output_interval = 0;
};

Args parse(int argc, char** argv) {
  Args a;
  for (int i = 1; i + 1 < argc; i += 2) {
    const std::string k = argv[i], v = argv[i + 1];
    if (k == "-d") a.depth = std::stoi(v);
    else if (k == "-dt") a.dt = std::stod(v);
    else if (k == "-tf") a.tf = std::stod(v);
    else if (k == "-seed") a.seed = v;
    else if (k == "-precision") a.precision = v == "mixed" ? LPM_MIXED : LPM_FP64;
    else if (k == "-renorm") a.renorm = std::stoi(v) != 0;
    else if (k == "-o") a.case_name = v;
    else if (k == "-f") a.output_interval = std::stoi(v);
    else throw std::runtime_error("unknown option " + k);
  }
  return a;
}

template <typename Seed>
int run(const Args& a) {
  const SolidBodyRotation rotation;
  constexpr Real OMG = SolidBodyRotation::OMEGA;
  MeshSeed<Seed> seed;
  Index nv, ne, nf;
  seed.set_max_allocations(nv, ne, nf, a.depth);
  BVESphere<Seed> sphere(nv, ne, nf);
  sphere.tree_init(a.depth, seed);
  sphere.set_omega(0);
  sphere.init_vorticity(rotation);
  sphere.init_velocity(a.precision);
  sphere.init_stream_fn();

  const Real courant = a.dt * 2 * constants::PI / sphere.appx_mesh_size();
  if (courant > 1.0) throw std::runtime_error("courant number exceeds 1");

  // initial velocity / stream-function errors (leaf faces and vertices)
  double ev = 0, nvv = 0, es = 0, ns = 0;
  for (Index i = 0; i < sphere.n_vertices_host(); ++i) {
    const Real x = sphere.vert_x.host(i, 0), y = sphere.vert_x.host(i, 1),
               z = sphere.vert_x.host(i, 2);
    const Real ex[3] = {-OMG * y, OMG * x, 0};
    for (int k = 0; k < 3; ++k) {
      const Real d = sphere.velocity_verts.host(i, k) - ex[k];
      ev += d * d;
      nvv += ex[k] * ex[k];
    }
    const Real ds = sphere.stream_fn_verts.host()[i] - OMG * z;
    es += ds * ds;
    ns += OMG * z * OMG * z;
  }
  // the same vertex velocity error with LPM's ErrNorms on the device, and
  // the largest |u . x| (SphereTangentFunctor)
  const int64_t nvh = sphere.n_vertices_host();
  VectorField uex(static_cast<size_t>(nvh));
  for (Index i = 0; i < sphere.n_vertices_host(); ++i)
    rotation.init_velocity(uex.host(i, 0), uex.host(i, 1), uex.host(i, 2),
                           sphere.vert_x.host(i, 0), sphere.vert_x.host(i, 1),
                           sphere.vert_x.host(i, 2));
  uex.update_device();
  const Context& ctx = *Context::default_context();
  const ErrNorms vel_norms(ctx, nullptr, sphere.velocity_verts, uex, nullptr,
                           nullptr, nvh);
  ScalarField udotx(static_cast<size_t>(nvh));
  sphere_tangency(ctx, udotx.data(), sphere.vert_x, sphere.velocity_verts, nvh);
  udotx.update_host();
  Real tangency = 0;
  for (Index i = 0; i < sphere.n_vertices_host(); ++i)
    tangency = std::fmax(tangency, std::fabs(udotx.host()[i]));
  const int nsteps = static_cast<int>(std::lround(a.tf / a.dt));
  std::vector<Real> x0(3 * sphere.n_vertices_host());
  for (Index i = 0; i < sphere.n_vertices_host(); ++i)
    for (int k = 0; k < 3; ++k) x0[3 * i + k] = sphere.vert_x.host(i, k);
  const unsigned flags =
      LPM_RK4_REFERENCE_COMPAT | (a.renorm ? LPM_RK4_RENORMALIZE : 0u);
  BVERK4 solver(a.dt, sphere, a.precision, flags);
  int frame = 0;
  auto output = [&]() {
    char name[256];
    std::snprintf(name, sizeof name, "%s_%s%d_dt%g_%04d.vtp",
                  a.case_name.c_str(), Seed::id_string().c_str(), a.depth, a.dt,
                  frame++);
    sphere.update_host();
    vtk_interface(sphere).write(name);
  };
  if (a.output_interval > 0) output();
  for (int s = 0; s < nsteps; ++s) {
    solver.advance_timestep(sphere);
    if (a.output_interval > 0 &&
        ((s + 1) % a.output_interval == 0 || s + 1 == nsteps))
      output();
  }
  sphere.update_host();
  const Real ang = OMG * sphere.t;
  Real pos_err = 0;
  for (Index i = 0; i < sphere.n_vertices_host(); ++i) {
    const Real ex[3] = {std::cos(ang) * x0[3 * i] - std::sin(ang) * x0[3 * i + 1],
                        std::sin(ang) * x0[3 * i] + std::cos(ang) * x0[3 * i + 1],
                        x0[3 * i + 2]};
    Real d = 0;
    for (int k = 0; k < 3; ++k) {
      const Real e = sphere.vert_x.host(i, k) - ex[k];
      d += e * e;
    }
    pos_err = std::fmax(pos_err, std::sqrt(d));
  }
  std::printf(
      "{\"seed\": \"%s\", \"depth\": %d, \"n_verts\": %d, \"n_faces\": %d, "
      "\"renorm\": %d, \"steps\": %d, \"t\": %.6f, \"courant\": %.4f, "
      "\"velocity_rel_l2\": %.6e, \"stream_fn_rel_l2\": %.6e, "
      "\"velocity_err_norms\": [%.6e, %.6e, %.6e], \"max_abs_u_dot_x\": %.6e, "
      "\"position_err_linf\": %.6e}\n",
      Seed::id_string().c_str(), a.depth, sphere.n_vertices_host(),
      sphere.n_faces_host(), a.renorm ? 1 : 0, nsteps, sphere.t, courant, std::sqrt(ev / nvv),
      std::sqrt(es / ns), vel_norms.l1, vel_norms.l2, vel_norms.linf, tangency,
      pos_err);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const Args a = parse(argc, argv);
    return a.seed == "icos" ? run<IcosTriSphereSeed>(a)
                            : run<CubedSphereSeed>(a);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "bve_rotation: %s\n", e.what());
    return 1;
  }
}
