// SWE<Seed> (sphere, zero bottom topography) and a sphere SWERK4 over the C
// ABI: the interface of src/lpm_swe.hpp:18-177 and src/lpm_swe_rk4.hpp:8-126.
#pragma once

#include "lpm/lpm.hpp"

namespace Lpm {

// CoriolisSphere (src/lpm_coriolis.hpp:154-195): f = 2 Omega z.
struct CoriolisSphere {
  Real Omega;
  explicit CoriolisSphere(const Real Omg = 2 * constants::PI) : Omega(Omg) {}
};

template <typename SeedType>
class SWE : public PolyMesh2d<SeedType> {
 public:
  SWE(Index nmaxverts, Index nmaxedges, Index nmaxfaces,
      const CoriolisSphere& coriolis, Real gravity = 1.0,
      std::shared_ptr<Context> ctx = Context::default_context())
      : PolyMesh2d<SeedType>(nmaxverts, nmaxedges, nmaxfaces),
        coriolis(coriolis),
        g(gravity),
        ctx_(std::move(ctx)) {}

  // The reference assigns eps/pse_eps only for planar meshes
  // (src/lpm_swe_impl.hpp:98-107); here both are explicit.
  void set_kernel_parameters(const Real vel_eps, const Real pse_eps_) {
    if (vel_eps < 0 || pse_eps_ <= 0)
      throw std::runtime_error("set_kernel_parameters: bad epsilon");
    eps = vel_eps;
    pse_eps = pse_eps_;
  }

  // surface s(x); zero bottom: depth = s, mass = s A
  // (SWE::init_surface, src/lpm_swe_impl.hpp:302-335)
  template <typename SurfaceFn>
  void init_surface(const SurfaceFn& sfc) {
    for (Index i = 0; i < this->n_vertices_host(); ++i) {
      const Real s = sfc(this->vert_x.host(i, 0), this->vert_x.host(i, 1),
                         this->vert_x.host(i, 2));
      surf_passive.host()[i] = s;
      depth_passive.host()[i] = s;
    }
    for (Index i = 0; i < this->n_faces_host(); ++i) {
      const Real s = sfc(this->face_x.host(i, 0), this->face_x.host(i, 1),
                         this->face_x.host(i, 2));
      surf_active.host()[i] = s;
      depth_active.host()[i] = s;
      mass_active.host()[i] = s * this->face_area.host()[i];
    }
    for (auto* a : {&surf_passive, &depth_passive, &surf_active, &depth_active,
                    &mass_active})
      a->update_device();
  }

  template <typename Fn>
  void init_vorticity(const Fn& fn) {
    fill(rel_vort_passive, rel_vort_active, fn);
  }
  template <typename Fn>
  void init_divergence(const Fn& fn) {
    fill(div_passive, div_active, fn);
  }

  // SWE::init_direct_sums (src/lpm_swe_impl.hpp:401-445), with the PSE
  // surface Laplacian.
  void init_direct_sums(int precision = LPM_FP64, unsigned flags = 0) {
    lpm_swe_fields f = fields();
    check(lpm_swe_init_sums(ctx_->get(), &f, eps, pse_eps, precision, flags,
                            nullptr),
          ctx_->get());
  }

  template <typename SolverType>
  void advance_timestep(SolverType& solver) {
    solver.advance_timestep_impl(*this);
    t = solver.t_idx * solver.dt;
  }

  lpm_swe_fields fields() {
    lpm_swe_fields f{};
    f.n_verts = this->n_vertices_host();
    f.vert_x = this->vert_x.data();
    f.vert_x_es = 1;
    f.vert_x_cs = this->vert_x.cs();
    f.vert_zeta = rel_vort_passive.data();
    f.vert_sigma = div_passive.data();
    f.vert_h = depth_passive.data();
    f.vert_u = velocity_passive.data();
    f.vert_u_es = 1;
    f.vert_u_cs = velocity_passive.cs();
    f.vert_ddot = double_dot_passive.data();
    f.vert_laps = surf_lap_passive.data();
    f.vert_surf = surf_passive.data();
    f.n_faces = this->n_faces_host();
    f.face_x = this->face_x.data();
    f.face_x_es = 1;
    f.face_x_cs = this->face_x.cs();
    f.face_zeta = rel_vort_active.data();
    f.face_sigma = div_active.data();
    f.face_area = this->face_area.data();
    f.face_mass = mass_active.data();
    f.face_h = depth_active.data();
    f.face_u = velocity_active.data();
    f.face_u_es = 1;
    f.face_u_cs = velocity_active.cs();
    f.face_ddot = double_dot_active.data();
    f.face_laps = surf_lap_active.data();
    f.face_surf = surf_active.data();
    f.face_mask = this->face_mask.data();
    return f;
  }

  void update_host() override {
    PolyMesh2d<SeedType>::update_host();
    for (auto* a : {&rel_vort_passive, &div_passive, &depth_passive,
                    &double_dot_passive, &surf_lap_passive, &surf_passive,
                    &rel_vort_active, &div_active, &depth_active,
                    &double_dot_active, &surf_lap_active, &surf_active})
      a->update_host();
    velocity_passive.update_host();
    velocity_active.update_host();
  }

  lpm_ctx* ctx() const { return ctx_->get(); }

  ScalarField rel_vort_passive, div_passive, depth_passive, double_dot_passive,
      surf_lap_passive, surf_passive;
  VectorField velocity_passive;
  ScalarField rel_vort_active, div_active, depth_active, double_dot_active,
      surf_lap_active, surf_active, mass_active;
  VectorField velocity_active;
  CoriolisSphere coriolis;
  Real g;
  Real eps = 0;
  Real pse_eps = 0;
  Real t = 0;

 protected:
  void init_fields() override {
    const size_t nv = this->n_vertices_host(), nf = this->n_faces_host();
    for (auto* a : {&rel_vort_passive, &div_passive, &depth_passive,
                    &double_dot_passive, &surf_lap_passive, &surf_passive})
      *a = ScalarField(nv);
    for (auto* a : {&rel_vort_active, &div_active, &depth_active,
                    &double_dot_active, &surf_lap_active, &surf_active,
                    &mass_active})
      *a = ScalarField(nf);
    velocity_passive = VectorField(nv);
    velocity_active = VectorField(nf);
  }

  template <typename Fn>
  void fill(ScalarField& pv, ScalarField& pf, const Fn& fn) {
    for (Index i = 0; i < this->n_vertices_host(); ++i)
      pv.host()[i] = fn(this->vert_x.host(i, 0), this->vert_x.host(i, 1),
                        this->vert_x.host(i, 2));
    for (Index i = 0; i < this->n_faces_host(); ++i)
      pf.host()[i] = fn(this->face_x.host(i, 0), this->face_x.host(i, 1),
                        this->face_x.host(i, 2));
    pv.update_device();
    pf.update_device();
  }

  std::shared_ptr<Context> ctx_;
};

// Sphere RK4 in the stage layout of SWERK4 (src/lpm_swe_rk4_impl.hpp).
template <typename SeedType>
class SWERK4 {
 public:
  SWERK4(const Real timestep, SWE<SeedType>& swe, int precision = LPM_FP64,
         unsigned flags = LPM_RK4_REFERENCE_COMPAT)
      : dt(timestep), t_idx(0), precision_(precision), flags_(flags) {
    (void)swe;
    warn_if_mixed_off_sphere(precision, flags, "SWERK4");
  }

  void advance_timestep_impl(SWE<SeedType>& swe) {
    lpm_swe_fields f = swe.fields();
    check(lpm_swe_rk4_step(swe.ctx(), &f, dt, swe.coriolis.Omega, swe.g,
                           swe.eps, swe.pse_eps, precision_, flags_, nullptr),
          swe.ctx());
    ++t_idx;
  }

  Real dt;
  int t_idx;

 private:
  int precision_;
  unsigned flags_;
};

// Sphere RK2 (Heun) in the stage layout of SWERK2
// (src/lpm_swe_rk2_impl.hpp:80-257); the surface Laplacian comes from the
// PSE sums.  The reference's topo (zero bottom here) and GMLS parameters
// have no counterpart.
template <typename SeedType>
class SWERK2 {
 public:
  SWERK2(const Real timestep, SWE<SeedType>& swe, int precision = LPM_FP64,
         unsigned flags = 0)
      : dt(timestep), t_idx(0), precision_(precision), flags_(flags) {
    (void)swe;
    warn_if_mixed_off_sphere(precision, flags, "SWERK2");
  }

  void advance_timestep_impl(SWE<SeedType>& swe) {
    lpm_swe_fields f = swe.fields();
    check(lpm_swe_rk2_step(swe.ctx(), &f, dt, swe.coriolis.Omega, swe.g,
                           swe.eps, swe.pse_eps, precision_, flags_, nullptr),
          swe.ctx());
    ++t_idx;
  }

  Real dt;
  int t_idx;

 private:
  int precision_;
  unsigned flags_;
};

}  // namespace Lpm
