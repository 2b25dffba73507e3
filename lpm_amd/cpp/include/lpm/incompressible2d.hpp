// Incompressible2D<Seed> and Incompressible2DRK2 on the sphere over the C
// ABI: the interface of src/lpm_incompressible2d.hpp and
// src/lpm_incompressible2d_rk2.hpp (fused velocity + stream function direct
// sums, Heun RK2 with CoriolisSphere).
#pragma once

#include "lpm/lpm.hpp"

namespace Lpm {

template <typename SeedType>
class Incompressible2D : public PolyMesh2d<SeedType> {
 public:
  Incompressible2D(Index nmaxverts, Index nmaxedges, Index nmaxfaces,
                   Real Omg = 2 * constants::PI, Real velocity_eps = 0,
                   std::shared_ptr<Context> ctx = Context::default_context())
      : PolyMesh2d<SeedType>(nmaxverts, nmaxedges, nmaxfaces),
        Omega(Omg),
        eps(velocity_eps),
        ctx_(std::move(ctx)) {}

  // Incompressible2D::init_vorticity (src/lpm_incompressible2d_impl.hpp:
  // 139-173): zeta = fn(x, y, z) at passive and active particles.
  template <typename Fn>
  void init_vorticity(const Fn& fn) {
    for (Index i = 0; i < this->n_vertices_host(); ++i)
      rel_vort_passive.host()[i] = fn(this->vert_x.host(i, 0),
                                      this->vert_x.host(i, 1),
                                      this->vert_x.host(i, 2));
    for (Index i = 0; i < this->n_faces_host(); ++i)
      rel_vort_active.host()[i] = fn(this->face_x.host(i, 0),
                                     this->face_x.host(i, 1),
                                     this->face_x.host(i, 2));
    rel_vort_passive.update_device();
    rel_vort_active.update_device();
  }

  // Incompressible2D::init_direct_sums
  // (src/lpm_incompressible2d_impl.hpp:235-254).
  void init_direct_sums(unsigned flags = 0, int precision = LPM_FP64) {
    lpm_ic2d_fields f = fields();
    check(lpm_ic2d_init_direct_sums(ctx_->get(), &f, eps, precision, flags,
                                    nullptr),
          ctx_->get());
  }

  template <typename SolverType>
  void advance_timestep(SolverType& solver) {
    solver.advance_timestep_impl(*this);
    t = solver.t_idx * solver.dt;
  }

  lpm_ic2d_fields fields() {
    lpm_ic2d_fields f{};
    f.n_verts = this->n_vertices_host();
    f.vert_x = this->vert_x.data();
    f.vert_x_es = 1;
    f.vert_x_cs = this->vert_x.cs();
    f.vert_zeta = rel_vort_passive.data();
    f.vert_u = velocity_passive.data();
    f.vert_u_es = 1;
    f.vert_u_cs = velocity_passive.cs();
    f.vert_psi = stream_fn_passive.data();
    f.n_faces = this->n_faces_host();
    f.face_x = this->face_x.data();
    f.face_x_es = 1;
    f.face_x_cs = this->face_x.cs();
    f.face_zeta = rel_vort_active.data();
    f.face_u = velocity_active.data();
    f.face_u_es = 1;
    f.face_u_cs = velocity_active.cs();
    f.face_psi = stream_fn_active.data();
    f.face_area = this->face_area.data();
    f.face_mask = this->face_mask.data();
    return f;
  }

  void update_host() override {
    PolyMesh2d<SeedType>::update_host();
    for (auto* a : {&rel_vort_passive, &stream_fn_passive, &rel_vort_active,
                    &stream_fn_active})
      a->update_host();
    velocity_passive.update_host();
    velocity_active.update_host();
  }

  lpm_ctx* ctx() const { return ctx_->get(); }

  // Leaf-face totals of Incompressible2D (src/lpm_incompressible2d_impl.hpp:
  // 91-137): sum zeta A, (1/2) sum zeta^2 A and (1/2) sum |u|^2 A.  They read
  // the current device fields (copied to their host mirrors).
  Real total_vorticity() {
    rel_vort_active.update_host();
    return leaf_sum([&](Index i) { return rel_vort_active.host()[i]; });
  }
  Real total_enstrophy() {
    rel_vort_active.update_host();
    return 0.5 * leaf_sum([&](Index i) {
      const Real z = rel_vort_active.host()[i];
      return z * z;
    });
  }
  Real total_kinetic_energy() {
    velocity_active.update_host();
    return 0.5 * leaf_sum([&](Index i) {
      const Real u0 = velocity_active.host(i, 0), u1 = velocity_active.host(i, 1),
                 u2 = velocity_active.host(i, 2);
      return u0 * u0 + u1 * u1 + u2 * u2;
    });
  }

  ScalarField rel_vort_passive, stream_fn_passive;
  VectorField velocity_passive;
  ScalarField rel_vort_active, stream_fn_active;
  VectorField velocity_active;
  Real Omega;
  Real eps;
  Real t = 0;

 protected:
  void init_fields() override {
    const size_t nv = this->n_vertices_host(), nf = this->n_faces_host();
    rel_vort_passive = ScalarField(nv);
    stream_fn_passive = ScalarField(nv);
    velocity_passive = VectorField(nv);
    rel_vort_active = ScalarField(nf);
    stream_fn_active = ScalarField(nf);
    velocity_active = VectorField(nf);
  }

  std::shared_ptr<Context> ctx_;

 private:
  template <typename Fn>
  Real leaf_sum(const Fn& f) const {
    Real s = 0;
    for (Index i = 0; i < this->n_faces_host(); ++i)
      if (!this->face_mask.host()[i]) s += f(i) * this->face_area.host()[i];
    return s;
  }
};

// Incompressible2DRK2 (src/lpm_incompressible2d_rk2_impl.hpp:75-172).
template <typename SeedType>
class Incompressible2DRK2 {
 public:
  Incompressible2DRK2(const Real timestep, Incompressible2D<SeedType>& ic2d,
                      unsigned flags = 0, int precision = LPM_FP64)
      : dt(timestep), t_idx(0), flags_(flags), precision_(precision) {
    (void)ic2d;
    warn_if_mixed_off_sphere(precision, flags, "Incompressible2DRK2");
  }

  void advance_timestep_impl(Incompressible2D<SeedType>& ic2d) {
    lpm_ic2d_fields f = ic2d.fields();
    check(lpm_ic2d_rk2_step(ic2d.ctx(), &f, dt, ic2d.Omega, ic2d.eps,
                            precision_, flags_, nullptr),
          ic2d.ctx());
    ++t_idx;
  }

  Real dt;
  int t_idx;

 private:
  unsigned flags_;
  int precision_;
};

}  // namespace Lpm
