// BVESphere<Seed> and BVERK4 over the C ABI: the interface of
// src/lpm_bve_sphere.hpp:24-109 and src/lpm_bve_rk4.hpp:10-77.
#pragma once

#include <vector>

#include "lpm/lpm.hpp"

namespace Lpm {

template <typename SeedType>
class BVESphere : public PolyMesh2d<SeedType> {
 public:
  using seed_type = SeedType;

  // Allocates for the given maximum sizes; tree_init() builds the mesh.
  BVESphere(Index nmaxverts, Index nmaxedges, Index nmaxfaces, int nq = 0,
            std::shared_ptr<Context> ctx = Context::default_context())
      : PolyMesh2d<SeedType>(nmaxverts, nmaxedges, nmaxfaces),
        ctx_(std::move(ctx)),
        ntracers_(nq) {}
  BVESphere(Index nmaxverts, Index nmaxedges, Index nmaxfaces,
            const std::vector<std::string>& tracers,
            std::shared_ptr<Context> ctx = Context::default_context())
      : BVESphere(nmaxverts, nmaxedges, nmaxfaces,
                  static_cast<int>(tracers.size()), std::move(ctx)) {
    tracer_names_ = tracers;
  }

  // zeta = fn(x, y, z); absolute vorticity zeta + 2 Omega z
  // (src/lpm_bve_sphere_impl.hpp:148-179).
  template <typename VorticityInitialCondition>
  void init_vorticity(const VorticityInitialCondition& fn) {
    for (Index i = 0; i < this->n_vertices_host(); ++i) {
      const Real z = fn(this->vert_x.host(i, 0), this->vert_x.host(i, 1),
                        this->vert_x.host(i, 2));
      rel_vort_verts.host()[i] = z;
      abs_vort_verts.host()[i] = z + 2 * Omega * this->vert_x.host(i, 2);
    }
    for (Index i = 0; i < this->n_faces_host(); ++i) {
      const Real z = fn(this->face_x.host(i, 0), this->face_x.host(i, 1),
                        this->face_x.host(i, 2));
      rel_vort_faces.host()[i] = z;
      abs_vort_faces.host()[i] = z + 2 * Omega * this->face_x.host(i, 2);
    }
    rel_vort_verts.update_device();
    abs_vort_verts.update_device();
    rel_vort_faces.update_device();
    abs_vort_faces.update_device();
  }

  // BVESphere::init_velocity (src/lpm_bve_sphere_impl.hpp:181-205)
  void init_velocity(int precision = LPM_FP64, unsigned flags = 0) {
    lpm_bve_fields f = fields();
    check(lpm_bve_init_velocity(ctx_->get(), &f, precision, flags, nullptr),
          ctx_->get());
    velocity_verts.update_host();
    velocity_faces.update_host();
  }

  // BVESphere::init_stream_fn (src/lpm_bve_sphere_impl.hpp:207-231)
  void init_stream_fn(unsigned flags = 0) {
    check(lpm_bve_stream_fn(ctx_->get(), this->vert_x.data(), 1,
                            this->vert_x.cs(), this->n_vertices_host(),
                            this->face_x.data(), 1, this->face_x.cs(),
                            rel_vort_faces.data(), this->face_area.data(),
                            this->face_mask.data(), this->n_faces_host(), 0,
                            LPM_FP64, flags, stream_fn_verts.data(), nullptr),
          ctx_->get());
    check(lpm_bve_stream_fn(ctx_->get(), this->face_x.data(), 1,
                            this->face_x.cs(), this->n_faces_host(),
                            this->face_x.data(), 1, this->face_x.cs(),
                            rel_vort_faces.data(), this->face_area.data(),
                            this->face_mask.data(), this->n_faces_host(), 1,
                            LPM_FP64, flags, stream_fn_faces.data(), nullptr),
          ctx_->get());
    stream_fn_verts.update_host();
    stream_fn_faces.update_host();
  }

  void set_omega(const Real& omg) {
    if (!omg_set_) {
      Omega = omg;
      omg_set_ = true;
    }
  }

  Real avg_mesh_size_radians() const { return this->appx_mesh_size(); }
  Real avg_mesh_size_degrees() const {
    return 180 / constants::PI * avg_mesh_size_radians();
  }

  void update_device() const override {
    PolyMesh2d<SeedType>::update_device();
    for (auto* a : {&rel_vort_verts, &abs_vort_verts, &stream_fn_verts,
                    &rel_vort_faces, &abs_vort_faces, &stream_fn_faces})
      a->update_device();
    velocity_verts.update_device();
    velocity_faces.update_device();
  }
  void update_host() override {
    PolyMesh2d<SeedType>::update_host();
    for (auto* a : {&rel_vort_verts, &abs_vort_verts, &stream_fn_verts,
                    &rel_vort_faces, &abs_vort_faces, &stream_fn_faces})
      a->update_host();
    velocity_verts.update_host();
    velocity_faces.update_host();
  }

  lpm_bve_fields fields() {
    lpm_bve_fields f{};
    f.n_verts = this->n_vertices_host();
    f.vert_x = this->vert_x.data();
    f.vert_x_es = 1;
    f.vert_x_cs = this->vert_x.cs();
    f.vert_zeta = rel_vort_verts.data();
    f.vert_u = velocity_verts.data();
    f.vert_u_es = 1;
    f.vert_u_cs = velocity_verts.cs();
    f.n_faces = this->n_faces_host();
    f.face_x = this->face_x.data();
    f.face_x_es = 1;
    f.face_x_cs = this->face_x.cs();
    f.face_zeta = rel_vort_faces.data();
    f.face_u = velocity_faces.data();
    f.face_u_es = 1;
    f.face_u_cs = velocity_faces.cs();
    f.face_area = this->face_area.data();
    f.face_mask = this->face_mask.data();
    return f;
  }

  lpm_ctx* ctx() const { return ctx_->get(); }

  ScalarField rel_vort_verts, abs_vort_verts, stream_fn_verts;
  VectorField velocity_verts;
  ScalarField rel_vort_faces, abs_vort_faces, stream_fn_faces;
  VectorField velocity_faces;
  std::vector<ScalarField> tracer_verts, tracer_faces;
  Real Omega = 2 * constants::PI;
  Real t = 0;

 protected:
  void init_fields() override {
    const size_t nv = this->n_vertices_host(), nf = this->n_faces_host();
    rel_vort_verts = ScalarField(nv);
    abs_vort_verts = ScalarField(nv);
    stream_fn_verts = ScalarField(nv);
    velocity_verts = VectorField(nv);
    rel_vort_faces = ScalarField(nf);
    abs_vort_faces = ScalarField(nf);
    stream_fn_faces = ScalarField(nf);
    velocity_faces = VectorField(nf);
    tracer_verts.clear();
    tracer_faces.clear();
    for (int k = 0; k < ntracers_; ++k) {
      tracer_verts.emplace_back(nv);
      tracer_faces.emplace_back(nf);
    }
  }

  std::shared_ptr<Context> ctx_;
  int ntracers_ = 0;
  std::vector<std::string> tracer_names_;
  bool omg_set_ = false;
};

// Classical RK4 (src/lpm_bve_rk4_impl.hpp:55-167).  Stage buffers are owned
// by the library context; flags select reference-compatible stage
// arithmetic, renormalisation, leaf-only evaluation and graph replay.
class BVERK4 {
 public:
  template <typename SeedType>
  BVERK4(const Real timestep, BVESphere<SeedType>& sph,
         int precision = LPM_FP64,
         unsigned flags = LPM_RK4_REFERENCE_COMPAT)
      : dt(timestep),
        Omega(sph.Omega),
        nverts(sph.n_vertices_host()),
        nfaces(sph.n_faces_host()),
        precision_(precision),
        flags_(flags) {
    warn_if_mixed_off_sphere(precision, flags, "BVERK4");
  }

  template <typename SeedType>
  void advance_timestep(BVESphere<SeedType>& sph) {
    lpm_bve_fields f = sph.fields();
    check(lpm_bve_rk4_step(sph.ctx(), &f, dt, Omega, precision_, flags_,
                           nullptr),
          sph.ctx());
    sph.t += dt;
  }

  // Stage buffers are sized by the library on first use.
  void init(const Index nv, const Index nf) {
    nverts = nv;
    nfaces = nf;
  }

  Real dt;
  Real Omega;
  Index nverts;
  Index nfaces;

 private:
  int precision_;
  unsigned flags_;
};

}  // namespace Lpm
