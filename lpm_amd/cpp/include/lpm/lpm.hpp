// C++ host API of the MI355X direct-sum library: the names and method
// meanings of LPM's BVESphere / BVERK4 / SWE / SWERK4 (pbosler/lpm), built
// on the C ABI in lpm_mi355x.h.  Header-only; link with liblpm_mi355x.so
// and the HIP runtime.
//
// Fields live in device memory as SoA arrays (x row, y row, z row), the
// layout of a Kokkos LayoutLeft View<double*[3]>; every array has a host
// mirror updated by update_host() / update_device(), as in the reference.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "lpm_mi355x.h"

namespace Lpm {

using Real = double;
using Index = int;

namespace constants {
constexpr Real PI = 3.1415926535897932384626433832795027975;
}

// Rethrows a negative lpm_status as std::runtime_error (LPM_REQUIRE's
// behaviour, src/lpm_assert.hpp:54).
inline void check(int rc, const lpm_ctx* ctx = nullptr) {
  if (rc != LPM_OK)
    throw std::runtime_error(std::string("lpm: ") + lpm_last_error(ctx) +
                             " (status " + std::to_string(rc) + ")");
}

// The mixed-precision pair kernels evaluate 1 - x.y as |x - y|^2 / 2, which
// holds for positions on the unit sphere only: a stepper that runs them
// without LPM_RK4_RENORMALIZE lets stage positions leave the sphere by
// O((dt |u|)^2) and its sums then differ from FP64 by that much.  Steppers
// print this once per process to stderr.
inline void warn_if_mixed_off_sphere(int precision, unsigned flags,
                                     const char* stepper) {
  static bool warned = false;
  if (precision == LPM_MIXED && !(flags & LPM_RK4_RENORMALIZE) && !warned) {
    warned = true;
    std::fprintf(stderr,
                 "lpm: %s with LPM_MIXED but without LPM_RK4_RENORMALIZE: the "
                 "mixed sums assume positions on the unit sphere; add "
                 "LPM_RK4_RENORMALIZE to the flags\n",
                 stepper);
  }
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// One lpm_ctx per process and device.
class Context {
 public:
  explicit Context(int device = 0) { check(lpm_ctx_create(device, &ctx_)); }
  ~Context() { lpm_ctx_destroy(ctx_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  lpm_ctx* get() const { return ctx_; }
  static std::shared_ptr<Context> default_context() {
    static std::shared_ptr<Context> c = std::make_shared<Context>(0);
    return c;
  }

 private:
  lpm_ctx* ctx_ = nullptr;
};

// Device array with a host mirror (Kokkos View + HostMirror).
template <typename T>
class DualArray {
 public:
  DualArray() = default;
  explicit DualArray(size_t n) : n_(n), host_(n, T(0)) {
    if (n > 0) hip_check(hipMalloc(&dev_, n * sizeof(T)), "hipMalloc");
    update_device();
  }
  ~DualArray() {
    if (dev_) (void)hipFree(dev_);
  }
  DualArray(const DualArray&) = delete;
  DualArray& operator=(const DualArray&) = delete;
  DualArray(DualArray&& o) noexcept { *this = std::move(o); }
  DualArray& operator=(DualArray&& o) noexcept {
    std::swap(dev_, o.dev_);
    std::swap(n_, o.n_);
    host_.swap(o.host_);
    return *this;
  }
  T* data() const { return dev_; }
  size_t size() const { return n_; }
  std::vector<T>& host() { return host_; }
  const std::vector<T>& host() const { return host_; }
  void update_device() const {
    if (n_)
      hip_check(hipMemcpy(dev_, host_.data(), n_ * sizeof(T),
                          hipMemcpyHostToDevice),
                "hipMemcpy H2D");
  }
  void update_host() {
    if (n_)
      hip_check(hipMemcpy(host_.data(), dev_, n_ * sizeof(T),
                          hipMemcpyDeviceToHost),
                "hipMemcpy D2H");
  }

 private:
  T* dev_ = nullptr;
  size_t n_ = 0;
  std::vector<T> host_;
};

// Coordinates / vectors: SoA, component k of element i at [k * n + i].
class VectorField {
 public:
  VectorField() = default;
  explicit VectorField(size_t n) : a_(3 * n), n_(n) {}
  double* data() const { return a_.data(); }
  size_t n() const { return n_; }
  int64_t es() const { return 1; }
  int64_t cs() const { return static_cast<int64_t>(n_); }
  double& host(size_t i, int k) { return a_.host()[k * n_ + i]; }
  double host(size_t i, int k) const { return a_.host()[k * n_ + i]; }
  void update_device() const { a_.update_device(); }
  void update_host() { a_.update_host(); }

 private:
  DualArray<double> a_;
  size_t n_ = 0;
};

using ScalarField = DualArray<double>;

// ------------------------------------------------------------ mesh seeds
struct IcosTriSphereSeed {
  static constexpr int id = LPM_SEED_ICOS_TRI;
  static std::string id_string() { return "IcosTriSphereSeed"; }
};
struct CubedSphereSeed {
  static constexpr int id = LPM_SEED_CUBED_SPHERE;
  static std::string id_string() { return "CubedSphereSeed"; }
};

template <typename SeedType>
struct MeshSeed {
  // MeshSeed::set_max_allocations (src/mesh/lpm_mesh_seed.cpp:266-279)
  void set_max_allocations(Index& nverts, Index& nedges, Index& nfaces,
                           int lev) const {
    int64_t v, e, f;
    check(lpm_mesh_max_allocations(SeedType::id, lev, &v, &e, &f));
    nverts = static_cast<Index>(v);
    nedges = static_cast<Index>(e);
    nfaces = static_cast<Index>(f);
  }
};

// Particles of a uniformly refined spherical mesh: vertices (passive) and
// all stored faces (active; divided faces masked).  PolyMesh2d<Seed> subset.
template <typename SeedType>
class PolyMesh2d {
 public:
  PolyMesh2d(Index nmaxverts, Index nmaxedges, Index nmaxfaces)
      : nmaxverts_(nmaxverts), nmaxedges_(nmaxedges), nmaxfaces_(nmaxfaces) {}
  virtual ~PolyMesh2d() = default;

  // PolyMesh2d::tree_init (src/mesh/lpm_polymesh2d_impl.hpp:25-42)
  void tree_init(int depth, const MeshSeed<SeedType>&) {
    lpm_mesh* m = nullptr;
    check(lpm_mesh_create(SeedType::id, depth, &m));
    int64_t nv, nf, nl, ne;
    check(lpm_mesh_sizes(m, &nv, &nf, &nl, &ne));
    if (nv > nmaxverts_ || nf > nmaxfaces_ || ne > nmaxedges_) {
      lpm_mesh_destroy(m);
      throw std::runtime_error("PolyMesh2d::tree_init: not enough memory");
    }
    nverts_ = nv;
    nfaces_ = nf;
    nleaves_ = nl;
    depth_ = depth;
    vert_x = VectorField(nv);
    face_x = VectorField(nf);
    face_area = ScalarField(nf);
    face_mask = DualArray<uint8_t>(nf);
    std::vector<double> vx(3 * nv), fx(3 * nf);
    check(lpm_mesh_copy_vertices(m, vx.data()));
    check(lpm_mesh_copy_faces(m, fx.data(), face_area.host().data(),
                              face_mask.host().data(), nullptr));
    int nfv = 0;
    check(lpm_mesh_copy_face_vertices(m, nullptr, &nfv));
    face_verts.assign(static_cast<size_t>(nf) * nfv, 0);
    check(lpm_mesh_copy_face_vertices(m, face_verts.data(), &nfv));
    nfaceverts = nfv;
    lpm_mesh_destroy(m);
    for (int k = 0; k < 3; ++k)
      for (int64_t i = 0; i < nv; ++i) vert_x.host(i, k) = vx[k * nv + i];
    for (int k = 0; k < 3; ++k)
      for (int64_t i = 0; i < nf; ++i) face_x.host(i, k) = fx[k * nf + i];
    vert_lag = std::move(vx);
    face_lag = std::move(fx);
    init_fields();
    PolyMesh2d::update_device();
  }

  Index n_vertices_host() const { return static_cast<Index>(nverts_); }
  Index n_faces_host() const { return static_cast<Index>(nfaces_); }
  Index n_leaves_host() const { return static_cast<Index>(nleaves_); }
  // mean leaf-panel size in radians
  Real appx_mesh_size() const {
    return std::sqrt(4 * constants::PI / static_cast<Real>(nleaves_));
  }
  virtual void update_device() const {
    vert_x.update_device();
    face_x.update_device();
    face_area.update_device();
    face_mask.update_device();
  }
  virtual void update_host() {
    vert_x.update_host();
    face_x.update_host();
    face_area.update_host();
  }

  VectorField vert_x;    // vertices.phys_crds
  VectorField face_x;    // faces.phys_crds
  ScalarField face_area; // faces.area
  DualArray<uint8_t> face_mask;
  // host only: face corners [n_faces][nfaceverts] (faces.verts) and the
  // initial positions, component-major [3][n] (vertices/faces.lag_crds)
  std::vector<int32_t> face_verts;
  int nfaceverts = 0;
  std::vector<Real> vert_lag, face_lag;

 protected:
  virtual void init_fields() {}
  Index nmaxverts_, nmaxedges_, nmaxfaces_;
  int64_t nverts_ = 0, nfaces_ = 0, nleaves_ = 0;
  int depth_ = 0;
};

}  // namespace Lpm
