// Uniformly refined spherical quadtree meshes: the input generator of the
// direct sum.  Produces the particle arrays the velocity kernels consume:
// passive particles (mesh vertices) and active particles (face centres with
// area and a "divided" mask), in the same order LPM's PolyMesh2d::tree_init
// creates them (src/mesh/lpm_polymesh2d_impl.hpp:25-42).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace lpm_amd {

enum class SeedKind : int { IcosTri = 0, CubedSphere = 1 };

struct SeedCounts {
  // Closed-form counts of src/mesh/lpm_mesh_seed.cpp:353-375.
  static int64_t vertices_at_level(SeedKind k, int lev);
  static int64_t leaf_faces_at_level(SeedKind k, int lev);
  static int64_t edges_at_level(SeedKind k, int64_t nv, int64_t nf);
  // MeshSeed::set_max_allocations (src/mesh/lpm_mesh_seed.cpp:266-279):
  // vertices at the finest level, and edge/face totals summed over all levels.
  static void max_allocations(SeedKind k, int lev, int64_t& nverts,
                              int64_t& nedges, int64_t& nfaces);
};

class SphereTreeMesh {
 public:
  SphereTreeMesh(SeedKind kind, int depth);

  SeedKind kind() const { return kind_; }
  int depth() const { return depth_; }
  int verts_per_face() const { return nfv_; }
  int64_t n_vertices() const { return static_cast<int64_t>(vx_.size()); }
  int64_t n_faces() const { return static_cast<int64_t>(fx_.size()); }
  int64_t n_edges() const { return static_cast<int64_t>(eo_.size()); }
  int64_t n_leaves() const { return n_leaves_; }

  // Divides leaf face f (PolyMesh2d::divide_face,
  // src/mesh/lpm_polymesh2d_impl.hpp:46-53); false if f is already divided.
  bool divide_face(int32_t f);
  // PolyMesh2d::divide_flagged_faces (src/mesh/lpm_polymesh2d_impl.hpp:
  // 129-173): faces i < n_faces() with flags[i] != 0, in index order, whose
  // level is <= max_level.  Returns the reference's refine count (flagged
  // faces within the level limit); *limit_reached is set when a flagged face
  // was beyond it.
  int64_t divide_flagged(const uint8_t* flags, int max_level,
                         bool* limit_reached);

  const std::vector<double>& vx() const { return vx_; }
  const std::vector<double>& vy() const { return vy_; }
  const std::vector<double>& vz() const { return vz_; }
  const std::vector<double>& fx() const { return fx_; }
  const std::vector<double>& fy() const { return fy_; }
  const std::vector<double>& fz() const { return fz_; }
  const std::vector<double>& farea() const { return farea_; }
  const std::vector<uint8_t>& fmask() const { return fmask_; }
  const std::vector<int32_t>& fverts() const { return fverts_; }
  const std::vector<int32_t>& fparent() const { return fparent_; }
  const std::vector<int32_t>& flevel() const { return flevel_; }

  double surface_area() const;  // sum over all stored faces (divided = 0)

 private:
  void load_seed();
  void divide_edge(int32_t e);
  void divide_tri(int32_t f);
  void divide_quad(int32_t f);
  int32_t add_vertex(const double p[3]);
  int32_t add_edge(int32_t o, int32_t d, int32_t l, int32_t r);
  void add_face(const int32_t* verts, const int32_t* edges, int32_t parent,
                const double ctr[3], double area);

  SeedKind kind_;
  int depth_;
  int nfv_;
  int64_t n_leaves_ = 0;

  std::vector<double> vx_, vy_, vz_;
  std::vector<int32_t> eo_, ed_, el_, er_, ek_;  // ek_: first kid or -1
  std::vector<double> fx_, fy_, fz_, farea_;
  std::vector<uint8_t> fmask_;
  std::vector<int32_t> fverts_, fedges_, fkid_, fparent_, flevel_;
};

// Sphere geometry helpers in the operation order of SphereGeometry
// (src/lpm_geometry.hpp:369-642); used by the mesh generator only.
namespace sphere {
double dot(const double a[3], const double b[3]);
void cross(double c[3], const double a[3], const double b[3]);
void normalize(double v[3]);
double distance(const double a[3], const double b[3]);
double tri_area(const double a[3], const double b[3], const double c[3]);
}  // namespace sphere

}  // namespace lpm_amd
