// VTK XML PolyData (.vtp) output of a particle mesh without the VTK library:
// the interface of VtkPolymeshInterface<Seed> (src/vtk/lpm_vtk_io.hpp:35-80,
// src/vtk/lpm_vtk_io_impl.hpp) and the field sets of vtk_interface(BVESphere)
// (src/lpm_bve_sphere_impl.hpp:106-125) and vtk_mesh_interface(SWE)
// (src/lpm_swe_impl.hpp:490-530).
//
// Points are the vertices' physical coordinates (host mirror: call
// update_host() on the mesh first), one polygon per leaf face with the
// mesh's corner order, one cell tuple per leaf face.  The constructor adds
// the cell array "area" and the point and cell arrays "lag_crds".  Arrays are
// written as inline base64 with a UInt64 byte count (VTK's "binary" format)
// or as ASCII.  With a height field h the points are (1 + h) x; the
// reference's sphere branch of that constructor computes those coordinates
// but never inserts them (src/vtk/lpm_vtk_io_impl.hpp:64-72).
#pragma once

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "lpm/bve_sphere.hpp"
#include "lpm/lpm.hpp"
#include "lpm/swe.hpp"

namespace Lpm {

namespace vtk_detail {

inline std::string base64(const unsigned char* p, size_t n) {
  static const char* tab =
      "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::string out;
  out.reserve(4 * ((n + 2) / 3));
  size_t i = 0;
  for (; i + 2 < n; i += 3) {
    const uint32_t v = (uint32_t(p[i]) << 16) | (uint32_t(p[i + 1]) << 8) | p[i + 2];
    out += tab[(v >> 18) & 63];
    out += tab[(v >> 12) & 63];
    out += tab[(v >> 6) & 63];
    out += tab[v & 63];
  }
  if (i < n) {
    uint32_t v = uint32_t(p[i]) << 16;
    if (i + 1 < n) v |= uint32_t(p[i + 1]) << 8;
    out += tab[(v >> 18) & 63];
    out += tab[(v >> 12) & 63];
    out += i + 1 < n ? tab[(v >> 6) & 63] : '=';
    out += '=';
  }
  return out;
}

struct Array {
  std::string name;
  int ncomp;
  std::vector<double> data;  // tuple-major
};

}  // namespace vtk_detail

template <typename SeedType>
class VtkPolymeshInterface {
 public:
  explicit VtkPolymeshInterface(const PolyMesh2d<SeedType>& pm) : mesh_(pm) {
    init(nullptr);
  }
  VtkPolymeshInterface(const PolyMesh2d<SeedType>& pm,
                       const std::vector<Real>& height_field)
      : mesh_(pm) {
    init(&height_field);
  }

  // Physical coordinates from the mesh's host mirror.
  void update_positions() { make_points(height_.empty() ? nullptr : &height_); }

  void add_scalar_point_data(const ScalarField& s, const std::string& name) {
    add_point(name, 1, s.host().data(), 0);
  }
  void add_scalar_point_data(const std::vector<Real>& s, const std::string& name) {
    add_point(name, 1, s.data(), 0);
  }
  void add_vector_point_data(const VectorField& v, const std::string& name) {
    add_point(name, 3, nullptr, 0, &v);
  }
  void add_scalar_cell_data(const ScalarField& s, const std::string& name) {
    add_cell(name, 1, s.host().data());
  }
  void add_scalar_cell_data(const std::vector<Real>& s, const std::string& name) {
    add_cell(name, 1, s.data());
  }
  void add_vector_cell_data(const VectorField& v, const std::string& name) {
    add_cell(name, 3, nullptr, &v);
  }
  // Vector data given component-major [3][n] on the host.
  void add_vector_point_data(const std::vector<Real>& v, const std::string& name) {
    add_point(name, 3, v.data(), mesh_.n_vertices_host());
  }
  void add_vector_cell_data(const std::vector<Real>& v, const std::string& name) {
    add_cell(name, 3, v.data(), nullptr, mesh_.n_faces_host());
  }

  void write(const std::string& ofilename, bool ascii = false) const {
    std::ofstream f(ofilename);
    if (!f) throw std::runtime_error("VtkPolymeshInterface: cannot open " + ofilename);
    const char* fmt = ascii ? "ascii" : "binary";
    f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PolyData\" version=\"1.0\" "
         "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n  <PolyData>\n"
      << "    <Piece NumberOfPoints=\"" << mesh_.n_vertices_host()
      << "\" NumberOfVerts=\"0\" NumberOfLines=\"0\" NumberOfStrips=\"0\" "
         "NumberOfPolys=\""
      << leaf_.size() << "\">\n      <PointData>\n";
    for (const auto& a : point_) write_array(f, a.name, "Float64", a.ncomp, a.data, fmt);
    f << "      </PointData>\n      <CellData>\n";
    for (const auto& a : cell_) write_array(f, a.name, "Float64", a.ncomp, a.data, fmt);
    f << "      </CellData>\n      <Points>\n";
    write_array(f, "Points", "Float64", 3, points_, fmt);
    f << "      </Points>\n      <Polys>\n";
    std::vector<int64_t> conn, offs;
    const int k = mesh_.nfaceverts;
    for (size_t c = 0; c < leaf_.size(); ++c) {
      for (int j = 0; j < k; ++j)
        conn.push_back(mesh_.face_verts[static_cast<size_t>(leaf_[c]) * k + j]);
      offs.push_back(static_cast<int64_t>(c + 1) * k);
    }
    write_array(f, "connectivity", "Int64", 1, conn, fmt);
    write_array(f, "offsets", "Int64", 1, offs, fmt);
    f << "      </Polys>\n    </Piece>\n  </PolyData>\n</VTKFile>\n";
  }

 private:
  void init(const std::vector<Real>* height) {
    if (mesh_.nfaceverts == 0)
      throw std::runtime_error("VtkPolymeshInterface: mesh has no face corners");
    for (Index i = 0; i < mesh_.n_faces_host(); ++i)
      if (!mesh_.face_mask.host()[i]) leaf_.push_back(i);
    if (height) height_ = *height;
    make_points(height);
    add_cell("area", 1, mesh_.face_area.host().data());
    add_point("lag_crds", 3, mesh_.vert_lag.data(), mesh_.n_vertices_host());
    add_cell("lag_crds", 3, mesh_.face_lag.data(), nullptr, mesh_.n_faces_host());
  }

  void make_points(const std::vector<Real>* h) {
    const Index n = mesh_.n_vertices_host();
    points_.assign(3 * static_cast<size_t>(n), 0.0);
    for (Index i = 0; i < n; ++i) {
      const Real s = h ? 1.0 + (*h)[i] : 1.0;
      for (int k = 0; k < 3; ++k) points_[3 * i + k] = s * mesh_.vert_x.host(i, k);
    }
  }

  // scalar: p[i]; vector: field v, or p component-major with leading
  // dimension ld
  void add_point(const std::string& name, int nc, const Real* p, Index ld,
                 const VectorField* v = nullptr) {
    const Index n = mesh_.n_vertices_host();
    vtk_detail::Array a{name, nc, std::vector<double>(static_cast<size_t>(n) * nc)};
    for (Index i = 0; i < n; ++i)
      for (int k = 0; k < nc; ++k)
        a.data[static_cast<size_t>(i) * nc + k] =
            v ? v->host(i, k) : (nc == 1 ? p[i] : p[static_cast<size_t>(k) * ld + i]);
    point_.push_back(std::move(a));
  }

  void add_cell(const std::string& name, int nc, const Real* p,
                const VectorField* v = nullptr, Index ld = 0) {
    vtk_detail::Array a{name, nc, std::vector<double>(leaf_.size() * nc)};
    for (size_t c = 0; c < leaf_.size(); ++c) {
      const Index i = leaf_[c];
      for (int k = 0; k < nc; ++k)
        a.data[c * nc + k] =
            v ? v->host(i, k) : (nc == 1 ? p[i] : p[static_cast<size_t>(k) * ld + i]);
    }
    cell_.push_back(std::move(a));
  }

  template <typename T>
  static void write_array(std::ostream& f, const std::string& name,
                          const char* type, int nc, const std::vector<T>& d,
                          const char* fmt) {
    f << "        <DataArray type=\"" << type << "\" Name=\"" << name
      << "\" NumberOfComponents=\"" << nc << "\" format=\"" << fmt << "\">";
    if (fmt[0] == 'a') {
      char buf[40];
      for (size_t i = 0; i < d.size(); ++i) {
        if (i) f << ' ';
        if (std::is_floating_point<T>::value)
          std::snprintf(buf, sizeof buf, "%.17g", static_cast<double>(d[i]));
        else
          std::snprintf(buf, sizeof buf, "%lld", static_cast<long long>(d[i]));
        f << buf;
      }
    } else {
      const uint64_t nbytes = d.size() * sizeof(T);
      f << vtk_detail::base64(reinterpret_cast<const unsigned char*>(&nbytes), 8)
        << vtk_detail::base64(reinterpret_cast<const unsigned char*>(d.data()),
                              nbytes);
    }
    f << "</DataArray>\n";
  }

  const PolyMesh2d<SeedType>& mesh_;
  std::vector<Index> leaf_;
  std::vector<double> points_;
  std::vector<Real> height_;
  std::vector<vtk_detail::Array> point_, cell_;
};

// The arrays of vtk_interface(BVESphere) (src/lpm_bve_sphere_impl.hpp:106-125),
// from the host mirrors (call bve.update_host() first).
template <typename SeedType>
VtkPolymeshInterface<SeedType> vtk_interface(const BVESphere<SeedType>& bve) {
  VtkPolymeshInterface<SeedType> vtk(bve);
  vtk.add_scalar_point_data(bve.rel_vort_verts, "rel_vort");
  vtk.add_scalar_point_data(bve.abs_vort_verts, "abs_vort");
  vtk.add_scalar_point_data(bve.stream_fn_verts, "stream_fn");
  vtk.add_vector_point_data(bve.velocity_verts, "velocity");
  vtk.add_scalar_cell_data(bve.rel_vort_faces, "rel_vort");
  vtk.add_scalar_cell_data(bve.abs_vort_faces, "abs_vort");
  vtk.add_scalar_cell_data(bve.stream_fn_faces, "stream_fn");
  vtk.add_vector_cell_data(bve.velocity_faces, "velocity");
  return vtk;
}

// The arrays of vtk_mesh_interface(SWE) (src/lpm_swe_impl.hpp:490-530) that
// this SWE holds, under the reference's field labels (:62-92, including its
// spelling "surface_laplaican" of the face array); bottom height is zero.
template <typename SeedType>
VtkPolymeshInterface<SeedType> vtk_mesh_interface(const SWE<SeedType>& swe) {
  VtkPolymeshInterface<SeedType> vtk(swe);
  const std::vector<Real> zv(swe.n_vertices_host(), 0.0), zf(swe.n_faces_host(), 0.0);
  vtk.add_scalar_point_data(swe.rel_vort_passive, "relative_vorticity");
  vtk.add_scalar_point_data(swe.div_passive, "divergence");
  vtk.add_scalar_point_data(swe.surf_passive, "surface_height");
  vtk.add_scalar_point_data(swe.surf_lap_passive, "surface_laplacian");
  vtk.add_scalar_point_data(swe.depth_passive, "depth");
  vtk.add_scalar_point_data(swe.double_dot_passive, "double_dot");
  vtk.add_scalar_point_data(zv, "bottom_height");
  vtk.add_vector_point_data(swe.velocity_passive, "velocity");
  vtk.add_scalar_cell_data(swe.rel_vort_active, "relative_vorticity");
  vtk.add_scalar_cell_data(swe.div_active, "divergence");
  vtk.add_scalar_cell_data(swe.surf_active, "surface_height");
  vtk.add_scalar_cell_data(swe.depth_active, "depth");
  vtk.add_scalar_cell_data(swe.surf_lap_active, "surface_laplaican");
  vtk.add_scalar_cell_data(swe.double_dot_active, "double_dot");
  vtk.add_scalar_cell_data(zf, "bottom_height");
  vtk.add_vector_cell_data(swe.velocity_active, "velocity");
  vtk.add_scalar_cell_data(swe.mass_active, "mass");
  return vtk;
}

}  // namespace Lpm
