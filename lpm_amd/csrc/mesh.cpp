// Uniform spherical quadtree refinement.  The division steps reproduce the
// face/edge/vertex creation order of LPM's FaceDivider<Geo,TriFace>,
// FaceDivider<Geo,QuadFace> (src/mesh/lpm_faces_impl.hpp:285-430, 434-574)
// and Edges::divide (src/mesh/lpm_edges.cpp:59-95), so that particle i of
// this mesh is particle i of the reference mesh.  The seed polyhedra are the
// tables of mesh_seeds/icosTriSphereSeed.dat and cubedSphereSeed.dat.
#include "mesh.hpp"

#include <cmath>
#include <stdexcept>

namespace lpm_amd {

namespace {

// ---- seed tables (data of mesh_seeds/*.dat: vertex coordinates followed by
// face-centre coordinates; edges as origin, destination, left, right face) --
const double kIcosCrds[32][3] = {
    {0.00000000000000000, 0.00000000000000000, 1.00000000000000000},
    {0.72360679774997894, 0.52573111211913359, 0.44721359549995793},
    {-0.27639320225002101, 0.85065080835203988, 0.44721359549995793},
    {-0.89442719099991586, 0.00000000000000000, 0.44721359549995793},
    {-0.27639320225002101, -0.85065080835203988, 0.44721359549995793},
    {0.72360679774997894, -0.52573111211913359, 0.44721359549995793},
    {0.89442719099991586, 0.00000000000000000, -0.44721359549995793},
    {0.27639320225002101, 0.85065080835203988, -0.44721359549995793},
    {-0.72360679774997894, 0.52573111211913359, -0.44721359549995793},
    {-0.72360679774997894, -0.52573111211913359, -0.44721359549995793},
    {0.27639320225002101, -0.85065080835203988, -0.44721359549995793},
    {0.00000000000000000, 0.00000000000000000, -1.00000000000000000},
    {0.18759247408507992, 0.57735026918962584, 0.79465447229176622},
    {-0.49112347318842298, 0.35682208977308993, 0.79465447229176611},
    {-0.49112347318842298, -0.35682208977308993, 0.79465447229176611},
    {0.18759247408507992, -0.57735026918962584, 0.79465447229176622},
    {0.60706199820668627, 0.00000000000000000, 0.79465447229176622},
    {0.79465447229176611, 0.57735026918962573, -0.18759247408507990},
    {0.30353099910334314, 0.93417235896271567, 0.18759247408507992},
    {-0.30353099910334314, 0.93417235896271567, -0.18759247408507992},
    {-0.79465447229176611, 0.57735026918962573, 0.18759247408507990},
    {-0.98224694637684595, 0.00000000000000000, -0.18759247408507990},
    {-0.79465447229176611, -0.57735026918962573, 0.18759247408507990},
    {-0.30353099910334314, -0.93417235896271567, -0.18759247408507992},
    {0.30353099910334314, -0.93417235896271567, 0.18759247408507992},
    {0.79465447229176611, -0.57735026918962573, -0.18759247408507990},
    {0.98224694637684595, 0.00000000000000000, 0.18759247408507990},
    {0.49112347318842298, 0.35682208977308993, -0.79465447229176611},
    {-0.18759247408507992, 0.57735026918962584, -0.79465447229176622},
    {-0.60706199820668627, 0.00000000000000000, -0.79465447229176622},
    {-0.18759247408507992, -0.57735026918962584, -0.79465447229176622},
    {0.49112347318842298, -0.35682208977308993, -0.79465447229176611}};
const int32_t kIcosEdges[30][4] = {
    {0, 1, 0, 4},    {1, 2, 0, 6},    {2, 0, 0, 1},    {2, 3, 1, 8},
    {0, 3, 2, 1},    {3, 4, 2, 10},   {4, 0, 2, 3},    {4, 5, 3, 12},
    {5, 0, 3, 4},    {5, 1, 4, 14},   {1, 6, 5, 14},   {6, 7, 5, 15},
    {7, 1, 5, 6},    {7, 2, 6, 7},    {7, 8, 7, 16},   {8, 2, 7, 8},
    {8, 3, 8, 9},    {8, 9, 9, 17},   {3, 9, 10, 9},   {9, 4, 10, 11},
    {9, 10, 11, 18}, {10, 4, 11, 12}, {10, 5, 12, 13}, {10, 6, 13, 19},
    {6, 5, 13, 14},  {11, 6, 19, 15}, {11, 7, 15, 16}, {8, 11, 17, 16},
    {11, 9, 17, 18}, {10, 11, 19, 18}};
const int32_t kIcosFaceVerts[20][3] = {
    {0, 1, 2},  {0, 2, 3},  {0, 3, 4},  {0, 4, 5},  {0, 5, 1},
    {1, 6, 7},  {7, 2, 1},  {2, 7, 8},  {8, 3, 2},  {3, 8, 9},
    {9, 4, 3},  {4, 9, 10}, {10, 5, 4}, {5, 10, 6}, {6, 1, 5},
    {11, 7, 6}, {11, 8, 7}, {11, 9, 8}, {11, 10, 9}, {11, 6, 10}};
const int32_t kIcosFaceEdges[20][3] = {
    {0, 1, 2},    {2, 3, 4},    {4, 5, 6},    {6, 7, 8},    {8, 9, 0},
    {10, 11, 12}, {13, 1, 12},  {13, 14, 15}, {16, 3, 15},  {16, 17, 18},
    {19, 5, 18},  {19, 20, 21}, {22, 7, 21},  {22, 23, 24}, {10, 9, 24},
    {26, 11, 25}, {27, 14, 26}, {28, 17, 27}, {29, 20, 28}, {25, 23, 29}};

const double kCubeCrds[14][3] = {
    {0.57735026918962584, -0.57735026918962584, 0.57735026918962584},
    {0.57735026918962584, -0.57735026918962584, -0.57735026918962584},
    {0.57735026918962584, 0.57735026918962584, -0.57735026918962584},
    {0.57735026918962584, 0.57735026918962584, 0.57735026918962584},
    {-0.57735026918962584, 0.57735026918962584, -0.57735026918962584},
    {-0.57735026918962584, 0.57735026918962584, 0.57735026918962584},
    {-0.57735026918962584, -0.57735026918962584, -0.57735026918962584},
    {-0.57735026918962584, -0.57735026918962584, 0.57735026918962584},
    {1.00000000000000000, 0.00000000000000000, 0.00000000000000000},
    {0.00000000000000000, 1.00000000000000000, 0.00000000000000000},
    {-1.00000000000000000, 0.00000000000000000, 0.00000000000000000},
    {0.00000000000000000, -1.00000000000000000, 0.00000000000000000},
    {0.00000000000000000, 0.00000000000000000, 1.00000000000000000},
    {0.00000000000000000, 0.00000000000000000, -1.00000000000000000}};
const int32_t kCubeEdges[12][4] = {{0, 1, 0, 3}, {1, 2, 0, 5}, {2, 3, 0, 1},
                                   {3, 0, 0, 4}, {2, 4, 1, 5}, {4, 5, 1, 2},
                                   {5, 3, 1, 4}, {4, 6, 2, 5}, {6, 7, 2, 3},
                                   {7, 5, 2, 4}, {6, 1, 3, 5}, {0, 7, 3, 4}};
const int32_t kCubeFaceVerts[6][4] = {{0, 1, 2, 3}, {3, 2, 4, 5},
                                      {5, 4, 6, 7}, {7, 6, 1, 0},
                                      {7, 0, 3, 5}, {1, 6, 4, 2}};
const int32_t kCubeFaceEdges[6][4] = {{0, 1, 2, 3}, {2, 4, 5, 6},
                                      {5, 7, 8, 9}, {8, 10, 0, 11},
                                      {11, 3, 6, 9}, {10, 7, 4, 1}};

int64_t ipow4(int lev) {
  int64_t r = 1;
  for (int i = 0; i < lev; ++i) r *= 4;
  return r;
}

}  // namespace

// ---------------------------------------------------------------- geometry
namespace sphere {

double dot(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

void cross(double c[3], const double a[3], const double b[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

void normalize(double v[3]) {
  const double s = 1.0 / std::sqrt(dot(v, v));
  v[0] *= s;
  v[1] *= s;
  v[2] *= s;
}

double distance(const double a[3], const double b[3]) {
  double c[3];
  cross(c, a, b);
  return std::atan2(std::sqrt(dot(c, c)), dot(a, b));
}

// Spherical excess by L'Huilier's theorem.
double tri_area(const double a[3], const double b[3], const double c[3]) {
  const double s1 = distance(a, b);
  const double s2 = distance(b, c);
  const double s3 = distance(c, a);
  const double hp = 0.5 * (s1 + s2 + s3);
  double zz = std::tan(0.5 * hp) * std::tan(0.5 * (hp - s1)) *
              std::tan(0.5 * (hp - s2)) * std::tan(0.5 * (hp - s3));
  if (std::abs(zz) < 2.220446049250313e-16) zz = 0;
  return 4 * std::atan(std::sqrt(zz));
}

}  // namespace sphere

namespace {

// Normalised vertex mean (SphereGeometry::barycenter) and the sum of the
// spherical triangles (ctr, v_i, v_i+1) (SphereGeometry::polygon_area).
void barycenter(double ctr[3], const double (*v)[3], int n) {
  ctr[0] = ctr[1] = ctr[2] = 0;
  for (int i = 0; i < n; ++i) {
    ctr[0] += v[i][0];
    ctr[1] += v[i][1];
    ctr[2] += v[i][2];
  }
  const double inv = 1.0 / n;
  ctr[0] *= inv;
  ctr[1] *= inv;
  ctr[2] *= inv;
  sphere::normalize(ctr);
}

double polygon_area(const double ctr[3], const double (*v)[3], int n) {
  double ar = 0;
  for (int i = 0; i < n; ++i) ar += sphere::tri_area(ctr, v[i], v[(i + 1) % n]);
  return ar;
}

}  // namespace

// ------------------------------------------------------------- seed counts
int64_t SeedCounts::vertices_at_level(SeedKind k, int lev) {
  return k == SeedKind::IcosTri ? 2 + 10 * ipow4(lev) : 2 + 6 * ipow4(lev);
}
int64_t SeedCounts::leaf_faces_at_level(SeedKind k, int lev) {
  return k == SeedKind::IcosTri ? 20 * ipow4(lev) : 6 * ipow4(lev);
}
int64_t SeedCounts::edges_at_level(SeedKind, int64_t nv, int64_t nf) {
  return nv + nf - 2;
}
void SeedCounts::max_allocations(SeedKind k, int lev, int64_t& nverts,
                                 int64_t& nedges, int64_t& nfaces) {
  nverts = vertices_at_level(k, lev);
  nedges = 0;
  nfaces = 0;
  for (int i = 0; i <= lev; ++i) {
    nfaces += leaf_faces_at_level(k, i);
    nedges += edges_at_level(k, vertices_at_level(k, i),
                             leaf_faces_at_level(k, i));
  }
}

// -------------------------------------------------------------------- mesh
SphereTreeMesh::SphereTreeMesh(SeedKind kind, int depth)
    : kind_(kind), depth_(depth), nfv_(kind == SeedKind::IcosTri ? 3 : 4) {
  if (depth < 0 || depth > 10)
    throw std::invalid_argument("SphereTreeMesh: depth must be in [0, 10]");
  int64_t nv, ne, nf;
  SeedCounts::max_allocations(kind, depth, nv, ne, nf);
  vx_.reserve(nv);
  vy_.reserve(nv);
  vz_.reserve(nv);
  eo_.reserve(ne);
  ed_.reserve(ne);
  el_.reserve(ne);
  er_.reserve(ne);
  ek_.reserve(ne);
  fx_.reserve(nf);
  fy_.reserve(nf);
  fz_.reserve(nf);
  farea_.reserve(nf);
  fmask_.reserve(nf);
  fkid_.reserve(nf);
  fparent_.reserve(nf);
  flevel_.reserve(nf);
  fverts_.reserve(nf * nfv_);
  fedges_.reserve(nf * nfv_);

  load_seed();
  // Breadth-first: every face of level l is divided before any of level l+1,
  // in index order (PolyMesh2d::tree_init).
  int64_t start = 0;
  for (int lev = 0; lev < depth; ++lev) {
    const int64_t stop = n_faces();
    for (int64_t j = start; j < stop; ++j) {
      if (fkid_[j] < 0) {
        if (nfv_ == 3)
          divide_tri(static_cast<int32_t>(j));
        else
          divide_quad(static_cast<int32_t>(j));
      }
    }
    start = stop;
  }
  if (n_vertices() != nv || n_faces() != nf || n_edges() != ne)
    throw std::logic_error("SphereTreeMesh: counts disagree with closed form");
}

void SphereTreeMesh::load_seed() {
  const bool icos = kind_ == SeedKind::IcosTri;
  const int nverts = icos ? 12 : 8;
  const int nfaces = icos ? 20 : 6;
  const int nedges = icos ? 30 : 12;
  const double (*crds)[3] = icos ? kIcosCrds : kCubeCrds;
  for (int i = 0; i < nverts; ++i) add_vertex(crds[i]);
  for (int e = 0; e < nedges; ++e) {
    const int32_t* r = icos ? kIcosEdges[e] : kCubeEdges[e];
    add_edge(r[0], r[1], r[2], r[3]);
  }
  for (int f = 0; f < nfaces; ++f) {
    const int32_t* fv = icos ? kIcosFaceVerts[f] : kCubeFaceVerts[f];
    const int32_t* fe = icos ? kIcosFaceEdges[f] : kCubeFaceEdges[f];
    double vc[4][3];
    for (int k = 0; k < nfv_; ++k)
      for (int c = 0; c < 3; ++c) vc[k][c] = crds[fv[k]][c];
    // MeshSeed::face_area uses the tabulated centre, not a recomputed one.
    const double* ctr = crds[nverts + f];
    add_face(fv, fe, -1, ctr, polygon_area(ctr, vc, nfv_));
  }
}

int32_t SphereTreeMesh::add_vertex(const double p[3]) {
  vx_.push_back(p[0]);
  vy_.push_back(p[1]);
  vz_.push_back(p[2]);
  return static_cast<int32_t>(vx_.size() - 1);
}

int32_t SphereTreeMesh::add_edge(int32_t o, int32_t d, int32_t l, int32_t r) {
  eo_.push_back(o);
  ed_.push_back(d);
  el_.push_back(l);
  er_.push_back(r);
  ek_.push_back(-1);
  return static_cast<int32_t>(eo_.size() - 1);
}

void SphereTreeMesh::add_face(const int32_t* verts, const int32_t* edges,
                              int32_t parent, const double ctr[3],
                              double area) {
  fx_.push_back(ctr[0]);
  fy_.push_back(ctr[1]);
  fz_.push_back(ctr[2]);
  farea_.push_back(area);
  fmask_.push_back(0);
  fkid_.push_back(-1);
  fparent_.push_back(parent);
  flevel_.push_back(parent < 0 ? 0 : flevel_[parent] + 1);
  for (int k = 0; k < nfv_; ++k) {
    fverts_.push_back(verts[k]);
    fedges_.push_back(edges[k]);
  }
  ++n_leaves_;
}

// Midpoint vertex, then the two half edges inheriting the parent's faces.
void SphereTreeMesh::divide_edge(int32_t e) {
  const int32_t o = eo_[e], d = ed_[e];
  double m[3] = {0.5 * (vx_[o] + vx_[d]), 0.5 * (vy_[o] + vy_[d]),
                 0.5 * (vz_[o] + vz_[d])};
  sphere::normalize(m);
  const int32_t mid = add_vertex(m);
  const int32_t l = el_[e], r = er_[e];
  const int32_t k0 = add_edge(o, mid, l, r);
  add_edge(mid, d, l, r);
  ek_[e] = k0;
}

void SphereTreeMesh::divide_tri(int32_t f) {
  const int32_t kid0 = static_cast<int32_t>(n_faces());
  int32_t nv[4][3], ne[4][3];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) nv[i][j] = ne[i][j] = -1;
  for (int i = 0; i < 3; ++i) nv[i][i] = fverts_[3 * f + i];

  for (int i = 0; i < 3; ++i) {
    const int32_t pe = fedges_[3 * f + i];
    if (ek_[pe] < 0) divide_edge(pe);
    const int32_t c0 = ek_[pe], c1 = ek_[pe] + 1;
    const int ip = (i + 1) % 3;
    if (el_[pe] == f) {  // parent edge positively oriented w.r.t. face f
      ne[i][i] = c0;
      el_[c0] = kid0 + i;
      ne[ip][i] = c1;
      el_[c1] = kid0 + ip;
    } else {
      ne[i][i] = c1;
      er_[c1] = kid0 + i;
      ne[ip][i] = c0;
      er_[c0] = kid0 + ip;
    }
    const int32_t mid = ed_[c0];
    if (i == 0) {
      nv[0][1] = nv[1][0] = nv[3][2] = mid;
    } else if (i == 1) {
      nv[1][2] = nv[2][1] = nv[3][0] = mid;
    } else {
      nv[2][0] = nv[0][2] = nv[3][1] = mid;
    }
  }
  // interior edges bound the central child 3
  const int32_t e0 = static_cast<int32_t>(n_edges());
  ne[3][0] = e0;
  ne[3][1] = e0 + 1;
  ne[3][2] = e0 + 2;
  ne[0][1] = e0 + 1;
  ne[1][2] = e0 + 2;
  ne[2][0] = e0;
  add_edge(nv[2][1], nv[2][0], kid0 + 3, kid0 + 2);
  add_edge(nv[0][2], nv[0][1], kid0 + 3, kid0 + 0);
  add_edge(nv[1][0], nv[1][2], kid0 + 3, kid0 + 1);

  double ctr[4][3], area[4];
  for (int i = 0; i < 4; ++i) {
    double vc[3][3];
    for (int j = 0; j < 3; ++j) {
      vc[j][0] = vx_[nv[i][j]];
      vc[j][1] = vy_[nv[i][j]];
      vc[j][2] = vz_[nv[i][j]];
    }
    barycenter(ctr[i], vc, 3);
    area[i] = polygon_area(ctr[i], vc, 3);
  }
  for (int i = 0; i < 4; ++i) add_face(nv[i], ne[i], f, ctr[i], area[i]);
  fkid_[f] = kid0;
  farea_[f] = 0.0;
  fmask_[f] = 1;
  --n_leaves_;
}

void SphereTreeMesh::divide_quad(int32_t f) {
  const int32_t kid0 = static_cast<int32_t>(n_faces());
  int32_t nv[4][4], ne[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) nv[i][j] = ne[i][j] = -1;
  for (int i = 0; i < 4; ++i) nv[i][i] = fverts_[4 * f + i];

  for (int i = 0; i < 4; ++i) {
    const int32_t pe = fedges_[4 * f + i];
    if (ek_[pe] < 0) divide_edge(pe);
    const int32_t c0 = ek_[pe], c1 = ek_[pe] + 1;
    const int ip = (i + 1) % 4;
    if (el_[pe] == f) {
      ne[i][i] = c0;
      el_[c0] = kid0 + i;
      ne[ip][i] = c1;
      el_[c1] = kid0 + ip;
    } else {
      ne[i][i] = c1;
      er_[c1] = kid0 + i;
      ne[ip][i] = c0;
      er_[c0] = kid0 + ip;
    }
    const int32_t mid = ed_[c0];
    nv[i][ip] = mid;
    nv[ip][i] = mid;
  }
  // the parent's centre becomes the shared corner of the four children
  const double pc[3] = {fx_[f], fy_[f], fz_[f]};
  const int32_t vc_idx = add_vertex(pc);
  for (int i = 0; i < 4; ++i) nv[i][(i + 2) % 4] = vc_idx;

  const int32_t e0 = static_cast<int32_t>(n_edges());
  add_edge(nv[0][1], nv[0][2], kid0 + 0, kid0 + 1);
  ne[0][1] = e0;
  ne[1][3] = e0;
  add_edge(nv[2][0], nv[2][3], kid0 + 3, kid0 + 2);
  ne[2][3] = e0 + 1;
  ne[3][1] = e0 + 1;
  add_edge(nv[2][1], nv[2][0], kid0 + 1, kid0 + 2);
  ne[1][2] = e0 + 2;
  ne[2][0] = e0 + 2;
  add_edge(nv[3][1], nv[3][0], kid0 + 0, kid0 + 3);
  ne[0][2] = e0 + 3;
  ne[3][0] = e0 + 3;

  double ctr[4][3], area[4];
  for (int i = 0; i < 4; ++i) {
    double vc[4][3];
    for (int j = 0; j < 4; ++j) {
      vc[j][0] = vx_[nv[i][j]];
      vc[j][1] = vy_[nv[i][j]];
      vc[j][2] = vz_[nv[i][j]];
    }
    barycenter(ctr[i], vc, 4);
    area[i] = polygon_area(ctr[i], vc, 4);
  }
  for (int i = 0; i < 4; ++i) add_face(nv[i], ne[i], f, ctr[i], area[i]);
  fkid_[f] = kid0;
  farea_[f] = 0.0;
  fmask_[f] = 1;
  --n_leaves_;
}

bool SphereTreeMesh::divide_face(int32_t f) {
  if (f < 0 || f >= n_faces())
    throw std::out_of_range("SphereTreeMesh::divide_face: bad face index");
  if (fkid_[f] >= 0) return false;
  if (nfv_ == 3)
    divide_tri(f);
  else
    divide_quad(f);
  return true;
}

int64_t SphereTreeMesh::divide_flagged(const uint8_t* flags, int max_level,
                                       bool* limit_reached) {
  const int64_t n_in = n_faces();
  int64_t count = 0;
  bool limit = false;
  for (int64_t i = 0; i < n_in; ++i) {
    if (!flags[i]) continue;
    if (flevel_[i] <= max_level) {
      divide_face(static_cast<int32_t>(i));
      ++count;
    } else {
      limit = true;
    }
  }
  if (limit_reached) *limit_reached = limit;
  return count;
}

double SphereTreeMesh::surface_area() const {
  double s = 0;
  for (double a : farea_) s += a;
  return s;
}

}  // namespace lpm_amd
