"""Spherical quadtree meshes (host arrays) from the library's generator.

``SphereMesh(seed, depth)`` mirrors what LPM's MeshSeed<Seed> + PolyMesh2d::
tree_init produce (src/mesh/lpm_polymesh2d_impl.hpp:25-42): vertices
(passive particles) and all stored faces (active particles; divided faces
have area 0 and mask 1), in the reference's creation order.
Arrays are SoA, shape (3, n), like a Kokkos LayoutLeft View<double*[3]>.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi

SEEDS = {"icos": _abi.LPM_SEED_ICOS_TRI, "IcosTriSphereSeed": _abi.LPM_SEED_ICOS_TRI,
         "cube": _abi.LPM_SEED_CUBED_SPHERE,
         "CubedSphereSeed": _abi.LPM_SEED_CUBED_SPHERE}


def seed_id(seed) -> int:
    if isinstance(seed, int):
        return seed
    return SEEDS[seed]


def max_allocations(seed, depth):
    """(n_verts, n_edges, n_faces) of MeshSeed::set_max_allocations."""
    L = _abi.lib()
    nv, ne, nf = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _abi.check(L.lpm_mesh_max_allocations(seed_id(seed), depth, ctypes.byref(nv),
                                          ctypes.byref(ne), ctypes.byref(nf)))
    return nv.value, ne.value, nf.value


@dataclass
class SphereMesh:
    seed: int
    depth: int
    vert_x: np.ndarray    # (3, n_verts)
    face_x: np.ndarray    # (3, n_faces)
    face_area: np.ndarray  # (n_faces,)
    face_mask: np.ndarray  # (n_faces,) uint8, 1 = divided
    face_parent: np.ndarray
    face_verts: np.ndarray  # (n_faces, verts_per_face)
    n_leaves: int
    n_edges: int

    face_level: np.ndarray = None  # (n_faces,) int32, 0 = seed face

    @classmethod
    def build(cls, seed, depth: int) -> "SphereMesh":
        """Uniform refinement to ``depth`` (PolyMesh2d::tree_init)."""
        return PolyMesh2d(seed, depth).sphere_mesh()

    @property
    def n_verts(self) -> int:
        return self.vert_x.shape[1]

    @property
    def n_faces(self) -> int:
        return self.face_x.shape[1]

    @property
    def leaf_slice(self) -> slice:
        """Leaves of a uniform refinement are the last n_leaves faces (not
        so after adaptive refinement: use face_mask == 0)."""
        return slice(self.n_faces - self.n_leaves, self.n_faces)

    def useful_pairs(self) -> int:
        """(N_V + N_F) * N_F: targets times undivided sources."""
        return (self.n_verts + self.n_leaves) * self.n_leaves

    def mean_spacing(self) -> float:
        """sqrt(4 pi / N_F), the mean leaf-panel size in radians."""
        return float(np.sqrt(4 * np.pi / self.n_leaves))


class PolyMesh2d:
    """The library's host quadtree mesh (PolyMesh2d<Seed>,
    src/mesh/lpm_polymesh2d.hpp, src/mesh/lpm_polymesh2d_impl.hpp):
    uniform refinement to ``depth`` (tree_init), then optional adaptive
    refinement (``divide_flagged_faces``); ``sphere_mesh()`` copies the
    arrays out.  ``amr_limit`` and ``max_faces`` play the reference's
    PolyMeshParameters amr_limit and nmaxfaces (0: no allocation limit)."""

    def __init__(self, seed, depth: int, amr_limit: int = 0,
                 max_faces: int = 0):
        self.seed = seed_id(seed)
        self.depth = int(depth)
        self.amr_limit = int(amr_limit)
        self.max_faces = int(max_faces)
        self._L = _abi.lib()
        self._h = ctypes.c_void_p()
        _abi.check(self._L.lpm_mesh_create(self.seed, self.depth,
                                           ctypes.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._L.lpm_mesh_destroy(h)
            self._h = None

    def sizes(self):
        """(n_verts, n_faces, n_leaves, n_edges)."""
        v = [ctypes.c_int64() for _ in range(4)]
        _abi.check(self._L.lpm_mesh_sizes(self._h, *map(ctypes.byref, v)))
        return tuple(x.value for x in v)

    def n_faces_host(self) -> int:
        return self.sizes()[1]

    def divide_flagged_faces(self, flags):
        """Divide the flagged faces (one flag per stored face, index order,
        within the level limit depth + amr_limit).  Returns (refine count,
        limit reached); refine count -1: not enough room under max_faces,
        nothing divided."""
        flags = np.ascontiguousarray(flags, dtype=np.uint8)
        n = ctypes.c_int64()
        lim = ctypes.c_int()
        _abi.check(self._L.lpm_mesh_divide_flagged_faces(
            self._h, flags.ctypes.data, len(flags), self.amr_limit,
            self.max_faces, ctypes.byref(n), ctypes.byref(lim)))
        return n.value, bool(lim.value)

    def sphere_mesh(self) -> SphereMesh:
        L, h = self._L, self._h
        nv, nf, nl, ne = self.sizes()
        vx = np.empty((3, nv))
        fx = np.empty((3, nf))
        fa = np.empty(nf)
        fm = np.empty(nf, np.uint8)
        fp = np.empty(nf, np.int32)
        fl = np.empty(nf, np.int32)
        _abi.check(L.lpm_mesh_copy_vertices(h, vx.ctypes.data))
        _abi.check(L.lpm_mesh_copy_faces(h, fx.ctypes.data, fa.ctypes.data,
                                         fm.ctypes.data, fp.ctypes.data))
        _abi.check(L.lpm_mesh_copy_face_levels(h, fl.ctypes.data))
        nfv = ctypes.c_int()
        _abi.check(L.lpm_mesh_copy_face_vertices(h, None, ctypes.byref(nfv)))
        fv = np.empty((nf, nfv.value), np.int32)
        _abi.check(L.lpm_mesh_copy_face_vertices(h, fv.ctypes.data,
                                                 ctypes.byref(nfv)))
        return SphereMesh(self.seed, self.depth, vx, fx, fa, fm, fp, fv, nl,
                          ne, fl)
