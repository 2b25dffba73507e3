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

    @classmethod
    def build(cls, seed, depth: int) -> "SphereMesh":
        L = _abi.lib()
        h = ctypes.c_void_p()
        _abi.check(L.lpm_mesh_create(seed_id(seed), depth, ctypes.byref(h)))
        try:
            nv, nf, nl, ne = (ctypes.c_int64() for _ in range(4))
            _abi.check(L.lpm_mesh_sizes(h, ctypes.byref(nv), ctypes.byref(nf),
                                        ctypes.byref(nl), ctypes.byref(ne)))
            vx = np.empty((3, nv.value))
            fx = np.empty((3, nf.value))
            fa = np.empty(nf.value)
            fm = np.empty(nf.value, np.uint8)
            fp = np.empty(nf.value, np.int32)
            _abi.check(L.lpm_mesh_copy_vertices(h, vx.ctypes.data))
            _abi.check(L.lpm_mesh_copy_faces(h, fx.ctypes.data, fa.ctypes.data,
                                             fm.ctypes.data, fp.ctypes.data))
            nfv = ctypes.c_int()
            _abi.check(L.lpm_mesh_copy_face_vertices(h, None, ctypes.byref(nfv)))
            fv = np.empty((nf.value, nfv.value), np.int32)
            _abi.check(L.lpm_mesh_copy_face_vertices(h, fv.ctypes.data,
                                                     ctypes.byref(nfv)))
        finally:
            L.lpm_mesh_destroy(h)
        return cls(seed_id(seed), depth, vx, fx, fa, fm, fp, fv, nl.value,
                   ne.value)

    @property
    def n_verts(self) -> int:
        return self.vert_x.shape[1]

    @property
    def n_faces(self) -> int:
        return self.face_x.shape[1]

    @property
    def leaf_slice(self) -> slice:
        """Leaves of a uniform refinement are the last n_leaves faces."""
        return slice(self.n_faces - self.n_leaves, self.n_faces)

    def useful_pairs(self) -> int:
        """(N_V + N_F) * N_F: targets times undivided sources."""
        return (self.n_verts + self.n_leaves) * self.n_leaves

    def mean_spacing(self) -> float:
        """sqrt(4 pi / N_F), the mean leaf-panel size in radians."""
        return float(np.sqrt(4 * np.pi / self.n_leaves))
