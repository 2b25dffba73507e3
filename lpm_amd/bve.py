"""BVESphere / BVERK4: the reference's barotropic-vorticity API over the ABI.

Mirrors ``BVESphere<Seed>`` (src/lpm_bve_sphere.hpp:24-109,
src/lpm_bve_sphere_impl.hpp:14-231) and ``BVERK4`` (src/lpm_bve_rk4.hpp:10-77,
src/lpm_bve_rk4_impl.hpp:55-167): same member names and method meanings.
Fields live in device memory as (3, n) / (n,) float64 tensors (SoA, the
Kokkos GPU LayoutLeft).  Every direct sum runs in the library's HIP kernels.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _abi
from .device import Context, crd_args
from .mesh import SphereMesh


class BVESphere:
    """Particles and fields of the BVE on the unit sphere.

    ``BVESphere(seed, depth)`` builds the uniformly refined mesh (the
    reference constructs with max allocations then calls tree_init).
    """

    def __init__(self, seed, depth, ctx: Context = None, device=None,
                 mesh: SphereMesh = None):
        self.ctx = ctx if ctx is not None else Context()
        dev = torch.device("cuda", self.ctx.device) if device is None else device
        self.mesh = mesh if mesh is not None else SphereMesh.build(seed, depth)
        m = self.mesh
        f64 = dict(dtype=torch.float64, device=dev)
        self.vert_x = torch.tensor(m.vert_x, **f64)
        self.face_x = torch.tensor(m.face_x, **f64)
        self.face_area = torch.tensor(m.face_area, **f64)
        self.face_mask = torch.tensor(m.face_mask, dtype=torch.uint8, device=dev)
        nv, nf = m.n_verts, m.n_faces
        self.rel_vort_verts = torch.zeros(nv, **f64)
        self.abs_vort_verts = torch.zeros(nv, **f64)
        self.stream_fn_verts = torch.zeros(nv, **f64)
        self.velocity_verts = torch.zeros((3, nv), **f64)
        self.rel_vort_faces = torch.zeros(nf, **f64)
        self.abs_vort_faces = torch.zeros(nf, **f64)
        self.stream_fn_faces = torch.zeros(nf, **f64)
        self.velocity_faces = torch.zeros((3, nf), **f64)
        self.Omega = 2 * np.pi
        self.t = 0.0
        self._omg_set = False

    # counts as in the reference
    def n_vertices_host(self) -> int:
        return self.mesh.n_verts

    def n_faces_host(self) -> int:
        return self.mesh.n_faces

    def set_omega(self, omg: float):
        """BVESphere::set_omega: only the first call takes effect."""
        if not self._omg_set:
            self.Omega = float(omg)
            self._omg_set = True

    def init_vorticity(self, vorticity_fn):
        """zeta = fn(x) at vertices and faces; absolute vorticity zeta + 2 Omega z
        (src/lpm_bve_sphere_impl.hpp:148-179).  ``fn`` maps a (3, n) numpy
        array to n values."""
        vx = self.vert_x.cpu().numpy()
        fx = self.face_x.cpu().numpy()
        zv = np.asarray(vorticity_fn(vx), dtype=np.float64)
        zf = np.asarray(vorticity_fn(fx), dtype=np.float64)
        self.rel_vort_verts.copy_(torch.from_numpy(zv))
        self.rel_vort_faces.copy_(torch.from_numpy(zf))
        self.abs_vort_verts.copy_(torch.from_numpy(zv + 2 * self.Omega * vx[2]))
        self.abs_vort_faces.copy_(torch.from_numpy(zf + 2 * self.Omega * fx[2]))

    def fields(self) -> _abi.BveFields:
        f = _abi.BveFields()
        vp, ves, vcs, nv = crd_args(self.vert_x)
        up, ues, ucs, _ = crd_args(self.velocity_verts)
        fp, fes, fcs, nf = crd_args(self.face_x)
        gp, ges, gcs, _ = crd_args(self.velocity_faces)
        f.n_verts, f.vert_x, f.vert_x_es, f.vert_x_cs = nv, vp, ves, vcs
        f.vert_zeta = self.rel_vort_verts.data_ptr()
        f.vert_u, f.vert_u_es, f.vert_u_cs = up, ues, ucs
        f.n_faces, f.face_x, f.face_x_es, f.face_x_cs = nf, fp, fes, fcs
        f.face_zeta = self.rel_vort_faces.data_ptr()
        f.face_u, f.face_u_es, f.face_u_cs = gp, ges, gcs
        f.face_area = self.face_area.data_ptr()
        f.face_mask = self.face_mask.data_ptr()
        return f

    def init_velocity(self, precision="fp64", flags=0):
        """BVESphere::init_velocity (src/lpm_bve_sphere_impl.hpp:181-205)."""
        self.ctx.bve_init_velocity(self.fields(), precision, flags)

    def init_stream_fn(self, flags=0):
        """BVESphere::init_stream_fn (src/lpm_bve_sphere_impl.hpp:207-231)."""
        c = self.ctx
        c.bve_stream_fn(self.vert_x, self.face_x, self.rel_vort_faces,
                        self.face_area, self.face_mask, collocated=False,
                        flags=flags, out=self.stream_fn_verts)
        c.bve_stream_fn(self.face_x, self.face_x, self.rel_vort_faces,
                        self.face_area, self.face_mask, collocated=True,
                        flags=flags, out=self.stream_fn_faces)

    def avg_mesh_size_radians(self) -> float:
        return self.mesh.mean_spacing()


class BVERK4:
    """Classical RK4 for BVESphere (4 direct-sum evaluations per step)."""

    def __init__(self, dt: float, sph: BVESphere, precision="fp64",
                 reference_compat=True, renormalize=False, leaf_only=False,
                 use_graph=False):
        self.dt = float(dt)
        self.Omega = sph.Omega  # copied at construction, as in the reference
        self.nverts = sph.n_vertices_host()
        self.nfaces = sph.n_faces_host()
        self.precision = precision
        flags = 0
        if reference_compat:
            flags |= _abi.LPM_RK4_REFERENCE_COMPAT
        if renormalize:
            flags |= _abi.LPM_RK4_RENORMALIZE
        if leaf_only:
            flags |= _abi.LPM_FLAG_SKIP_MASKED_TARGETS
        if use_graph:
            flags |= _abi.LPM_RK4_USE_GRAPH
        self.flags = flags
        self._fields = None

    def advance_timestep(self, sph: BVESphere):
        if self._fields is None:
            self._fields = sph.fields()
        sph.ctx.bve_rk4_step(self._fields, self.dt, self.Omega, self.precision,
                             self.flags)
        sph.t += self.dt
