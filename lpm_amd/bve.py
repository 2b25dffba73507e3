"""BVESphere / BVERK4: the reference's barotropic-vorticity API over the ABI.

Mirrors ``BVESphere<Seed>`` (src/lpm_bve_sphere.hpp:24-109,
src/lpm_bve_sphere_impl.hpp:14-231) and ``BVERK4`` (src/lpm_bve_rk4.hpp:10-77,
src/lpm_bve_rk4_impl.hpp:55-167): same member names and method meanings.
Fields live in device memory as (3, n) / (n,) float64 tensors (SoA, the
Kokkos GPU LayoutLeft).  Every direct sum runs in the library's HIP kernels.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _abi
from .device import Context, crd_args, warn_if_mixed_off_sphere
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
        warn_if_mixed_off_sphere(precision, renormalize)
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


class BVEEnsemble:
    """Independent BVE problems on the unit sphere stored back to back and
    evaluated by one persistent launch per direct sum
    (``lpm_bve_ensemble_*``).  Member m owns vertices
    ``vert_offsets[m]:vert_offsets[m+1]`` and faces
    ``face_offsets[m]:face_offsets[m+1]`` of the concatenated (3, n) / (n,)
    tensors.  No reference counterpart: the reference evolves one BVESphere
    at a time; each member here behaves as a BVESphere of its own mesh."""

    def __init__(self, meshes, ctx: Context = None, device=None):
        self.ctx = ctx if ctx is not None else Context()
        dev = torch.device("cuda", self.ctx.device) if device is None else device
        self.meshes = list(meshes)
        assert self.meshes, "an ensemble needs at least one member"
        self.vert_offsets = np.concatenate(
            [[0], np.cumsum([m.n_verts for m in self.meshes])]).astype(np.int64)
        self.face_offsets = np.concatenate(
            [[0], np.cumsum([m.n_faces for m in self.meshes])]).astype(np.int64)
        f64 = dict(dtype=torch.float64, device=dev)
        self.vert_x = torch.tensor(
            np.concatenate([m.vert_x for m in self.meshes], axis=1), **f64)
        self.face_x = torch.tensor(
            np.concatenate([m.face_x for m in self.meshes], axis=1), **f64)
        self.face_area = torch.tensor(
            np.concatenate([m.face_area for m in self.meshes]), **f64)
        self.face_mask = torch.tensor(
            np.concatenate([m.face_mask for m in self.meshes]),
            dtype=torch.uint8, device=dev)
        nv, nf = int(self.vert_offsets[-1]), int(self.face_offsets[-1])
        self.rel_vort_verts = torch.zeros(nv, **f64)
        self.rel_vort_faces = torch.zeros(nf, **f64)
        self.velocity_verts = torch.zeros((3, nv), **f64)
        self.velocity_faces = torch.zeros((3, nf), **f64)
        self.t = 0.0
        self._c_voff = (ctypes.c_int64 * len(self.vert_offsets))(
            *self.vert_offsets.tolist())
        self._c_foff = (ctypes.c_int64 * len(self.face_offsets))(
            *self.face_offsets.tolist())

    @property
    def n_members(self) -> int:
        return len(self.meshes)

    def vert_slice(self, m) -> slice:
        return slice(int(self.vert_offsets[m]), int(self.vert_offsets[m + 1]))

    def face_slice(self, m) -> slice:
        return slice(int(self.face_offsets[m]), int(self.face_offsets[m + 1]))

    def useful_pairs(self) -> int:
        """Sum over members of (vertices + leaf faces) x leaf faces."""
        return sum(m.useful_pairs() for m in self.meshes)

    def init_vorticity(self, fns):
        """zeta of member m = fns[m](x) for x (3, n) numpy points."""
        vx = self.vert_x.cpu().numpy()
        fx = self.face_x.cpu().numpy()
        zv = np.empty(vx.shape[1])
        zf = np.empty(fx.shape[1])
        for m, fn in enumerate(fns):
            vs, fs = self.vert_slice(m), self.face_slice(m)
            zv[vs] = fn(vx[:, vs])
            zf[fs] = fn(fx[:, fs])
        self.rel_vort_verts.copy_(torch.from_numpy(zv))
        self.rel_vort_faces.copy_(torch.from_numpy(zf))

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

    def descriptor(self) -> _abi.BveEnsemble:
        e = _abi.BveEnsemble()
        e.n_members = self.n_members
        e.vert_offsets = ctypes.cast(self._c_voff, _abi.c_i64p)
        e.face_offsets = ctypes.cast(self._c_foff, _abi.c_i64p)
        e.fields = self.fields()
        return e

    def init_velocity(self, precision="fp64", flags=0):
        self.ctx.bve_ensemble_init_velocity(self.descriptor(), precision, flags)

    def autotune(self, precision="fp64", flags=0, repeats=1):
        return self.ctx.bve_ensemble_autotune(self.descriptor(), precision,
                                              flags, repeats)


class BVEEnsembleRK4:
    """RK4 of every member of a BVEEnsemble (4 ensemble launches per step);
    flags as BVERK4."""

    def __init__(self, dt: float, ens: BVEEnsemble, Omega=2 * np.pi,
                 precision="fp64", reference_compat=True, renormalize=False,
                 leaf_only=False, use_graph=False, stream=None):
        self.dt = float(dt)
        self.Omega = float(Omega)
        self.precision = precision
        warn_if_mixed_off_sphere(precision, renormalize)
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
        self.stream = stream
        self._desc = None

    def advance_timestep(self, ens: BVEEnsemble):
        if self._desc is None:
            self._desc = ens.descriptor()
        ens.ctx.bve_ensemble_rk4_step(self._desc, self.dt, self.Omega,
                                      self.precision, self.flags, self.stream)
        ens.t += self.dt
