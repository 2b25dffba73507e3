"""SWE / SWERK4 / SWERK2: the reference's sphere shallow-water API over the
ABI.

Mirrors ``SWE<Seed>`` (src/lpm_swe.hpp:18-177, src/lpm_swe_impl.hpp) for the
sphere with zero bottom topography, and a sphere RK4 stepper in the stage
layout of ``SWERK4`` (src/lpm_swe_rk4.hpp:8-126,
src/lpm_swe_rk4_impl.hpp:204-445), and the RK2 stepper of ``SWERK2``
(src/lpm_swe_rk2_impl.hpp:80-257).  Fields are device tensors; the direct
sums, tendencies and stage updates run in the library's HIP kernels.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _abi
from .device import Context, crd_args, warn_if_mixed_off_sphere
from .mesh import SphereMesh


class SWE:
    """Particles and fields of the sphere SWE (passive = vertices, active =
    faces), with CoriolisSphere(Omega) and gravity g."""

    def __init__(self, seed, depth, Omega=2 * np.pi, g=1.0, ctx: Context = None,
                 mesh: SphereMesh = None):
        self.ctx = ctx if ctx is not None else Context()
        dev = torch.device("cuda", self.ctx.device)
        self.mesh = mesh if mesh is not None else SphereMesh.build(seed, depth)
        m = self.mesh
        f64 = dict(dtype=torch.float64, device=dev)
        nv, nf = m.n_verts, m.n_faces
        self.vert_x = torch.tensor(m.vert_x, **f64)
        self.face_x = torch.tensor(m.face_x, **f64)
        self.face_area = torch.tensor(m.face_area, **f64)
        self.face_mask = torch.tensor(m.face_mask, dtype=torch.uint8, device=dev)
        for kind, n in (("passive", nv), ("active", nf)):
            for name in ("rel_vort", "div", "depth", "double_dot", "surf_lap",
                         "surf"):
                setattr(self, f"{name}_{kind}", torch.zeros(n, **f64))
            setattr(self, f"velocity_{kind}", torch.zeros((3, n), **f64))
        self.mass_active = torch.zeros(nf, **f64)
        self.Omega = float(Omega)
        self.g = float(g)
        self.eps = 0.0
        self.pse_eps = 0.0
        self.t = 0.0

    def n_vertices_host(self):
        return self.mesh.n_verts

    def n_faces_host(self):
        return self.mesh.n_faces

    def set_kernel_parameters(self, vel_eps: float, pse_eps: float):
        """SWE::set_kernel_parameters (src/lpm_swe_impl.hpp:98-107); here the
        sphere values are set explicitly (the reference leaves them unset)."""
        assert vel_eps >= 0 and pse_eps > 0
        self.eps = float(vel_eps)
        self.pse_eps = float(pse_eps)

    def init_surface(self, surface_fn):
        """Surface s(x) with zero bottom: depth = s, mass = s A
        (SWE::init_surface, src/lpm_swe_impl.hpp:302-335)."""
        sv = np.asarray(surface_fn(self.mesh.vert_x), dtype=np.float64)
        sf = np.asarray(surface_fn(self.mesh.face_x), dtype=np.float64)
        self.surf_passive.copy_(torch.from_numpy(sv))
        self.depth_passive.copy_(torch.from_numpy(sv))
        self.surf_active.copy_(torch.from_numpy(sf))
        self.depth_active.copy_(torch.from_numpy(sf))
        self.mass_active.copy_(torch.from_numpy(sf * self.mesh.face_area))

    def init_vorticity(self, fn):
        self.rel_vort_passive.copy_(torch.from_numpy(
            np.asarray(fn(self.mesh.vert_x), dtype=np.float64)))
        self.rel_vort_active.copy_(torch.from_numpy(
            np.asarray(fn(self.mesh.face_x), dtype=np.float64)))

    def init_divergence(self, fn):
        self.div_passive.copy_(torch.from_numpy(
            np.asarray(fn(self.mesh.vert_x), dtype=np.float64)))
        self.div_active.copy_(torch.from_numpy(
            np.asarray(fn(self.mesh.face_x), dtype=np.float64)))

    def fields(self) -> _abi.SweFields:
        f = _abi.SweFields()
        vp, ves, vcs, nv = crd_args(self.vert_x)
        up, ues, ucs, _ = crd_args(self.velocity_passive)
        fp, fes, fcs, nf = crd_args(self.face_x)
        gp, ges, gcs, _ = crd_args(self.velocity_active)
        f.n_verts, f.vert_x, f.vert_x_es, f.vert_x_cs = nv, vp, ves, vcs
        f.vert_zeta = self.rel_vort_passive.data_ptr()
        f.vert_sigma = self.div_passive.data_ptr()
        f.vert_h = self.depth_passive.data_ptr()
        f.vert_u, f.vert_u_es, f.vert_u_cs = up, ues, ucs
        f.vert_ddot = self.double_dot_passive.data_ptr()
        f.vert_laps = self.surf_lap_passive.data_ptr()
        f.vert_surf = self.surf_passive.data_ptr()
        f.n_faces, f.face_x, f.face_x_es, f.face_x_cs = nf, fp, fes, fcs
        f.face_zeta = self.rel_vort_active.data_ptr()
        f.face_sigma = self.div_active.data_ptr()
        f.face_area = self.face_area.data_ptr()
        f.face_mass = self.mass_active.data_ptr()
        f.face_h = self.depth_active.data_ptr()
        f.face_u, f.face_u_es, f.face_u_cs = gp, ges, gcs
        f.face_ddot = self.double_dot_active.data_ptr()
        f.face_laps = self.surf_lap_active.data_ptr()
        f.face_surf = self.surf_active.data_ptr()
        f.face_mask = self.face_mask.data_ptr()
        return f

    def init_direct_sums(self, precision="fp64", flags=0):
        """Velocity, double dot and surface Laplacian at the current state
        (SWE::init_direct_sums, src/lpm_swe_impl.hpp:401-445)."""
        self.ctx.swe_init_sums(self.fields(), self.eps, self.pse_eps, precision,
                               flags)

    def advance_timestep(self, solver: "SWERK4"):
        solver.advance_timestep_impl(self)
        self.t = solver.t_idx * solver.dt


class SWERK4:
    """Sphere RK4 for SWE (4 direct-sum evaluations per step)."""

    def __init__(self, dt: float, swe: SWE, precision="fp64",
                 reference_compat=True, leaf_only=False, use_graph=False,
                 stream=None, renormalize=False):
        self.dt = float(dt)
        self.t_idx = 0
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
        self._fields = None

    def advance_timestep_impl(self, swe: SWE):
        if self._fields is None:
            self._fields = swe.fields()
        swe.ctx.swe_rk4_step(self._fields, self.dt, swe.Omega, swe.g, swe.eps,
                             swe.pse_eps, self.precision, self.flags,
                             self.stream)
        self.t_idx += 1


class SWERK2(SWERK4):
    """Sphere RK2 (Heun) for SWE in the stage layout of ``SWERK2``
    (src/lpm_swe_rk2.hpp:17-60, src/lpm_swe_rk2_impl.hpp:80-257): 2 direct-sum
    evaluations per step, the surface Laplacian from the PSE sums where the
    reference evaluates GMLS.  The reference's ``topo`` (zero bottom here) and
    ``gmls_params`` have no counterpart."""

    def __init__(self, dt: float, swe: SWE, precision="fp64", leaf_only=False,
                 use_graph=False, stream=None, renormalize=False):
        super().__init__(dt, swe, precision, reference_compat=False,
                         leaf_only=leaf_only, use_graph=use_graph,
                         stream=stream, renormalize=renormalize)

    def advance_timestep_impl(self, swe: SWE):
        if self._fields is None:
            self._fields = swe.fields()
        swe.ctx.swe_rk2_step(self._fields, self.dt, swe.Omega, swe.g, swe.eps,
                             swe.pse_eps, self.precision, self.flags,
                             self.stream)
        self.t_idx += 1


class SWEEnsemble(SWE):
    """Independent sphere SWE problems stored back to back (same field names
    as SWE, concatenated), evaluated by one persistent launch per direct sum
    (``lpm_swe_ensemble_*``).  Member m owns vertices
    ``vert_offsets[m]:vert_offsets[m+1]`` and faces
    ``face_offsets[m]:face_offsets[m+1]``.  No reference counterpart."""

    def __init__(self, meshes, Omega=2 * np.pi, g=1.0, ctx: Context = None):
        self.meshes = list(meshes)
        assert self.meshes, "an ensemble needs at least one member"
        ms = self.meshes
        # concatenated particles (parents and corner indices stay per member)
        cat = SphereMesh(
            seed=-1, depth=-1,
            vert_x=np.concatenate([m.vert_x for m in ms], axis=1),
            face_x=np.concatenate([m.face_x for m in ms], axis=1),
            face_area=np.concatenate([m.face_area for m in ms]),
            face_mask=np.concatenate([m.face_mask for m in ms]),
            face_parent=None, face_verts=None,
            n_leaves=sum(m.n_leaves for m in ms),
            n_edges=sum(m.n_edges for m in ms))
        super().__init__(None, None, Omega, g, ctx, mesh=cat)
        self.vert_offsets = np.concatenate(
            [[0], np.cumsum([m.n_verts for m in self.meshes])]).astype(np.int64)
        self.face_offsets = np.concatenate(
            [[0], np.cumsum([m.n_faces for m in self.meshes])]).astype(np.int64)
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
        return sum(m.useful_pairs() for m in self.meshes)

    def _per_member(self, fns, x, slicer):
        out = np.empty(x.shape[1])
        for m, fn in enumerate(fns):
            sl = slicer(m)
            out[sl] = fn(x[:, sl])
        return out

    def init_surface(self, fns):
        """Surface of member m = fns[m](x); zero bottom (depth = surface)."""
        sv = self._per_member(fns, self.mesh.vert_x, self.vert_slice)
        sf = self._per_member(fns, self.mesh.face_x, self.face_slice)
        self.surf_passive.copy_(torch.from_numpy(sv))
        self.depth_passive.copy_(torch.from_numpy(sv))
        self.surf_active.copy_(torch.from_numpy(sf))
        self.depth_active.copy_(torch.from_numpy(sf))
        self.mass_active.copy_(torch.from_numpy(sf * self.mesh.face_area))

    def init_vorticity(self, fns):
        self.rel_vort_passive.copy_(torch.from_numpy(
            self._per_member(fns, self.mesh.vert_x, self.vert_slice)))
        self.rel_vort_active.copy_(torch.from_numpy(
            self._per_member(fns, self.mesh.face_x, self.face_slice)))

    def init_divergence(self, fns):
        self.div_passive.copy_(torch.from_numpy(
            self._per_member(fns, self.mesh.vert_x, self.vert_slice)))
        self.div_active.copy_(torch.from_numpy(
            self._per_member(fns, self.mesh.face_x, self.face_slice)))

    def descriptor(self) -> _abi.SweEnsemble:
        e = _abi.SweEnsemble()
        e.n_members = self.n_members
        e.vert_offsets = ctypes.cast(self._c_voff, _abi.c_i64p)
        e.face_offsets = ctypes.cast(self._c_foff, _abi.c_i64p)
        e.fields = self.fields()
        return e

    def init_direct_sums(self, precision="fp64", flags=0):
        self.ctx.swe_ensemble_init_sums(self.descriptor(), self.eps,
                                        self.pse_eps, precision, flags)


class SWEEnsembleRK4(SWERK4):
    """RK4 of every member of an SWEEnsemble (flags as SWERK4)."""

    def advance_timestep_impl(self, swe: SWEEnsemble):
        if self._fields is None:
            self._fields = swe.descriptor()
        swe.ctx.swe_ensemble_rk4_step(self._fields, self.dt, swe.Omega, swe.g,
                                      swe.eps, swe.pse_eps, self.precision,
                                      self.flags, self.stream)
        self.t_idx += 1


class SWEEnsembleRK2(SWERK2):
    """RK2 of every member of an SWEEnsemble (flags as SWERK2)."""

    def advance_timestep_impl(self, swe: SWEEnsemble):
        if self._fields is None:
            self._fields = swe.descriptor()
        swe.ctx.swe_ensemble_rk2_step(self._fields, self.dt, swe.Omega, swe.g,
                                      swe.eps, swe.pse_eps, self.precision,
                                      self.flags, self.stream)
        self.t_idx += 1
