"""SWE / SWERK4: the reference's sphere shallow-water API over the ABI.

Mirrors ``SWE<Seed>`` (src/lpm_swe.hpp:18-177, src/lpm_swe_impl.hpp) for the
sphere with zero bottom topography, and a sphere RK4 stepper in the stage
layout of ``SWERK4`` (src/lpm_swe_rk4.hpp:8-126,
src/lpm_swe_rk4_impl.hpp:204-445).  Fields are device tensors; the direct
sums, tendencies and stage updates run in the library's HIP kernels.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _abi
from .device import Context, crd_args
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
                 reference_compat=True, leaf_only=False):
        self.dt = float(dt)
        self.t_idx = 0
        self.precision = precision
        flags = 0
        if reference_compat:
            flags |= _abi.LPM_RK4_REFERENCE_COMPAT
        if leaf_only:
            flags |= _abi.LPM_FLAG_SKIP_MASKED_TARGETS
        self.flags = flags
        self._fields = None

    def advance_timestep_impl(self, swe: SWE):
        if self._fields is None:
            self._fields = swe.fields()
        swe.ctx.swe_rk4_step(self._fields, self.dt, swe.Omega, swe.g, swe.eps,
                             swe.pse_eps, self.precision, self.flags)
        self.t_idx += 1
