"""Incompressible2D / Incompressible2DRK2 on the sphere over the ABI.

Mirrors ``Incompressible2D<Seed>`` (src/lpm_incompressible2d.hpp,
src/lpm_incompressible2d_impl.hpp:18-254) with its fused velocity + stream
function direct sums, and ``Incompressible2DRK2``
(src/lpm_incompressible2d_rk2.hpp, src/lpm_incompressible2d_rk2_impl.hpp:
75-172).  Passive particles are the mesh vertices, active ones the faces.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _abi
from .device import Context, crd_args, warn_if_mixed_off_sphere
from .mesh import SphereMesh


class Incompressible2D:
    """2-D incompressible flow on the unit sphere with CoriolisSphere(Omega)
    and velocity kernel smoothing eps."""

    def __init__(self, seed, depth, Omega=2 * np.pi, eps=0.0,
                 ctx: Context = None, mesh: SphereMesh = None):
        self.ctx = ctx if ctx is not None else Context()
        dev = torch.device("cuda", self.ctx.device)
        self.mesh = mesh if mesh is not None else SphereMesh.build(seed, depth)
        m = self.mesh
        f64 = dict(dtype=torch.float64, device=dev)
        self.vert_x = torch.tensor(m.vert_x, **f64)
        self.face_x = torch.tensor(m.face_x, **f64)
        self.face_area = torch.tensor(m.face_area, **f64)
        self.face_mask = torch.tensor(m.face_mask, dtype=torch.uint8, device=dev)
        for kind, n in (("passive", m.n_verts), ("active", m.n_faces)):
            setattr(self, f"rel_vort_{kind}", torch.zeros(n, **f64))
            setattr(self, f"stream_fn_{kind}", torch.zeros(n, **f64))
            setattr(self, f"velocity_{kind}", torch.zeros((3, n), **f64))
        self.Omega = float(Omega)
        self.eps = float(eps)
        self.t = 0.0

    def init_vorticity(self, fn):
        """Relative vorticity zeta = fn(x) (Incompressible2D::init_vorticity,
        src/lpm_incompressible2d_impl.hpp:139-173)."""
        self.rel_vort_passive.copy_(torch.from_numpy(
            np.asarray(fn(self.mesh.vert_x), dtype=np.float64)))
        self.rel_vort_active.copy_(torch.from_numpy(
            np.asarray(fn(self.mesh.face_x), dtype=np.float64)))

    def fields(self) -> _abi.Ic2dFields:
        f = _abi.Ic2dFields()
        vp, ves, vcs, nv = crd_args(self.vert_x)
        up, ues, ucs, _ = crd_args(self.velocity_passive)
        fp, fes, fcs, nf = crd_args(self.face_x)
        gp, ges, gcs, _ = crd_args(self.velocity_active)
        f.n_verts, f.vert_x, f.vert_x_es, f.vert_x_cs = nv, vp, ves, vcs
        f.vert_zeta = self.rel_vort_passive.data_ptr()
        f.vert_u, f.vert_u_es, f.vert_u_cs = up, ues, ucs
        f.vert_psi = self.stream_fn_passive.data_ptr()
        f.n_faces, f.face_x, f.face_x_es, f.face_x_cs = nf, fp, fes, fcs
        f.face_zeta = self.rel_vort_active.data_ptr()
        f.face_u, f.face_u_es, f.face_u_cs = gp, ges, gcs
        f.face_psi = self.stream_fn_active.data_ptr()
        f.face_area = self.face_area.data_ptr()
        f.face_mask = self.face_mask.data_ptr()
        return f

    def init_direct_sums(self, flags=0, precision="fp64"):
        """Incompressible2D::init_direct_sums
        (src/lpm_incompressible2d_impl.hpp:235-254); precision "fp64" or
        "mixed"."""
        self.ctx.ic2d_init_direct_sums(self.fields(), self.eps, flags,
                                       precision=precision)

    def advance_timestep(self, solver: "Incompressible2DRK2"):
        solver.advance_timestep_impl(self)
        self.t = solver.t_idx * solver.dt

    # Leaf-face totals (src/lpm_incompressible2d_impl.hpp:91-137), reduced
    # on the device in FP64.
    def _leaf_sum(self, v):
        leaf = (self.face_mask == 0).to(torch.float64)
        return float((v * self.face_area * leaf).sum())

    def total_vorticity(self) -> float:
        """sum zeta A over the leaf faces."""
        return self._leaf_sum(self.rel_vort_active)

    def total_enstrophy(self) -> float:
        """(1/2) sum zeta^2 A over the leaf faces."""
        return 0.5 * self._leaf_sum(self.rel_vort_active ** 2)

    def total_kinetic_energy(self) -> float:
        """(1/2) sum |u|^2 A over the leaf faces."""
        return 0.5 * self._leaf_sum((self.velocity_active ** 2).sum(dim=0))


class Incompressible2DRK2:
    """Heun's method with the direct sums at the predictor and the new
    state (2 evaluations per step)."""

    def __init__(self, dt: float, ic2d: Incompressible2D, leaf_only=False,
                 renormalize=False, use_graph=False, stream=None,
                 precision="fp64"):
        self.dt = float(dt)
        self.t_idx = 0
        self.precision = precision
        warn_if_mixed_off_sphere(precision, renormalize)
        flags = 0
        if leaf_only:
            flags |= _abi.LPM_FLAG_SKIP_MASKED_TARGETS
        if renormalize:
            flags |= _abi.LPM_RK4_RENORMALIZE
        if use_graph:
            flags |= _abi.LPM_RK4_USE_GRAPH
        self.flags = flags
        self.stream = stream
        self._fields = None

    def advance_timestep_impl(self, ic2d: Incompressible2D):
        if self._fields is None:
            self._fields = ic2d.fields()
        ic2d.ctx.ic2d_rk2_step(self._fields, self.dt, ic2d.Omega, ic2d.eps,
                               self.flags, self.stream,
                               precision=self.precision)
        self.t_idx += 1
