"""Device-side entry points over the C ABI.

Device memory and streams come from PyTorch (ROCm build); all arithmetic of
the direct sums runs in the library's HIP kernels.  Coordinate tensors have
shape (3, n); their strides are passed through, so SoA (3, n) tensors and
transposed (n, 3) AoS tensors both work without copies.
"""
from __future__ import annotations

import ctypes

import torch

from . import _abi

PRECISIONS = {"fp64": _abi.LPM_FP64, "mixed": _abi.LPM_MIXED,
              _abi.LPM_FP64: _abi.LPM_FP64, _abi.LPM_MIXED: _abi.LPM_MIXED}


def precision_id(p) -> int:
    return PRECISIONS[p]


def crd_args(t: torch.Tensor):
    """(pointer, element stride, component stride, n) of a (3, n) tensor."""
    if t.dim() != 2 or t.shape[0] != 3:
        raise ValueError(f"expected a (3, n) coordinate tensor, got {tuple(t.shape)}")
    if t.dtype != torch.float64:
        raise TypeError("coordinates must be float64")
    if not t.is_cuda:
        raise ValueError("coordinates must be device tensors")
    return t.data_ptr(), t.stride(1), t.stride(0), t.shape[1]


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("arrays must be device tensors")
    if not t.is_contiguous():
        raise ValueError("scalar arrays must be contiguous")
    return t.data_ptr()


def current_stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Context:
    """An lpm_ctx bound to one HIP device (one per process and device)."""

    def __init__(self, device: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("lpm_amd.Context needs a HIP device "
                               "(there is no CPU fallback)")
        torch.cuda.set_device(device)
        self.L = _abi.lib()
        h = ctypes.c_void_p()
        _abi.check(self.L.lpm_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.L.lpm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        return _abi.check(rc, self.h)

    # ----------------------------------------------------------- BVE sums
    def bve_velocity(self, tgt_x, src_x, src_zeta, src_area, src_mask=None,
                     collocated=False, eps=0.0, precision="fp64", flags=0,
                     out=None, stream=None):
        tp, tes, tcs, nt = crd_args(tgt_x)
        sp, ses, scs, ns = crd_args(src_x)
        if out is None:
            out = torch.zeros((3, nt), dtype=torch.float64, device=tgt_x.device)
        op, oes, ocs, no = crd_args(out)
        assert no == nt
        self.check(self.L.lpm_bve_velocity(
            self.h, tp, tes, tcs, nt, sp, ses, scs, _ptr(src_zeta),
            _ptr(src_area), _ptr(src_mask), ns, int(collocated), float(eps),
            precision_id(precision), flags, op, oes, ocs,
            stream if stream is not None else current_stream()))
        return out

    def bve_stream_fn(self, tgt_x, src_x, src_zeta, src_area, src_mask=None,
                      collocated=False, precision="fp64", flags=0, out=None,
                      stream=None):
        tp, tes, tcs, nt = crd_args(tgt_x)
        sp, ses, scs, ns = crd_args(src_x)
        if out is None:
            out = torch.zeros(nt, dtype=torch.float64, device=tgt_x.device)
        self.check(self.L.lpm_bve_stream_fn(
            self.h, tp, tes, tcs, nt, sp, ses, scs, _ptr(src_zeta),
            _ptr(src_area), _ptr(src_mask), ns, int(collocated),
            precision_id(precision), flags, _ptr(out),
            stream if stream is not None else current_stream()))
        return out

    def bve_solve(self, tgt_x, src_x, src_zeta, src_area, src_mask=None,
                  collocated=False, precision="fp64", flags=0, stream=None):
        """Velocity (3, n) and stream function (n) in one pass
        (BVEVertexSolve / BVEFaceSolve)."""
        tp, tes, tcs, nt = crd_args(tgt_x)
        sp, ses, scs, ns = crd_args(src_x)
        u = torch.zeros((3, nt), dtype=torch.float64, device=tgt_x.device)
        psi = torch.zeros(nt, dtype=torch.float64, device=tgt_x.device)
        self.check(self.L.lpm_bve_solve(
            self.h, tp, tes, tcs, nt, sp, ses, scs, _ptr(src_zeta),
            _ptr(src_area), _ptr(src_mask), ns, int(collocated),
            precision_id(precision), flags, u.data_ptr(), u.stride(1),
            u.stride(0), psi.data_ptr(),
            stream if stream is not None else current_stream()))
        return u, psi

    def bve_init_velocity(self, fields, precision="fp64", flags=0, stream=None):
        self.check(self.L.lpm_bve_init_velocity(
            self.h, ctypes.byref(fields), precision_id(precision), flags,
            stream if stream is not None else current_stream()))

    def bve_rk4_step(self, fields, dt, Omega, precision="fp64", flags=0,
                     stream=None):
        self.check(self.L.lpm_bve_rk4_step(
            self.h, ctypes.byref(fields), float(dt), float(Omega),
            precision_id(precision), flags,
            stream if stream is not None else current_stream()))

    # ---------------------------------------------------------- ensembles
    def bve_ensemble_init_velocity(self, ens, precision="fp64", flags=0,
                                   stream=None):
        self.check(self.L.lpm_bve_ensemble_init_velocity(
            self.h, ctypes.byref(ens), precision_id(precision), flags,
            stream if stream is not None else current_stream()))

    def bve_ensemble_rk4_step(self, ens, dt, Omega, precision="fp64", flags=0,
                              stream=None):
        self.check(self.L.lpm_bve_ensemble_rk4_step(
            self.h, ctypes.byref(ens), float(dt), float(Omega),
            precision_id(precision), flags,
            stream if stream is not None else current_stream()))

    def bve_ensemble_autotune(self, ens, precision="fp64", flags=0, repeats=2,
                              stream=None):
        """Returns (best variant, its ms, list of ms of every variant)."""
        kid = (_abi.LPM_KERNEL_BVE_ENSEMBLE_F64
               if precision_id(precision) == _abi.LPM_FP64
               else _abi.LPM_KERNEL_BVE_ENSEMBLE_MIXED)
        nv = self.L.lpm_variant_count(kid)
        best = ctypes.c_int()
        ms = ctypes.c_double()
        all_ms = (ctypes.c_double * nv)()
        self.check(self.L.lpm_bve_ensemble_autotune(
            self.h, ctypes.byref(ens), precision_id(precision), flags,
            repeats, ctypes.byref(best), ctypes.byref(ms), all_ms,
            stream if stream is not None else current_stream()))
        return best.value, ms.value, list(all_ms)

    # --------------------------------------------------- Incompressible2D
    def ic2d_sums(self, tgt_x, src_x, src_zeta, src_area, src_mask=None,
                  collocated=False, eps=0.0, flags=0, out_u=None, out_psi=None,
                  stream=None, precision="fp64"):
        """Fused velocity (3, n) and stream function (n) at the targets."""
        tp, tes, tcs, nt = crd_args(tgt_x)
        sp, ses, scs, ns = crd_args(src_x)
        if out_u is None:
            out_u = torch.zeros((3, nt), dtype=torch.float64, device=tgt_x.device)
        if out_psi is None:
            out_psi = torch.zeros(nt, dtype=torch.float64, device=tgt_x.device)
        op, oes, ocs, _ = crd_args(out_u)
        self.check(self.L.lpm_ic2d_sums(
            self.h, tp, tes, tcs, nt, sp, ses, scs, _ptr(src_zeta),
            _ptr(src_area), _ptr(src_mask), ns, int(collocated), float(eps),
            precision_id(precision), flags, op, oes, ocs, _ptr(out_psi),
            stream if stream is not None else current_stream()))
        return out_u, out_psi

    def ic2d_init_direct_sums(self, fields, eps, flags=0, stream=None,
                              precision="fp64"):
        self.check(self.L.lpm_ic2d_init_direct_sums(
            self.h, ctypes.byref(fields), float(eps), precision_id(precision),
            flags, stream if stream is not None else current_stream()))

    def ic2d_rk2_step(self, fields, dt, Omega, eps, flags=0, stream=None,
                      precision="fp64"):
        self.check(self.L.lpm_ic2d_rk2_step(
            self.h, ctypes.byref(fields), float(dt), float(Omega), float(eps),
            precision_id(precision), flags,
            stream if stream is not None else current_stream()))

    # ----------------------------------------------------------- SWE sums
    def swe_sums(self, tgt_x, tgt_surf, src_x, src_zeta, src_sigma, src_area,
                 src_surf, src_mask=None, collocated=False, eps=0.0,
                 pse_eps=0.0, precision="fp64", flags=0, stream=None):
        """Returns (u (3, n), ddot (n), laps (n))."""
        tp, tes, tcs, nt = crd_args(tgt_x)
        sp, ses, scs, ns = crd_args(src_x)
        dev = tgt_x.device
        u = torch.zeros((3, nt), dtype=torch.float64, device=dev)
        dd = torch.zeros(nt, dtype=torch.float64, device=dev)
        lap = torch.zeros(nt, dtype=torch.float64, device=dev)
        op, oes, ocs, _ = crd_args(u)
        self.check(self.L.lpm_swe_sphere_sums(
            self.h, tp, tes, tcs, nt, _ptr(tgt_surf), sp, ses, scs,
            _ptr(src_zeta), _ptr(src_sigma), _ptr(src_area), _ptr(src_mask),
            _ptr(src_surf), ns, int(collocated), float(eps), float(pse_eps),
            precision_id(precision), flags, op, oes, ocs, _ptr(dd), _ptr(lap),
            stream if stream is not None else current_stream()))
        return u, dd, lap

    def swe_init_sums(self, fields, eps, pse_eps, precision="fp64", flags=0,
                      stream=None):
        self.check(self.L.lpm_swe_init_sums(
            self.h, ctypes.byref(fields), float(eps), float(pse_eps),
            precision_id(precision), flags,
            stream if stream is not None else current_stream()))

    def swe_rk4_step(self, fields, dt, Omega, g, eps, pse_eps,
                     precision="fp64", flags=0, stream=None):
        """One sphere SWE RK4 step (LPM_RK4_USE_GRAPH needs a stream)."""
        self.check(self.L.lpm_swe_rk4_step(
            self.h, ctypes.byref(fields), float(dt), float(Omega), float(g),
            float(eps), float(pse_eps), precision_id(precision), flags,
            stream if stream is not None else current_stream()))

    def swe_rk2_step(self, fields, dt, Omega, g, eps, pse_eps,
                     precision="fp64", flags=0, stream=None):
        """One sphere SWE RK2 (Heun) step, PSE surface Laplacian."""
        self.check(self.L.lpm_swe_rk2_step(
            self.h, ctypes.byref(fields), float(dt), float(Omega), float(g),
            float(eps), float(pse_eps), precision_id(precision), flags,
            stream if stream is not None else current_stream()))

    def swe_ensemble_init_sums(self, ens, eps, pse_eps, precision="fp64",
                               flags=0, stream=None):
        self.check(self.L.lpm_swe_ensemble_init_sums(
            self.h, ctypes.byref(ens), float(eps), float(pse_eps),
            precision_id(precision), flags,
            stream if stream is not None else current_stream()))

    def swe_ensemble_rk4_step(self, ens, dt, Omega, g, eps, pse_eps,
                              precision="fp64", flags=0, stream=None):
        self.check(self.L.lpm_swe_ensemble_rk4_step(
            self.h, ctypes.byref(ens), float(dt), float(Omega), float(g),
            float(eps), float(pse_eps), precision_id(precision), flags,
            stream if stream is not None else current_stream()))

    def swe_ensemble_rk2_step(self, ens, dt, Omega, g, eps, pse_eps,
                              precision="fp64", flags=0, stream=None):
        self.check(self.L.lpm_swe_ensemble_rk2_step(
            self.h, ctypes.byref(ens), float(dt), float(Omega), float(g),
            float(eps), float(pse_eps), precision_id(precision), flags,
            stream if stream is not None else current_stream()))

    # ----------------------------------------------------------- variants
    def variants(self, kernel_id):
        L = self.L
        return [L.lpm_variant_name(kernel_id, v).decode()
                for v in range(L.lpm_variant_count(kernel_id))]

    def set_variant(self, kernel_id, variant):
        self.check(self.L.lpm_set_variant(self.h, kernel_id, variant))

    def bve_autotune(self, fields, precision="fp64", flags=0, repeats=2,
                     stream=None):
        """Returns (best variant, its ms, list of ms of every variant)."""
        kid = (_abi.LPM_KERNEL_BVE_VEL_F64 if precision_id(precision) == _abi.LPM_FP64
               else _abi.LPM_KERNEL_BVE_VEL_MIXED)
        nv = self.L.lpm_variant_count(kid)
        best = ctypes.c_int()
        ms = ctypes.c_double()
        all_ms = (ctypes.c_double * nv)()
        self.check(self.L.lpm_bve_autotune(
            self.h, ctypes.byref(fields), precision_id(precision), flags,
            repeats, ctypes.byref(best), ctypes.byref(ms), all_ms,
            stream if stream is not None else current_stream()))
        return best.value, ms.value, list(all_ms)

    def tune_save(self, path):
        self.check(self.L.lpm_tune_save(self.h, path.encode()))

    def tune_load(self, path):
        self.check(self.L.lpm_tune_load(self.h, path.encode()))

    # ------------------------------------------------------------- timing
    def timing_enable(self, enable=True):
        self.check(self.L.lpm_timing_enable(self.h, int(enable)))

    # -------------------------------------------------------- diagnostics
    def err_norms(self, appx, exact, weight=None, mask=None, err=None,
                  stream=None):
        """ErrNorms of appx against exact: (l1, l2, linf).  appx / exact /
        err are (ncomp, n) device tensors (or (n,) for scalars); mask marks
        elements whose error is taken as zero; err receives appx - exact."""
        def fld(t):
            if t.dim() == 1:
                return t.data_ptr(), t.stride(0), 1, 1, t.shape[0]
            return t.data_ptr(), t.stride(1), t.stride(0), t.shape[0], t.shape[1]
        for t in (appx, exact) + ((err,) if err is not None else ()):
            if t.dtype != torch.float64 or not t.is_cuda:
                raise TypeError("fields must be float64 device tensors")
        ap, aes, acs, nc, n = fld(appx)
        xp, xes, xcs, nc2, n2 = fld(exact)
        if (nc, n) != (nc2, n2):
            raise ValueError("appx and exact shapes differ")
        ep, ees, ecs = (None, 1, 1)
        if err is not None:
            ep, ees, ecs, nc3, n3 = fld(err)
            if (nc3, n3) != (nc, n):
                raise ValueError("err shape differs")
        out = (ctypes.c_double * 3)()
        self.check(self.L.lpm_err_norms(
            self.h, ap, aes, acs, xp, xes, xcs, _ptr(weight), _ptr(mask), n, nc,
            ep, ees, ecs, out, stream if stream is not None else current_stream()))
        return out[0], out[1], out[2]

    def sphere_tangency(self, u, x, out=None, stream=None):
        """u . x per particle of (3, n) device tensors (SphereTangentFunctor)."""
        up, ues, ucs, n = crd_args(u)
        xp, xes, xcs, n2 = crd_args(x)
        if n != n2:
            raise ValueError("u and x sizes differ")
        if out is None:
            out = torch.zeros(n, dtype=torch.float64, device=u.device)
        self.check(self.L.lpm_sphere_tangency(
            self.h, up, ues, ucs, xp, xes, xcs, n, _ptr(out),
            stream if stream is not None else current_stream()))
        return out

    def timing_query(self, kernel_id):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        pairs = ctypes.c_double()
        self.check(self.L.lpm_timing_query(self.h, kernel_id, ctypes.byref(ms),
                                           ctypes.byref(n), ctypes.byref(pairs)))
        return ms.value, n.value, pairs.value


def warn_if_mixed_off_sphere(precision, renormalize):
    """Mixed-precision sums evaluate 1 - x.y as |x - y|^2 / 2, exact only for
    positions on the unit sphere; a stepper that does not project its stage
    positions back drifts off it by O((dt |u|)^2), and the mixed sums then
    differ from FP64 by about that much."""
    if precision == "mixed" and not renormalize:
        import warnings
        warnings.warn("mixed precision assumes positions on the unit sphere; "
                      "pass renormalize=True to keep them there", stacklevel=3)
