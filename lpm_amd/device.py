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

    # ------------------------------------------------------------- timing
    def timing_enable(self, enable=True):
        self.check(self.L.lpm_timing_enable(self.h, int(enable)))

    def timing_query(self, kernel_id):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        pairs = ctypes.c_double()
        self.check(self.L.lpm_timing_query(self.h, kernel_id, ctypes.byref(ms),
                                           ctypes.byref(n), ctypes.byref(pairs)))
        return ms.value, n.value, pairs.value
