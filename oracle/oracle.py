"""ORACLE — test infrastructure only.

ctypes wrapper around ``liblpm_oracle.so``, the plain-C FP64 restatement of
LPM's direct sums (see ``lpm_oracle.c``).  Only ``tests/``, the smoke check and
the benchmark's CPU-baseline leg import this module; the product never does.

Coordinates are numpy arrays of shape (n, 3) (AoS, the Kokkos-OpenMP layout).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblpm_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i64 = ctypes.c_int64


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, a second or two)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def load(path: str):
    """Use the oracle library at ``path`` (e.g. a build for the host's own
    instruction set) instead of the default one."""
    global _lib, _LIB_PATH
    _LIB_PATH = path
    _lib = None
    return lib()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.ora_max_threads.restype = ctypes.c_int
        L.ora_greens_fn.restype = ctypes.c_double
        L.ora_pse_laplacian_pair.restype = ctypes.c_double
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _u8(a):
    return a.ctypes.data_as(_u8p) if a is not None else None


def _c3(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert x.ndim == 2 and x.shape[1] == 3
    return x


def _c1(a, dtype=np.float64):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def max_threads() -> int:
    return lib().ora_max_threads()


# ------------------------------------------------------------------ pair fns
def biot_savart(x, y, zeta, area):
    u = np.zeros(3)
    lib().ora_biot_savart(_d(u), _d(_c1(x)), _d(_c1(y)), ctypes.c_double(zeta),
                          ctypes.c_double(area))
    return u


def greens_fn(x, y, zeta, area):
    return lib().ora_greens_fn(_d(_c1(x)), _d(_c1(y)), ctypes.c_double(zeta),
                               ctypes.c_double(area))


def kzeta_sphere(x, y, vort, area, eps=0.0):
    u = np.zeros(3)
    lib().ora_kzeta_sphere(_d(u), _d(_c1(x)), _d(_c1(y)), ctypes.c_double(vort),
                           ctypes.c_double(area), ctypes.c_double(eps))
    return u


def ksigma_sphere(x, y, div, area, eps=0.0):
    u = np.zeros(3)
    lib().ora_ksigma_sphere(_d(u), _d(_c1(x)), _d(_c1(y)), ctypes.c_double(div),
                            ctypes.c_double(area), ctypes.c_double(eps))
    return u


def grad_kzeta(x, y, eps=0.0):
    g = np.zeros(9)
    lib().ora_grad_kzeta(_d(g), _d(_c1(x)), _d(_c1(y)), ctypes.c_double(eps))
    return g


def grad_ksigma(x, y, eps=0.0):
    g = np.zeros(9)
    lib().ora_grad_ksigma(_d(g), _d(_c1(x)), _d(_c1(y)), ctypes.c_double(eps))
    return g


def sphere_swe_velocity_sums(x, y, zeta, sigma, area, eps=0.0):
    r = np.zeros(12)
    lib().ora_sphere_swe_velocity_sums(_d(r), _d(_c1(x)), _d(_c1(y)),
                                       ctypes.c_double(zeta),
                                       ctypes.c_double(sigma),
                                       ctypes.c_double(area),
                                       ctypes.c_double(eps))
    return r


def pse_laplacian_pair(x, y, fy, fx, area, pse_eps):
    return lib().ora_pse_laplacian_pair(_d(_c1(x)), _d(_c1(y)),
                                        ctypes.c_double(fy), ctypes.c_double(fx),
                                        ctypes.c_double(area),
                                        ctypes.c_double(pse_eps))


# ------------------------------------------------------------------ BVE sums
def bve_velocity(tx, sx, szeta, sarea, smask=None, collocated=False,
                 tgt_mask=None, nthreads=0, out=None):
    tx, sx = _c3(tx), _c3(sx)
    szeta, sarea = _c1(szeta), _c1(sarea)
    smask, tgt_mask = _c1(smask, np.uint8), _c1(tgt_mask, np.uint8)
    u = np.zeros_like(tx) if out is None else out
    lib().ora_bve_velocity(_d(tx), _i64(len(tx)), _d(sx), _d(szeta), _d(sarea),
                           _u8(smask), _i64(len(sx)), ctypes.c_int(int(collocated)),
                           _u8(tgt_mask), _d(u), ctypes.c_int(nthreads))
    return u


def bve_stream_fn(tx, sx, szeta, sarea, smask=None, collocated=False,
                  tgt_mask=None, nthreads=0):
    tx, sx = _c3(tx), _c3(sx)
    szeta, sarea = _c1(szeta), _c1(sarea)
    smask, tgt_mask = _c1(smask, np.uint8), _c1(tgt_mask, np.uint8)
    psi = np.zeros(len(tx))
    lib().ora_bve_stream_fn(_d(tx), _i64(len(tx)), _d(sx), _d(szeta), _d(sarea),
                            _u8(smask), _i64(len(sx)),
                            ctypes.c_int(int(collocated)), _u8(tgt_mask),
                            _d(psi), ctypes.c_int(nthreads))
    return psi


def ic2d_sums(tx, sx, szeta, sarea, smask=None, collocated=False, eps=0.0,
              nthreads=0):
    """Incompressible2D sphere sums: (velocity (n, 3), stream function (n))."""
    tx, sx = _c3(tx), _c3(sx)
    szeta, sarea = _c1(szeta), _c1(sarea)
    smask = _c1(smask, np.uint8)
    u = np.zeros((len(tx), 3))
    psi = np.zeros(len(tx))
    lib().ora_ic2d_sums(_d(tx), _i64(len(tx)), _d(sx), _d(szeta), _d(sarea),
                        _u8(smask), _i64(len(sx)), ctypes.c_int(int(collocated)),
                        ctypes.c_double(eps), _d(u), _d(psi),
                        ctypes.c_int(nthreads))
    return u, psi


def ic2d_rk2_step(state: dict, dt, Omega, eps, nthreads=0):
    """One Incompressible2DRK2 step of ``state`` (modified in place).

    Keys: vx (nv,3), vzeta, vu (nv,3), vpsi, fx (nf,3), fzeta, fu (nf,3),
    fpsi, farea, fmask.  vu/fu must hold the velocity at the current state.
    """
    s = state
    for k in ("vx", "vu", "fx", "fu"):
        assert s[k].flags["C_CONTIGUOUS"] and s[k].dtype == np.float64
    lib().ora_ic2d_rk2_step(
        _d(s["vx"]), _d(s["vzeta"]), _d(s["vu"]), _d(s["vpsi"]),
        _i64(len(s["vx"])), _d(s["fx"]), _d(s["fzeta"]), _d(s["fu"]),
        _d(s["fpsi"]), _d(s["farea"]), _u8(s.get("fmask")), _i64(len(s["fx"])),
        ctypes.c_double(dt), ctypes.c_double(Omega), ctypes.c_double(eps),
        ctypes.c_int(nthreads))
    return state


def bve_velocity_xp(tx, sx, szeta, sarea, smask=None, collocated=False,
                    nthreads=0):
    tx, sx = _c3(tx), _c3(sx)
    szeta, sarea, smask = _c1(szeta), _c1(sarea), _c1(smask, np.uint8)
    u = np.zeros_like(tx)
    lib().ora_bve_velocity_xp(_d(tx), _i64(len(tx)), _d(sx), _d(szeta),
                              _d(sarea), _u8(smask), _i64(len(sx)),
                              ctypes.c_int(int(collocated)), _d(u),
                              ctypes.c_int(nthreads))
    return u


def bve_rk4_step(state: dict, dt, Omega, compat=True, renorm=False,
                 nthreads=0):
    """Advance ``state`` (dict of numpy arrays, modified in place) one step.

    Keys: vx (nv,3), vzeta, vu (nv,3), fx (nf,3), fzeta, fu (nf,3), farea,
    fmask.  vu/fu must hold the velocity at the current positions.
    """
    s = state
    for k in ("vx", "vu", "fx", "fu"):
        assert s[k].flags["C_CONTIGUOUS"] and s[k].dtype == np.float64
    lib().ora_bve_rk4_step(
        _d(s["vx"]), _d(s["vzeta"]), _d(s["vu"]), _i64(len(s["vx"])),
        _d(s["fx"]), _d(s["fzeta"]), _d(s["fu"]), _d(s["farea"]),
        _u8(s.get("fmask")), _i64(len(s["fx"])), ctypes.c_double(dt),
        ctypes.c_double(Omega), ctypes.c_int(int(compat)),
        ctypes.c_int(int(renorm)), ctypes.c_int(nthreads))
    return state


# ------------------------------------------------------------------ SWE sums
def swe_sums(tx, ts, sx, szeta, ssigma, sarea, ss, smask=None,
             collocated=False, eps=0.0, pse_eps=0.0, nthreads=0):
    tx, sx = _c3(tx), _c3(sx)
    ts, ss = _c1(ts), _c1(ss)
    szeta, ssigma, sarea = _c1(szeta), _c1(ssigma), _c1(sarea)
    smask = _c1(smask, np.uint8)
    nt = len(tx)
    u = np.zeros((nt, 3))
    dd = np.zeros(nt)
    lap = np.zeros(nt)
    lib().ora_swe_sums(_d(tx), _d(ts), _i64(nt), _d(sx), _d(szeta), _d(ssigma),
                       _d(sarea), _u8(smask), _d(ss), _i64(len(sx)),
                       ctypes.c_int(int(collocated)), ctypes.c_double(eps),
                       ctypes.c_double(pse_eps), _d(u), _d(dd), _d(lap),
                       ctypes.c_int(nthreads))
    return u, dd, lap


def swe_rk4_step(s: dict, dt, Omega, g, eps, pse_eps, compat=True,
                 renorm=False, nthreads=0):
    """Advance the sphere SWE state ``s`` (dict of numpy arrays) one step."""
    lib().ora_swe_rk4_step(
        _d(s["vx"]), _d(s["vzeta"]), _d(s["vsigma"]), _d(s["vh"]),
        _d(s["vu"]), _d(s["vdd"]), _d(s["vlap"]), _d(s["vs"]),
        _i64(len(s["vx"])), _d(s["fx"]), _d(s["fzeta"]), _d(s["fsigma"]),
        _d(s["farea"]), _d(s["fmass"]), _d(s["fh"]), _d(s["fu"]), _d(s["fdd"]),
        _d(s["flap"]), _d(s["fs"]), _u8(s.get("fmask")), _i64(len(s["fx"])),
        ctypes.c_double(dt), ctypes.c_double(Omega), ctypes.c_double(g),
        ctypes.c_double(eps), ctypes.c_double(pse_eps),
        ctypes.c_int(int(compat)), ctypes.c_int(int(renorm)),
        ctypes.c_int(nthreads))
    return s


def swe_rk2_step(s: dict, dt, Omega, g, eps, pse_eps, renorm=False,
                 nthreads=0):
    """Advance the sphere SWE state ``s`` one RK2 (Heun) step."""
    lib().ora_swe_rk2_step(
        _d(s["vx"]), _d(s["vzeta"]), _d(s["vsigma"]), _d(s["vh"]),
        _d(s["vu"]), _d(s["vdd"]), _d(s["vlap"]), _d(s["vs"]),
        _i64(len(s["vx"])), _d(s["fx"]), _d(s["fzeta"]), _d(s["fsigma"]),
        _d(s["farea"]), _d(s["fmass"]), _d(s["fh"]), _d(s["fu"]), _d(s["fdd"]),
        _d(s["flap"]), _d(s["fs"]), _u8(s.get("fmask")), _i64(len(s["fx"])),
        ctypes.c_double(dt), ctypes.c_double(Omega), ctypes.c_double(g),
        ctypes.c_double(eps), ctypes.c_double(pse_eps),
        ctypes.c_int(int(renorm)), ctypes.c_int(nthreads))
    return s


# ---------------------------------------------------------------- galleries
def rh54_vorticity(x, u0, amp):
    x = _c3(x)
    z = np.zeros(len(x))
    lib().ora_rh54_vorticity(_d(x), _i64(len(x)), ctypes.c_double(u0),
                             ctypes.c_double(amp), _d(z))
    return z


def rh54_velocity(x, u0, amp):
    x = _c3(x)
    u = np.zeros_like(x)
    lib().ora_rh54_velocity(_d(x), _i64(len(x)), ctypes.c_double(u0),
                            ctypes.c_double(amp), _d(u))
    return u


def gaussian_vorticity(x, ctr, strength, b, gauss_const=0.0):
    x = _c3(x)
    z = np.zeros(len(x))
    lib().ora_gaussian_vorticity(_d(x), _i64(len(x)), _d(_c1(ctr)),
                                 ctypes.c_double(strength), ctypes.c_double(b),
                                 ctypes.c_double(gauss_const), _d(z))
    return z


def err_norms(appx, exact, wt=None):
    """Relative (l1, l2, linf) of ErrNorms; rows are vectors when 2-D."""
    appx = np.ascontiguousarray(appx, dtype=np.float64)
    exact = np.ascontiguousarray(exact, dtype=np.float64)
    dim = 1 if appx.ndim == 1 else appx.shape[1]
    out = np.zeros(3)
    wt = _c1(wt)
    lib().ora_err_norms(_d(appx), _d(exact), _d(wt), _i64(len(appx)),
                        ctypes.c_int(dim), _d(out))
    return tuple(out)
