"""ctypes bindings of liblpm_mi355x.so (declarations in include/lpm_mi355x.h).

The library is loaded from this package directory only; if it is missing the
import fails loudly — there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblpm_mi355x.so")

LPM_OK = 0
LPM_ERR_ARG = -1
LPM_ERR_HIP = -2
LPM_ERR_NONFINITE = -3
LPM_ERR_STATE = -4
LPM_ERR_NOMEM = -5
LPM_ERR_INTERNAL = -6

LPM_FP64 = 0
LPM_MIXED = 1

LPM_FLAG_CHECK_FINITE = 0x1
LPM_FLAG_SKIP_MASKED_TARGETS = 0x2
LPM_RK4_REFERENCE_COMPAT = 0x10
LPM_RK4_RENORMALIZE = 0x20
LPM_RK4_USE_GRAPH = 0x40
LPM_GRAPH_CACHE_MAX = 16

LPM_SEED_ICOS_TRI = 0
LPM_SEED_CUBED_SPHERE = 1

LPM_KERNEL_BVE_VEL_F64 = 0
LPM_KERNEL_BVE_VEL_MIXED = 1
LPM_KERNEL_BVE_PSI = 2
LPM_KERNEL_SWE_F64 = 3
LPM_KERNEL_SWE_MIXED = 4
LPM_KERNEL_BVE_ENSEMBLE_F64 = 5
LPM_KERNEL_BVE_ENSEMBLE_MIXED = 6
LPM_KERNEL_IC2D = 7

c_int = ctypes.c_int
c_uint = ctypes.c_uint
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_dblp = ctypes.POINTER(ctypes.c_double)
c_intp = ctypes.POINTER(ctypes.c_int)


class LpmError(RuntimeError):
    """Raised when an entry point returns a negative lpm_status."""

    def __init__(self, code, message):
        super().__init__(f"[lpm_status {code}] {message}")
        self.code = code


class BveFields(ctypes.Structure):
    """struct lpm_bve_fields."""
    _fields_ = [
        ("n_verts", c_i64), ("vert_x", c_vp), ("vert_x_es", c_i64),
        ("vert_x_cs", c_i64), ("vert_zeta", c_vp), ("vert_u", c_vp),
        ("vert_u_es", c_i64), ("vert_u_cs", c_i64),
        ("n_faces", c_i64), ("face_x", c_vp), ("face_x_es", c_i64),
        ("face_x_cs", c_i64), ("face_zeta", c_vp), ("face_u", c_vp),
        ("face_u_es", c_i64), ("face_u_cs", c_i64), ("face_area", c_vp),
        ("face_mask", c_vp),
    ]


class Ic2dFields(ctypes.Structure):
    """struct lpm_ic2d_fields."""
    _fields_ = [
        ("n_verts", c_i64), ("vert_x", c_vp), ("vert_x_es", c_i64),
        ("vert_x_cs", c_i64), ("vert_zeta", c_vp), ("vert_u", c_vp),
        ("vert_u_es", c_i64), ("vert_u_cs", c_i64), ("vert_psi", c_vp),
        ("n_faces", c_i64), ("face_x", c_vp), ("face_x_es", c_i64),
        ("face_x_cs", c_i64), ("face_zeta", c_vp), ("face_u", c_vp),
        ("face_u_es", c_i64), ("face_u_cs", c_i64), ("face_psi", c_vp),
        ("face_area", c_vp), ("face_mask", c_vp),
    ]


class BveEnsemble(ctypes.Structure):
    """struct lpm_bve_ensemble (offsets are host int64 arrays)."""
    _fields_ = [
        ("n_members", ctypes.c_int32), ("vert_offsets", c_i64p),
        ("face_offsets", c_i64p), ("fields", BveFields),
    ]


class SweFields(ctypes.Structure):
    """struct lpm_swe_fields."""
    _fields_ = [
        ("n_verts", c_i64), ("vert_x", c_vp), ("vert_x_es", c_i64),
        ("vert_x_cs", c_i64), ("vert_zeta", c_vp), ("vert_sigma", c_vp),
        ("vert_h", c_vp), ("vert_u", c_vp), ("vert_u_es", c_i64),
        ("vert_u_cs", c_i64), ("vert_ddot", c_vp), ("vert_laps", c_vp),
        ("vert_surf", c_vp),
        ("n_faces", c_i64), ("face_x", c_vp), ("face_x_es", c_i64),
        ("face_x_cs", c_i64), ("face_zeta", c_vp), ("face_sigma", c_vp),
        ("face_area", c_vp), ("face_mass", c_vp), ("face_h", c_vp),
        ("face_u", c_vp), ("face_u_es", c_i64), ("face_u_cs", c_i64),
        ("face_ddot", c_vp), ("face_laps", c_vp), ("face_surf", c_vp),
        ("face_mask", c_vp),
    ]


class SweEnsemble(ctypes.Structure):
    """struct lpm_swe_ensemble (offsets are host int64 arrays)."""
    _fields_ = [
        ("n_members", ctypes.c_int32), ("vert_offsets", c_i64p),
        ("face_offsets", c_i64p), ("fields", SweFields),
    ]


# name -> (restype, argtypes)
_SIGS = {
    "lpm_abi_version": (c_int, []),
    "lpm_ctx_create": (c_int, [c_int, ctypes.POINTER(c_vp)]),
    "lpm_ctx_destroy": (c_int, [c_vp]),
    "lpm_last_error": (ctypes.c_char_p, [c_vp]),
    "lpm_stream_synchronize": (c_int, [c_vp, c_vp]),
    "lpm_graph_cache_size": (c_int, [c_vp, c_intp]),
    "lpm_mesh_create": (c_int, [c_int, c_int, ctypes.POINTER(c_vp)]),
    "lpm_mesh_destroy": (c_int, [c_vp]),
    "lpm_mesh_sizes": (c_int, [c_vp, c_i64p, c_i64p, c_i64p, c_i64p]),
    "lpm_mesh_max_allocations": (c_int, [c_int, c_int, c_i64p, c_i64p, c_i64p]),
    "lpm_mesh_copy_vertices": (c_int, [c_vp, c_vp]),
    "lpm_mesh_copy_faces": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "lpm_mesh_copy_face_vertices": (c_int, [c_vp, c_vp, c_intp]),
    "lpm_mesh_copy_face_levels": (c_int, [c_vp, c_vp]),
    "lpm_mesh_divide_flagged_faces": (c_int, [c_vp, c_vp, c_i64, c_int, c_i64,
                                              ctypes.POINTER(c_i64), c_intp]),
    "lpm_bve_velocity": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64,
                                 c_i64, c_vp, c_vp, c_vp, c_i64, c_int, c_dbl,
                                 c_int, c_uint, c_vp, c_i64, c_i64, c_vp]),
    "lpm_bve_solve": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64,
                              c_i64, c_vp, c_vp, c_vp, c_i64, c_int, c_int,
                              c_uint, c_vp, c_i64, c_i64, c_vp, c_vp]),
    "lpm_bve_velocity_noalloc": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp,
                                         c_i64, c_i64, c_vp, c_vp, c_vp, c_i64,
                                         c_int, c_dbl, c_int, c_uint, c_vp,
                                         c_i64, c_i64, c_vp]),
    "lpm_bve_velocity_noalloc_max_sources": (c_i64, []),
    "lpm_bve_stream_fn": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64,
                                  c_i64, c_vp, c_vp, c_vp, c_i64, c_int, c_int,
                                  c_uint, c_vp, c_vp]),
    "lpm_bve_init_velocity": (c_int, [c_vp, ctypes.POINTER(BveFields), c_int,
                                      c_uint, c_vp]),
    "lpm_bve_rk4_step": (c_int, [c_vp, ctypes.POINTER(BveFields), c_dbl, c_dbl,
                                 c_int, c_uint, c_vp]),
    "lpm_bve_ensemble_init_velocity": (c_int, [c_vp, ctypes.POINTER(BveEnsemble),
                                               c_int, c_uint, c_vp]),
    "lpm_bve_ensemble_rk4_step": (c_int, [c_vp, ctypes.POINTER(BveEnsemble),
                                          c_dbl, c_dbl, c_int, c_uint, c_vp]),
    "lpm_bve_ensemble_autotune": (c_int, [c_vp, ctypes.POINTER(BveEnsemble),
                                          c_int, c_uint, c_int, c_intp, c_dblp,
                                          c_dblp, c_vp]),
    "lpm_swe_sphere_sums": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp,
                                    c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp,
                                    c_i64, c_int, c_dbl, c_dbl, c_int, c_uint,
                                    c_vp, c_i64, c_i64, c_vp, c_vp, c_vp]),
    "lpm_swe_init_sums": (c_int, [c_vp, ctypes.POINTER(SweFields), c_dbl, c_dbl,
                                  c_int, c_uint, c_vp]),
    "lpm_swe_rk4_step": (c_int, [c_vp, ctypes.POINTER(SweFields), c_dbl, c_dbl,
                                 c_dbl, c_dbl, c_dbl, c_int, c_uint, c_vp]),
    "lpm_swe_rk2_step": (c_int, [c_vp, ctypes.POINTER(SweFields), c_dbl, c_dbl,
                                 c_dbl, c_dbl, c_dbl, c_int, c_uint, c_vp]),
    "lpm_swe_ensemble_init_sums": (c_int, [c_vp, ctypes.POINTER(SweEnsemble),
                                           c_dbl, c_dbl, c_int, c_uint, c_vp]),
    "lpm_swe_ensemble_rk4_step": (c_int, [c_vp, ctypes.POINTER(SweEnsemble),
                                          c_dbl, c_dbl, c_dbl, c_dbl, c_dbl,
                                          c_int, c_uint, c_vp]),
    "lpm_swe_ensemble_rk2_step": (c_int, [c_vp, ctypes.POINTER(SweEnsemble),
                                          c_dbl, c_dbl, c_dbl, c_dbl, c_dbl,
                                          c_int, c_uint, c_vp]),
    "lpm_ic2d_sums": (c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64,
                              c_i64, c_vp, c_vp, c_vp, c_i64, c_int, c_dbl,
                              c_int, c_uint, c_vp, c_i64, c_i64, c_vp, c_vp]),
    "lpm_ic2d_init_direct_sums": (c_int, [c_vp, ctypes.POINTER(Ic2dFields),
                                          c_dbl, c_int, c_uint, c_vp]),
    "lpm_ic2d_rk2_step": (c_int, [c_vp, ctypes.POINTER(Ic2dFields), c_dbl,
                                  c_dbl, c_dbl, c_int, c_uint, c_vp]),
    "lpm_variant_count": (c_int, [c_int]),
    "lpm_variant_default": (c_int, [c_int]),
    "lpm_variant_name": (ctypes.c_char_p, [c_int, c_int]),
    "lpm_set_variant": (c_int, [c_vp, c_int, c_int]),
    "lpm_bve_autotune": (c_int, [c_vp, ctypes.POINTER(BveFields), c_int, c_uint,
                                 c_int, c_intp, c_dblp, c_dblp, c_vp]),
    "lpm_tune_save": (c_int, [c_vp, ctypes.c_char_p]),
    "lpm_tune_load": (c_int, [c_vp, ctypes.c_char_p]),
    "lpm_timing_enable": (c_int, [c_vp, c_int]),
    "lpm_timing_query": (c_int, [c_vp, c_int, c_dblp, c_i64p, c_dblp]),
    "lpm_err_norms": (c_int, [c_vp, c_vp, c_i64, c_i64, c_vp, c_i64, c_i64,
                              c_vp, c_vp, c_i64, c_int, c_vp, c_i64, c_i64,
                              c_dblp, c_vp]),
    "lpm_sphere_tangency": (c_int, [c_vp, c_vp, c_i64, c_i64, c_vp, c_i64,
                                    c_i64, c_i64, c_vp, c_vp]),
}

_lib = None


def lib():
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} not found: build it with `python -m lpm_amd.build` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def check(rc, ctx=None):
    """Raise LpmError for a negative status."""
    if rc != LPM_OK:
        msg = lib().lpm_last_error(ctx)
        raise LpmError(rc, msg.decode() if msg else "")
    return rc
