"""The allocation-free velocity entry point (lpm_bve_velocity_noalloc),
called through the C ABI, against lpm_bve_velocity on the same inputs.

Both run the same kernel variant on the same packed sources, so the outputs
must be bit-equal (NaN where the other has NaN: a divided icosahedral face
whose corners are equidistant shares its centre with its central child, and
its collocated velocity is 0/0 in both).  Each case makes two calls in a
row on one stream with no host synchronisation between them, the second on
a smaller input, so the second call runs over what the first left in the
shared scratch array.
"""
import ctypes

import numpy as np
import pytest

from conftest import mesh

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from lpm_amd import _abi  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    from lpm_amd.device import Context
    c = Context(0)
    yield c
    c.close()


def _same_bits(p, q):
    return bool(((p == q) | (p.isnan() & q.isnan())).all())


def _problem(seed, depth, strided=False):
    from lpm_amd import galleries as G
    m = mesh(seed, depth)
    f64 = dict(dtype=torch.float64, device="cuda")
    z = G.rh54_vorticity(m.face_x, 0.1, -0.1)
    if strided:
        # AoS (n, 3) storage: element stride 3, component stride 1
        vx = torch.tensor(np.ascontiguousarray(m.vert_x.T), **f64).t()
        fx = torch.tensor(np.ascontiguousarray(m.face_x.T), **f64).t()
    else:
        vx = torch.tensor(m.vert_x, **f64)
        fx = torch.tensor(m.face_x, **f64)
    return dict(vx=vx, fx=fx, zeta=torch.tensor(z, **f64),
                area=torch.tensor(m.face_area, **f64),
                mask=torch.tensor(m.face_mask, dtype=torch.uint8, device="cuda"))


def _call(fn, ctx, p, collocated, eps, precision, flags, use_mask, stream):
    tgt = p["fx"] if collocated else p["vx"]
    n_tgt = tgt.shape[1]
    out = torch.full((n_tgt, 3), 7.0, dtype=torch.float64, device="cuda").t()
    src = p["fx"]
    rc = fn(ctx.h, ctypes.c_void_p(tgt.data_ptr()), tgt.stride(1),
            tgt.stride(0), n_tgt, ctypes.c_void_p(src.data_ptr()),
            src.stride(1), src.stride(0), ctypes.c_void_p(p["zeta"].data_ptr()),
            ctypes.c_void_p(p["area"].data_ptr()),
            ctypes.c_void_p(p["mask"].data_ptr()) if use_mask else None,
            src.shape[1], int(collocated), float(eps),
            _abi.LPM_MIXED if precision == "mixed" else _abi.LPM_FP64, flags,
            ctypes.c_void_p(out.data_ptr()), out.stride(1), out.stride(0),
            ctypes.c_void_p(stream.cuda_stream))
    ctx.check(rc)
    return out


CASES = [
    # precision, collocated, eps, use_mask, flags
    ("fp64", False, 0.0, True, 0),
    ("fp64", True, 0.0, True, 0),
    ("fp64", True, 0.0, True, _abi.LPM_FLAG_SKIP_MASKED_TARGETS),
    ("fp64", False, 0.05, False, 0),
    ("mixed", False, 0.0, True, 0),
    ("mixed", True, 0.0, True, _abi.LPM_FLAG_SKIP_MASKED_TARGETS),
    ("mixed", True, 0.02, False, 0),
]


@pytest.mark.parametrize("precision,collocated,eps,use_mask,flags", CASES)
def test_noalloc_matches_velocity_twice_shrinking(ctx, precision, collocated,
                                                  eps, use_mask, flags):
    L = _abi.lib()
    big = _problem(0, 5)
    small = _problem(1, 3, strided=True)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ref_big = _call(L.lpm_bve_velocity, ctx, big, collocated, eps,
                        precision, flags, use_mask, stream)
        ref_small = _call(L.lpm_bve_velocity, ctx, small, collocated, eps,
                          precision, flags, use_mask, stream)
        # two allocation-free calls in a row, the second smaller
        got_big = _call(L.lpm_bve_velocity_noalloc, ctx, big, collocated, eps,
                        precision, flags, use_mask, stream)
        got_small = _call(L.lpm_bve_velocity_noalloc, ctx, small, collocated,
                          eps, precision, flags, use_mask, stream)
    stream.synchronize()
    assert _same_bits(got_big, ref_big)
    assert _same_bits(got_small, ref_small)
    # the outputs were written (the 7.0 fill is gone where targets are live)
    if not (flags & _abi.LPM_FLAG_SKIP_MASKED_TARGETS):
        assert not bool((got_small == 7.0).all(dim=0).any())


def test_noalloc_rejects_sizes_beyond_the_maximum(ctx):
    L = _abi.lib()
    nmax = L.lpm_bve_velocity_noalloc_max_sources()
    assert nmax == 1 << 19
    p = _problem(1, 2)
    x = p["fx"]
    u = torch.zeros_like(x)
    # rejected by the argument check before any device work
    rc = L.lpm_bve_velocity_noalloc(
        ctx.h, ctypes.c_void_p(x.data_ptr()), 1, x.stride(0), 4,
        ctypes.c_void_p(x.data_ptr()), 1, x.stride(0),
        ctypes.c_void_p(p["zeta"].data_ptr()),
        ctypes.c_void_p(p["area"].data_ptr()), None, nmax + 1, 0, 0.0,
        _abi.LPM_FP64, 0, ctypes.c_void_p(u.data_ptr()), 1, u.stride(0), None)
    assert rc == _abi.LPM_ERR_ARG
    assert "allocation-free maximum" in L.lpm_last_error(ctx.h).decode()


def test_noalloc_at_the_largest_workload_size(ctx):
    """Icosahedral level 7 (436,900 stored faces, under the maximum):
    vertex targets of the FP64 sum against lpm_bve_velocity, then a smaller
    call."""
    L = _abi.lib()
    big = _problem(0, 7)
    small = _problem(0, 4)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ref = _call(L.lpm_bve_velocity, ctx, big, False, 0.0, "fp64", 0, True,
                    stream)
        got = _call(L.lpm_bve_velocity_noalloc, ctx, big, False, 0.0, "fp64",
                    0, True, stream)
        ref_s = _call(L.lpm_bve_velocity, ctx, small, True, 0.0, "mixed",
                      _abi.LPM_FLAG_SKIP_MASKED_TARGETS, True, stream)
        got_s = _call(L.lpm_bve_velocity_noalloc, ctx, small, True, 0.0,
                      "mixed", _abi.LPM_FLAG_SKIP_MASKED_TARGETS, True, stream)
    stream.synchronize()
    assert _same_bits(got, ref)
    assert _same_bits(got_s, ref_s)
