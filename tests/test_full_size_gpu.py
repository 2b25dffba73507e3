"""The direct sums at the benchmark sizes, checked through properties that do
not depend on the size: mixed against FP64, FP64 against the oracle on a
sample of targets, tangency of the velocity, the analytic Rossby-Haurwitz
5-4 velocity and, for the shallow-water sums, the oracle on a sample.

Sizes: icosahedral level 7 (163,842 vertices + 327,680 faces as targets,
327,680 sources: 1.6e11 pairs), cubed sphere level 8 (393,218 + 393,216
targets, 393,216 sources: 3.1e11 pairs), shallow water on icosahedral level
6 (40,962 + 81,920 targets, 81,920 sources).

Tolerances as in the smaller tests: FP64 against the oracle 1e-12 relative
L2 and L-inf on the sample, mixed against FP64 1e-6 relative L2; tangency
|u.x| <= 1e-12 max|u| (the velocity is x cross S, orthogonal to x up to
rounding); the analytic RH velocity at level 7 is matched to the
discretisation error, 2e-4 relative L2 at the vertices.
"""
import numpy as np
import pytest

from conftest import aos, mesh, rel_l2, rel_linf

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from lpm_amd import _abi  # noqa: E402
from lpm_amd import galleries as G  # noqa: E402

TOL64 = 1e-12
TOL_MIXED = 1e-6


@pytest.fixture(scope="module")
def ctx():
    from lpm_amd.device import Context
    c = Context(0)
    yield c
    c.close()


def dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _velocities(ctx, m, z):
    fx, vx = dev(m.face_x), dev(m.vert_x)
    args = (dev(z), dev(m.face_area), dev(m.face_mask, torch.uint8))
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    out = {}
    for prec in ("fp64", "mixed"):
        uv = ctx.bve_velocity(vx, fx, *args, precision=prec).cpu().numpy()
        uf = ctx.bve_velocity(fx, fx, *args, collocated=True, flags=skip,
                              precision=prec).cpu().numpy()
        out[prec] = (uv, uf)
    return out


def _tangent(u, x):
    return np.abs((u * x).sum(axis=0)).max() / np.abs(u).max()


@pytest.mark.parametrize("seed,depth,stride", [("icos", 7, 1024),
                                               ("cube", 8, 2048)])
def test_bve_velocity_at_benchmark_size(ctx, ora, seed, depth, stride):
    m = mesh(seed, depth)
    if seed == "icos":
        z = G.rh54_vorticity(m.face_x, 0.1, -0.1)
    else:
        z = G.gaussian_vortex_vorticity(m.face_x, np.array([0.0, 0.0, 1.0]),
                                        4 * np.pi, 4.0)
    v = _velocities(ctx, m, z)
    leaf = m.face_mask == 0
    uv64, uf64 = v["fp64"]
    uv32, uf32 = v["mixed"]
    # mixed against FP64
    assert rel_l2(uv32, uv64) < TOL_MIXED
    assert rel_l2(uf32[:, leaf], uf64[:, leaf]) < TOL_MIXED
    # FP64 against the oracle on a sample of vertex targets ...
    sv = slice(0, m.n_verts, stride)
    ref = ora.bve_velocity(aos(m.vert_x[:, sv]), aos(m.face_x), z, m.face_area,
                           m.face_mask, nthreads=16)
    assert rel_l2(uv64[:, sv].T, ref) < TOL64
    assert rel_linf(uv64[:, sv].T, ref) < TOL64
    # ... and of leaf-face targets (collocated: each skips its own source;
    # the oracle evaluates only the unmasked targets)
    fl = np.flatnonzero(leaf)[::stride]
    skip = np.ones(m.n_faces, np.uint8)
    skip[fl] = 0
    ref = ora.bve_velocity(aos(m.face_x), aos(m.face_x), z, m.face_area,
                           m.face_mask, collocated=True, tgt_mask=skip,
                           nthreads=16)
    assert rel_l2(uf64[:, fl].T, ref[fl]) < TOL64
    assert rel_linf(uf64[:, fl].T, ref[fl]) < TOL64
    # tangency
    assert _tangent(uv64, m.vert_x) < TOL64
    assert _tangent(uf64[:, leaf], m.face_x[:, leaf]) < TOL64
    if seed == "icos":
        # discretisation error against the analytic RH 5-4 velocity
        ue = G.rh54_velocity(m.vert_x, 0.1, -0.1)
        assert rel_l2(uv64, ue) < 2e-4
        assert rel_l2(uv32, ue) < 2e-4


def test_swe_sums_at_benchmark_size(ctx, ora):
    m = mesh("icos", 6)
    rng = np.random.default_rng(7)
    st_f = G.tc2_state(m.face_x)
    st_v = G.tc2_state(m.vert_x)
    zf = st_f["zeta"] + 0.1 * rng.normal(size=m.n_faces)
    sf = 0.05 * rng.normal(size=m.n_faces)
    hf = st_f["surface"] + 1e-3 * rng.normal(size=m.n_faces)
    hv = st_v["surface"]
    h = m.mean_spacing()
    eps, pse = h, 2 * h
    mask = dev(m.face_mask, torch.uint8)
    args = (dev(m.face_x), dev(zf), dev(sf), dev(m.face_area), dev(hf))
    out = {}
    for prec in ("fp64", "mixed"):
        out[prec] = [t.cpu().numpy() for t in ctx.swe_sums(
            dev(m.vert_x), dev(hv), *args, src_mask=mask, eps=eps,
            pse_eps=pse, precision=prec)]
    (u64, d64, l64), (u32, d32, l32) = out["fp64"], out["mixed"]
    assert rel_l2(u32, u64) < TOL_MIXED
    assert rel_l2(d32, d64) < TOL_MIXED
    assert rel_l2(l32, l64) < TOL_MIXED
    sv = slice(0, m.n_verts, 512)
    ru, rdd, rlap = ora.swe_sums(aos(m.vert_x[:, sv]), hv[sv], aos(m.face_x),
                                 zf, sf, m.face_area, hf, m.face_mask, eps=eps,
                                 pse_eps=pse, nthreads=16)
    assert rel_l2(u64[:, sv].T, ru) < TOL64
    assert rel_l2(d64[sv], rdd) < TOL64
    assert rel_l2(l64[sv], rlap) < TOL64
    assert _tangent(u64, m.vert_x) < TOL64


@pytest.mark.parametrize("seed,depth,stride", [(0, 7, 1024), (1, 8, 2048)])
def test_autotuned_variants_at_benchmark_size(ctx, ora, seed, depth, stride):
    """The variants the autotuner picks at these sizes, through
    init_velocity (where the source-split forms run): the FP64 sum in 32
    source chunks per target block, and the matrix-core mixed sum in 8
    chunks with one reciprocal per four distances.  FP64 against the oracle
    on a sample of vertices and against the fixed unsplit FP64 variant,
    mixed against FP64."""
    from lpm_amd.bve import BVESphere
    k64 = _abi.LPM_KERNEL_BVE_VEL_F64
    kmx = _abi.LPM_KERNEL_BVE_VEL_MIXED
    n64, nmx = ctx.variants(k64), ctx.variants(kmx)
    if seed == 0:
        zfn = lambda x: G.rh54_vorticity(x, 0.1, -0.1)  # noqa: E731
    else:
        zfn = lambda x: G.gaussian_vortex_vorticity(  # noqa: E731
            x, np.array([0.0, 0.0, 1.0]), 4 * np.pi, 4.0)
    s = BVESphere(seed, depth, ctx=ctx)
    s.init_vorticity(zfn)
    leaf = s.mesh.face_mask == 0
    out = {}
    try:
        for name, kid, v, prec in (
                ("fixed", k64, ctx.L.lpm_variant_default(k64), "fp64"),
                ("split", k64, n64.index(
                    "block256_tpl4_nr1_quad_fused_sy_u2_tile512_split32"), "fp64"),
                ("mixed", kmx, nmx.index(
                    "mfma16_occ5_acc2_qrcp_block256_near3_split8"), "mixed")):
            ctx.set_variant(kid, v)
            s.init_velocity(prec)
            torch.cuda.synchronize()
            out[name] = (s.velocity_verts.cpu().numpy().copy(),
                         s.velocity_faces.cpu().numpy()[:, leaf].copy())
    finally:
        ctx.set_variant(k64, -1)
        ctx.set_variant(kmx, -1)
    for a, b in zip(out["split"], out["fixed"]):
        assert rel_l2(a, b) < 1e-13
    for a, b in zip(out["mixed"], out["split"]):
        assert rel_l2(a, b) < TOL_MIXED
    m = s.mesh
    zf = s.rel_vort_faces.cpu().numpy()
    sv = slice(0, m.n_verts, stride)
    ref = ora.bve_velocity(aos(m.vert_x[:, sv]), aos(m.face_x), zf,
                           m.face_area, m.face_mask, nthreads=16)
    assert rel_l2(out["split"][0][:, sv].T, ref) < TOL64
    assert rel_linf(out["split"][0][:, sv].T, ref) < TOL64


def test_stream_function_at_benchmark_size(ctx, ora):
    """The level-7 stream function, whose launches split every target
    block's sources into chunks (26 for the 641 vertex blocks, 10 for the
    1,707 face blocks): against the oracle on a sample of vertex and leaf
    face targets (1e-12) and against the analytic RH 5-4 stream function
    (discretisation error)."""
    m = mesh("icos", 7)
    z = G.rh54_vorticity(m.face_x, 0.1, -0.1)
    fx, vx = dev(m.face_x), dev(m.vert_x)
    args = (dev(z), dev(m.face_area), dev(m.face_mask, torch.uint8))
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    pv = ctx.bve_stream_fn(vx, fx, *args).cpu().numpy()
    pf = ctx.bve_stream_fn(fx, fx, *args, collocated=True,
                           flags=skip).cpu().numpy()
    sv = slice(0, m.n_verts, 1024)
    ref = ora.bve_stream_fn(aos(m.vert_x[:, sv]), aos(m.face_x), z,
                            m.face_area, m.face_mask, nthreads=16)
    assert rel_l2(pv[sv], ref) < TOL64
    assert rel_linf(pv[sv], ref) < TOL64
    leaf = m.face_mask == 0
    fl = np.flatnonzero(leaf)[::1024]
    tmask = np.ones(m.n_faces, np.uint8)
    tmask[fl] = 0
    ref = ora.bve_stream_fn(aos(m.face_x), aos(m.face_x), z, m.face_area,
                            m.face_mask, collocated=True, tgt_mask=tmask,
                            nthreads=16)
    assert rel_l2(pf[fl], ref[fl]) < TOL64
    assert rel_linf(pf[fl], ref[fl]) < TOL64
    # the analytic stream function of the RH 5-4 state (2.4e-5 measured)
    assert rel_l2(pv, G.rh54_stream_function(m.vert_x, 0.1, -0.1)) < 1e-4
