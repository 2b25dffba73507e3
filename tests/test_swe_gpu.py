"""GPU parity of the sphere SWE direct sums (velocity, grad(u):grad(u)^T,
PSE surface Laplacian) and the sphere SWE RK4 stepper against the FP64
oracle.

Tolerances: FP64 <= 1e-12 relative L2 (the PSE sum omits terms below
exp(-50) relative, far under that); mixed precision: velocity, double dot
and Laplacian <= 1e-6 relative L2 against the FP64 kernel (near tiles and
the PSE sum run in FP64 in the mixed kernel).
"""
import numpy as np
import pytest

from conftest import aos, mesh, rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from lpm_amd import _abi  # noqa: E402
from lpm_amd import galleries as G  # noqa: E402

OMEGA = 2 * np.pi
TOL64 = 1e-12


@pytest.fixture(scope="module")
def ctx():
    from lpm_amd.device import Context
    c = Context(0)
    yield c
    c.close()


def dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _fields(seed, depth, rng_seed=3):
    m = mesh(seed, depth)
    rng = np.random.default_rng(rng_seed)
    st_f = G.tc2_state(m.face_x)
    st_v = G.tc2_state(m.vert_x)
    zf = st_f["zeta"] + 0.1 * rng.normal(size=m.n_faces)
    sf = 0.05 * rng.normal(size=m.n_faces)  # nonzero divergence
    hf = st_f["surface"] + 1e-3 * rng.normal(size=m.n_faces)
    hv = st_v["surface"]
    return m, zf, sf, hf, hv


@pytest.mark.parametrize("seed,depth", [(0, 3), (1, 3)])
@pytest.mark.parametrize("eps_h", [0.0, 1.0])
def test_swe_sums_fp64_match_oracle(ctx, ora, seed, depth, eps_h):
    m, zf, sf, hf, hv = _fields(seed, depth)
    h = m.mean_spacing()
    eps, pse = eps_h * h, 2 * h
    mask = dev(m.face_mask, torch.uint8)
    args = (dev(zf), dev(sf), dev(m.face_area), dev(hf), mask)
    # vertex targets
    u, dd, lap = ctx.swe_sums(dev(m.vert_x), dev(hv), dev(m.face_x), *args[:4],
                              src_mask=mask, eps=eps, pse_eps=pse)
    ru, rdd, rlap = ora.swe_sums(aos(m.vert_x), hv, aos(m.face_x), zf, sf,
                                 m.face_area, hf, m.face_mask, eps=eps,
                                 pse_eps=pse)
    assert rel_l2(u.cpu().numpy().T, ru) < TOL64
    assert rel_l2(dd.cpu().numpy(), rdd) < TOL64
    assert rel_l2(lap.cpu().numpy(), rlap) < TOL64
    # face targets
    fx = dev(m.face_x)
    u, dd, lap = ctx.swe_sums(fx, dev(hf), fx, *args[:4], src_mask=mask,
                              collocated=True, eps=eps, pse_eps=pse)
    ru, rdd, rlap = ora.swe_sums(aos(m.face_x), hf, aos(m.face_x), zf, sf,
                                 m.face_area, hf, m.face_mask, collocated=True,
                                 eps=eps, pse_eps=pse)
    leaf = m.face_mask == 0
    assert rel_l2(u.cpu().numpy().T[leaf], ru[leaf]) < TOL64
    assert rel_l2(dd.cpu().numpy()[leaf], rdd[leaf]) < TOL64
    assert rel_l2(lap.cpu().numpy()[leaf], rlap[leaf]) < TOL64


def test_swe_sums_mixed_match_fp64(ctx):
    m, zf, sf, hf, hv = _fields(0, 5)
    h = m.mean_spacing()
    mask = dev(m.face_mask, torch.uint8)
    fx = dev(m.face_x)
    out = {}
    for prec in ("fp64", "mixed"):
        out[prec] = ctx.swe_sums(fx, dev(hf), fx, dev(zf), dev(sf),
                                 dev(m.face_area), dev(hf), src_mask=mask,
                                 collocated=True, eps=h, pse_eps=2 * h,
                                 precision=prec)
    leaf = torch.from_numpy(m.face_mask == 0)
    u64, d64, l64 = (t.cpu().numpy() for t in out["fp64"])
    u32, d32, l32 = (t.cpu().numpy() for t in out["mixed"])
    lf = leaf.numpy()
    assert rel_l2(u32[:, lf], u64[:, lf]) < 1e-6
    assert rel_l2(d32[lf], d64[lf]) < 1e-6
    assert rel_l2(l32[lf], l64[lf]) < 1e-6


def _variant_names(ctx, kid):
    return ctx.variants(kid)


@pytest.mark.parametrize("eps_h", [0.0, 1.0])
def test_mixed_variants_match_fp64_at_level6(ctx, eps_h):
    """Every mixed shape against FP64 on the icosahedral mesh of level 6
    (81,920 sources in 320 tiles, so the matrix-core split into far tiles,
    far 64-source sub-tiles and FP64 near sub-tiles all occur): vertex
    targets and collocated face targets (self pairs skipped at eps = 0),
    velocity, double dot and Laplacian <= 1e-6 relative L2."""
    m, zf, sf, hf, hv = _fields(0, 6)
    h = m.mean_spacing()
    eps, pse = eps_h * h, 2 * h
    mask = dev(m.face_mask, torch.uint8)
    fx, vx = dev(m.face_x), dev(m.vert_x)
    src = (dev(zf), dev(sf), dev(m.face_area), dev(hf))
    leaf = m.face_mask == 0

    def run(prec):
        v = ctx.swe_sums(vx, dev(hv), fx, *src, src_mask=mask, eps=eps,
                         pse_eps=pse, precision=prec)
        f = ctx.swe_sums(fx, dev(hf), fx, *src, src_mask=mask,
                         collocated=True, eps=eps, pse_eps=pse,
                         precision=prec)
        return ([t.cpu().numpy() for t in v], [t.cpu().numpy() for t in f])

    kmx = _abi.LPM_KERNEL_SWE_MIXED
    (v64, f64) = run("fp64")
    names = _variant_names(ctx, kmx)
    assert any(n.startswith("mfma") for n in names)
    try:
        for v, name in enumerate(names):
            ctx.set_variant(kmx, v)
            vm, fm = run("mixed")
            for k, what in enumerate(("velocity", "double dot", "laplacian")):
                assert rel_l2(vm[k], v64[k]) < 1e-6, (name, "vertices", what)
                assert rel_l2(fm[k][..., leaf], f64[k][..., leaf]) < 1e-6, (
                    name, "faces", what)
    finally:
        ctx.set_variant(kmx, -1)


@pytest.mark.parametrize("eps_h", [0.0, 1.0])
def test_every_swe_variant_matches(ctx, ora, eps_h):
    """Each compiled shape of the sums (reciprocal sharing, occupancy):
    FP64 against the oracle at 1e-12, mixed against FP64 at 1e-6, on the
    collocated face sums of a masked cubed-sphere mesh."""
    m, zf, sf, hf, hv = _fields(1, 4)
    h = m.mean_spacing()
    eps, pse = eps_h * h, 2 * h
    mask = dev(m.face_mask, torch.uint8)
    fx = dev(m.face_x)
    leaf = m.face_mask == 0
    ru, rdd, rlap = ora.swe_sums(aos(m.face_x), hf, aos(m.face_x), zf, sf,
                                 m.face_area, hf, m.face_mask, collocated=True,
                                 eps=eps, pse_eps=pse)
    k64, kmx = _abi.LPM_KERNEL_SWE_F64, _abi.LPM_KERNEL_SWE_MIXED
    assert len(_variant_names(ctx, k64)) >= 2
    assert len(_variant_names(ctx, kmx)) >= 2
    ref64 = None
    try:
        for v, name in enumerate(_variant_names(ctx, k64)):
            ctx.set_variant(k64, v)
            u, dd, lap = ctx.swe_sums(fx, dev(hf), fx, dev(zf), dev(sf),
                                      dev(m.face_area), dev(hf), src_mask=mask,
                                      collocated=True, eps=eps, pse_eps=pse)
            u, dd, lap = u.cpu().numpy().T, dd.cpu().numpy(), lap.cpu().numpy()
            assert rel_l2(u[leaf], ru[leaf]) < TOL64, name
            assert rel_l2(dd[leaf], rdd[leaf]) < TOL64, name
            assert rel_l2(lap[leaf], rlap[leaf]) < TOL64, name
            if ref64 is None:
                ref64 = (u, dd, lap)
        for v, name in enumerate(_variant_names(ctx, kmx)):
            ctx.set_variant(kmx, v)
            u, dd, lap = ctx.swe_sums(fx, dev(hf), fx, dev(zf), dev(sf),
                                      dev(m.face_area), dev(hf), src_mask=mask,
                                      collocated=True, eps=eps, pse_eps=pse,
                                      precision="mixed")
            u, dd, lap = u.cpu().numpy().T, dd.cpu().numpy(), lap.cpu().numpy()
            assert rel_l2(u[leaf], ref64[0][leaf]) < 1e-6, name
            assert rel_l2(dd[leaf], ref64[1][leaf]) < 1e-6, name
            assert rel_l2(lap[leaf], ref64[2][leaf]) < 1e-6, name
    finally:
        ctx.set_variant(k64, -1)
        ctx.set_variant(kmx, -1)


def test_swe_sums_tc2_closed_forms(ctx):
    m = mesh(0, 5)
    h = m.mean_spacing()
    st_f, st_v = G.tc2_state(m.face_x), G.tc2_state(m.vert_x)
    u, dd, lap = ctx.swe_sums(dev(m.vert_x), dev(st_v["surface"]),
                              dev(m.face_x), dev(st_f["zeta"]),
                              dev(st_f["sigma"]), dev(m.face_area),
                              dev(st_f["surface"]),
                              src_mask=dev(m.face_mask, torch.uint8), eps=h,
                              pse_eps=2 * h)
    assert rel_l2(u.cpu().numpy(), st_v["u"]) < 5e-3
    assert rel_l2(dd.cpu().numpy(), st_v["ddot"]) < 5e-2
    assert rel_l2(lap.cpu().numpy(), st_v["lap_surface"]) < 5e-2


def _swe(ctx, seed, depth):
    from lpm_amd.swe import SWE
    m, zf, sf, hf, hv = _fields(seed, depth)
    s = SWE(seed, depth, ctx=ctx)
    rng = np.random.default_rng(9)
    s.init_surface(lambda x: G.tc2_state(x)["surface"])
    s.init_vorticity(lambda x: G.tc2_state(x)["zeta"])
    s.init_divergence(lambda x: 0.01 * np.sin(3 * x[0]) * x[2])
    h = m.mean_spacing()
    s.set_kernel_parameters(h, 2 * h)
    return s


def _oracle_state(s):
    g = lambda t: t.cpu().numpy().copy()  # noqa: E731
    return {"vx": aos(g(s.vert_x)), "vzeta": g(s.rel_vort_passive),
            "vsigma": g(s.div_passive), "vh": g(s.depth_passive),
            "vu": aos(g(s.velocity_passive)), "vdd": g(s.double_dot_passive),
            "vlap": g(s.surf_lap_passive), "vs": g(s.surf_passive),
            "fx": aos(g(s.face_x)), "fzeta": g(s.rel_vort_active),
            "fsigma": g(s.div_active), "farea": g(s.face_area),
            "fmass": g(s.mass_active), "fh": g(s.depth_active),
            "fu": aos(g(s.velocity_active)), "fdd": g(s.double_dot_active),
            "flap": g(s.surf_lap_active), "fs": g(s.surf_active),
            "fmask": g(s.face_mask)}


@pytest.mark.parametrize("seed,compat", [(1, True), (1, False), (0, False)])
def test_swe_rk4_matches_oracle(ctx, ora, seed, compat):
    from lpm_amd.swe import SWERK4
    s = _swe(ctx, seed, 2)
    s.init_direct_sums()
    torch.cuda.synchronize()
    st = _oracle_state(s)
    # the oracle's initial sums must agree with the GPU's
    leaf = st["fmask"] == 0
    solver = SWERK4(0.01, s, reference_compat=compat)
    for _ in range(2):
        s.advance_timestep(solver)
        ora.swe_rk4_step(st, 0.01, OMEGA, s.g, s.eps, s.pse_eps, compat=compat)
    torch.cuda.synchronize()
    got = _oracle_state(s)
    for k in ("vx", "vzeta", "vsigma", "vh", "vu", "vdd", "vlap"):
        assert rel_l2(got[k], st[k]) < 1e-11, k
    for k in ("fx", "fzeta", "fsigma", "farea", "fu", "fdd", "flap", "fh"):
        assert rel_l2(got[k][leaf], st[k][leaf]) < 1e-11, k


@pytest.mark.parametrize("seed", [0, 1])
def test_swe_rk2_matches_oracle(ctx, ora, seed):
    from lpm_amd.swe import SWERK2
    s = _swe(ctx, seed, 2)
    s.init_direct_sums()
    torch.cuda.synchronize()
    st = _oracle_state(s)
    leaf = st["fmask"] == 0
    solver = SWERK2(0.01, s)
    for _ in range(3):
        s.advance_timestep(solver)
        ora.swe_rk2_step(st, 0.01, OMEGA, s.g, s.eps, s.pse_eps)
    torch.cuda.synchronize()
    got = _oracle_state(s)
    for k in ("vx", "vzeta", "vsigma", "vh", "vu", "vdd", "vlap"):
        assert rel_l2(got[k], st[k]) < 1e-11, k
    for k in ("fx", "fzeta", "fsigma", "farea", "fu", "fdd", "flap", "fh"):
        assert rel_l2(got[k][leaf], st[k][leaf]) < 1e-11, k


@pytest.mark.parametrize("order", [4, 2])
def test_swe_renormalized_steps_match_oracle(ctx, ora, order):
    """LPM_RK4_RENORMALIZE projects work and updated positions onto the
    sphere in both SWE steppers, as the oracle does."""
    from lpm_amd.swe import SWERK2, SWERK4
    s = _swe(ctx, 1, 2)
    s.init_direct_sums()
    torch.cuda.synchronize()
    st = _oracle_state(s)
    leaf = st["fmask"] == 0
    solver = (SWERK4(0.02, s, reference_compat=False, renormalize=True)
              if order == 4 else SWERK2(0.02, s, renormalize=True))
    for _ in range(3):
        s.advance_timestep(solver)
        if order == 4:
            ora.swe_rk4_step(st, 0.02, OMEGA, s.g, s.eps, s.pse_eps,
                             compat=False, renorm=True)
        else:
            ora.swe_rk2_step(st, 0.02, OMEGA, s.g, s.eps, s.pse_eps,
                             renorm=True)
    torch.cuda.synchronize()
    got = _oracle_state(s)
    for k in ("vx", "vzeta", "vsigma", "vh", "vu", "vdd", "vlap"):
        assert rel_l2(got[k], st[k]) < 1e-11, k
    for k in ("fx", "fzeta", "fsigma", "farea", "fu", "fdd", "flap", "fh"):
        assert rel_l2(got[k][leaf], st[k][leaf]) < 1e-11, k
    r = np.linalg.norm(got["fx"][leaf], axis=1)
    assert np.max(np.abs(r - 1.0)) < 1e-14
    assert np.max(np.abs(np.linalg.norm(got["vx"], axis=1) - 1.0)) < 1e-14


def test_swe_mixed_renormalized_trajectory_tracks_fp64(ctx):
    """Mixed-precision SWE RK4 with renormalisation stays within 1e-6 of the
    FP64 trajectory (velocity) over several steps; positions stay on the
    sphere."""
    from lpm_amd.swe import SWERK4
    runs = {}
    for prec in ("fp64", "mixed"):
        s = _swe(ctx, 0, 4)
        s.init_direct_sums(precision=prec)
        solver = SWERK4(0.01, s, precision=prec, reference_compat=False,
                        renormalize=True, leaf_only=True)
        for _ in range(5):
            s.advance_timestep(solver)
        torch.cuda.synchronize()
        runs[prec] = _oracle_state(s)
    a, b = runs["fp64"], runs["mixed"]
    leaf = a["fmask"] == 0
    assert rel_l2(b["vu"], a["vu"]) < 1e-6
    assert rel_l2(b["fu"][leaf], a["fu"][leaf]) < 1e-6
    assert rel_l2(b["vx"], a["vx"]) < 1e-8
    assert rel_l2(b["fh"][leaf], a["fh"][leaf]) < 1e-8
    assert np.max(np.abs(np.linalg.norm(b["fx"][leaf], axis=1) - 1)) < 1e-14


def test_swe_argument_errors(ctx):
    L = _abi.lib()
    x = dev(np.eye(3))
    one = dev(np.ones(3))
    p = x.data_ptr()
    # PSE without surfaces
    rc = L.lpm_swe_sphere_sums(ctx.h, p, 1, 3, 3, None, p, 1, 3, one.data_ptr(),
                               one.data_ptr(), one.data_ptr(), None, None, 3, 0,
                               0.1, 0.1, 0, 0, p, 1, 3, None, None, None)
    assert rc == _abi.LPM_ERR_ARG
    assert b"PSE" in L.lpm_last_error(ctx.h)
