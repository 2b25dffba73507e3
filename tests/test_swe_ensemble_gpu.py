"""GPU tests of the sphere SWE ensemble path (one persistent launch for the
sums of many problems) and of graph-captured SWE steps, through the C ABI.

Each member must give exactly (bit for bit) what the one-problem entry
points give for that member alone -- the work item runs the same block loop
over the same target slots and source tiles -- and the FP64 ensemble sums
must match the FP64 oracle per member to 1e-12 relative.
"""
import numpy as np
import pytest

from conftest import aos, mesh, rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from lpm_amd import _abi  # noqa: E402
from lpm_amd import galleries as G  # noqa: E402

TOL64 = 1e-12
MEMBERS = [(0, 2), (1, 3), (0, 3), (1, 2), (0, 1)]
EPS, PSE_EPS = 0.12, 0.25


@pytest.fixture(scope="module")
def ctx():
    from lpm_amd.device import Context
    c = Context(0)
    yield c
    c.close()


def _state_fns(k):
    """Seeded TC2-like state of member k with a divergence perturbation."""
    rng = np.random.default_rng(100 + k)
    a, b = rng.uniform(0.5, 1.5, 2)
    surf = lambda x: G.tc2_state(x)["surface"] + 1e-3 * a * x[0] * x[1]  # noqa: E731
    zeta = lambda x: G.tc2_state(x)["zeta"] + 0.05 * b * x[1]  # noqa: E731
    div = lambda x: 0.01 * a * np.sin(3 * x[0]) * x[2]  # noqa: E731
    return surf, zeta, div


def _ensemble(ctx, members=MEMBERS):
    from lpm_amd.swe import SWEEnsemble
    ens = SWEEnsemble([mesh(s, d) for s, d in members], ctx=ctx)
    fns = [_state_fns(k) for k in range(len(members))]
    ens.init_surface([f[0] for f in fns])
    ens.init_vorticity([f[1] for f in fns])
    ens.init_divergence([f[2] for f in fns])
    ens.set_kernel_parameters(EPS, PSE_EPS)
    return ens


def _single(ctx, ens, m):
    from lpm_amd.swe import SWE
    s = SWE(None, None, ctx=ctx, mesh=ens.meshes[m])
    surf, zeta, div = _state_fns(m)
    s.init_surface(surf)
    s.init_vorticity(zeta)
    s.init_divergence(div)
    s.set_kernel_parameters(EPS, PSE_EPS)
    return s


FIELDS_V = ("velocity_passive", "double_dot_passive", "surf_lap_passive")
FIELDS_F = ("velocity_active", "double_dot_active", "surf_lap_active")
STATE_V = ("vert_x", "rel_vort_passive", "div_passive", "depth_passive",
           "surf_passive")
STATE_F = ("face_x", "rel_vort_active", "div_active", "face_area",
           "depth_active", "surf_active")


def _same(a, b):
    return np.array_equal(a.cpu().numpy(), b.cpu().numpy(), equal_nan=True)


def _member_view(t, sl):
    return t[:, sl] if t.dim() == 2 else t[sl]


def _check_member(ens, s, m, names_v, names_f, leaf_only):
    vs, fs = ens.vert_slice(m), ens.face_slice(m)
    for n in names_v:
        assert _same(_member_view(getattr(ens, n), vs), getattr(s, n)), (m, n)
    lf = torch.tensor(ens.meshes[m].face_mask == 0, device="cuda")
    for n in names_f:
        a = _member_view(getattr(ens, n), fs)
        b = getattr(s, n)
        if leaf_only:
            a = a[:, lf] if a.dim() == 2 else a[lf]
            b = b[:, lf] if b.dim() == 2 else b[lf]
        assert _same(a, b), (m, n)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("skip", [False, True])
def test_sums_equal_members_alone(ctx, precision, skip):
    flags = _abi.LPM_FLAG_SKIP_MASKED_TARGETS if skip else 0
    ens = _ensemble(ctx)
    ens.init_direct_sums(precision, flags)
    for m in range(ens.n_members):
        s = _single(ctx, ens, m)
        s.init_direct_sums(precision, flags)
        torch.cuda.synchronize()
        _check_member(ens, s, m, FIELDS_V, FIELDS_F, leaf_only=skip)


def test_mixed_larger_members_equal_members_alone(ctx):
    """Members large enough (icosahedral levels 4 and 5, cubed sphere level
    4; up to 80 source tiles) that the mixed sums take every path: far
    tiles and far 64-source sub-tiles on the matrix cores, FP64 near
    sub-tiles, with each member's tile records and far-sum slots at its own
    offsets.  Bit-equal to the members alone, and within 1e-6 of FP64."""
    members = [(0, 4), (1, 4), (0, 5)]
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    out = {}
    for precision in ("fp64", "mixed"):
        ens = _ensemble(ctx, members)
        ens.init_direct_sums(precision, skip)
        torch.cuda.synchronize()
        out[precision] = {n: getattr(ens, n).cpu().numpy()
                          for n in FIELDS_V + FIELDS_F}
        if precision == "mixed":
            for m in range(ens.n_members):
                s = _single(ctx, ens, m)
                s.init_direct_sums(precision, skip)
                torch.cuda.synchronize()
                _check_member(ens, s, m, FIELDS_V, FIELDS_F, leaf_only=True)
    leaf = np.concatenate([mesh(s, d).face_mask == 0 for s, d in members])
    for n in FIELDS_V + FIELDS_F:
        a, b = out["mixed"][n], out["fp64"][n]
        if n in FIELDS_F:
            a, b = a[..., leaf], b[..., leaf]
        assert rel_l2(a, b) < 1e-6, n


def test_fp64_ensemble_matches_oracle(ctx, ora):
    ens = _ensemble(ctx)
    ens.init_direct_sums("fp64", _abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    torch.cuda.synchronize()
    g = lambda t: t.cpu().numpy()  # noqa: E731
    for m, msh in enumerate(ens.meshes):
        vs, fs = ens.vert_slice(m), ens.face_slice(m)
        zf, sf = g(ens.rel_vort_active)[fs], g(ens.div_active)[fs]
        hf, hv = g(ens.surf_active)[fs], g(ens.surf_passive)[vs]
        ru, rdd, rlap = ora.swe_sums(aos(msh.vert_x), hv, aos(msh.face_x), zf,
                                     sf, msh.face_area, hf, msh.face_mask,
                                     eps=EPS, pse_eps=PSE_EPS)
        assert rel_l2(g(ens.velocity_passive)[:, vs].T, ru) < TOL64
        assert rel_l2(g(ens.double_dot_passive)[vs], rdd) < TOL64
        assert rel_l2(g(ens.surf_lap_passive)[vs], rlap) < TOL64


def _steppers(order):
    from lpm_amd.swe import SWEEnsembleRK2, SWEEnsembleRK4, SWERK2, SWERK4
    return (SWEEnsembleRK4, SWERK4) if order == 4 else (SWEEnsembleRK2, SWERK2)


@pytest.mark.parametrize("order", [4, 2])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_rk_equals_members_alone(ctx, precision, order):
    SWEEnsembleRK4, SWERK4 = _steppers(order)
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    ens = _ensemble(ctx)
    ens.init_direct_sums(precision, skip)
    singles = [_single(ctx, ens, m) for m in range(ens.n_members)]
    for s in singles:
        s.init_direct_sums(precision, skip)
    se = SWEEnsembleRK4(0.01, ens, precision, leaf_only=True)
    ss = [SWERK4(0.01, s, precision, leaf_only=True) for s in singles]
    for _ in range(2):
        ens.advance_timestep(se)
        for s, r in zip(singles, ss):
            s.advance_timestep(r)
    torch.cuda.synchronize()
    for m, s in enumerate(singles):
        _check_member(ens, s, m, FIELDS_V + STATE_V, FIELDS_F + STATE_F,
                      leaf_only=True)


@pytest.mark.parametrize("order", [4, 2])
@pytest.mark.parametrize("ensemble", [False, True])
def test_graph_replays_match_eager(ctx, ensemble, order):
    SWEEnsembleRK4, SWERK4 = _steppers(order)
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    if ensemble:
        a, b = _ensemble(ctx), _ensemble(ctx)
        cls = SWEEnsembleRK4
    else:
        holder = _ensemble(ctx, [(1, 3)])
        a, b = _single(ctx, holder, 0), _single(ctx, holder, 0)
        cls = SWERK4
    a.init_direct_sums("mixed", skip)
    b.init_direct_sums("mixed", skip)
    stream = torch.cuda.Stream()
    ra = cls(0.01, a, "mixed", leaf_only=True)
    rb = cls(0.01, b, "mixed", leaf_only=True, use_graph=True,
             stream=stream.cuda_stream)
    for step in range(3):
        a.advance_timestep(ra)
        torch.cuda.synchronize()
        b.advance_timestep(rb)
        stream.synchronize()
        for n in ("vert_x", "velocity_passive", "surf_lap_passive",
                  "rel_vort_active", "depth_active"):
            assert _same(getattr(a, n), getattr(b, n)), (step, n)


def test_ensemble_argument_errors(ctx):
    import ctypes
    ens = _ensemble(ctx, [(0, 1), (1, 1)])
    e = ens.descriptor()
    bad = (ctypes.c_int64 * 3)(0, 3, int(ens.vert_offsets[-1]) - 1)
    e.vert_offsets = ctypes.cast(bad, _abi.c_i64p)
    with pytest.raises(_abi.LpmError) as ex:
        ctx.swe_ensemble_init_sums(e, EPS, PSE_EPS)
    assert ex.value.code == _abi.LPM_ERR_ARG
