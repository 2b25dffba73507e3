"""GPU tests of the BVE ensemble path: many independent problems of uneven
sizes evaluated by one persistent launch per direct sum, through the C ABI.

Each member must give exactly (bit for bit) what the one-problem entry
points give for that member alone with the kernel variant of the same
shape -- the work item runs the same block loop over the same target slots
and source tiles -- and the FP64 ensemble must match the FP64 oracle per
member to 1e-12 relative.
"""
import numpy as np
import pytest

from conftest import aos, mesh, rel_l2, rel_linf

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from lpm_amd import _abi  # noqa: E402
from lpm_amd import galleries as G  # noqa: E402

TOL64 = 1e-12
OMEGA = 2 * np.pi
# uneven members, both seed kinds, divided faces present in every member
MEMBERS = [(0, 2), (1, 3), (0, 3), (1, 2), (0, 4), (1, 1)]
KIDS = {"fp64": (_abi.LPM_KERNEL_BVE_VEL_F64, _abi.LPM_KERNEL_BVE_ENSEMBLE_F64),
        "mixed": (_abi.LPM_KERNEL_BVE_VEL_MIXED,
                  _abi.LPM_KERNEL_BVE_ENSEMBLE_MIXED)}


@pytest.fixture(scope="module")
def ctx():
    from lpm_amd.device import Context
    c = Context(0)
    yield c
    c.close()


def _ensemble(ctx, members=MEMBERS, seed_base=G.ENSEMBLE_SEED):
    from lpm_amd.bve import BVEEnsemble
    ens = BVEEnsemble([mesh(s, d) for s, d in members], ctx=ctx)
    params = [G.ensemble_member(k, seed_base=seed_base)
              for k in range(len(members))]
    ens.init_vorticity([
        (lambda x, p=p: G.rotated_rh54_vorticity(x, p["omega"], p["K"],
                                                 p["rotation"]))
        for p in params])
    return ens


def _single(ctx, ens, m):
    from lpm_amd.bve import BVESphere
    s = BVESphere(None, None, ctx=ctx, mesh=ens.meshes[m])
    s.rel_vort_verts.copy_(ens.rel_vort_verts[ens.vert_slice(m)])
    s.rel_vort_faces.copy_(ens.rel_vort_faces[ens.face_slice(m)])
    return s


def _force(ctx, precision, ens_variant):
    """Force ensemble variant ens_variant and its one-problem twin."""
    k1, ke = KIDS[precision]
    name = ctx.variants(ke)[ens_variant]
    assert name.startswith("persistent_")
    ctx.set_variant(k1, ctx.variants(k1).index(name[len("persistent_"):]))
    ctx.set_variant(ke, ens_variant)


def _unforce(ctx):
    for k in (0, 1, 5, 6):
        ctx.set_variant(k, -1)


def _same(a, b):
    return np.array_equal(a.cpu().numpy(), b.cpu().numpy(), equal_nan=True)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("skip", [False, True])
def test_init_velocity_equals_members_alone(ctx, precision, skip):
    flags = _abi.LPM_FLAG_SKIP_MASKED_TARGETS if skip else 0
    kid = KIDS[precision][1]
    try:
        for v in range(len(ctx.variants(kid))):
            _force(ctx, precision, v)
            ens = _ensemble(ctx)
            ens.init_velocity(precision, flags)
            for m in range(ens.n_members):
                s = _single(ctx, ens, m)
                s.init_velocity(precision, flags)
                torch.cuda.synchronize()
                assert _same(ens.velocity_verts[:, ens.vert_slice(m)],
                             s.velocity_verts), (v, m)
                fs = ens.face_slice(m)
                if skip:
                    leaf = torch.tensor(ens.meshes[m].face_mask == 0,
                                        device="cuda")
                    assert _same(ens.velocity_faces[:, fs][:, leaf],
                                 s.velocity_faces[:, leaf]), (v, m)
                else:
                    assert _same(ens.velocity_faces[:, fs],
                                 s.velocity_faces), (v, m)
    finally:
        _unforce(ctx)


def test_fp64_ensemble_matches_oracle_per_member(ctx, ora):
    ens = _ensemble(ctx)
    ens.init_velocity("fp64", _abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    torch.cuda.synchronize()
    zf = ens.rel_vort_faces.cpu().numpy()
    for m, msh in enumerate(ens.meshes):
        z = zf[ens.face_slice(m)]
        ref_v = ora.bve_velocity(aos(msh.vert_x), aos(msh.face_x), z,
                                 msh.face_area, msh.face_mask)
        got_v = ens.velocity_verts[:, ens.vert_slice(m)].cpu().numpy().T
        assert rel_l2(got_v, ref_v) < TOL64
        assert rel_linf(got_v, ref_v) < TOL64
        lf = msh.face_mask == 0
        ref_f = ora.bve_velocity(aos(msh.face_x), aos(msh.face_x), z,
                                 msh.face_area, msh.face_mask, collocated=True)
        got_f = ens.velocity_faces[:, ens.face_slice(m)].cpu().numpy().T
        assert rel_l2(got_f[lf], ref_f[lf]) < TOL64
        assert rel_linf(got_f[lf], ref_f[lf]) < TOL64


def test_mixed_ensemble_close_to_fp64(ctx):
    ens = _ensemble(ctx)
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    ens.init_velocity("fp64", skip)
    u64 = ens.velocity_verts.clone()
    ens.init_velocity("mixed", skip)
    assert rel_l2(ens.velocity_verts.cpu().numpy(), u64.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_rk4_equals_members_alone(ctx, precision):
    from lpm_amd.bve import BVEEnsembleRK4, BVERK4
    flags = dict(reference_compat=True, renormalize=precision == "mixed",
                 leaf_only=True)
    ens = _ensemble(ctx)
    ens.init_velocity(precision, _abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    singles = [_single(ctx, ens, m) for m in range(ens.n_members)]
    for s in singles:
        s.init_velocity(precision, _abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    st = BVEEnsembleRK4(0.01, ens, OMEGA, precision, **flags)
    ss = [BVERK4(0.01, s, precision, **flags) for s in singles]
    for _ in range(2):
        st.advance_timestep(ens)
        for s, r in zip(singles, ss):
            r.advance_timestep(s)
    torch.cuda.synchronize()
    for m, s in enumerate(singles):
        vs, fs = ens.vert_slice(m), ens.face_slice(m)
        lf = torch.tensor(ens.meshes[m].face_mask == 0, device="cuda")
        assert _same(ens.vert_x[:, vs], s.vert_x), m
        assert _same(ens.velocity_verts[:, vs], s.velocity_verts), m
        assert _same(ens.rel_vort_verts[vs], s.rel_vort_verts), m
        assert _same(ens.face_x[:, fs][:, lf], s.face_x[:, lf]), m
        assert _same(ens.rel_vort_faces[fs][lf], s.rel_vort_faces[lf]), m


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_eager(ctx, precision):
    from lpm_amd.bve import BVEEnsembleRK4
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    a = _ensemble(ctx)
    b = _ensemble(ctx)
    a.init_velocity(precision, skip)
    b.init_velocity(precision, skip)
    stream = torch.cuda.Stream()
    ra = BVEEnsembleRK4(0.01, a, OMEGA, precision, leaf_only=True)
    rb = BVEEnsembleRK4(0.01, b, OMEGA, precision, leaf_only=True,
                        use_graph=True, stream=stream.cuda_stream)
    for step in range(3):
        ra.advance_timestep(a)
        torch.cuda.synchronize()
        rb.advance_timestep(b)
        stream.synchronize()
        assert _same(a.vert_x, b.vert_x), step
        assert _same(a.velocity_verts, b.velocity_verts), step
        assert _same(a.rel_vort_faces, b.rel_vort_faces), step


def test_autotune_and_timing(ctx):
    ens = _ensemble(ctx)
    best, ms, all_ms = ens.autotune("mixed", _abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    kid = _abi.LPM_KERNEL_BVE_ENSEMBLE_MIXED
    assert 0 <= best < len(ctx.variants(kid))
    assert ms > 0 and all(t > 0 for t in all_ms)
    assert ms == min(all_ms)
    ctx.timing_enable(True)
    ens.init_velocity("mixed", _abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    total, launches, pairs = ctx.timing_query(kid)
    ctx.timing_enable(False)
    assert launches == 1 and total > 0
    expect = sum((m.n_verts + m.n_faces) * m.n_faces for m in ens.meshes)
    assert pairs == pytest.approx(expect)


def test_argument_errors(ctx):
    import ctypes
    ens = _ensemble(ctx, members=[(0, 1), (1, 1)])
    e = ens.descriptor()
    bad = (ctypes.c_int64 * 3)(0, 5, int(ens.face_offsets[-1]) + 1)
    e.face_offsets = ctypes.cast(bad, _abi.c_i64p)
    with pytest.raises(_abi.LpmError) as ex:
        ctx.bve_ensemble_init_velocity(e)
    assert ex.value.code == _abi.LPM_ERR_ARG
    e = ens.descriptor()
    empty = (ctypes.c_int64 * 3)(0, 0, int(ens.face_offsets[-1]))
    e.face_offsets = ctypes.cast(empty, _abi.c_i64p)
    with pytest.raises(_abi.LpmError):
        ctx.bve_ensemble_init_velocity(e)
    e = ens.descriptor()
    e.n_members = 0
    with pytest.raises(_abi.LpmError):
        ctx.bve_ensemble_init_velocity(e)
