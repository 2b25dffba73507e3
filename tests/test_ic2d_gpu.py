"""GPU parity of the Incompressible2D sphere sums (fused velocity + stream
function) and the Incompressible2DRK2 stepper against the FP64 oracle.

Tolerances: 1e-12 relative for the sums (FP64 throughout; tiles, the
factored cross product and the Newton reciprocal reorder the arithmetic),
1e-11 after RK2 steps; the mixed path stays within 1e-6 of FP64.
"""
import numpy as np
import pytest

from conftest import aos, mesh, rel_l2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from lpm_amd import _abi  # noqa: E402
from lpm_amd import galleries as G  # noqa: E402

TOL = 1e-12


@pytest.fixture(scope="module")
def ctx():
    from lpm_amd.device import Context
    c = Context(0)
    yield c
    c.close()


def dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


# level 3 runs each target block over all sources; at level 5 (80 or 24
# source tiles of 256 behind 148 or 57 target blocks) the launch splits
# every block's sources into chunks and adds the chunk partials in a second
# kernel
@pytest.mark.parametrize("seed,depth", [(0, 3), (1, 3), (0, 5), (1, 5)])
@pytest.mark.parametrize("eps", [0.0, 0.05])
def test_sums_match_oracle(ctx, ora, seed, depth, eps):
    m = mesh(seed, depth)
    z = G.rh54_vorticity(m.face_x, 0.1, -0.1)
    mask = dev(m.face_mask, torch.uint8)
    u, psi = ctx.ic2d_sums(dev(m.vert_x), dev(m.face_x), dev(z),
                           dev(m.face_area), mask, eps=eps)
    ru, rpsi = ora.ic2d_sums(aos(m.vert_x), aos(m.face_x), z, m.face_area,
                             m.face_mask, eps=eps)
    assert rel_l2(u.cpu().numpy().T, ru) < TOL
    assert rel_l2(psi.cpu().numpy(), rpsi) < TOL
    fx = dev(m.face_x)
    u, psi = ctx.ic2d_sums(fx, fx, dev(z), dev(m.face_area), mask,
                           collocated=True, eps=eps,
                           flags=_abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    ru, rpsi = ora.ic2d_sums(aos(m.face_x), aos(m.face_x), z, m.face_area,
                             m.face_mask, collocated=True, eps=eps)
    lf = m.face_mask == 0
    assert rel_l2(u.cpu().numpy().T[lf], ru[lf]) < TOL
    assert rel_l2(psi.cpu().numpy()[lf], rpsi[lf]) < TOL


def _model(ctx, seed, depth, eps, Omega):
    from lpm_amd.incompressible2d import Incompressible2D
    tilt = np.array([0.6, 0.0, 0.8])
    s = Incompressible2D(seed, depth, Omega=Omega, eps=eps, ctx=ctx)
    s.init_vorticity(lambda x: 3.0 * (tilt @ np.asarray(x)) +
                     G.rh54_vorticity(x, 0.1, -0.1))
    return s


def _state(s):
    g = lambda t: t.cpu().numpy().copy()  # noqa: E731
    return {"vx": aos(g(s.vert_x)), "vzeta": g(s.rel_vort_passive),
            "vu": aos(g(s.velocity_passive)), "vpsi": g(s.stream_fn_passive),
            "fx": aos(g(s.face_x)), "fzeta": g(s.rel_vort_active),
            "fu": aos(g(s.velocity_active)), "fpsi": g(s.stream_fn_active),
            "farea": g(s.face_area), "fmask": g(s.face_mask)}


@pytest.mark.parametrize("eps", [0.0, 0.05])
def test_rk2_matches_oracle(ctx, ora, eps):
    from lpm_amd.incompressible2d import Incompressible2DRK2
    s = _model(ctx, 1, 2, eps, 2 * np.pi)
    s.init_direct_sums()
    torch.cuda.synchronize()
    st = _state(s)
    solver = Incompressible2DRK2(0.01, s)
    for _ in range(2):
        s.advance_timestep(solver)
        ora.ic2d_rk2_step(st, 0.01, s.Omega, eps)
    torch.cuda.synchronize()
    got = _state(s)
    for k in ("vx", "vzeta", "vu", "vpsi", "fx", "fzeta", "fu", "fpsi"):
        assert rel_l2(got[k], st[k]) < 1e-11, k


@pytest.mark.parametrize("depth", [3, 5])
def test_rk2_graph_replays_match_eager(ctx, depth):
    from lpm_amd.incompressible2d import Incompressible2DRK2
    a = _model(ctx, 0, depth, 0.0, 2 * np.pi)
    b = _model(ctx, 0, depth, 0.0, 2 * np.pi)
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    a.init_direct_sums(skip)
    b.init_direct_sums(skip)
    stream = torch.cuda.Stream()
    ra = Incompressible2DRK2(0.01, a, leaf_only=True, renormalize=True)
    rb = Incompressible2DRK2(0.01, b, leaf_only=True, renormalize=True,
                             use_graph=True, stream=stream.cuda_stream)
    for step in range(3):
        a.advance_timestep(ra)
        torch.cuda.synchronize()
        b.advance_timestep(rb)
        stream.synchronize()
        for n in ("vert_x", "velocity_passive", "stream_fn_active",
                  "rel_vort_active"):
            assert torch.equal(getattr(a, n), getattr(b, n)), (step, n)


def test_argument_errors(ctx):
    L = _abi.lib()
    x = dev(np.eye(3))
    one = dev(np.ones(3))
    p = x.data_ptr()
    rc = L.lpm_ic2d_sums(ctx.h, p, 1, 3, 3, p, 1, 3, one.data_ptr(),
                         one.data_ptr(), None, 2, 1, 0.0, 0, 0, p, 1, 3, None,
                         None)
    assert rc == _abi.LPM_ERR_ARG
    assert b"collocated" in L.lpm_last_error(ctx.h)
    rc = L.lpm_ic2d_sums(ctx.h, p, 1, 3, 3, p, 1, 3, one.data_ptr(),
                         one.data_ptr(), None, 3, 0, -1.0, 0, 0, p, 1, 3, None,
                         None)
    assert rc == _abi.LPM_ERR_ARG
    rc = L.lpm_ic2d_sums(ctx.h, p, 1, 3, 3, p, 1, 3, one.data_ptr(),
                         one.data_ptr(), None, 3, 0, 0.0, 7, 0, p, 1, 3, None,
                         None)
    assert rc == _abi.LPM_ERR_ARG
    assert b"precision" in L.lpm_last_error(ctx.h)


@pytest.mark.parametrize("eps", [0.0, 0.05])
@pytest.mark.parametrize("seed", [0, 1])
def test_mixed_sums_match_fp64(ctx, seed, eps):
    # FP32 pairs on tile-relative offsets, native FP32 log, FP64 carries
    m = mesh(seed, 4)
    z = G.rh54_vorticity(m.face_x, 0.1, -0.1)
    mask = dev(m.face_mask, torch.uint8)
    leaf = m.face_mask == 0
    for tgt, col in ((dev(m.vert_x), False), (dev(m.face_x), True)):
        u64, p64 = ctx.ic2d_sums(tgt, dev(m.face_x), dev(z), dev(m.face_area),
                                 mask, collocated=col, eps=eps)
        u32, p32 = ctx.ic2d_sums(tgt, dev(m.face_x), dev(z), dev(m.face_area),
                                 mask, collocated=col, eps=eps,
                                 precision="mixed")
        a, b = u32.cpu().numpy(), u64.cpu().numpy()
        pa, pb = p32.cpu().numpy(), p64.cpu().numpy()
        if col:
            a, b, pa, pb = a[:, leaf], b[:, leaf], pa[leaf], pb[leaf]
        assert rel_l2(a, b) < 1e-6, (col, rel_l2(a, b))
        assert rel_l2(pa, pb) < 1e-6, (col, rel_l2(pa, pb))


def test_mixed_rk2_tracks_fp64(ctx):
    from lpm_amd.incompressible2d import Incompressible2DRK2
    a = _model(ctx, 0, 4, 0.0, 2 * np.pi)
    b = _model(ctx, 0, 4, 0.0, 2 * np.pi)
    a.init_direct_sums(flags=_abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    b.init_direct_sums(flags=_abi.LPM_FLAG_SKIP_MASKED_TARGETS, precision="mixed")
    ra = Incompressible2DRK2(0.01, a, leaf_only=True, renormalize=True)
    rb = Incompressible2DRK2(0.01, b, leaf_only=True, renormalize=True,
                             precision="mixed")
    for _ in range(5):
        a.advance_timestep(ra)
        b.advance_timestep(rb)
    torch.cuda.synchronize()
    assert rel_l2(b.velocity_passive.cpu().numpy(),
                  a.velocity_passive.cpu().numpy()) < 1e-6
    # positions integrate the ~1e-7 velocity difference: dt |u| 5 steps
    assert rel_l2(b.vert_x.cpu().numpy(), a.vert_x.cpu().numpy()) < 1e-8


def test_leaf_face_totals(ctx):
    """total_vorticity / total_enstrophy / total_kinetic_energy
    (src/lpm_incompressible2d_impl.hpp:91-137) over the leaf faces."""
    s = _model(ctx, 1, 3, 0.0, 2 * np.pi)
    s.init_direct_sums()
    st = _state(s)
    leaf = st["fmask"] == 0
    a = st["farea"] * leaf
    assert np.isclose(s.total_vorticity(), np.sum(st["fzeta"] * a),
                      rtol=1e-13, atol=1e-13)
    assert np.isclose(s.total_enstrophy(), 0.5 * np.sum(st["fzeta"] ** 2 * a),
                      rtol=1e-13)
    assert np.isclose(s.total_kinetic_energy(),
                      0.5 * np.sum((st["fu"] ** 2).sum(axis=1) * a), rtol=1e-13)
    # divided faces (zero area) do not count even with garbage values
    s.rel_vort_active[torch.from_numpy(~leaf).to(s.rel_vort_active.device)] = 1e30
    assert np.isclose(s.total_vorticity(), np.sum(st["fzeta"] * a),
                      rtol=1e-13, atol=1e-13)
