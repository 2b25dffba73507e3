"""GPU parity of the BVE direct sums and RK4 stepper (HIP, through the C ABI)
against the FP64 oracle on the same inputs.

Tolerances: FP64 kernels <= 1e-12 relative (L2 and Linf) against the oracle
-- the order of summation differs (tiles, factored cross product, Newton
reciprocal) but the arithmetic is FP64 throughout; the mixed-precision
kernel <= 1e-6 relative L2 against the FP64 kernel.
"""
import ctypes

import numpy as np
import pytest

from conftest import aos, mesh, rel_l2, rel_linf

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from lpm_amd import _abi  # noqa: E402

OMEGA = 2 * np.pi
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


def rh(x):
    from lpm_amd import galleries as G
    return G.rh54_vorticity(x, 0.1, -0.1)


def _inputs(seed, depth):
    m = mesh(seed, depth)
    z = rh(m.face_x)
    return m, z


@pytest.mark.parametrize("seed,depth", [(0, 3), (1, 3), (0, 4)])
def test_velocity_fp64_vertices_match_oracle(ctx, ora, seed, depth):
    m, z = _inputs(seed, depth)
    u = ctx.bve_velocity(dev(m.vert_x), dev(m.face_x), dev(z), dev(m.face_area),
                         dev(m.face_mask, torch.uint8))
    ref = ora.bve_velocity(aos(m.vert_x), aos(m.face_x), z, m.face_area,
                           m.face_mask)
    got = u.cpu().numpy().T
    assert rel_l2(got, ref) < TOL64
    assert rel_linf(got, ref) < TOL64


@pytest.mark.parametrize("seed,depth", [(0, 3), (1, 3), (1, 4)])
def test_velocity_fp64_faces_match_oracle(ctx, ora, seed, depth):
    m, z = _inputs(seed, depth)
    fx = dev(m.face_x)
    u = ctx.bve_velocity(fx, fx, dev(z), dev(m.face_area),
                         dev(m.face_mask, torch.uint8), collocated=True)
    ref = ora.bve_velocity(aos(m.face_x), aos(m.face_x), z, m.face_area,
                           m.face_mask, collocated=True)
    got = u.cpu().numpy().T
    leaf = m.face_mask == 0
    assert rel_l2(got[leaf], ref[leaf]) < TOL64
    assert rel_linf(got[leaf], ref[leaf]) < TOL64
    if seed == 1:
        # cubed sphere: no divided face coincides with a source, all finite
        assert rel_l2(got, ref) < TOL64
    else:
        # icosahedral seed faces share their central child's centre: the
        # reference's parent-face velocity is singular, so is ours
        assert not np.all(np.isfinite(got[:20]))


def test_strided_layouts_and_unmasked_sources(ctx):
    m, z = _inputs(1, 3)
    vx = dev(m.vert_x)
    fz, fa = dev(z), dev(m.face_area)
    fm = dev(m.face_mask, torch.uint8)
    soa = ctx.bve_velocity(vx, dev(m.face_x), fz, fa, fm)
    # AoS (n,3) buffers passed as transposed views: element stride 3
    vx_aos = dev(aos(m.vert_x)).T
    fx_aos = dev(aos(m.face_x)).T
    out_aos = torch.zeros((m.n_verts, 3), dtype=torch.float64, device="cuda").T
    ctx.bve_velocity(vx_aos, fx_aos, fz, fa, fm, out=out_aos)
    assert torch.equal(soa, out_aos)
    # leaf faces only, no mask: same sources, same order -> identical
    ls = m.leaf_slice
    leaf_only = ctx.bve_velocity(vx, dev(m.face_x[:, ls]), dev(z[ls]),
                                 dev(m.face_area[ls]), None)
    assert torch.equal(soa, leaf_only)


def test_velocity_mixed_matches_fp64(ctx):
    for seed, depth in [(0, 5), (1, 5)]:
        m, z = _inputs(seed, depth)
        vx, fx = dev(m.vert_x), dev(m.face_x)
        args = (dev(z), dev(m.face_area), dev(m.face_mask, torch.uint8))
        u64 = ctx.bve_velocity(vx, fx, *args)
        u32 = ctx.bve_velocity(vx, fx, *args, precision="mixed")
        e = rel_l2(u32.cpu().numpy(), u64.cpu().numpy())
        assert e < TOL_MIXED, e
        f64 = ctx.bve_velocity(fx, fx, *args, collocated=True)
        f32 = ctx.bve_velocity(fx, fx, *args, collocated=True, precision="mixed")
        leaf = torch.from_numpy(m.face_mask == 0)
        e = rel_l2(f32[:, leaf].cpu().numpy(), f64[:, leaf].cpu().numpy())
        assert e < TOL_MIXED, e


# level 5: the vertex launch has 41 target blocks, so every block's sources
# are split into chunks whose partials a second kernel adds
@pytest.mark.parametrize("seed,depth", [(0, 3), (0, 5), (1, 5)])
def test_stream_function_matches_oracle(ctx, ora, seed, depth):
    m, z = _inputs(seed, depth)
    psi = ctx.bve_stream_fn(dev(m.vert_x), dev(m.face_x), dev(z),
                            dev(m.face_area), dev(m.face_mask, torch.uint8))
    ref = ora.bve_stream_fn(aos(m.vert_x), aos(m.face_x), z, m.face_area,
                            m.face_mask)
    assert rel_l2(psi.cpu().numpy(), ref) < TOL64
    fx = dev(m.face_x)
    psi = ctx.bve_stream_fn(fx, fx, dev(z), dev(m.face_area),
                            dev(m.face_mask, torch.uint8), collocated=True)
    ref = ora.bve_stream_fn(aos(m.face_x), aos(m.face_x), z, m.face_area,
                            m.face_mask, collocated=True)
    leaf = m.face_mask == 0
    assert rel_l2(psi.cpu().numpy()[leaf], ref[leaf]) < TOL64


@pytest.mark.parametrize("seed,depth", [(0, 3), (1, 3)])
def test_solve_matches_oracle(ctx, ora, seed, depth):
    """lpm_bve_solve (BVEVertexSolve / BVEFaceSolve): velocity and stream
    function against the oracle's two separate sums."""
    m, z = _inputs(seed, depth)
    args = (dev(m.face_x), dev(z), dev(m.face_area),
            dev(m.face_mask, torch.uint8))
    u, psi = ctx.bve_solve(dev(m.vert_x), *args)
    ref_u = ora.bve_velocity(aos(m.vert_x), aos(m.face_x), z, m.face_area,
                             m.face_mask)
    ref_p = ora.bve_stream_fn(aos(m.vert_x), aos(m.face_x), z, m.face_area,
                              m.face_mask)
    assert rel_l2(u.cpu().numpy().T, ref_u) < TOL64
    assert rel_linf(u.cpu().numpy().T, ref_u) < TOL64
    assert rel_l2(psi.cpu().numpy(), ref_p) < TOL64
    u, psi = ctx.bve_solve(dev(m.face_x), *args, collocated=True,
                           flags=_abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    ref_u = ora.bve_velocity(aos(m.face_x), aos(m.face_x), z, m.face_area,
                             m.face_mask, collocated=True)
    ref_p = ora.bve_stream_fn(aos(m.face_x), aos(m.face_x), z, m.face_area,
                              m.face_mask, collocated=True)
    lf = m.face_mask == 0
    assert rel_l2(u.cpu().numpy().T[lf], ref_u[lf]) < TOL64
    assert rel_l2(psi.cpu().numpy()[lf], ref_p[lf]) < TOL64


def test_stream_function_log_accuracy_and_special_values(ctx):
    # one source at the pole: psi_i = s log(1 - z_i) with s = -zeta A / 4 pi,
    # so the kernel's table-driven FP64 log is compared directly against the
    # host libm log of the same double argument (1 - z is rounded once on
    # both sides), over arguments from 1e-15 to 1e300 (1 - z cannot reach
    # below 2^-53), both sides of 1 and of powers of two, and the
    # coincident (log 0 = -inf) and just-past-the-pole (NaN) cases
    rng = np.random.default_rng(5)
    arg = np.concatenate([
        10.0 ** rng.uniform(-15, 300, 4000), rng.uniform(0, 2, 4000),
        1 + rng.uniform(-1e-6, 1e-6, 500), 2.0 ** np.arange(-53, 1000, 7.0),
        1 - 2.0 ** -np.arange(1, 53.0), 1 + 2.0 ** -np.arange(1, 53.0),
        np.nextafter(2.0 ** np.arange(-40, 40.0), 0)])
    z = np.concatenate([1 - arg, [1.0, 1.0 + 2.0 ** -52]])
    x = np.zeros((3, len(z)))
    x[2] = z
    y = np.array([[0.0], [0.0], [1.0]])
    zeta, area = np.array([-4 * np.pi]), np.array([1.0])
    psi = ctx.bve_stream_fn(dev(x), dev(y), dev(zeta), dev(area),
                            None).cpu().numpy()
    s = -zeta[0] * area[0] / (4 * np.pi)
    with np.errstate(divide="ignore", invalid="ignore"):
        want = s * np.log(1 - z)
    fin = np.isfinite(want)
    np.testing.assert_allclose(psi[fin], want[fin], rtol=1e-15, atol=0)
    assert psi[-2] == -np.inf and np.isnan(psi[-1])
    assert fin[:-2].all()


def test_skip_masked_targets_leaves_them_untouched(ctx):
    m, z = _inputs(0, 3)
    fx = dev(m.face_x)
    out = torch.full((3, m.n_faces), 7.0, dtype=torch.float64, device="cuda")
    ctx.bve_velocity(fx, fx, dev(z), dev(m.face_area),
                     dev(m.face_mask, torch.uint8), collocated=True,
                     flags=_abi.LPM_FLAG_SKIP_MASKED_TARGETS, out=out)
    full = ctx.bve_velocity(fx, fx, dev(z), dev(m.face_area),
                            dev(m.face_mask, torch.uint8), collocated=True)
    div = torch.from_numpy(m.face_mask == 1)
    assert torch.all(out[:, div] == 7.0)
    # compacted targets are grouped into different waves, which may take the
    # self-exclusion path for different tiles: equal to rounding
    assert rel_l2(out[:, ~div].cpu().numpy(), full[:, ~div].cpu().numpy()) < 1e-14


def test_edge_cases(ctx, ora):
    rng = np.random.default_rng(11)
    for ns in (1, 3, 255, 256, 257, 1000):
        y = rng.normal(size=(3, ns))
        y /= np.linalg.norm(y, axis=0)
        x = rng.normal(size=(3, 37))
        x /= np.linalg.norm(x, axis=0)
        zz = rng.normal(size=ns)
        aa = rng.uniform(0.1, 1.0, size=ns)
        for prec in ("fp64", "mixed"):
            u = ctx.bve_velocity(dev(x), dev(y), dev(zz), dev(aa), None,
                                 precision=prec).cpu().numpy().T
            ref = ora.bve_velocity(aos(x), aos(y), zz, aa)
            if prec == "fp64":
                # random clouds hold close pairs where 1 - x.y cancels; the
                # reference formula itself then strays from the exact sum,
                # so compare with the extended-precision sum and allow the
                # reference formula's own error when comparing to it
                xp = ora.bve_velocity_xp(aos(x), aos(y), zz, aa)
                assert rel_l2(u, xp) < TOL64, (ns, prec)
                assert rel_l2(u, ref) < TOL64 + rel_l2(ref, xp), (ns, prec)
            else:
                assert rel_l2(u, ref) < 1e-5, (ns, prec)
    # no targets
    y = dev(np.eye(3))
    out = ctx.bve_velocity(torch.zeros((3, 0), dtype=torch.float64, device="cuda"),
                           y, dev(np.ones(3)), dev(np.ones(3)), None)
    assert out.shape == (3, 0)
    # every source divided: zero velocity
    x = dev(np.array([[1.0], [0.0], [0.0]]))
    u = ctx.bve_velocity(x, y, dev(np.ones(3)), dev(np.ones(3)),
                         dev(np.ones(3), torch.uint8))
    assert torch.all(u == 0)


def test_argument_errors(ctx):
    L = _abi.lib()
    x = dev(np.eye(3))
    p = x.data_ptr()
    one = dev(np.ones(3))
    # negative size
    rc = L.lpm_bve_velocity(ctx.h, p, 1, 3, -1, p, 1, 3, one.data_ptr(),
                            one.data_ptr(), None, 3, 0, 0.0, 0, 0, p, 1, 3, None)
    assert rc == _abi.LPM_ERR_ARG
    # collocated with n_tgt != n_src
    rc = L.lpm_bve_velocity(ctx.h, p, 1, 3, 2, p, 1, 3, one.data_ptr(),
                            one.data_ptr(), None, 3, 1, 0.0, 0, 0, p, 1, 3, None)
    assert rc == _abi.LPM_ERR_ARG
    assert b"collocated" in L.lpm_last_error(ctx.h)
    # unknown precision
    rc = L.lpm_bve_velocity(ctx.h, p, 1, 3, 3, p, 1, 3, one.data_ptr(),
                            one.data_ptr(), None, 3, 0, 0.0, 9, 0, p, 1, 3, None)
    assert rc == _abi.LPM_ERR_ARG


def test_nonfinite_is_reported(ctx):
    m, z = _inputs(1, 2)
    z = z.copy()
    z[m.n_faces - 1] = np.nan
    with pytest.raises(_abi.LpmError) as ei:
        ctx.bve_velocity(dev(m.vert_x), dev(m.face_x), dev(z), dev(m.face_area),
                         dev(m.face_mask, torch.uint8),
                         flags=_abi.LPM_FLAG_CHECK_FINITE)
    assert ei.value.code == _abi.LPM_ERR_NONFINITE


# ------------------------------------------------------------------ RK4
def _sphere(ctx, seed, depth, zeta_fn=rh):
    from lpm_amd.bve import BVESphere
    s = BVESphere(seed, depth, ctx=ctx)
    s.init_vorticity(zeta_fn)
    return s


def _oracle_state(s):
    m = s.mesh
    return {"vx": aos(s.vert_x.cpu().numpy()), "vzeta": s.rel_vort_verts.cpu().numpy().copy(),
            "vu": aos(s.velocity_verts.cpu().numpy()),
            "fx": aos(s.face_x.cpu().numpy()), "fzeta": s.rel_vort_faces.cpu().numpy().copy(),
            "fu": aos(s.velocity_faces.cpu().numpy()),
            "farea": m.face_area.copy(), "fmask": m.face_mask.copy()}


@pytest.mark.parametrize("compat,renorm", [(True, False), (False, False), (False, True)])
def test_rk4_steps_match_oracle(ctx, ora, compat, renorm):
    from lpm_amd.bve import BVERK4
    s = _sphere(ctx, 1, 3)
    s.init_velocity()
    st = _oracle_state(s)
    solver = BVERK4(0.01, s, reference_compat=compat, renormalize=renorm)
    for _ in range(3):
        solver.advance_timestep(s)
        ora.bve_rk4_step(st, 0.01, OMEGA, compat=compat, renorm=renorm)
    torch.cuda.synchronize()
    assert rel_l2(aos(s.vert_x.cpu().numpy()), st["vx"]) < TOL64
    assert rel_l2(aos(s.face_x.cpu().numpy()), st["fx"]) < TOL64
    assert rel_l2(s.rel_vort_verts.cpu().numpy(), st["vzeta"]) < TOL64
    assert rel_l2(s.rel_vort_faces.cpu().numpy(), st["fzeta"]) < TOL64
    assert rel_l2(aos(s.velocity_verts.cpu().numpy()), st["vu"]) < 1e-11
    assert rel_l2(aos(s.velocity_faces.cpu().numpy()), st["fu"]) < 1e-11


def test_small_rh54_run_matches_oracle(ctx, ora):
    # the CPU-runnable configuration: icosahedral level 4, RH 5-4
    # (u0 = 0.1, amp = -0.1), Omega = 2 pi, eps = 0, dt = 0.01, 10 RK4 steps
    # in the reference's stage layout (compat), every stored face a target
    from lpm_amd.bve import BVERK4
    s = _sphere(ctx, 0, 4)
    s.init_velocity()
    st = _oracle_state(s)
    solver = BVERK4(0.01, s, reference_compat=True)
    for _ in range(10):
        solver.advance_timestep(s)
        ora.bve_rk4_step(st, 0.01, OMEGA, compat=True)
    torch.cuda.synchronize()
    leaf = s.mesh.face_mask == 0
    assert rel_l2(aos(s.vert_x.cpu().numpy()), st["vx"]) < TOL64
    assert rel_l2(aos(s.face_x.cpu().numpy())[leaf], st["fx"][leaf]) < TOL64
    assert rel_l2(s.rel_vort_verts.cpu().numpy(), st["vzeta"]) < TOL64
    assert rel_l2(s.rel_vort_faces.cpu().numpy()[leaf], st["fzeta"][leaf]) < TOL64
    assert rel_l2(aos(s.velocity_verts.cpu().numpy()), st["vu"]) < TOL64
    assert rel_l2(aos(s.velocity_faces.cpu().numpy())[leaf], st["fu"][leaf]) < TOL64


def test_rk4_graph_replays_match_eager(ctx):
    from lpm_amd.bve import BVERK4
    a = _sphere(ctx, 0, 3)
    b = _sphere(ctx, 0, 3)
    flags = dict(reference_compat=True, leaf_only=True)
    a.init_velocity(flags=_abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    b.init_velocity(flags=_abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    ra = BVERK4(0.01, a, **flags)
    rb = BVERK4(0.01, b, use_graph=True, **flags)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for k in range(4):
            ra.advance_timestep(a)
            rb.advance_timestep(b)
            # eager work of the program between replays
            torch.cuda.synchronize()
            _ = (a.vert_x * 2).sum().item()
            assert torch.equal(a.vert_x, b.vert_x), k
            assert torch.equal(a.face_x, b.face_x), k
            assert torch.equal(a.rel_vort_faces, b.rel_vort_faces), k
            assert torch.equal(a.velocity_verts, b.velocity_verts), k


def test_rk4_graph_replays_are_asynchronous(ctx):
    """Ten graph-replayed steps issued back to back with no host
    synchronisation in between equal ten eager steps bit for bit."""
    from lpm_amd.bve import BVERK4
    a = _sphere(ctx, 1, 3)
    b = _sphere(ctx, 1, 3)
    a.init_velocity()
    b.init_velocity()
    ra = BVERK4(0.01, a, reference_compat=True)
    stream = torch.cuda.Stream()
    rb = BVERK4(0.01, b, reference_compat=True, use_graph=True)
    torch.cuda.synchronize()
    for _ in range(10):
        ra.advance_timestep(a)
    with torch.cuda.stream(stream):
        for _ in range(10):
            rb.advance_timestep(b)
    stream.synchronize()
    torch.cuda.synchronize()
    assert torch.equal(a.vert_x, b.vert_x)
    assert torch.equal(a.face_x, b.face_x)
    assert torch.equal(a.rel_vort_faces, b.rel_vort_faces)
    assert torch.equal(a.velocity_faces, b.velocity_faces)


@pytest.mark.parametrize("compat", [True, False])
def test_rk4_graph_replay_matches_oracle(ctx, ora, compat):
    """Graph-replayed steps against the oracle directly (not only against
    eager steps)."""
    from lpm_amd.bve import BVERK4
    s = _sphere(ctx, 0, 3)
    s.init_velocity()
    st = _oracle_state(s)
    solver = BVERK4(0.01, s, reference_compat=compat, use_graph=True)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for _ in range(3):
            solver.advance_timestep(s)
    stream.synchronize()
    for _ in range(3):
        ora.bve_rk4_step(st, 0.01, OMEGA, compat=compat)
    leaf = s.mesh.face_mask == 0
    assert rel_l2(aos(s.vert_x.cpu().numpy()), st["vx"]) < TOL64
    assert rel_l2(aos(s.face_x.cpu().numpy())[leaf], st["fx"][leaf]) < TOL64
    assert rel_l2(s.rel_vort_faces.cpu().numpy()[leaf], st["fzeta"][leaf]) < TOL64
    assert rel_l2(aos(s.velocity_verts.cpu().numpy()), st["vu"]) < 1e-11


def test_graph_cache_stays_bounded(ctx):
    """Adaptive time steps (a new dt every step) capture a new graph each
    time; the context keeps at most LPM_GRAPH_CACHE_MAX of them and the
    results still equal eager steps."""
    from lpm_amd.bve import BVERK4
    a = _sphere(ctx, 1, 2)
    b = _sphere(ctx, 1, 2)
    a.init_velocity()
    b.init_velocity()
    stream = torch.cuda.Stream()
    L = _abi.lib()
    n = ctypes.c_int()
    with torch.cuda.stream(stream):
        for k in range(_abi.LPM_GRAPH_CACHE_MAX + 6):
            dt = 0.001 * (1 + k)
            BVERK4(dt, a).advance_timestep(a)
            BVERK4(dt, b, use_graph=True).advance_timestep(b)
            ctx.check(L.lpm_graph_cache_size(ctx.h, ctypes.byref(n)))
            assert n.value <= _abi.LPM_GRAPH_CACHE_MAX
    stream.synchronize()
    assert n.value == _abi.LPM_GRAPH_CACHE_MAX
    assert torch.equal(a.vert_x, b.vert_x)
    assert torch.equal(a.velocity_faces, b.velocity_faces)


def test_graph_recaptured_after_variant_change():
    """A captured step records the kernels of the variant in force when it
    was captured: after set_variant the next graph step captures again, and
    graph steps stay bit-equal to eager steps under the same choice."""
    from lpm_amd.bve import BVERK4
    from lpm_amd.device import Context
    c = Context(0)
    kid = _abi.LPM_KERNEL_BVE_VEL_F64
    L = _abi.lib()
    n = ctypes.c_int()
    sizes = []
    try:
        a, b = _sphere(c, 0, 3), _sphere(c, 0, 3)
        a.init_velocity()
        b.init_velocity()
        ra, rb = BVERK4(0.01, a), BVERK4(0.01, b, use_graph=True)
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            for v in (-1, 0, len(c.variants(kid)) - 1):
                c.set_variant(kid, v)
                for _ in range(2):
                    ra.advance_timestep(a)
                    rb.advance_timestep(b)
                c.check(L.lpm_graph_cache_size(c.h, ctypes.byref(n)))
                sizes.append(n.value)
        stream.synchronize()
        assert sizes == [1, 2, 3]
        # divided faces hold NaN unless leaf_only is off in both
        assert _same_bits(a.vert_x, b.vert_x)
        assert _same_bits(a.face_x, b.face_x)
        assert _same_bits(a.velocity_verts, b.velocity_verts)
    finally:
        c.close()


def test_split_variant_under_a_callers_capture_runs_its_twin():
    """A source-split variant cannot grow its workspace inside a capture the
    library did not start: in a fresh context, a split init_velocity
    captured by the caller runs the unsplit twin (bit-equal to it on every
    replay), and the next eager call allocates the workspace and splits
    (same sums to 1e-13, another summation order)."""
    from lpm_amd.device import Context
    c = Context(0)
    kid = _abi.LPM_KERNEL_BVE_VEL_F64
    names = c.variants(kid)
    v = names.index("block256_tpl4_nr1_quad_fused_sy_u2_tile512_split32")
    twin = names.index("block256_tpl4_nr1_quad_fused_sy_u2_tile512")
    try:
        # level 5: 40 source tiles, so some of the 32 chunks hold two and
        # the split sum's association differs from the unsplit one
        a, b = _sphere(c, 0, 5), _sphere(c, 0, 5)
        c.set_variant(kid, twin)
        a.init_velocity()
        b.init_velocity()  # sizes every workspace but the split one
        c.set_variant(kid, v)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.graph(g, stream=side):
            b.init_velocity()
        for _ in range(3):
            b.velocity_verts.zero_()
            b.velocity_faces.zero_()
            g.replay()
            torch.cuda.synchronize()
            _ = (a.vert_x * 2).sum().item()  # eager work between replays
            assert torch.equal(a.velocity_verts, b.velocity_verts)
            assert _same_bits(a.velocity_faces, b.velocity_faces)
        del g
        b.init_velocity()
        torch.cuda.synchronize()
        assert not torch.equal(a.velocity_verts, b.velocity_verts)
        assert rel_l2(b.velocity_verts.cpu().numpy(),
                      a.velocity_verts.cpu().numpy()) < 1e-13
    finally:
        c.close()


def _same_bits(p, q):
    """Equal values, with NaN where the other has NaN."""
    return bool(((p == q) | (p.isnan() & q.isnan())).all())


def test_context_switching_streams_is_ordered(ctx):
    """One context driven from two streams in turn, with no host
    synchronisation: the library orders each call after the previous
    stream's work, so the results equal one-stream steps."""
    from lpm_amd.bve import BVERK4
    a = _sphere(ctx, 0, 4)
    b = _sphere(ctx, 0, 4)
    a.init_velocity()
    b.init_velocity()
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    ra = BVERK4(0.01, a)
    for _ in range(4):
        ra.advance_timestep(a)
    rb = BVERK4(0.01, b)
    for k in range(4):
        with torch.cuda.stream(s1 if k % 2 == 0 else s2):
            rb.advance_timestep(b)
    s1.synchronize()
    s2.synchronize()
    torch.cuda.synchronize()
    assert torch.equal(a.vert_x, b.vert_x)
    # a divided icosahedral face whose corners are equidistant shares its
    # centre with its central child, so its collocated velocity is 0/0 = NaN,
    # as in the reference's sum over every stored face: NaNs must sit at the
    # same faces, everything else bit-equal
    assert _same_bits(a.velocity_faces, b.velocity_faces)
    assert _same_bits(a.face_x, b.face_x)


def test_rk4_solid_body_rotation(ctx):
    from lpm_amd.bve import BVERK4
    Om_r = 0.5
    s = _sphere(ctx, 0, 4, lambda x: 2 * Om_r * x[2])
    s.init_velocity(flags=_abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    x0 = s.vert_x.cpu().numpy().copy()
    solver = BVERK4(0.05, s, leaf_only=True, precision="mixed", renormalize=True)
    for _ in range(8):
        solver.advance_timestep(s)
    ang = Om_r * 0.4
    c, sn = np.cos(ang), np.sin(ang)
    ex = np.stack([c * x0[0] - sn * x0[1], sn * x0[0] + c * x0[1], x0[2]])
    err = np.max(np.linalg.norm(s.vert_x.cpu().numpy() - ex, axis=0))
    assert err < 1e-2 * ang


@pytest.mark.parametrize("seed,tag", [(0, "icos"), (1, "cube")])
def test_frozen_oracle_fixture(ctx, oracle_fixtures, seed, tag):
    # level 2, RH 5-4: velocity, stream function and one classical RK4 step
    # against the committed oracle outputs (tests/golden)
    from lpm_amd.bve import BVERK4
    F = {k[len(tag) + 1:]: v for k, v in oracle_fixtures.items()
         if k.startswith(tag + "_")}
    s = _sphere(ctx, seed, 2)
    s.init_velocity()
    s.init_stream_fn()
    torch.cuda.synchronize()
    leaf = s.mesh.face_mask == 0
    assert rel_l2(aos(s.velocity_verts.cpu().numpy()), F["vert_u"]) < TOL64
    assert rel_l2(aos(s.velocity_faces.cpu().numpy())[leaf],
                  F["face_u"][leaf]) < TOL64
    assert rel_l2(s.stream_fn_verts.cpu().numpy(), F["vert_psi"]) < TOL64
    solver = BVERK4(0.01, s, reference_compat=False)
    solver.advance_timestep(s)
    torch.cuda.synchronize()
    assert rel_l2(aos(s.vert_x.cpu().numpy()), F["step_vert_x"]) < TOL64
    assert rel_l2(aos(s.face_x.cpu().numpy())[leaf], F["step_face_x"][leaf]) < TOL64
    assert rel_l2(s.rel_vort_verts.cpu().numpy(), F["step_vert_zeta"]) < TOL64
    assert rel_l2(s.rel_vort_faces.cpu().numpy()[leaf],
                  F["step_face_zeta"][leaf]) < TOL64
    assert rel_l2(aos(s.velocity_verts.cpu().numpy()), F["step_vert_u"]) < 1e-11


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_rh54_wave_moves_at_its_phase_speed(ctx, prec):
    # RH 5-4 is an exact solution of the nonlinear BVE that rotates rigidly
    # in longitude at nu = -0.3255 (u0 = 0.1, Omega = 2 pi): after 20 RK4
    # steps the velocity at the particles matches the rotated pattern to
    # the t = 0 quadrature error and is far from the unrotated pattern
    from lpm_amd import galleries as G
    from lpm_amd.bve import BVERK4
    s = _sphere(ctx, 0, 5)
    s.init_velocity(precision=prec, flags=_abi.LPM_FLAG_SKIP_MASKED_TARGETS)
    nu = G.rh54_wave_speed(0.1, OMEGA)

    def errs(t):
        x = s.vert_x.cpu().numpy()
        u = s.velocity_verts.cpu().numpy()
        ue = G.rotate_z(G.rh54_velocity(G.rotate_z(x, -nu * t), 0.1, -0.1), nu * t)
        us = G.rh54_velocity(x, 0.1, -0.1)
        return rel_l2(u, ue), rel_l2(u, us)

    e0, _ = errs(0.0)
    solver = BVERK4(0.01, s, precision=prec, reference_compat=False,
                    leaf_only=True, renormalize=True)
    for _ in range(20):
        solver.advance_timestep(s)
    torch.cuda.synchronize()
    e_rot, e_static = errs(0.2)
    assert e_rot < 1.6 * e0, (e0, e_rot)
    assert e_static > 15 * e_rot, (e_rot, e_static)


@pytest.mark.parametrize("kid,prec", [(_abi.LPM_KERNEL_BVE_VEL_F64, "fp64"),
                                      (_abi.LPM_KERNEL_BVE_VEL_MIXED, "mixed")])
def test_every_variant_matches_oracle(ctx, ora, kid, prec):
    m, z = _inputs(1, 3)
    vx, fx = dev(m.vert_x), dev(m.face_x)
    args = (dev(z), dev(m.face_area), dev(m.face_mask, torch.uint8))
    ref_v = ora.bve_velocity(aos(m.vert_x), aos(m.face_x), z, m.face_area,
                             m.face_mask)
    ref_f = ora.bve_velocity(aos(m.face_x), aos(m.face_x), z, m.face_area,
                             m.face_mask, collocated=True)
    tol = TOL64 if prec == "fp64" else TOL_MIXED
    names = ctx.variants(kid)
    assert len(names) >= 4
    try:
        for v in range(len(names)):
            ctx.set_variant(kid, v)
            for eps in (0.0, 0.05):
                uv = ctx.bve_velocity(vx, fx, *args, precision=prec, eps=eps)
                uf = ctx.bve_velocity(fx, fx, *args, collocated=True,
                                      precision=prec, eps=eps)
                if eps == 0.0:
                    assert rel_l2(uv.cpu().numpy().T, ref_v) < tol, names[v]
                    assert rel_l2(uf.cpu().numpy().T, ref_f) < tol, names[v]
                else:
                    ctx.set_variant(kid, 0)
                    ev = ctx.bve_velocity(vx, fx, *args, precision=prec, eps=eps)
                    ctx.set_variant(kid, v)
                    assert rel_l2(uv.cpu().numpy(), ev.cpu().numpy()) < tol, names[v]
    finally:
        ctx.set_variant(kid, -1)


@pytest.mark.parametrize("kid,prec", [(_abi.LPM_KERNEL_BVE_VEL_F64, "fp64"),
                                      (_abi.LPM_KERNEL_BVE_VEL_MIXED, "mixed")])
def test_adaptively_refined_mesh_matches_oracle(ctx, ora, kid, prec):
    # a locally refined mesh (cap refined twice): leaves of three levels,
    # divided faces interleaved with leaves in the face order
    from lpm_amd.mesh import PolyMesh2d
    pm = PolyMesh2d("cube", 3, amr_limit=2)
    for _ in range(2):
        m = pm.sphere_mesh()
        pm.divide_flagged_faces((m.face_mask == 0) & (m.face_x[2] > 0.5))
    m = pm.sphere_mesh()
    assert len(np.unique(m.face_level[m.face_mask == 0])) == 3
    z = rh(m.face_x)
    vx, fx = dev(m.vert_x), dev(m.face_x)
    args = (dev(z), dev(m.face_area), dev(m.face_mask, torch.uint8))
    ref_v = ora.bve_velocity(aos(m.vert_x), aos(m.face_x), z, m.face_area,
                             m.face_mask)
    ref_f = ora.bve_velocity(aos(m.face_x), aos(m.face_x), z, m.face_area,
                             m.face_mask, collocated=True)
    leaf = m.face_mask == 0
    tol = TOL64 if prec == "fp64" else TOL_MIXED
    names = ctx.variants(kid)
    try:
        for v in range(len(names)):
            ctx.set_variant(kid, v)
            uv = ctx.bve_velocity(vx, fx, *args, precision=prec).cpu().numpy()
            uf = ctx.bve_velocity(fx, fx, *args, collocated=True,
                                  precision=prec).cpu().numpy()
            assert rel_l2(uv.T, ref_v) < tol, names[v]
            assert rel_l2(uf.T[leaf], ref_f[leaf]) < tol, names[v]
    finally:
        ctx.set_variant(kid, -1)


@pytest.mark.parametrize("seed", [0, 1])
def test_matrix_core_variant_on_a_larger_mesh(ctx, seed):
    # level 5: most source tiles are far from every wave, so most pairs take
    # the matrix-core path; the near tiles take the difference form
    m, z = _inputs(seed, 5)
    vx, fx = dev(m.vert_x), dev(m.face_x)
    args = (dev(z), dev(m.face_area), dev(m.face_mask, torch.uint8))
    kid = _abi.LPM_KERNEL_BVE_VEL_MIXED
    names = ctx.variants(kid)
    mv = [i for i, n in enumerate(names) if n.startswith("mfma")]
    assert mv, names
    u64v = ctx.bve_velocity(vx, fx, *args).cpu().numpy()
    u64f = ctx.bve_velocity(fx, fx, *args, collocated=True).cpu().numpy()
    leaf = m.face_mask == 0
    try:
        for v in mv:
            ctx.set_variant(kid, v)
            uv = ctx.bve_velocity(vx, fx, *args, precision="mixed").cpu().numpy()
            uf = ctx.bve_velocity(fx, fx, *args, collocated=True,
                                  precision="mixed").cpu().numpy()
            assert rel_l2(uv, u64v) < TOL_MIXED, names[v]
            assert rel_l2(uf[:, leaf], u64f[:, leaf]) < TOL_MIXED, names[v]
            assert rel_linf(uv, u64v) < 10 * TOL_MIXED, names[v]
    finally:
        ctx.set_variant(kid, -1)


@pytest.mark.parametrize("kid,prec", [(_abi.LPM_KERNEL_BVE_VEL_F64, "fp64"),
                                      (_abi.LPM_KERNEL_BVE_VEL_MIXED, "mixed")])
def test_source_split_variants_match_oracle(ctx, ora, kid, prec):
    """The source-split variants (each target block's sources in K chunks,
    partial sums added in chunk order by a second kernel) run where the
    context may keep their workspace: init_velocity and the RK4 step.
    Level 4 has 5120 leaf sources, so the chunks hold a few source tiles
    (or none).  Against the oracle (FP64 1e-12, mixed 1e-6), and within
    1e-12 of the unsplit twin (same shape, one summation level less); the
    one-shot velocity entry point runs the twin in their place (FP64:
    bit-equal to it; the mixed one-shot sum keeps the mesh's target order,
    so its waves differ)."""
    from lpm_amd.bve import BVERK4
    names = ctx.variants(kid)
    split = [v for v, n in enumerate(names) if "_split" in n]
    assert split, names
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    tol = TOL64 if prec == "fp64" else TOL_MIXED
    try:
        for v in split:
            twin = names.index(names[v].rsplit("_split", 1)[0])
            for flags in (0, skip):
                ctx.set_variant(kid, twin)
                t = _sphere(ctx, 0, 4)
                t.init_velocity(prec, flags=flags)
                ctx.set_variant(kid, v)
                s = _sphere(ctx, 0, 4)
                s.init_velocity(prec, flags=flags)
                if prec == "fp64":
                    u1 = ctx.bve_velocity(s.vert_x, s.face_x,
                                          s.rel_vort_faces, s.face_area,
                                          s.face_mask)
                    assert torch.equal(u1, t.velocity_verts), names[v]
                st = _oracle_state(s)
                ref_v = ora.bve_velocity(st["vx"], st["fx"], st["fzeta"],
                                         st["farea"], st["fmask"])
                ref_f = ora.bve_velocity(st["fx"], st["fx"], st["fzeta"],
                                         st["farea"], st["fmask"],
                                         collocated=True)
                leaf = s.mesh.face_mask == 0
                assert rel_l2(st["vu"], ref_v) < tol, names[v]
                assert rel_l2(st["fu"][leaf], ref_f[leaf]) < tol, names[v]
                for a, b in ((s.velocity_verts, t.velocity_verts),
                             (s.velocity_faces[:, leaf],
                              t.velocity_faces[:, leaf])):
                    assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-12, \
                        names[v]
                if flags:
                    # masked faces untouched (their initial zeros)
                    assert not s.velocity_faces.cpu().numpy()[:, ~leaf].any()
            s = _sphere(ctx, 1, 4)
            s.init_velocity(prec)
            st = _oracle_state(s)
            # positions kept on the sphere, where the mixed sums are exact
            r = BVERK4(0.01, s, precision=prec, reference_compat=True,
                       renormalize=True)
            for _ in range(2):
                r.advance_timestep(s)
                ora.bve_rk4_step(st, 0.01, OMEGA, compat=True, renorm=True)
            torch.cuda.synchronize()
            xtol = TOL64 if prec == "fp64" else 1e-8
            assert rel_l2(aos(s.vert_x.cpu().numpy()), st["vx"]) < xtol
            assert rel_l2(s.rel_vort_faces.cpu().numpy(), st["fzeta"]) < xtol
            assert rel_l2(aos(s.velocity_verts.cpu().numpy()),
                          st["vu"]) < tol, names[v]
    finally:
        ctx.set_variant(kid, -1)


def test_regularised_kernel_matches_pair_formula(ctx, ora):
    # eps > 0: the pair term is kzeta_sphere with eps (the BVE kernel with
    # 1 - x.y + eps^2 in the denominator)
    m, z = _inputs(0, 2)
    eps = 0.05
    u = ctx.bve_velocity(dev(m.vert_x), dev(m.face_x), dev(z), dev(m.face_area),
                         dev(m.face_mask, torch.uint8), eps=eps).cpu().numpy().T
    vx, fx = aos(m.vert_x), aos(m.face_x)
    ref = np.zeros_like(vx)
    for j in np.where(m.face_mask == 0)[0]:
        for i in range(len(vx)):
            ref[i] += ora.kzeta_sphere(vx[i], fx[j], z[j], m.face_area[j], eps)
    assert rel_l2(u, ref) < TOL64


def test_autotune_records_a_variant(ctx, tmp_path):
    s = _sphere(ctx, 0, 4)
    for prec, kid in (("fp64", _abi.LPM_KERNEL_BVE_VEL_F64),
                      ("mixed", _abi.LPM_KERNEL_BVE_VEL_MIXED)):
        best, ms, all_ms = ctx.bve_autotune(
            s.fields(), prec, flags=_abi.LPM_FLAG_SKIP_MASKED_TARGETS, repeats=1)
        assert 0 <= best < len(ctx.variants(kid))
        assert ms > 0 and ms == min(all_ms)
    path = str(tmp_path / "tune.txt")
    ctx.tune_save(path)
    assert "gfx950" in open(path).read() or len(open(path).read()) > 0
    ctx.tune_load(path)
    # a table entry naming no compiled variant falls back to the default
    m, z = _inputs(0, 3)
    args = (dev(m.vert_x), dev(m.face_x), dev(z), dev(m.face_area),
            dev(m.face_mask, torch.uint8))
    for kid in (_abi.LPM_KERNEL_BVE_VEL_F64, _abi.LPM_KERNEL_BVE_VEL_MIXED):
        ctx.set_variant(kid, ctx.L.lpm_variant_default(kid))
    want = ctx.bve_velocity(*args).cpu()
    for kid in (_abi.LPM_KERNEL_BVE_VEL_F64, _abi.LPM_KERNEL_BVE_VEL_MIXED):
        ctx.set_variant(kid, -1)
    # keys are arch/family/size-bucket: poison every bucket of both families
    arch = open(path).readline().split()[0].rsplit("/", 2)[0]
    fams = {ln.split()[0].rsplit("/", 2)[1] for ln in open(path) if ln.strip()}
    bad = str(tmp_path / "bad.txt")
    with open(bad, "w") as f:
        for fam in fams:
            for b in range(63):
                f.write(f"{arch}/{fam}/{b} {-3 if b % 2 else 999}\n")
    ctx.tune_load(bad)
    got = ctx.bve_velocity(*args).cpu()
    ctx.tune_load(path)
    assert torch.equal(got, want)


def test_timing_records_launches(ctx):
    m, z = _inputs(0, 3)
    ctx.timing_enable(True)
    for _ in range(3):
        ctx.bve_velocity(dev(m.vert_x), dev(m.face_x), dev(z), dev(m.face_area),
                         dev(m.face_mask, torch.uint8))
    ms, n, pairs = ctx.timing_query(_abi.LPM_KERNEL_BVE_VEL_F64)
    ctx.timing_enable(False)
    assert n == 3 and ms > 0
    assert pairs == 3 * m.n_verts * m.n_faces
