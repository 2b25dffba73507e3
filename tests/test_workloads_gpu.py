"""GPU parity at the benchmark workloads' own sizes, through the C ABI.

* 64-member BVE ensemble (icosahedral levels 3-6, Rossby-Haurwitz 5-4 with
  seeded omega, K and pattern rotation) in one persistent launch: FP64
  against the oracle on sampled targets of one member per level (<= 1e-12
  relative L2 and Linf), mixed against FP64 (<= 1e-6 relative L2), every
  member bit-equal to its own one-problem run, and two graph-captured RK4
  steps equal to eager steps, with mixed against FP64 after them.
* 16 x level-6 shallow-water batch (steady zonal flow, seeded height noise):
  FP64 sums against the oracle on sampled targets of two members, mixed
  against FP64 (velocity, double dot, Laplacian <= 1e-6), and one
  graph-captured RK4 step with mixed against FP64.
* Cubed sphere level 8 Gaussian vortex with seeded noise: mixed against FP64
  after 1 and 50 RK4 steps (velocity <= 1e-6 relative L2).
* Icosahedral level 7 Rossby-Haurwitz 5-4: 20 RK4 steps, mixed against FP64
  (<= 1e-6 relative L2) and FP64 against the analytic pattern rotated at
  the Rossby-Haurwitz phase speed (a discretisation-error bound).
"""
import numpy as np
import pytest

from conftest import mesh, rel_l2, rel_linf

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from lpm_amd import _abi  # noqa: E402
from lpm_amd import galleries as G  # noqa: E402

TOL64 = 1e-12
TOL_MIXED = 1e-6
OMEGA = 2 * np.pi
SKIP = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
# the fastest shapes of the ensemble kernels (and their one-problem twins)
ENS_VARIANT = {"fp64": "persistent_block256_tpl4_nr1_quad_fused_sy",
               "mixed": "persistent_mfma16_occ5_block256_near3"}
KIDS = {"fp64": (_abi.LPM_KERNEL_BVE_VEL_F64, _abi.LPM_KERNEL_BVE_ENSEMBLE_F64),
        "mixed": (_abi.LPM_KERNEL_BVE_VEL_MIXED,
                  _abi.LPM_KERNEL_BVE_ENSEMBLE_MIXED)}


@pytest.fixture(scope="module")
def ctx():
    from lpm_amd.device import Context
    c = Context(0)
    yield c
    for k in range(8):
        c.set_variant(k, -1)
    c.close()


def _aos(t):
    return np.ascontiguousarray(t.cpu().numpy().T)


def _same(a, b):
    return np.array_equal(a.cpu().numpy(), b.cpu().numpy(), equal_nan=True)


def _force(ctx, precision):
    k1, ke = KIDS[precision]
    name = ENS_VARIANT[precision]
    ctx.set_variant(ke, ctx.variants(ke).index(name))
    ctx.set_variant(k1, ctx.variants(k1).index(name[len("persistent_"):]))


# ------------------------------------------------------------ BVE batch --
def _c4(ctx):
    from lpm_amd.bve import BVEEnsemble
    params = [G.ensemble_member(m) for m in range(64)]
    ens = BVEEnsemble([mesh(0, p["level"]) for p in params], ctx=ctx)
    ens.init_vorticity([
        (lambda x, p=p: G.rotated_rh54_vorticity(x, p["omega"], p["K"],
                                                 p["rotation"]))
        for p in params])
    return ens, params


def test_c4_batch_in_one_launch(ctx, ora):
    from lpm_amd.bve import BVESphere
    ens, params = _c4(ctx)
    levels = sorted({p["level"] for p in params})
    assert levels == [3, 4, 5, 6]
    out = {}
    for prec in ("fp64", "mixed"):
        _force(ctx, prec)
        ens.init_velocity(prec, SKIP)
        torch.cuda.synchronize()
        out[prec] = (ens.velocity_verts.clone(), ens.velocity_faces.clone())
    zf = ens.rel_vort_faces.cpu().numpy()
    # FP64 against the oracle: sampled vertices and leaf faces of the first
    # member of each level
    for L in levels:
        m = next(k for k, p in enumerate(params) if p["level"] == L)
        msh = ens.meshes[m]
        z = zf[ens.face_slice(m)]
        vs = slice(0, msh.n_verts, max(1, msh.n_verts // 256))
        leaf = np.flatnonzero(msh.face_mask == 0)[::max(1, msh.n_leaves // 256)]
        ref_v = ora.bve_velocity(np.ascontiguousarray(msh.vert_x[:, vs].T),
                                 np.ascontiguousarray(msh.face_x.T), z,
                                 msh.face_area, msh.face_mask)
        got_v = out["fp64"][0][:, ens.vert_slice(m)][:, vs].cpu().numpy().T
        assert rel_l2(got_v, ref_v) < TOL64, L
        assert rel_linf(got_v, ref_v) < TOL64, L
        # collocated face targets: only the sampled leaves are evaluated
        fx = np.ascontiguousarray(msh.face_x.T)
        skip_t = np.ones(msh.n_faces, dtype=np.uint8)
        skip_t[leaf] = 0
        ref_f = ora.bve_velocity(fx, fx, z, msh.face_area, msh.face_mask,
                                 collocated=True, tgt_mask=skip_t)[leaf]
        got_f = out["fp64"][1][:, ens.face_slice(m)].cpu().numpy().T[leaf]
        assert rel_l2(got_f, ref_f) < TOL64, L
        assert rel_linf(got_f, ref_f) < TOL64, L
    # mixed against FP64 over every member's vertices and leaf faces
    lf = torch.tensor(np.concatenate([m.face_mask == 0 for m in ens.meshes]),
                      device="cuda")
    u64v, u64f = (t.cpu().numpy() for t in out["fp64"])
    umxv, umxf = (t.cpu().numpy() for t in out["mixed"])
    lfn = lf.cpu().numpy()
    assert rel_l2(umxv, u64v) < TOL_MIXED
    assert rel_l2(umxf[:, lfn], u64f[:, lfn]) < TOL_MIXED
    # every member bit-equal to its one-problem run (same kernel shape)
    for prec in ("fp64", "mixed"):
        _force(ctx, prec)
        for m in range(ens.n_members):
            s = BVESphere(None, None, ctx=ctx, mesh=ens.meshes[m])
            s.rel_vort_verts.copy_(ens.rel_vort_verts[ens.vert_slice(m)])
            s.rel_vort_faces.copy_(ens.rel_vort_faces[ens.face_slice(m)])
            s.init_velocity(prec, SKIP)
            torch.cuda.synchronize()
            assert _same(out[prec][0][:, ens.vert_slice(m)],
                         s.velocity_verts), (prec, m)
            leaf_m = torch.tensor(ens.meshes[m].face_mask == 0, device="cuda")
            assert _same(out[prec][1][:, ens.face_slice(m)][:, leaf_m],
                         s.velocity_faces[:, leaf_m]), (prec, m)


def test_c4_graph_captured_rk4(ctx, ora):
    from lpm_amd.bve import BVEEnsembleRK4
    stream = torch.cuda.Stream()
    res = {}
    for prec in ("fp64", "mixed"):
        _force(ctx, prec)
        runs = []
        for graph in (False, True):
            ens, params = _c4(ctx)
            ens.init_velocity(prec, SKIP)
            st = BVEEnsembleRK4(0.01, ens, OMEGA, prec, reference_compat=True,
                                renormalize=True, leaf_only=True,
                                use_graph=graph, stream=stream.cuda_stream)
            for _ in range(2):
                st.advance_timestep(ens)
            stream.synchronize()
            runs.append(ens)
        a, b = runs
        assert _same(a.vert_x, b.vert_x), prec
        assert _same(a.velocity_verts, b.velocity_verts), prec
        res[prec] = b
    # mixed against FP64 after two steps (velocity at the new positions)
    assert rel_l2(res["mixed"].velocity_verts.cpu().numpy(),
                  res["fp64"].velocity_verts.cpu().numpy()) < TOL_MIXED
    # the smallest member against two oracle steps from the same state
    ens, params = _c4(ctx)
    _force(ctx, "fp64")
    ens.init_velocity("fp64", SKIP)
    torch.cuda.synchronize()
    m = next(k for k, p in enumerate(params) if p["level"] == 3)
    msh = ens.meshes[m]
    vs, fs = ens.vert_slice(m), ens.face_slice(m)
    st_o = {"vx": _aos(ens.vert_x[:, vs]),
            "vzeta": ens.rel_vort_verts[vs].cpu().numpy().copy(),
            "vu": _aos(ens.velocity_verts[:, vs]),
            "fx": _aos(ens.face_x[:, fs]),
            "fzeta": ens.rel_vort_faces[fs].cpu().numpy().copy(),
            "fu": _aos(ens.velocity_faces[:, fs]),
            "farea": msh.face_area.copy(), "fmask": msh.face_mask.copy()}
    # divided faces are neither evaluated nor advanced on the GPU (leaf
    # only); the oracle's values there do not reach the leaves
    st_o["fu"][msh.face_mask != 0] = 0.0
    for _ in range(2):
        ora.bve_rk4_step(st_o, 0.01, OMEGA, compat=True, renorm=True)
    g = res["fp64"]
    assert rel_l2(_aos(g.vert_x[:, vs]), st_o["vx"]) < TOL64
    assert rel_l2(_aos(g.velocity_verts[:, vs]), st_o["vu"]) < TOL64
    leaf = msh.face_mask == 0
    assert rel_l2(_aos(g.velocity_faces[:, fs])[leaf], st_o["fu"][leaf]) < TOL64


# ------------------------------------------------------------ SWE batch --
def _c5(ctx, members=16, level=6):
    from lpm_amd.swe import SWEEnsemble
    msh = mesh(0, level)
    h = msh.mean_spacing()
    st_v, st_f = G.tc2_state(msh.vert_x), G.tc2_state(msh.face_x)
    noise = []
    for b in range(members):
        rng = G.SplitMix64(0x4C504D35 + b)
        noise.append((1e-4 * (2 * rng.uniform_array(msh.n_verts) - 1),
                      1e-4 * (2 * rng.uniform_array(msh.n_faces) - 1)))
    e = SWEEnsemble([msh] * members, ctx=ctx)
    sv = np.concatenate([st_v["surface"] + nv for nv, _ in noise])
    sf = np.concatenate([st_f["surface"] + nf for _, nf in noise])
    area = np.tile(msh.face_area, members)
    for t, v in ((e.surf_passive, sv), (e.depth_passive, sv),
                 (e.surf_active, sf), (e.depth_active, sf),
                 (e.mass_active, sf * area),
                 (e.rel_vort_passive, np.tile(st_v["zeta"], members)),
                 (e.rel_vort_active, np.tile(st_f["zeta"], members))):
        t.copy_(torch.from_numpy(v))
    e.set_kernel_parameters(h, 2 * h)
    return e, msh, noise, st_v, st_f


def test_c5_swe_batch(ctx, ora):
    from lpm_amd.swe import SWEEnsembleRK4
    out = {}
    for prec in ("fp64", "mixed"):
        e, msh, noise, st_v, st_f = _c5(ctx)
        e.init_direct_sums(prec, SKIP)
        torch.cuda.synchronize()
        out[prec] = {k: getattr(e, k).cpu().numpy().copy() for k in (
            "velocity_passive", "double_dot_passive", "surf_lap_passive")}
    h = msh.mean_spacing()
    # FP64 against the oracle on sampled vertices of members 0 and 15
    sl = slice(0, msh.n_verts, 128)
    for b in (0, 15):
        sv = st_v["surface"] + noise[b][0]
        sf = st_f["surface"] + noise[b][1]
        ru, rdd, rlap = ora.swe_sums(
            np.ascontiguousarray(msh.vert_x[:, sl].T), sv[sl],
            np.ascontiguousarray(msh.face_x.T), st_f["zeta"], st_f["sigma"],
            msh.face_area, sf, msh.face_mask, eps=h, pse_eps=2 * h)
        off = b * msh.n_verts
        o = out["fp64"]
        u = o["velocity_passive"][:, off:off + msh.n_verts][:, sl].T
        assert rel_l2(u, ru) < TOL64, b
        assert rel_linf(u, ru) < TOL64, b
        assert rel_l2(o["double_dot_passive"][off:off + msh.n_verts][sl],
                      rdd) < TOL64, b
        assert rel_l2(o["surf_lap_passive"][off:off + msh.n_verts][sl],
                      rlap) < TOL64, b
    for k in out["fp64"]:
        assert rel_l2(out["mixed"][k], out["fp64"][k]) < TOL_MIXED, k
    # one graph-captured RK4 step of each precision
    stream = torch.cuda.Stream()
    after = {}
    for prec in ("fp64", "mixed"):
        e, *_ = _c5(ctx)
        e.init_direct_sums(prec, SKIP)
        r = SWEEnsembleRK4(0.005, e, prec, reference_compat=True,
                           leaf_only=True, use_graph=True,
                           stream=stream.cuda_stream, renormalize=True)
        e.advance_timestep(r)
        stream.synchronize()
        after[prec] = {k: getattr(e, k).cpu().numpy().copy() for k in (
            "velocity_passive", "surf_passive", "vert_x")}
    for k in after["fp64"]:
        assert rel_l2(after["mixed"][k], after["fp64"][k]) < TOL_MIXED, k


# --------------------------------------------------- precision workload --
def test_c3_cubed_sphere_level8_mixed_vs_fp64(ctx, ora):
    from lpm_amd.bve import BVERK4, BVESphere
    msh = mesh(1, 8)
    ctr = np.array([np.sqrt(0.5), 0.0, np.sqrt(0.5)])
    rng = G.SplitMix64(0x4C504D33)
    noise_f = 1e-3 * (2 * rng.uniform_array(msh.n_faces) - 1)
    noise_v = 1e-3 * (2 * rng.uniform_array(msh.n_verts) - 1)
    zf = G.gaussian_vortex_vorticity(msh.face_x, ctr, 4 * np.pi, 4.0) + noise_f
    zv = G.gaussian_vortex_vorticity(msh.vert_x, ctr, 4 * np.pi, 4.0) + noise_v
    snaps = {}
    for prec in ("fp64", "mixed"):
        s = BVESphere(None, None, ctx=ctx, mesh=msh)
        s.rel_vort_faces.copy_(torch.from_numpy(zf))
        s.rel_vort_verts.copy_(torch.from_numpy(zv))
        s.init_velocity(prec, SKIP)
        torch.cuda.synchronize()
        if prec == "fp64":
            sl = slice(0, msh.n_verts, 1024)
            ref = ora.bve_velocity(np.ascontiguousarray(msh.vert_x[:, sl].T),
                                   np.ascontiguousarray(msh.face_x.T), zf,
                                   msh.face_area, msh.face_mask)
            got = s.velocity_verts.cpu().numpy()[:, sl].T
            assert rel_l2(got, ref) < TOL64
            assert rel_linf(got, ref) < TOL64
        solver = BVERK4(0.005, s, precision=prec, reference_compat=True,
                        renormalize=True, leaf_only=True)
        for n in range(1, 51):
            solver.advance_timestep(s)
            if n in (1, 50):
                torch.cuda.synchronize()
                snaps[(prec, n)] = (s.velocity_verts.cpu().numpy(),
                                    s.vert_x.cpu().numpy())
    for n in (1, 50):
        um, xm = snaps[("mixed", n)]
        uf, xf = snaps[("fp64", n)]
        assert rel_l2(um, uf) < TOL_MIXED, n
        assert rel_linf(um, uf) < 10 * TOL_MIXED, n
        assert rel_l2(xm, xf) < TOL_MIXED, n


# ------------------------------------------------------- headline run --
def test_c2_icosahedral_level7_twenty_steps(ctx):
    from lpm_amd.bve import BVERK4, BVESphere
    msh = mesh(0, 7)
    dt, steps = 0.005, 20
    res = {}
    for prec, renorm in (("fp64", False), ("fp64", True), ("mixed", True)):
        s = BVESphere(None, None, ctx=ctx, mesh=msh)
        s.init_vorticity(lambda x: G.rh54_vorticity(x, 0.1, -0.1))
        s.init_velocity(prec, SKIP)
        solver = BVERK4(dt, s, precision=prec, reference_compat=True,
                        renormalize=renorm, leaf_only=True)
        for _ in range(steps):
            solver.advance_timestep(s)
        torch.cuda.synchronize()
        res[(prec, renorm)] = (s.velocity_verts.cpu().numpy(),
                               s.vert_x.cpu().numpy())
    um, xm = res[("mixed", True)]
    uf, xf = res[("fp64", True)]
    assert rel_l2(um, uf) < TOL_MIXED
    assert rel_l2(xm, xf) < TOL_MIXED
    # against the reference layout's FP64 run (no renormalisation): the
    # positions differ by the off-sphere drift of the FP64 run as well
    u0, x0 = res[("fp64", False)]
    assert rel_l2(um, u0) < 2 * TOL_MIXED
    # the RH 5-4 pattern rotates rigidly at the phase speed nu: the FP64
    # velocity after the run against the rotated analytic velocity is
    # bounded by the quadrature error of the level-7 mesh
    t = steps * dt
    nu = G.rh54_wave_speed(0.1, OMEGA)
    ua = G.rotate_z(G.rh54_velocity(G.rotate_z(x0, -nu * t), 0.1, -0.1),
                    nu * t)
    assert rel_l2(u0, ua) < 5e-4
