#!/usr/bin/env python3
"""Benchmark of the BVE direct sum + RK4 on one MI355X.

Workload (the plain-FP64-at-scale configuration): barotropic vorticity on the
unit sphere, icosahedral mesh refined 7 times (163,842 vertices + 327,680
leaf faces as targets, 327,680 leaf faces as sources; 1.6106e11 pair
interactions per velocity evaluation), Rossby-Haurwitz 5-4 vorticity
(omega = K = 0.1), Omega = 2 pi, RK4 with dt = 0.005.  A step is one RK4
step (4 direct-sum evaluations over all targets plus the O(N) stage work).

Reported: pair-interactions/s (useful pairs (N_V + N_F) N_F per evaluation)
and RK4 steps/s for the FP64 path (the headline value) and for the mixed
path, relative L2/Linf velocity errors against the FP64 path, against the
CPU oracle on a sample of targets and against the analytic RH velocity, the
roofline of the dominant kernel (HIP events around its launches) and the
CPU oracle's rate on a bounded sample of the same evaluation.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
Prints exactly one JSON line on stdout; progress goes to stderr.

``--copy-calibration MIB`` instead runs a plain device copy (read MIB, write
MIB) a few times: the HBM byte counters (FETCH_SIZE, WRITE_SIZE) of that run
scale the counters of the direct-sum kernel to bytes.  The measured traffic
per launch of the FP64 kernel is read from
profiles/bve_vel_f64_icos_l7_traffic.md when that summary is present.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Gpair-interactions/s + RK4 steps/s, mixed vs FP64 at stated error, "
          "1 GPU")
F_BVE = 22  # flops per pair (canonical count)
F_SWE = 80  # flops per pair of the shallow-water sums (canonical count)
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector, AMD public figure


# Comparisons with the CPU oracle are deferred to the CPU-baseline leg: the
# GPU legs register closures over their own outputs here, and that leg (the
# only one that loads the oracle) runs them with the oracle module.
CPU_CHECKS = []


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr,
          flush=True)


def errors(a, b):
    """Relative L2 and Linf of vector fields (3, n) a against b."""
    import numpy as np
    e = np.linalg.norm(a - b, axis=0)
    r = np.linalg.norm(b, axis=0)
    return float(np.sqrt((e ** 2).sum() / (r ** 2).sum())), float(e.max() / r.max())


def errors1(a, b):
    """Relative L2 and Linf of scalar fields."""
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    return (float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())),
            float(np.abs(a - b).max() / np.abs(b).max()))


def run_precision(ctx, mesh, precision, steps, warmup, dt, flags_extra,
                  autotune=True):
    import numpy as np
    import torch
    from lpm_amd import _abi
    from lpm_amd import galleries as G
    from lpm_amd.bve import BVERK4, BVESphere

    s = BVESphere("icos", mesh.depth, ctx=ctx, mesh=mesh)
    s.init_vorticity(lambda x: G.rh54_vorticity(x, 0.1, -0.1))
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    kid = (_abi.LPM_KERNEL_BVE_VEL_F64 if precision == "fp64"
           else _abi.LPM_KERNEL_BVE_VEL_MIXED)
    tune = None
    if autotune:
        names = ctx.variants(kid)
        best, best_ms, all_ms = ctx.bve_autotune(s.fields(), precision,
                                                 flags=skip, repeats=1)
        d = ctx.L.lpm_variant_default(kid)
        tune = {"variant": names[best], "ms": round(best_ms, 3),
                "fixed_tile_baseline": {"variant": names[d],
                                        "ms": round(all_ms[d], 3)},
                "all_ms": {n: round(t, 3) for n, t in zip(names, all_ms)}}
        log(f"autotune {precision}: {names[best]} {best_ms:.2f} ms "
            f"(fixed {names[d]} {all_ms[d]:.2f} ms)")
    s.init_velocity(precision, flags=skip)
    torch.cuda.synchronize()
    u0_v = s.velocity_verts.cpu().numpy()
    u0_f = s.velocity_faces.cpu().numpy()[:, mesh.leaf_slice]
    solver = BVERK4(dt, s, precision=precision, leaf_only=True, **flags_extra)
    for _ in range(warmup):
        solver.advance_timestep(s)
    torch.cuda.synchronize()
    ctx.timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.advance_timestep(s)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kms, launches, _ = ctx.timing_query(kid)
    ctx.timing_enable(False)
    elapsed = t1 - t0
    out = {"elapsed_s": elapsed, "kernel_ms_total": kms, "launches": launches,
           "autotune": tune,
           "u0_v": u0_v, "u0_f": u0_f,
           "uT_v": s.velocity_verts.cpu().numpy(),
           "xT_v": s.vert_x.cpu().numpy()}
    del s, solver
    return out


def ensemble_run(ctx, steps, warmup, dt=0.01, autotune=True):
    """64-member ensemble (icosahedral levels 3-6, RH 5-4 with random
    omega, K and pattern rotation), one persistent launch per evaluation,
    graph-captured RK4; mixed precision timed, FP64 for the error."""
    import numpy as np
    import torch
    from lpm_amd import _abi
    from lpm_amd import galleries as G
    from lpm_amd.bve import BVEEnsemble, BVEEnsembleRK4
    from lpm_amd.mesh import SphereMesh

    params = [G.ensemble_member(m) for m in range(64)]
    cache = {}
    meshes = []
    for p in params:
        if p["level"] not in cache:
            cache[p["level"]] = SphereMesh.build("icos", p["level"])
        meshes.append(cache[p["level"]])
    fns = [(lambda x, p=p: G.rotated_rh54_vorticity(x, p["omega"], p["K"],
                                                    p["rotation"]))
           for p in params]
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    out = {"members": 64,
           "levels": {str(L): sum(1 for p in params if p["level"] == L)
                      for L in sorted(cache)}}
    pairs = None
    res = {}
    stream = torch.cuda.Stream()
    for prec in ("mixed", "fp64"):
        ens = BVEEnsemble(meshes, ctx=ctx)
        ens.init_vorticity(fns)
        pairs = ens.useful_pairs()
        kid = (_abi.LPM_KERNEL_BVE_ENSEMBLE_MIXED if prec == "mixed"
               else _abi.LPM_KERNEL_BVE_ENSEMBLE_F64)
        names = ctx.variants(kid)
        tune = None
        if autotune:
            best, best_ms, all_ms = ens.autotune(prec, skip, repeats=2)
            d = ctx.L.lpm_variant_default(kid)
            tune = {"variant": names[best], "ms": round(best_ms, 3),
                    "fixed_tile_baseline": {"variant": names[d],
                                            "ms": round(all_ms[d], 3)},
                    "all_ms": {n: round(t, 3) for n, t in zip(names, all_ms)}}
            log(f"ensemble autotune {prec}: {names[best]} {best_ms:.2f} ms "
                f"(fixed {names[d]} {all_ms[d]:.2f} ms)")
        ens.init_velocity(prec, skip)
        torch.cuda.synchronize()
        u0 = ens.velocity_verts.cpu().numpy()
        st = BVEEnsembleRK4(dt, ens, 2 * np.pi, prec, reference_compat=True,
                            renormalize=True, leaf_only=True, use_graph=True,
                            stream=stream.cuda_stream)
        for _ in range(warmup):
            st.advance_timestep(ens)
        stream.synchronize()
        ctx.timing_enable(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            st.advance_timestep(ens)
        stream.synchronize()
        el = time.perf_counter() - t0
        ctx.timing_enable(False)
        res[prec] = {"elapsed": el, "u0": u0, "tune": tune,
                     "uT": ens.velocity_verts.cpu().numpy(),
                     "xT": ens.vert_x.cpu().numpy()}
        log(f"ensemble {prec}: {el / steps * 1e3:.1f} ms/step, "
            f"{4 * pairs * steps / el / 1e9:.1f} Gpairs/s")
        if prec == "mixed":
            # analytic RH velocity of each member's rotated pattern at t = 0
            ue = []
            for m, p in enumerate(params):
                xm = meshes[m].vert_x
                ue.append(p["rotation"] @ G.rh54_velocity(
                    p["rotation"].T @ xm, p["omega"], -p["K"]))
            out["mixed_vs_analytic_rh54_t0_vertices_rel_l2_linf"] = errors(
                u0, np.concatenate(ue, axis=1))
        del ens, st
    m, f = res["mixed"], res["fp64"]
    out.update({
        "pairs_per_evaluation": pairs,
        "steps": steps, "warmup": warmup, "dt": dt,
        "value": round(4 * pairs * steps / m["elapsed"] / 1e9, 3),
        "unit": "Gpair-interactions/s",
        "ms_per_step": round(m["elapsed"] / steps * 1e3, 3),
        "rk4_steps_per_s": round(steps / m["elapsed"], 4),
        "dtype": "f32 pair arithmetic, f64 carries/state",
        "graph": True,
        "fp64_value": round(4 * pairs * steps / f["elapsed"] / 1e9, 3),
        "fp64_ms_per_step": round(f["elapsed"] / steps * 1e3, 3),
        "mixed_vs_fp64_t0_vertices_rel_l2_linf": errors(m["u0"], f["u0"]),
        "mixed_vs_fp64_after_steps_vertices_rel_l2_linf": errors(m["uT"], f["uT"]),
        "mixed_vs_fp64_after_steps_positions_rel_l2_linf": errors(m["xT"], f["xT"]),
        "autotune_mixed": m["tune"], "autotune_fp64": f["tune"],
    })
    return out


def precision_study(ctx, steps_list=(1, 50), dt=0.005, level=8,
                    oracle_stride=256):
    """Gaussian vortex on the cubed sphere, level 8 (393,216 faces): mixed
    against FP64 velocity and positions after 1 and 50 RK4 steps."""
    import numpy as np
    import torch
    from lpm_amd import galleries as G
    from lpm_amd.bve import BVERK4, BVESphere
    from lpm_amd.mesh import SphereMesh

    mesh = SphereMesh.build("cube", level)
    ctr = np.array([np.sqrt(0.5), 0.0, np.sqrt(0.5)])
    rng = G.SplitMix64(0x4C504D33)
    noise_f = 1e-3 * (2 * rng.uniform_array(mesh.n_faces) - 1)
    noise_v = 1e-3 * (2 * rng.uniform_array(mesh.n_verts) - 1)

    def init(s):
        zf = G.gaussian_vortex_vorticity(mesh.face_x, ctr, 4 * np.pi, 4.0)
        zv = G.gaussian_vortex_vorticity(mesh.vert_x, ctr, 4 * np.pi, 4.0)
        s.rel_vort_faces.copy_(torch.from_numpy(zf + noise_f))
        s.rel_vort_verts.copy_(torch.from_numpy(zv + noise_v))

    snaps = {}
    times = {}
    for prec in ("fp64", "mixed"):
        s = BVESphere("cube", level, ctx=ctx, mesh=mesh)
        init(s)
        s.init_velocity(prec, flags=2)
        if prec == "fp64":
            torch.cuda.synchronize()
            u_t0 = s.velocity_verts.cpu().numpy()
        solver = BVERK4(dt, s, precision=prec, reference_compat=True,
                        renormalize=True, leaf_only=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = 0
        for n in sorted(steps_list):
            while done < n:
                solver.advance_timestep(s)
                done += 1
            torch.cuda.synchronize()
            snaps[(prec, n)] = (s.velocity_verts.cpu().numpy(),
                                s.vert_x.cpu().numpy())
        times[prec] = (time.perf_counter() - t0) / done
        log(f"precision study {prec}: {times[prec] * 1e3:.1f} ms/step")
        del s, solver
    out = {"workload": (f"cubed sphere level {level} ({mesh.n_verts} vertices, "
                        f"{mesh.n_leaves} faces), Gaussian vortex 4 pi exp(-16 "
                        f"(1 - x.x0)) at 45N 0E + U(-1e-3, 1e-3) noise "
                        f"(splitmix64 0x4C504D33), Omega = 2 pi, dt = {dt}, "
                        f"stage positions renormalised"),
           "ms_per_step": {k: round(v * 1e3, 3) for k, v in times.items()}}
    for n in sorted(steps_list):
        uf, xf = snaps[("fp64", n)]
        um, xm = snaps[("mixed", n)]
        out[f"after_{n}_steps"] = {
            "velocity_rel_l2_linf": errors(um, uf),
            "positions_rel_l2_linf": errors(xm, xf)}
    # FP64 initial velocity against the CPU oracle on a sample of vertices
    sl = slice(0, mesh.n_verts, oracle_stride)
    zf = G.gaussian_vortex_vorticity(mesh.face_x, ctr, 4 * np.pi, 4.0) + noise_f

    def check(oracle):
        ru = oracle.bve_velocity(np.ascontiguousarray(mesh.vert_x[:, sl].T),
                                 np.ascontiguousarray(mesh.face_x.T), zf,
                                 mesh.face_area, mesh.face_mask,
                                 nthreads=min(16, os.cpu_count() or 1))
        out["fp64_vs_cpu_oracle_sample_t0"] = {
            "targets": f"every {oracle_stride}th vertex",
            "velocity_rel_l2_linf": errors(u_t0[:, sl], ru.T)}
    CPU_CHECKS.append(check)
    return out


def swe_variant_sweep(ctx, e, prec, flags, reps=1):
    """Times every kernel variant of the SWE ensemble sums on the current
    state (one launch each after a warm-up) and keeps the fastest."""
    import torch
    from lpm_amd import _abi
    kid = (_abi.LPM_KERNEL_SWE_F64 if prec == "fp64"
           else _abi.LPM_KERNEL_SWE_MIXED)
    names = ctx.variants(kid)
    ms = {}
    for v, name in enumerate(names):
        ctx.set_variant(kid, v)
        e.init_direct_sums(prec, flags)
        torch.cuda.synchronize()
        ctx.timing_enable(True)
        for _ in range(reps):
            e.init_direct_sums(prec, flags)
        torch.cuda.synchronize()
        t, n, _ = ctx.timing_query(kid)
        ctx.timing_enable(False)
        ms[name] = t / max(n, 1)
    best = min(ms, key=ms.get)
    ctx.set_variant(kid, names.index(best))
    d = names[ctx.L.lpm_variant_default(kid)]
    log(f"swe variants {prec}: " + ", ".join(f"{k} {v:.1f} ms"
                                            for k, v in ms.items()))
    return {"variant": best, "ms": round(ms[best], 3),
            "fixed_tile_baseline": {"variant": d, "ms": round(ms[d], 3)},
            "all_ms": {k: round(v, 3) for k, v in ms.items()}}


def swe_ensemble_run(ctx, steps=10, warmup=1, dt=0.005, level=6, members=16,
                     oracle_stride=64, tune=True):
    """Shallow water: 16 steady zonal (TC2) states on the icosahedral mesh of
    level 6 with seeded 1e-4 height noise, one persistent launch per
    evaluation, graph-captured RK4; mixed timed, FP64 for the error."""
    import numpy as np
    import torch
    from lpm_amd import _abi
    from lpm_amd import galleries as G
    from lpm_amd.mesh import SphereMesh
    from lpm_amd.swe import SWEEnsemble, SWEEnsembleRK4

    mesh = SphereMesh.build("icos", level)
    h = mesh.mean_spacing()
    eps, pse_eps = h, 2 * h
    noise = []
    for b in range(members):
        rng = G.SplitMix64(0x4C504D35 + b)
        noise.append((1e-4 * (2 * rng.uniform_array(mesh.n_verts) - 1),
                      1e-4 * (2 * rng.uniform_array(mesh.n_faces) - 1)))
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    st_v, st_f = G.tc2_state(mesh.vert_x), G.tc2_state(mesh.face_x)

    def build():
        e = SWEEnsemble([mesh] * members, ctx=ctx)
        sv = np.concatenate([st_v["surface"] + nv for nv, _ in noise])
        sf = np.concatenate([st_f["surface"] + nf for _, nf in noise])
        area = np.tile(mesh.face_area, members)
        for t, v in ((e.surf_passive, sv), (e.depth_passive, sv),
                     (e.surf_active, sf), (e.depth_active, sf),
                     (e.mass_active, sf * area),
                     (e.rel_vort_passive, np.tile(st_v["zeta"], members)),
                     (e.rel_vort_active, np.tile(st_f["zeta"], members))):
            t.copy_(torch.from_numpy(v))
        e.set_kernel_parameters(eps, pse_eps)
        return e

    stream = torch.cuda.Stream()
    res = {}
    pairs = None
    tuned = {}
    for prec in ("mixed", "fp64"):
        e = build()
        pairs = e.useful_pairs()
        if tune:
            tuned[prec] = swe_variant_sweep(ctx, e, prec, skip)
        e.init_direct_sums(prec, skip)
        torch.cuda.synchronize()
        t0 = {k: getattr(e, k).cpu().numpy().copy() for k in (
            "velocity_passive", "double_dot_passive", "surf_lap_passive",
            "surf_passive", "rel_vort_active", "div_active")}
        r = SWEEnsembleRK4(dt, e, prec, reference_compat=True, leaf_only=True,
                           use_graph=True, stream=stream.cuda_stream,
                           renormalize=True)
        for _ in range(warmup):
            e.advance_timestep(r)
        stream.synchronize()
        tb = time.perf_counter()
        for _ in range(steps):
            e.advance_timestep(r)
        stream.synchronize()
        el = time.perf_counter() - tb
        tT = {k: getattr(e, k).cpu().numpy().copy() for k in (
            "velocity_passive", "surf_passive", "vert_x", "div_passive")}
        # kernel time of one evaluation (for the roofline of the sums)
        ctx.timing_enable(True)
        e.init_direct_sums(prec, skip)
        torch.cuda.synchronize()
        kms, kn, _ = ctx.timing_query(_abi.LPM_KERNEL_SWE_F64 if prec == "fp64"
                                      else _abi.LPM_KERNEL_SWE_MIXED)
        ctx.timing_enable(False)
        res[prec] = {"elapsed": el, "t0": t0, "tT": tT,
                     "kernel_ms": kms / max(kn, 1)}
        log(f"swe ensemble {prec}: {el / steps * 1e3:.1f} ms/step, "
            f"{4 * pairs * steps / el / 1e9:.1f} Gpairs/s")
        del e, r
    m, f = res["mixed"], res["fp64"]
    uex = np.tile(st_v["u"], (1, members))
    out = {
        "workload": (f"{members} x shallow water on the icosahedral mesh level "
                     f"{level} ({mesh.n_verts} vertices, {mesh.n_leaves} faces "
                     f"each), steady zonal flow u0 = 0.2, g = 1, h0 = 1 (polar "
                     f"depth), Omega = 2 pi, height noise U(-1e-4, 1e-4) "
                     f"(splitmix64 0x4C504D35 + b), eps = {eps:.5f}, PSE "
                     f"width {pse_eps:.5f}, dt = {dt}, stage positions "
                     f"renormalised in both precisions"),
        "pairs_per_evaluation": pairs,
        "steps": steps, "warmup": warmup,
        "value": round(4 * pairs * steps / m["elapsed"] / 1e9, 3),
        "unit": "Gpair-interactions/s",
        "ms_per_step": round(m["elapsed"] / steps * 1e3, 3),
        "rk4_steps_per_s": round(steps / m["elapsed"], 4),
        "dtype": ("far tiles and 64-source sub-tiles: FP32 distances and "
                  "f16 hi/lo weight x feature products on the matrix cores "
                  "(FP32 accumulation, FP64 carry); near sub-tiles and PSE "
                  "terms f64; f64 state"),
        "graph": True,
        "fp64_value": round(4 * pairs * steps / f["elapsed"] / 1e9, 3),
        "fp64_ms_per_step": round(f["elapsed"] / steps * 1e3, 3),
        "kernel_ms_per_evaluation": {"mixed": round(m["kernel_ms"], 3),
                                     "fp64": round(f["kernel_ms"], 3)},
        "roofline_fp64_sums": {
            "flops_per_pair": F_SWE,
            "achieved": round(pairs * F_SWE / (f["kernel_ms"] * 1e-3) / 1e12, 3),
            "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(pairs * F_SWE / (f["kernel_ms"] * 1e-3) / 1e12
                          / PEAK_FP64_TFLOPS, 4)},
        "mixed_speedup_vs_fp64_sums": round(f["kernel_ms"] / m["kernel_ms"], 4),
        "variants": tuned or None,
        "counters_fp64_default_shape": valu_counters("swe_ens_f64",
                                                     "share4_occ3", "6x16"),
        "counters_mixed_default_shape": {
            "far_pass": valu_counters("swe_ens_mf_far", "mfma_f2n3", "6x16"),
            "near_pass": valu_counters("swe_ens_mf_near", "mfma_f2n3",
                                       "6x16")},
        "mixed_vs_fp64_t0": {
            "velocity_rel_l2_linf": errors(m["t0"]["velocity_passive"],
                                           f["t0"]["velocity_passive"]),
            "ddot_rel_l2_linf": errors1(m["t0"]["double_dot_passive"],
                                        f["t0"]["double_dot_passive"]),
            "laps_rel_l2_linf": errors1(m["t0"]["surf_lap_passive"],
                                        f["t0"]["surf_lap_passive"])},
        "mixed_vs_fp64_after_steps": {
            "velocity_rel_l2_linf": errors(m["tT"]["velocity_passive"],
                                           f["tT"]["velocity_passive"]),
            "surface_rel_l2_linf": errors1(m["tT"]["surf_passive"],
                                           f["tT"]["surf_passive"])},
        "vs_steady_state_t0": {
            "velocity_rel_l2_linf": errors(m["t0"]["velocity_passive"], uex),
            "ddot_rel_l2_linf": errors1(m["t0"]["double_dot_passive"],
                                        np.tile(st_v["ddot"], members)),
            "laps_rel_l2_linf": errors1(m["t0"]["surf_lap_passive"],
                                        np.tile(st_v["lap_surface"], members))},
        "vs_steady_state_after_steps": {
            "surface_rel_l2_linf": errors1(
                m["tT"]["surf_passive"], G.tc2_state(m["tT"]["vert_x"])["surface"]),
            "max_abs_divergence": float(np.abs(m["tT"]["div_passive"]).max())},
    }
    # FP64 GPU against the CPU oracle on a sample of member 0's vertices
    sl = slice(0, mesh.n_verts, oracle_stride)
    sv0 = st_v["surface"] + noise[0][0]
    sf0 = st_f["surface"] + noise[0][1]
    f0 = f["t0"]

    def check(oracle):
        ru, rdd, rlap = oracle.swe_sums(
            np.ascontiguousarray(mesh.vert_x[:, sl].T), sv0[sl],
            np.ascontiguousarray(mesh.face_x.T), st_f["zeta"], st_f["sigma"],
            mesh.face_area, sf0, mesh.face_mask, eps=eps, pse_eps=pse_eps,
            nthreads=min(16, os.cpu_count() or 1))
        out["fp64_vs_cpu_oracle_sample"] = {
            "targets": f"every {oracle_stride}th vertex of member 0",
            "velocity_rel_l2_linf": errors(f0["velocity_passive"][:, sl], ru.T),
            "ddot_rel_l2_linf": errors1(f0["double_dot_passive"][sl], rdd),
            "laps_rel_l2_linf": errors1(f0["surf_lap_passive"][sl], rlap)}
    CPU_CHECKS.append(check)
    return out


def ic2d_run(ctx, steps=10, warmup=1, dt=0.005, level=6, oracle_stride=64):
    """Incompressible2D on the sphere: fused FP64 velocity + stream function
    sums, Heun RK2 (2 evaluations per step), RH 5-4, icosahedral level 6."""
    import numpy as np
    import torch
    from lpm_amd import _abi
    from lpm_amd import galleries as G
    from lpm_amd.incompressible2d import Incompressible2D, Incompressible2DRK2
    from lpm_amd.mesh import SphereMesh

    mesh = SphereMesh.build("icos", level)
    s = Incompressible2D("icos", level, ctx=ctx, mesh=mesh)
    s.init_vorticity(lambda x: G.rh54_vorticity(x, 0.1, -0.1))
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    s.init_direct_sums(skip)
    torch.cuda.synchronize()
    u0 = s.velocity_passive.cpu().numpy()
    p0 = s.stream_fn_passive.cpu().numpy()
    stream = torch.cuda.Stream()
    r = Incompressible2DRK2(dt, s, leaf_only=True, renormalize=True,
                            use_graph=True, stream=stream.cuda_stream)
    for _ in range(warmup):
        s.advance_timestep(r)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.advance_timestep(r)
    stream.synchronize()
    el = time.perf_counter() - t0
    pairs = mesh.useful_pairs()
    sl = slice(0, mesh.n_verts, oracle_stride)
    z = G.rh54_vorticity(mesh.face_x, 0.1, -0.1)
    out = {
        "workload": (f"Incompressible2D, icosahedral mesh level {level} "
                     f"({mesh.n_verts} vertices + {mesh.n_leaves} faces), RH "
                     f"5-4, Omega = 2 pi, eps = 0, RK2 dt = {dt}, FP64 velocity "
                     f"+ stream function (log per pair)"),
        "pairs_per_evaluation": pairs,
        "evaluations_per_step": 2,
        "value": round(2 * pairs * steps / el / 1e9, 3),
        "unit": "Gpair-interactions/s",
        "ms_per_step": round(el / steps * 1e3, 3),
        "rk2_steps_per_s": round(steps / el, 4),
        "dtype": "f64",
        "graph": True,
        "vs_analytic_rh54_t0_vertices_rel_l2_linf": errors(
            u0, G.rh54_velocity(mesh.vert_x, 0.1, -0.1)),
        "stream_fn_vs_analytic_rh54_t0_rel_l2_linf": errors1(
            p0, G.rh54_stream_function(mesh.vert_x, 0.1, -0.1)),
    }

    def check(oracle):
        ru, rpsi = oracle.ic2d_sums(np.ascontiguousarray(mesh.vert_x[:, sl].T),
                                    np.ascontiguousarray(mesh.face_x.T), z,
                                    mesh.face_area, mesh.face_mask,
                                    nthreads=min(16, os.cpu_count() or 1))
        out["fp64_vs_cpu_oracle_sample"] = {
            "targets": f"every {oracle_stride}th vertex",
            "velocity_rel_l2_linf": errors(u0[:, sl], ru.T),
            "stream_fn_rel_l2_linf": errors1(p0[sl], rpsi)}
    CPU_CHECKS.append(check)
    log(f"incompressible2d: {el / steps * 1e3:.1f} ms/step, "
        f"{out['value']:.1f} Gpairs/s")
    uT = s.velocity_passive.cpu().numpy()
    del s, r

    # mixed precision: same run with FP32 pairs and the native FP32 log
    m = Incompressible2D("icos", level, ctx=ctx, mesh=mesh)
    m.init_vorticity(lambda x: G.rh54_vorticity(x, 0.1, -0.1))
    m.init_direct_sums(skip, precision="mixed")
    torch.cuda.synchronize()
    mu0 = m.velocity_passive.cpu().numpy()
    mp0 = m.stream_fn_passive.cpu().numpy()
    rm = Incompressible2DRK2(dt, m, leaf_only=True, renormalize=True,
                             use_graph=True, stream=stream.cuda_stream,
                             precision="mixed")
    for _ in range(warmup):
        m.advance_timestep(rm)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        m.advance_timestep(rm)
    stream.synchronize()
    elm = time.perf_counter() - t0
    out["mixed"] = {
        "value": round(2 * pairs * steps / elm / 1e9, 3),
        "unit": "Gpair-interactions/s",
        "ms_per_step": round(elm / steps * 1e3, 3),
        "speedup_vs_fp64": round(el / elm, 4),
        "dtype": "f32 pair arithmetic and log, f64 carries/state",
        "velocity_vs_fp64_t0_rel_l2_linf": errors(mu0, u0),
        "stream_fn_vs_fp64_t0_rel_l2_linf": errors1(mp0, p0),
        "velocity_vs_fp64_after_steps_rel_l2_linf": errors(
            m.velocity_passive.cpu().numpy(), uT),
    }
    log(f"incompressible2d mixed: {elm / steps * 1e3:.1f} ms/step, "
        f"{out['mixed']['value']:.1f} Gpairs/s ({el / elm:.2f}x)")
    del m, rm
    return out


def diagnostics_run(ctx, n=50_000_000, reps=10):
    """ErrNorms and the tangency dot product (lpm_err_norms,
    lpm_sphere_tangency) over n three-component FP64 fields with weights:
    HBM-bound, 56 B per element read (80 B with the error field written).
    Wall time per blocking call (err_norms returns its norms to the host)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    kw = dict(dtype=torch.float64, device="cuda", generator=g)
    a = torch.randn(3, n, **kw)
    x = torch.randn(3, n, **kw)
    w = torch.rand(n, **kw)
    err = torch.empty_like(a)
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    res = {"elements": n, "reps": reps}
    for name, fn, nbytes in (
            ("err_norms", lambda: ctx.err_norms(a, x, weight=w), 56 * n),
            ("err_norms_with_error_field",
             lambda: ctx.err_norms(a, x, weight=w, err=err), 80 * n),
            ("sphere_tangency", lambda: ctx.sphere_tangency(a, x, out=out),
             56 * n)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        res[name] = {"ms_per_call": round(dt * 1e3, 4),
                     "algorithmic_GB_per_s": round(nbytes / dt / 1e9, 1)}
    del a, x, w, err, out
    torch.cuda.empty_cache()
    return res


def solid_body_run(ctx, mesh, omega=2 * 3.141592653589793):
    """Solid-body rotation on the workload mesh (the reference driver's
    check, examples/bve_rotation.cpp:147-167): zeta = 2 Omega z, exact
    u = Omega (-y, x, 0) and psi = Omega z; FP64 and mixed velocity and the
    FP64 stream function from one evaluation each."""
    import numpy as np
    import torch
    from lpm_amd import _abi
    from lpm_amd import galleries as G
    from lpm_amd.bve import BVESphere

    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS
    s = BVESphere("icos", mesh.depth, ctx=ctx, mesh=mesh)
    s.init_vorticity(lambda x: G.solid_body_vorticity(x, omega))
    fx = mesh.face_x[:, mesh.leaf_slice]
    out = {"workload": (f"solid-body rotation zeta = 2 Omega z, Omega = "
                        f"{omega:.6f}, icosahedral level {mesh.depth}; exact "
                        f"u = Omega (-y, x, 0), psi = Omega z")}
    for prec in ("fp64", "mixed"):
        s.init_velocity(prec, flags=skip)
        torch.cuda.synchronize()
        uv = s.velocity_verts.cpu().numpy()
        uf = s.velocity_faces.cpu().numpy()[:, mesh.leaf_slice]
        out[prec] = {
            "vs_exact_velocity_vertices_rel_l2_linf": errors(
                uv, G.solid_body_velocity(mesh.vert_x, omega)),
            "vs_exact_velocity_faces_rel_l2_linf": errors(
                uf, G.solid_body_velocity(fx, omega))}
    s.init_stream_fn(flags=skip)
    torch.cuda.synchronize()
    out["fp64"]["vs_exact_stream_fn_vertices_rel_l2_linf"] = errors1(
        s.stream_fn_verts.cpu().numpy(), omega * mesh.vert_x[2])
    log("solid body: FP64 velocity rel L2 "
        f"{out['fp64']['vs_exact_velocity_vertices_rel_l2_linf'][0]:.2e}, "
        f"mixed {out['mixed']['vs_exact_velocity_vertices_rel_l2_linf'][0]:.2e}")
    del s
    return out


def stream_fn_run(ctx, mesh, reps=3, oracle_stride=512):
    """BVE stream function (BVESphere::init_stream_fn): the FP64 sum of
    s log(1 - x.y) at every vertex and leaf face of the workload mesh, one
    log per pair (table-driven FP64 log)."""
    import numpy as np
    import torch
    from lpm_amd import _abi
    from lpm_amd import galleries as G

    def dev(a, dtype=torch.float64):
        return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")

    z = G.rh54_vorticity(mesh.face_x, 0.1, -0.1)
    vx, fx = dev(mesh.vert_x), dev(mesh.face_x)
    fz, fa = dev(z), dev(mesh.face_area)
    fm = dev(mesh.face_mask, torch.uint8)
    pv = torch.zeros(mesh.n_verts, dtype=torch.float64, device="cuda")
    pf = torch.zeros(mesh.n_faces, dtype=torch.float64, device="cuda")
    skip = _abi.LPM_FLAG_SKIP_MASKED_TARGETS

    def call():
        ctx.bve_stream_fn(vx, fx, fz, fa, fm, out=pv)
        ctx.bve_stream_fn(fx, fx, fz, fa, fm, collocated=True, flags=skip,
                          out=pf)
    call()
    torch.cuda.synchronize()
    ctx.timing_enable(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    kms, _, _ = ctx.timing_query(_abi.LPM_KERNEL_BVE_PSI)
    ctx.timing_enable(False)
    pairs = mesh.useful_pairs()
    psi_v = pv.cpu().numpy()
    sl = slice(0, mesh.n_verts, oracle_stride)
    out = {
        "workload": (f"BVE stream function, icosahedral mesh level "
                     f"{mesh.depth} ({mesh.n_verts} vertices + "
                     f"{mesh.n_leaves} faces), RH 5-4, FP64 with one log "
                     f"per pair"),
        "pairs_per_evaluation": pairs,
        "value": round(pairs / el / 1e9, 3),
        "unit": "Gpair-interactions/s",
        "ms_per_evaluation": round(el * 1e3, 3),
        "kernel_ms_per_evaluation": round(kms / reps, 3),
        "dtype": "f64",
        "vs_analytic_rh54_vertices_rel_l2_linf": errors1(
            psi_v, G.rh54_stream_function(mesh.vert_x, 0.1, -0.1)),
    }

    def check(oracle):
        ref = oracle.bve_stream_fn(np.ascontiguousarray(mesh.vert_x[:, sl].T),
                                   np.ascontiguousarray(mesh.face_x.T), z,
                                   mesh.face_area, mesh.face_mask,
                                   nthreads=min(16, os.cpu_count() or 1))
        out["fp64_vs_cpu_oracle_sample"] = {
            "targets": f"every {oracle_stride}th vertex",
            "stream_fn_rel_l2_linf": errors1(psi_v[sl], ref)}
    CPU_CHECKS.append(check)
    log(f"stream function: {el * 1e3:.1f} ms/evaluation, "
        f"{pairs / el / 1e9:.1f} Gpairs/s")
    return out


def cpu_baseline(mesh, z_faces, budget_s=12.0):
    """The oracle (reference-faithful loop: every stored face, masked ones
    skipped) on the host's cores, in a child process with its OpenMP threads
    pinned one per physical core: the icosahedral level-4 run in full and a
    bounded sample of this workload's evaluation (oracle/cpu_baseline.py).
    Threads: the physical cores of this process's affinity set, at most
    OMP_NUM_THREADS (the CPU share a one-GPU job is given)."""
    import subprocess
    import tempfile

    import numpy as np
    from lpm_amd.mesh import SphereMesh
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cpu_baseline as CB

    host = CB.host_cpu()
    cap = int(os.environ.get("OMP_NUM_THREADS") or host["physical_cores_in_affinity"])
    threads = max(1, min(host["physical_cores_in_affinity"], cap))
    c1 = SphereMesh.build("icos", 4)
    with tempfile.TemporaryDirectory(prefix="lpm_bench_cpu_") as tmp:
        meshes = os.path.join(tmp, "meshes.npz")
        out = os.path.join(tmp, "sample.npz")
        np.savez(meshes, c1_vert_x=c1.vert_x, c1_face_x=c1.face_x,
                 c1_face_area=c1.face_area, c1_face_mask=c1.face_mask,
                 c2_vert_x=mesh.vert_x, c2_face_x=mesh.face_x,
                 c2_face_area=mesh.face_area, c2_face_mask=mesh.face_mask)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads),
                   OMP_PROC_BIND="spread", OMP_PLACES="cores")
        r = subprocess.run(
            [sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"),
             "--meshes", meshes, "--out", out, "--threads", str(threads),
             "--budget", str(budget_s)],
            env=env, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"CPU baseline failed:\n{r.stderr[-2000:]}")
        cb = json.loads(r.stdout.strip().splitlines()[-1])
        with np.load(out) as f:
            u, stride, n = f["u"], int(f["stride"]), int(f["n"])
    cb["cores_note"] = (f"{threads} OpenMP threads, one per physical core of "
                        f"this process's affinity set, capped at "
                        f"OMP_NUM_THREADS={cap} (the CPU share of a one-GPU "
                        f"job)")
    return cb, n, stride, u


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles",
                               "bve_vel_f64_icos_l7_traffic.md")


def measured_traffic(level, variant):
    """HBM bytes per launch of the FP64 kernel from the committed rocprofv3
    counter summary (line ``traffic_bytes_per_launch: <variant> <level>
    <bytes>``), or None when it does not cover this variant and level."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            text = f.read()
    except OSError:
        return None
    for m in re.finditer(r"^traffic_bytes_per_launch:\s+(\S+)\s+(\d+)\s+(\d+)",
                         text, re.M):
        if m.group(1) == variant and int(m.group(2)) == level:
            return int(m.group(3))
    return None


COUNTER_SUMMARIES = [os.path.join(ROOT, "profiles", f) for f in (
    "bve_vel_icos_l7_valu_counters.md", "swe_ens_icos_l6x16_valu_counters.md")]


def valu_counters(kernel, variant, workload):
    """Counter figures of one kernel shape and workload from the committed
    rocprofv3 summaries (lines ``valu_counters: <kernel> <variant>
    <workload> key=value ...``), or None."""
    for path in COUNTER_SUMMARIES:
        try:
            with open(path) as f:
                text = f.read()
        except OSError:
            continue
        for m in re.finditer(r"^valu_counters:\s+(\S+)\s+(\S+)\s+(\S+)\s+(.*)$",
                             text, re.M):
            if (m.group(1), m.group(2), m.group(3)) == (kernel, variant,
                                                        workload):
                out = {"source": os.path.relpath(path, ROOT)}
                for kv in m.group(4).split():
                    k, _, v = kv.partition("=")
                    out[k] = float(v)
                return out
    return None


def copy_calibration(mib, reps=5):
    """A plain elementwise copy of known size for scaling byte counters."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {
        "bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8,
        "launches": reps, "kernel": "torch.neg (elementwise, f64)"}}),
        flush=True)
    return 0


# legs that --only selects (each has a --no-<leg> switch)
LEGS = ["mixed", "solid-body", "stream-fn", "diagnostics", "ensemble",
        "precision-study", "swe", "ic2d", "cpu-baseline"]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--level", type=int, default=7,
                    help="icosahedral refinement level (7 = the workload)")
    ap.add_argument("--dt", type=float, default=0.005)
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of CPU oracle work for the baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the fixed default kernel variants")
    ap.add_argument("--variant-f64", default=None,
                    help="force this FP64 kernel variant (implies no autotune "
                         "of the FP64 run)")
    ap.add_argument("--variant-mixed", default=None,
                    help="force this mixed kernel variant (implies no autotune "
                         "of the mixed run)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run only a plain device copy of MIB MiB")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 64-member ensemble")
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--no-precision-study", action="store_true",
                    help="skip the cubed-sphere mixed-vs-FP64 study")
    ap.add_argument("--no-swe", action="store_true",
                    help="skip the 16-state shallow-water ensemble")
    ap.add_argument("--no-ic2d", action="store_true",
                    help="skip the Incompressible2D run")
    ap.add_argument("--no-solid-body", action="store_true",
                    help="skip the solid-body rotation check")
    ap.add_argument("--no-stream-fn", action="store_true",
                    help="skip the BVE stream-function evaluation")
    ap.add_argument("--no-diagnostics", action="store_true",
                    help="skip the error-norm / tangency bandwidth run")
    ap.add_argument("--swe-steps", type=int, default=10,
                    help="RK4 steps of each shallow-water ensemble run")
    ap.add_argument("--only", default=None,
                    help="comma-separated legs to run after the FP64 run "
                         f"(of {','.join(LEGS)}); the others are skipped")
    a = ap.parse_args(argv)
    if a.only is not None:
        keep = {x.strip() for x in a.only.split(",") if x.strip()}
        unknown = keep - set(LEGS)
        if unknown:
            ap.error(f"unknown legs {sorted(unknown)}; have {LEGS}")
        for leg in LEGS:
            if leg not in keep:
                setattr(a, "no_" + leg.replace("-", "_"), True)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding "
              "of the direct sum is a design note (targets split across GPUs "
              "with an all-gather of sources per RK stage), not implemented",
              file=sys.stderr)
        return 2
    if a.copy_calibration:
        return copy_calibration(a.copy_calibration)

    import numpy as np
    import torch
    from lpm_amd import galleries as G
    from lpm_amd.device import Context
    from lpm_amd.mesh import SphereMesh

    log(f"building icosahedral mesh, level {a.level}")
    mesh = SphereMesh.build("icos", a.level)
    pairs_eval = mesh.useful_pairs()
    log(f"N_V={mesh.n_verts} N_F={mesh.n_leaves} stored={mesh.n_faces} "
        f"pairs/eval={pairs_eval:.4e}")
    ctx = Context(0)
    props = torch.cuda.get_device_properties(0)

    from lpm_amd import _abi
    kid64 = _abi.LPM_KERNEL_BVE_VEL_F64
    names64 = ctx.variants(kid64)
    if a.variant_f64 is not None:
        if a.variant_f64 not in names64:
            log(f"unknown FP64 variant {a.variant_f64}; have {names64}")
            return 2
        ctx.set_variant(kid64, names64.index(a.variant_f64))
    kidm = _abi.LPM_KERNEL_BVE_VEL_MIXED
    if a.variant_mixed is not None:
        namesm = ctx.variants(kidm)
        if a.variant_mixed not in namesm:
            log(f"unknown mixed variant {a.variant_mixed}; have {namesm}")
            return 2
        ctx.set_variant(kidm, namesm.index(a.variant_mixed))
    log("FP64 run")
    r64 = run_precision(ctx, mesh, "fp64", a.steps, a.warmup, a.dt,
                        dict(reference_compat=True),
                        autotune=not a.no_autotune and a.variant_f64 is None)
    v64 = (r64["autotune"] or {}).get("variant") or (
        a.variant_f64 or names64[ctx.L.lpm_variant_default(kid64)])
    step_s = r64["elapsed_s"] / a.steps
    value = 4 * pairs_eval * a.steps / r64["elapsed_s"] / 1e9
    avg_ms = r64["kernel_ms_total"] / max(r64["launches"], 1)
    achieved = pairs_eval * F_BVE / (avg_ms * 1e-3) / 1e12
    log(f"FP64: {step_s*1e3:.1f} ms/step, {value:.1f} Gpairs/s, kernel "
        f"{avg_ms:.2f} ms/launch -> {achieved:.1f} TFLOP/s")

    # analytic RH5-4 velocity at t = 0
    ue_v = G.rh54_velocity(mesh.vert_x, 0.1, -0.1)
    ue_f = G.rh54_velocity(mesh.face_x[:, mesh.leaf_slice], 0.1, -0.1)
    an_v = errors(r64["u0_v"], ue_v)
    an_f = errors(r64["u0_f"], ue_f)
    # the RH 5-4 wave rotates rigidly at nu (exact nonlinear solution): the
    # velocity at the particles after the run vs the rotated pattern
    t_end = (a.warmup + a.steps) * a.dt
    nu = G.rh54_wave_speed(0.1, 2 * np.pi)

    def rh54_at(x, t):
        return G.rotate_z(G.rh54_velocity(G.rotate_z(x, -nu * t), 0.1, -0.1),
                          nu * t)
    an_t = errors(r64["uT_v"], rh54_at(r64["xT_v"], t_end))

    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpair-interactions/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"BVE RK4 direct sum, icosahedral mesh level {a.level}"
                         f" ({mesh.n_verts} vertices + {mesh.n_leaves} leaf "
                         f"faces as targets, {mesh.n_leaves} sources), "
                         f"Rossby-Haurwitz 5-4 (omega=K=0.1), Omega=2pi, "
                         f"dt={a.dt}, FP64 pair arithmetic"),
            "pairs_per_evaluation": pairs_eval,
            "evaluations_per_step": 4,
        },
        "rk4_steps_per_s": round(1.0 / step_s, 4),
        "errors_fp64": {
            "vs_analytic_rh54_vertices_rel_l2_linf": an_v,
            "vs_analytic_rh54_faces_rel_l2_linf": an_f,
            "vs_analytic_rh54_after_steps_vertices_rel_l2_linf": an_t,
            "analytic_after_steps": f"t = {t_end:g}, pattern rotated by "
                                    f"nu t = {nu * t_end:.5f} rad",
        },
        "roofline": {
            # bound by FP64 issue on the vector ALUs (78.6 TFLOP/s)
            "bound": "valu",
            "compute_unit": "FP64 vector ALU",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
            "traffic": measured_traffic(a.level, v64),
            # a source-split variant is two launches (partial sums, then
            # their reduction); avg_launch_ms times both
            "kernel": ("bve_vel_f64_split_kernel + bve_split_reduce_kernel"
                       if "_split" in v64 else "bve_vel_f64_kernel"),
            "variant": v64,
            "avg_launch_ms": round(avg_ms, 4),
            "flops_per_pair": F_BVE,
            "pairs_per_launch": pairs_eval,
        },
        "device": {"name": props.name, "cus": props.multi_processor_count},
        "autotune_fp64": r64["autotune"],
    }

    vc = valu_counters("bve_vel_f64", v64, f"{a.level}")
    if vc is not None:
        # hardware-counted FP64 flops per pair from the committed profile,
        # at this run's kernel time
        counted = vc["fp64_flops_per_pair"] * pairs_eval / (avg_ms * 1e-3) / 1e12
        result["roofline"]["counters"] = {
            "source": vc["source"],
            "valu_busy": vc["valu_busy"],
            "fp64_flops_per_pair_counted": vc["fp64_flops_per_pair"],
            "fp64_instr_per_pair_counted": vc["fp64_instr_per_pair"],
            "counted_tflops": round(counted, 3),
            "counted_frac": round(counted / PEAK_FP64_TFLOPS, 4),
            "profiled_avg_launch_ms": vc["avg_ms"],
        }

    if not a.no_mixed:
        log("mixed run")
        rmx = run_precision(ctx, mesh, "mixed", a.steps, a.warmup, a.dt,
                            dict(reference_compat=True, renormalize=True),
                            autotune=not a.no_autotune and a.variant_mixed is None)
        mstep = rmx["elapsed_s"] / a.steps
        mval = 4 * pairs_eval * a.steps / rmx["elapsed_s"] / 1e9
        mavg = rmx["kernel_ms_total"] / max(rmx["launches"], 1)
        result["mixed"] = {
            "value": round(mval, 3),
            "unit": "Gpair-interactions/s",
            "ms_per_step": round(mstep * 1e3, 3),
            "rk4_steps_per_s": round(1.0 / mstep, 4),
            "speedup_vs_fp64": round(mval / value, 4),
            "avg_launch_ms": round(mavg, 4),
            "dtype": "f32 pair arithmetic, f64 carries/state",
            "velocity_error_vs_fp64_t0_vertices_rel_l2_linf":
                errors(rmx["u0_v"], r64["u0_v"]),
            "velocity_error_vs_fp64_t0_faces_rel_l2_linf":
                errors(rmx["u0_f"], r64["u0_f"]),
            "velocity_error_vs_fp64_after_steps_vertices_rel_l2_linf":
                errors(rmx["uT_v"], r64["uT_v"]),
            "vs_analytic_rh54_vertices_rel_l2_linf": errors(rmx["u0_v"], ue_v),
            "vs_analytic_rh54_after_steps_vertices_rel_l2_linf":
                errors(rmx["uT_v"], rh54_at(rmx["xT_v"], t_end)),
            "autotune": rmx["autotune"],
        }
        vm = (rmx["autotune"] or {}).get("variant") or a.variant_mixed
        vcm = valu_counters("bve_vel_mfma16", vm, f"{a.level}") if vm else None
        if vcm is not None:
            result["mixed"]["counters"] = vcm
        log(f"mixed: {mstep*1e3:.1f} ms/step, {mval:.1f} Gpairs/s "
            f"({mval/value:.2f}x FP64)")

    if not a.no_solid_body:
        log("solid-body rotation")
        result["solid_body"] = solid_body_run(ctx, mesh)
    if not a.no_stream_fn:
        log("stream function")
        result["stream_fn"] = stream_fn_run(ctx, mesh)
    if not a.no_diagnostics:
        log("diagnostics (error norms, tangency)")
        result["diagnostics"] = diagnostics_run(ctx)
    if not a.no_ensemble:
        log("ensemble run (64 members)")
        result["ensemble"] = ensemble_run(ctx, a.ensemble_steps, 1,
                                          autotune=not a.no_autotune)
        result["ensemble"]["speedup_vs_fp64_single_problem"] = round(
            result["ensemble"]["value"] / value, 4)
    if not a.no_precision_study:
        log("precision study (cubed sphere level 8)")
        result["precision_study"] = precision_study(ctx)
    if not a.no_swe:
        log("shallow-water ensemble (16 x level 6)")
        result["swe_ensemble"] = swe_ensemble_run(ctx, steps=a.swe_steps,
                                                  tune=not a.no_autotune)
    if not a.no_ic2d:
        log("Incompressible2D (level 6)")
        result["incompressible2d"] = ic2d_run(ctx)

    if not a.no_cpu_baseline:
        log("CPU baseline (oracle) on a bounded sample")
        z = G.rh54_vorticity(mesh.face_x, 0.1, -0.1)
        cb, n_sample, stride, u_cpu = cpu_baseline(mesh, z, a.cpu_budget)
        result["cpu_baseline"] = cb
        # parity of the GPU FP64 initial velocity with the oracle, same sample
        gpu_sample = r64["u0_v"][:, ::stride][:, :n_sample]
        result["errors_fp64"]["vs_cpu_oracle_sample_rel_l2_linf"] = errors(
            gpu_sample, u_cpu.T)
        log(f"CPU: {cb['value']:.3f} Gpairs/s on {cb['cores']} threads "
            f"({cb['host']['model']}), level-4 run "
            f"{cb['small_workload_full']['seconds']:.2f} s")
        # both FP64 sums against an extended-precision evaluation of the
        # same pair formula (long double products, compensated sums): the
        # GPU kernel and the reference-faithful loop round 1 - x.y and the
        # sum differently, and both stay this close to the exact sum
        import oracle
        xs = 256
        vs = np.ascontiguousarray(mesh.vert_x.T[::xs])
        fxt = np.ascontiguousarray(mesh.face_x.T)
        thr = min(16, os.cpu_count() or 1)
        uxp = oracle.bve_velocity_xp(vs, fxt, z, mesh.face_area, mesh.face_mask,
                                     nthreads=thr)
        uor = oracle.bve_velocity(vs, fxt, z, mesh.face_area, mesh.face_mask,
                                  nthreads=thr)
        result["errors_fp64"]["extended_precision_sample"] = {
            "targets": f"every {xs}th vertex ({len(vs)})",
            "gpu_vs_extended_rel_l2_linf": errors(r64["u0_v"][:, ::xs], uxp.T),
            "oracle_vs_extended_rel_l2_linf": errors(uor.T, uxp.T),
        }
        log(f"oracle comparisons of the other legs ({len(CPU_CHECKS)})")
        for check in CPU_CHECKS:
            check(oracle)

    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
