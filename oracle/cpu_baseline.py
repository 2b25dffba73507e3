"""ORACLE — test infrastructure only: the CPU baseline of the benchmark.

Run as a child process by the benchmark's CPU-baseline leg (never by the
product):

    python oracle/cpu_baseline.py --meshes meshes.npz --out sample.npz \
        --threads N --budget SECONDS

It times the plain-C FP64 restatement of the reference's direct sum
(``lpm_oracle.c``, the reference-faithful loop of
src/lpm_bve_sphere_kernels.hpp:53-82, 179-211, 249-274, 365-394: one target
per thread, every stored face visited, divided faces skipped by the mask,
a serial inner sum) on the host's own cores, as the reference's driver loop
times its steps (examples/bve_rotation.cpp:214-241 with the Timer of
src/util/lpm_timer.hpp:32-51):

* the small workload in full: BVE on the icosahedral mesh of level 4, RH 5-4
  (omega = K = 0.1), Omega = 2 pi, dt = 0.01, the initial velocity and 10
  RK4 steps (41 evaluations);
* one velocity evaluation of the benchmark workload on a bounded sample of its
  vertex targets (about --budget seconds), from which the full evaluation
  and the RK4 step rate are estimated as 1 / (4 t_eval + t_O(N)), with the
  O(N) stage work t_O(N) measured on the small workload and scaled by the
  particle count.

The library is compiled here with ``gcc -O3 -march=native`` (the host's own
instruction set; contraction into FMAs allowed, as GCC compiles the
reference's C++ by default), falling back to the prebuilt portable build.
OpenMP threads are pinned one per physical core (OMP_PROC_BIND=spread,
OMP_PLACES=cores, set by the parent before this process starts).  Prints one
JSON object.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import oracle  # noqa: E402


def host_cpu():
    """CPU model, sockets, cores, SMT and the affinity set of this process."""
    model, phys = "unknown", {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                if k.strip() == "model name":
                    model = v.strip()
                    break
    except OSError:
        pass
    logical = os.cpu_count() or 1
    affinity = sorted(os.sched_getaffinity(0))

    def topo(c, name):
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/{name}") as f:
                return int(f.read().strip())
        except (OSError, ValueError):
            return None

    for c in range(logical):
        pkg, core = topo(c, "physical_package_id"), topo(c, "core_id")
        if pkg is not None and core is not None:
            phys[c] = (pkg, core)
    sockets = len({p for p, _ in phys.values()}) or None
    cores = len(set(phys.values())) or None
    in_aff = len({phys[c] for c in affinity if c in phys}) or len(affinity)
    return {
        "model": model,
        "sockets": sockets,
        "physical_cores": cores,
        "threads_per_core": (logical // cores) if cores else None,
        "logical_cpus": logical,
        "affinity_cpus": len(affinity),
        "physical_cores_in_affinity": in_aff,
    }


def build_native(tmp):
    """gcc -O3 -march=native build of the oracle in tmp, or None."""
    cc = shutil.which("gcc")
    if cc is None:
        return None, "gcc not found"
    out = os.path.join(tmp, "liblpm_oracle_native.so")
    flags = ["-O3", "-march=native", "-ffp-contract=fast", "-fopenmp",
             "-fPIC", "-std=c11"]
    cmd = [cc, *flags, "-shared", "-o", out,
           os.path.join(HERE, "lpm_oracle.c"), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return None, r.stderr[-500:]
    return out, "gcc " + " ".join(flags) + " (compiled on this host)"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--meshes", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--threads", type=int, required=True)
    ap.add_argument("--budget", type=float, default=12.0)
    a = ap.parse_args(argv)

    # before the OpenMP runtime pins this thread to one core
    host = host_cpu()
    tmp = tempfile.mkdtemp(prefix="lpm_cpu_baseline_")
    try:
        lib_path, build = build_native(tmp)
        if lib_path is not None:
            oracle.load(lib_path)
        else:
            build = f"prebuilt portable library (native build failed: {build})"
        return run(a, build, host)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def run(a, build, host):
    thr = a.threads
    with np.load(a.meshes) as f:
        M = {k: f[k] for k in f.files}

    # ---- small workload in full: icosahedral level 4, 10 RK4 steps
    vx = np.ascontiguousarray(M["c1_vert_x"].T)
    fx = np.ascontiguousarray(M["c1_face_x"].T)
    area, mask = M["c1_face_area"], M["c1_face_mask"]
    nv, nf, nl = len(vx), len(fx), int((mask == 0).sum())
    st = {"vx": vx.copy(), "vzeta": oracle.rh54_vorticity(vx, 0.1, -0.1),
          "fx": fx.copy(), "fzeta": oracle.rh54_vorticity(fx, 0.1, -0.1),
          "farea": area, "fmask": mask}
    def evaluate():
        st["vu"] = oracle.bve_velocity(st["vx"], st["fx"], st["fzeta"], area,
                                       mask, nthreads=thr)
        st["fu"] = oracle.bve_velocity(st["fx"], st["fx"], st["fzeta"], area,
                                       mask, collocated=True, nthreads=thr)

    evaluate()  # thread start-up and caches
    t0 = time.perf_counter()
    evaluate()
    t_eval1 = time.perf_counter() - t0
    steps, dt = 10, 0.01
    t1 = time.perf_counter()
    for _ in range(steps):
        oracle.bve_rk4_step(st, dt, 2 * np.pi, compat=True, nthreads=thr)
    t_run1 = time.perf_counter() - t1
    t_step1 = t_run1 / steps
    t_on1 = max(t_step1 - 4 * t_eval1, 0.0)
    pairs1 = (nv + nl) * nl

    # ---- level-7 evaluation on a bounded sample of vertex targets
    vx7 = np.ascontiguousarray(M["c2_vert_x"].T)
    fx7 = np.ascontiguousarray(M["c2_face_x"].T)
    area7, mask7 = M["c2_face_area"], M["c2_face_mask"]
    z7 = oracle.rh54_vorticity(fx7, 0.1, -0.1)
    nv7, nf7, nl7 = len(vx7), len(fx7), int((mask7 == 0).sum())
    n = 256
    while True:
        tp = time.perf_counter()
        oracle.bve_velocity(vx7[:n], fx7, z7, area7, mask7, nthreads=thr)
        dtp = time.perf_counter() - tp
        if dtp > 0.5 or n >= nv7:
            break
        n = min(nv7, n * 4)
    n2 = int(min(nv7, max(n, n * a.budget / max(dtp, 1e-3))))
    stride = max(1, nv7 // n2)
    sample = np.ascontiguousarray(vx7[::stride][:n2])
    tp = time.perf_counter()
    u = oracle.bve_velocity(sample, fx7, z7, area7, mask7, nthreads=thr)
    t_sample = time.perf_counter() - tp
    rate = len(sample) * nl7 / t_sample           # useful pairs / s
    pairs7 = (nv7 + nl7) * nl7
    t_eval7 = pairs7 / rate
    t_on7 = t_on1 * (nv7 + nf7) / (nv + nf)
    np.savez(a.out, u=u, stride=np.int64(stride), n=np.int64(len(sample)))
    res = {
        "value": rate / 1e9,
        "unit": "Gpair-interactions/s",
        "cores": thr,
        "kind": "port",
        "sample": (f"one FP64 velocity evaluation of the workload "
                   f"at {len(sample)} of {nv7} vertex targets ("
                   + ("the first" if stride == 1 else f"every {stride}th")
                   + f") over all {nf7} stored faces ({nl7} undivided), "
                   f"reference-faithful loop, {t_sample:.1f} s; rate counts "
                   f"useful pairs ({len(sample) * nl7:.3e}); loop iterations "
                   f"{len(sample) * nf7:.3e}"),
        "host": host,
        "omp": {k: os.environ.get(k) for k in (
            "OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
        "build": build,
        "small_workload_full": {
            "workload": (f"icosahedral level 4 ({nv} vertices + {nl} leaf "
                         f"faces as targets, {nl} sources, {nf} stored "
                         f"faces), RH 5-4, initial velocity + {steps} RK4 "
                         f"steps, dt = {dt}"),
            "evaluations": 1 + 4 * steps,
            "seconds": round(t_eval1 + t_run1, 4),
            "s_per_evaluation": round(t_eval1, 5),
            "Gpair_interactions_per_s": round(pairs1 / t_eval1 / 1e9, 4),
            "rk4_steps_per_s": round(1.0 / t_step1, 4),
            "o_n_s_per_step": round(t_on1, 5),
            "o_n_note": "step time minus four steady-state evaluations",
        },
        "workload_estimate": {
            "s_per_evaluation": round(t_eval7, 3),
            "o_n_s_per_step": round(t_on7, 4),
            "rk4_steps_per_s": round(1.0 / (4 * t_eval7 + t_on7), 6),
            "formula": "1 / (4 t_eval + t_O(N)), t_O(N) scaled from level 4",
        },
    }
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
