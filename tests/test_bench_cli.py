"""bench.py's command-line contract that needs no GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_summary_lines_are_machine_readable():
    # every traffic line names a compiled FP64 variant and a level
    sys.path.insert(0, ROOT)
    import bench
    text = open(bench.TRAFFIC_SUMMARY).read()
    lines = [l for l in text.splitlines() if l.startswith("traffic_bytes_per_launch:")]
    assert lines
    from lpm_amd import _abi
    L = _abi.lib()
    kid = _abi.LPM_KERNEL_BVE_VEL_F64
    names = [L.lpm_variant_name(kid, v).decode()
             for v in range(L.lpm_variant_count(kid))]
    for l in lines:
        _, v, lev, b = l.split()
        assert v in names, v
        assert bench.measured_traffic(int(lev), v) == int(b)


def test_cpu_baseline_child_reports_host_and_small_run():
    """The CPU-baseline leg: a child process builds the oracle for this
    host, times the icosahedral level-4 run in full and a sample of the
    given mesh's evaluation, and reports the host's CPU."""
    sys.path.insert(0, ROOT)
    import bench
    import numpy as np
    from lpm_amd import galleries as G
    from lpm_amd.mesh import SphereMesh
    m = SphereMesh.build("icos", 3)
    cb, n, stride, u = bench.cpu_baseline(
        m, G.rh54_vorticity(m.face_x, 0.1, -0.1), budget_s=0.2)
    assert cb["kind"] == "port" and cb["cores"] >= 1
    assert cb["host"]["physical_cores_in_affinity"] >= 1
    assert cb["omp"]["OMP_PROC_BIND"] == "spread"
    assert cb["omp"]["OMP_PLACES"] == "cores"
    s = cb["small_workload_full"]
    assert s["evaluations"] == 41 and s["rk4_steps_per_s"] > 0
    assert cb["workload_estimate"]["rk4_steps_per_s"] > 0
    assert u.shape == (n, 3) and np.all(np.isfinite(u))
