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
