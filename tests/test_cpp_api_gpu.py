"""The C++ host API (BVESphere/BVERK4, SWE/SWERK4 over the C ABI) through
its example drivers, checked against closed forms."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "lpm_amd", "_build")


def run(name, *args):
    exe = os.path.join(BUILD, name)
    assert os.path.exists(exe), f"{exe} not built (python -m lpm_amd.build)"
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("renorm", [0, 1])
@pytest.mark.parametrize("seed", ["cube", "icos"])
def test_bve_rotation_driver(seed, renorm):
    # dt = 0.005 keeps the un-normalised stage positions of the reference
    # stepper far enough from the sphere's nearest-neighbour singularity at
    # level 4 (at dt = 0.01 the icosahedral run diverges with or without
    # this library: the FP64 CPU restatement gives the same blow-up).
    out = run("bve_rotation", "-d", 4, "-dt", 0.005, "-tf", 0.05, "-seed", seed,
              "-renorm", renorm)
    assert out["steps"] == 10
    assert out["velocity_rel_l2"] < 1e-2
    assert out["stream_fn_rel_l2"] < 1e-2
    # Rigid rotation by 2 pi t: the position error is the vertex velocity's
    # max relative discretisation error at level 4 (0.28 % icos, 0.67 %
    # cube, from the quadrature) times the arc 2 pi t, plus the O(dt^2)
    # off-sphere stage error when not renormalising.
    assert out["position_err_linf"] < 1e-2 * 2 * 3.14159 * out["t"]


def test_bve_rotation_mixed_matches_fp64():
    a = run("bve_rotation", "-d", 4, "-seed", "cube")
    b = run("bve_rotation", "-d", 4, "-seed", "cube", "-precision", "mixed")
    assert abs(a["velocity_rel_l2"] - b["velocity_rel_l2"]) < 1e-6


def test_sphere_swe_tc2_driver():
    out = run("sphere_swe_tc2", "-d", 4, "-n", 2)
    assert out["velocity_rel_l2"] < 2e-2
    assert out["ddot_rel_l2"] < 7e-2
    assert out["laps_rel_l2"] < 5e-2
    assert out["max_abs_divergence"] < 0.05
