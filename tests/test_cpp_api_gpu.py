"""The C++ host API (BVESphere/BVERK4, SWE/SWERK4 over the C ABI) through
its example drivers, checked against closed forms."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "lpm_amd", "_build")


def run(name, *args, cwd=None):
    exe = os.path.join(BUILD, name)
    assert os.path.exists(exe), f"{exe} not built (python -m lpm_amd.build)"
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True,
                       timeout=120, cwd=cwd)
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
    # ErrNorms on the device agrees with the host sum; the velocity is
    # tangent to the sphere (SphereTangentFunctor) to rounding
    l1, l2, linf = out["velocity_err_norms"]
    assert abs(l2 - out["velocity_rel_l2"]) <= 1e-12 * out["velocity_rel_l2"]
    assert 0 < l1 and 0 < linf
    assert out["max_abs_u_dot_x"] < 1e-12 * 2 * 3.14159


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


def test_bve_rotation_vtk_frames(tmp_path):
    import numpy as np
    from lpm_amd.mesh import SphereMesh
    from lpm_amd.vtk import VtkPolymeshInterface, read_vtp
    out = run("bve_rotation", "-d", 3, "-dt", 0.01, "-tf", 0.03, "-seed",
              "icos", "-o", "rot", "-f", 2, cwd=tmp_path)
    frames = sorted(p.name for p in tmp_path.glob("rot_*.vtp"))
    # t = 0, after step 2 and after the last step (3)
    assert len(frames) == 3, frames
    f0 = read_vtp(str(tmp_path / frames[0]))
    f2 = read_vtp(str(tmp_path / frames[-1]))
    # the C++ writer's polygons and areas are the Python writer's
    py = VtkPolymeshInterface(SphereMesh.build("icos", 3))
    np.testing.assert_array_equal(f0["connectivity"], py.polys.reshape(-1))
    np.testing.assert_array_equal(f0["cell_data"]["area"], py.cell_data["area"])
    np.testing.assert_array_equal(f0["points"], f0["point_data"]["lag_crds"])
    np.testing.assert_array_equal(f2["point_data"]["lag_crds"],
                                  f0["point_data"]["lag_crds"])
    x = f0["points"]
    omg = 2 * np.pi
    exact = np.stack([-omg * x[:, 1], omg * x[:, 0], 0 * x[:, 0]], 1)
    u = f0["point_data"]["velocity"]
    assert np.linalg.norm(u - exact) / np.linalg.norm(exact) < 2e-2
    # the last frame holds the rotated positions the example program reports
    ang = omg * out["t"]
    c, s = np.cos(ang), np.sin(ang)
    xr = np.stack([c * x[:, 0] - s * x[:, 1], s * x[:, 0] + c * x[:, 1],
                   x[:, 2]], 1)
    err = np.linalg.norm(f2["points"] - xr, axis=1).max()
    assert err == pytest.approx(out["position_err_linf"], rel=1e-5)


@pytest.mark.parametrize("seed", ["cube", "icos"])
def test_sphere_rh54_driver(seed):
    """Stationary RH 5-4 wave (u0 = Omega / 14) with Incompressible2D: the
    velocity matches RossbyWave54Velocity to the discretisation error, and
    the relative vorticity carried by the particles stays the exact
    RossbyHaurwitz54 at their new positions."""
    out = run("sphere_rh54", "-d", 4, "-n", 5, "-tf", 0.025, "-seed", seed)
    assert out["steps"] == 5 and abs(out["t"] - 0.025) < 1e-15
    assert abs(out["u0"] - 2 * 3.141592653589793 / 14) < 1e-15
    i, f = out["initial"], out["final"]
    assert i["vorticity_err"] == [0.0, 0.0, 0.0]
    # level-4 discretisation error of the velocity (1.7e-2 cube, 9e-3 icos)
    # and its effect on the carried vorticity over five steps (6e-4, 7e-4)
    for r in (i, f):
        assert 0 < r["velocity_err"][1] < 2.5e-2
        assert 0 < r["velocity_err"][2] < 5e-2
    assert 0 < f["vorticity_err"][1] < 2e-3
    assert 0 < f["vorticity_err"][2] < 5e-3
    # totals: the wave integrates to zero vorticity; enstrophy and energy
    # of a stationary wave do not change beyond the time-stepping error
    assert abs(i["total_vorticity"]) < 1e-12 * i["total_enstrophy"]
    assert abs(f["total_enstrophy"] / i["total_enstrophy"] - 1) < 1e-6
    assert abs(f["total_kinetic_energy"] / i["total_kinetic_energy"] - 1) < 1e-3
    mixed = run("sphere_rh54", "-d", 4, "-n", 5, "-tf", 0.025, "-seed", seed,
                "-precision", "mixed")
    # the mixed sums (FP32 pairs and logarithm, ~2e-7 relative to FP64 at
    # level 4, positions kept on the sphere) move the norms by about that
    assert abs(mixed["final"]["velocity_err"][1] - f["velocity_err"][1]) < 1e-6
    assert abs(mixed["final"]["vorticity_err"][1] - f["vorticity_err"][1]) < 1e-6


def test_sphere_gaussian_vortex_driver():
    """Gaussian vortex with Incompressible2D: zero mean vorticity after
    set_gauss_const; absolute vorticity is carried exactly, so the relative
    totals move only with the discrete integral of z; the velocity is
    tangent to the sphere."""
    out = run("sphere_gaussian_vortex", "-d", 4, "-n", 5, "-tf", 0.025)
    vort, ens, ke = (out["total_vorticity"], out["total_enstrophy"],
                     out["total_kinetic_energy"])
    assert len(vort) == 6 and out["steps"] == 5
    assert abs(vort[0]) < 1e-12 * ens[0]
    assert max(abs(v) for v in vort) < 1e-5 * ens[0]
    assert max(abs(e / ens[0] - 1) for e in ens) < 1e-5
    assert max(abs(k / ke[0] - 1) for k in ke) < 1e-3
    assert out["max_abs_u_dot_x"] < 1e-13
