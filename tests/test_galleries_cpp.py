"""The C++ gallery functors (lpm_amd/cpp/include/lpm/galleries.hpp) against
the Python galleries and against what the mathematics fixes: the spherical
Laplacian of the Rossby-Haurwitz 5-4 vorticity by finite differences, the
tangency of the analytic velocities, the stationary background rotation
Omega / 14 and the two forms of the test case 2 surface.

The functors are host code: tests/galleries_eval.cpp is compiled with hipcc
(host only) and run on points on the unit sphere, including the poles and
the longitude branch cut.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from lpm_amd import galleries as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
NOUT = 18
GRAVITY = 9.80616


@pytest.fixture(scope="module")
def evaluate(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("galleries")
    exe = str(d / "galleries_eval")
    r = subprocess.run([HIPCC, "-std=c++17", "-O2",
                        f"-I{os.path.join(ROOT, 'include')}",
                        f"-I{os.path.join(ROOT, 'lpm_amd', 'cpp', 'include')}",
                        os.path.join(ROOT, "tests", "galleries_eval.cpp"),
                        "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

    def run(x):
        fin, fout = str(d / "in.bin"), str(d / "out.bin")
        np.ascontiguousarray(x.T, dtype=np.float64).tofile(fin)
        rc = subprocess.run([exe, fin, fout], capture_output=True).returncode
        assert rc == 0, f"galleries_eval exited with {rc}"
        return np.fromfile(fout).reshape(-1, NOUT).T
    return run


def _points(n=4000, seed=3):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(3, n))
    extra = np.array([[0, 0, 1], [0, 0, -1], [1, 0, 0], [-1, 0, 0],
                      [0.6, -1e-17, 0.8], [0.6, 1e-17, -0.8]], float).T
    p = np.concatenate([p, extra], axis=1)
    return p / np.linalg.norm(p, axis=0)


def close(a, b, rtol=1e-14):
    return np.max(np.abs(a - b)) <= rtol * max(np.max(np.abs(b)), 1.0)


def test_vorticity_functors_match_python(evaluate):
    x = _points()
    o = evaluate(x)
    assert close(o[0], G.solid_body_vorticity(x))
    assert close(o[1:4], G.solid_body_velocity(x))
    lon, lat = 0.7, -0.4
    ctr = np.array([np.cos(lon) * np.cos(lat), np.sin(lon) * np.cos(lat),
                    np.sin(lat)])
    assert close(o[4], G.gaussian_vortex_vorticity(x, ctr, 2.5, 3.0,
                                                   1.3 / (4 * np.pi)))
    assert close(o[5], G.rh54_vorticity(x, 0.1, -0.1))
    assert close(o[10], G.rh54_vorticity(x, 2 * np.pi / 14, 1.0))


def test_rh54_velocity_matches_python_and_is_tangent(evaluate):
    x = _points()
    o = evaluate(x)
    assert close(o[7:10], G.rh54_velocity(x, 0.1, -0.1))
    assert np.max(np.abs((o[7:10] * x).sum(axis=0))) < 1e-14


def test_rh54_laplacian_by_finite_differences(evaluate):
    """Laplace-Beltrami of f on the unit sphere = the 3-D Laplacian of its
    degree-0 extension f(x / |x|) at |x| = 1; central differences, h = 1e-3
    (truncation ~h^2 times fourth derivatives of a degree-5 harmonic)."""
    x = _points(300, seed=5)
    # keep away from the poles, where the longitude is singular
    x = x[:, np.abs(x[2]) < 0.95]
    lap = evaluate(x)[6]

    def f(p):
        return G.rh54_vorticity(p / np.linalg.norm(p, axis=0), 0.1, -0.1)
    h = 1e-3
    fd = -6 * f(x)
    for k in range(3):
        e = np.zeros((3, 1))
        e[k] = h
        fd = fd + f(x + e) + f(x - e)
    fd /= h * h
    scale = np.max(np.abs(lap))
    assert np.max(np.abs(fd - lap)) < 1e-4 * scale
    # and the eigenvalues: -2 on the degree-1 term, -30 on the degree-5 term
    z = x[2]
    zeta = G.rh54_vorticity(x, 0.1, -0.1)
    assert close(lap, -2 * (0.2 * z) - 30 * (zeta - 0.2 * z), 1e-13)


def test_test_case_2(evaluate):
    x = _points()
    o = evaluate(x)
    u0max, Omega, h0max = np.pi / 6, 2 * np.pi, 2.0
    st = G.tc2_state(x, u0=u0max, Omega=Omega, g=GRAVITY,
                     h0=h0max - (Omega * u0max + 0.5 * u0max ** 2) / GRAVITY)
    assert close(o[11:14], st["u"])
    assert close(o[14], st["zeta"])
    assert close(o[15], st["surface"])
    assert np.all(o[16] == 0)
    # zonal flow u0max cos(lat): |u| = u0max hypot(x, y), tangent
    assert close(np.linalg.norm(o[11:14], axis=0),
                 u0max * np.hypot(x[0], x[1]))
    assert np.max(np.abs((o[11:14] * x).sum(axis=0))) < 1e-15
    assert close(o[17], 10 + Omega * (2 * np.pi / 12) * (1 - x[2] ** 2))


def test_mixed_without_renormalisation_warns_once(tmp_path):
    """The C++ steppers' guard: LPM_MIXED without LPM_RK4_RENORMALIZE prints
    one warning per process to stderr; FP64, or mixed with the flag, none."""
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    src = tmp_path / "warn.cpp"
    src.write_text(
        '#include "lpm/lpm.hpp"\n'
        "int main() {\n"
        "  Lpm::warn_if_mixed_off_sphere(LPM_FP64, 0, \"A\");\n"
        "  Lpm::warn_if_mixed_off_sphere(LPM_MIXED, LPM_RK4_RENORMALIZE, \"B\");\n"
        "  Lpm::warn_if_mixed_off_sphere(LPM_MIXED, LPM_RK4_REFERENCE_COMPAT, \"C\");\n"
        "  Lpm::warn_if_mixed_off_sphere(LPM_MIXED, 0, \"D\");\n"
        "  return 0;\n}\n")
    exe = str(tmp_path / "warn")
    r = subprocess.run([HIPCC, "-std=c++17", f"-I{os.path.join(ROOT, 'include')}",
                        f"-I{os.path.join(ROOT, 'lpm_amd', 'cpp', 'include')}",
                        str(src), "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0
    lines = [l for l in r.stderr.splitlines() if l.startswith("lpm:")]
    assert len(lines) == 1 and lines[0].startswith("lpm: C with LPM_MIXED")
