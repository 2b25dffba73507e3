"""The Morton keys of the mixed paths' target order (lpm_amd/csrc/order.cpp):
the bit interleaving against a bit-by-bit construction (in the driver), the
quantisation of the coordinates, and that ordering mesh vertices by the key
puts each wave of 64 close together, which is what the order is for.

tests/order_keys.cpp includes the key functions (order_keys.hpp) in a host
program (hipcc, host only); no GPU is needed.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from lpm_amd.mesh import SphereMesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def keys(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("order")
    exe = str(d / "order_keys")
    r = subprocess.run([HIPCC, "-std=c++17", "-O2",
                        f"-I{os.path.join(ROOT, 'include')}",
                        f"-I{os.path.join(ROOT, 'lpm_amd', 'csrc')}",
                        os.path.join(ROOT, "tests", "order_keys.cpp"),
                        "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

    def run(x):
        fin, fout = str(d / "in.bin"), str(d / "out.bin")
        np.ascontiguousarray(x.T, dtype=np.float64).tofile(fin)
        rc = subprocess.run([exe, fin, fout], capture_output=True).returncode
        assert rc == 0, f"order_keys exited with {rc}"
        return np.fromfile(fout, dtype=np.uint64)
    return run


def _naive_key(x):
    q = np.clip(np.floor((x + 1.0) * 0.5 * (1 << 20)), 0, (1 << 20) - 1)
    q = q.astype(np.uint64)
    k = np.zeros(x.shape[1], dtype=np.uint64)
    for b in range(20):
        for d in range(3):
            k |= ((q[d] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + d)
    return k


def test_keys_interleave_quantised_coordinates(keys):
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, size=(3, 5000))
    x[:, :6] = [[-1, 1, 0, 0.5, -0.999999, 1.5],
                [1, -1, 0, -0.5, 2.0, -3.0],
                [0, 0, 0, 0.25, 1e-300, 0.0]]
    assert np.array_equal(keys(x), _naive_key(np.clip(x, -1, 1)))


def _near_subtiles(x, src):
    """Per wave of 64 targets x, the number of 64-source sub-tiles of src
    (tree order) with a target within three sub-tile radii of the
    centroid: the sub-tiles the mixed kernels would sum on the exact
    path."""
    ns = src.shape[1] // 64
    sub = src[:, :ns * 64].reshape(3, ns, 64)
    cen = sub.mean(2)
    r2 = ((sub - cen[:, :, None]) ** 2).sum(0).max(1)
    out = []
    for w in range(0, x.shape[1], 64):
        p = x[:, w:w + 64]
        d2 = ((p[:, :, None] - cen[:, None, :]) ** 2).sum(0)
        out.append(int((d2 <= 9 * r2[None]).any(0).sum()))
    return np.array(out)


def test_morton_ordered_vertex_waves_meet_few_near_subtiles(keys):
    m = SphereMesh.build("icos", 5)
    src = m.face_x[:, m.face_mask == 0]
    order = np.argsort(keys(m.vert_x), kind="stable")
    mesh_order = _near_subtiles(m.vert_x, src)
    morton = _near_subtiles(m.vert_x[:, order], src)
    # refinement order: the first vertex waves span the sphere (320 of the
    # 320 sub-tiles near at level 5); Morton order keeps every wave local
    assert mesh_order.max() > 200
    assert morton.max() < 0.3 * mesh_order.max()
    assert morton.mean() < 0.8 * mesh_order.mean()
    # a permutation of the vertices
    assert np.array_equal(np.sort(order), np.arange(m.n_verts))
