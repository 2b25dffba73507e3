import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE_DIR = os.path.join(ROOT, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a HIP device (MI355X); runs the HIP kernels")


@pytest.fixture(scope="session")
def fixtures():
    with open(os.path.join(GOLDEN, "reference_fixtures.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle_fixtures():
    """Frozen oracle outputs (tests/golden/make_oracle_fixtures.py)."""
    with np.load(os.path.join(GOLDEN, "oracle_fixtures.npz")) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="session")
def ora():
    import oracle
    oracle.lib()
    return oracle


_MESHES = {}


def mesh(seed, depth):
    """Cached host mesh from the library's generator."""
    from lpm_amd.mesh import SphereMesh
    key = (seed, depth)
    if key not in _MESHES:
        _MESHES[key] = SphereMesh.build(seed, depth)
    return _MESHES[key]


def aos(x):
    """(3, n) -> contiguous (n, 3)."""
    return np.ascontiguousarray(np.asarray(x).T)


def rel_l2(a, b, w=None):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim == 2 and a.shape[0] == 3 and a.shape[1] != 3:
        a, b = a.T, b.T
    e = a - b
    if e.ndim == 2:
        en = np.sum(e * e, axis=1)
        bn = np.sum(b * b, axis=1)
    else:
        en, bn = e * e, b * b
    if w is None:
        w = np.ones_like(en)
    return float(np.sqrt(np.sum(en * w) / np.sum(bn * w)))


def rel_linf(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim == 2 and a.shape[0] == 3 and a.shape[1] != 3:
        a, b = a.T, b.T
    if a.ndim == 2:
        return float(np.max(np.linalg.norm(a - b, axis=1)) /
                     np.max(np.linalg.norm(b, axis=1)))
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
