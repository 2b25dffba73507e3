"""CPU tests of the input generator and of the C-ABI library surface.

Mesh invariants follow tests/lpm_polymesh_tests.cpp:76-117 and
tests/lpm_mesh_seed_tests.cpp:70-83 of the reference (counts equal
MeshSeed::set_max_allocations, total area 4 pi); seed geometry is checked
against the reference's seed files (fixture) and against closed forms.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, mesh

from lpm_amd import _abi
from lpm_amd.mesh import max_allocations


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("depth", [0, 1, 3, 5])
def test_counts_match_max_allocations(seed, depth):
    m = mesh(seed, depth)
    nv, ne, nf = max_allocations(seed, depth)
    assert m.n_verts == nv
    assert m.n_faces == nf
    assert m.n_edges == ne
    base = 20 if seed == 0 else 6
    assert m.n_leaves == base * 4 ** depth
    assert m.n_verts == (2 + 10 * 4 ** depth if seed == 0 else 2 + 6 * 4 ** depth)


def test_sizes_of_benchmark_meshes():
    # the configurations the benchmark runs at (icos L4, L7; cube L8)
    nv, _, nf = max_allocations(0, 4)
    assert (nv, nf) == (2562, 6820)
    nv, _, nf = max_allocations(0, 7)
    assert (nv, nf) == (163842, 436900)
    nv, _, nf = max_allocations(1, 8)
    assert (nv, nf) == (393218, 524286)


@pytest.mark.parametrize("seed,tol_key", [(0, "icos_area_tol"), (1, "cube_area_tol")])
def test_total_area(fixtures, seed, tol_key):
    m = mesh(seed, fixtures["mesh"]["tree_lev"])
    assert abs(m.face_area.sum() - 4 * np.pi) < fixtures["mesh"][tol_key]
    for d in (0, 5):
        m = mesh(seed, d)
        assert abs(m.face_area.sum() - 4 * np.pi) < 1e-12


def test_seed_coordinates_match_reference_seed_files(fixtures):
    for seed, key, nverts in [(0, "icos_crds", 12), (1, "cube_crds", 8)]:
        m = mesh(seed, 0)
        crds = np.asarray(fixtures["seeds"][key])
        np.testing.assert_array_equal(m.vert_x.T, crds[:nverts])
        np.testing.assert_array_equal(m.face_x.T, crds[nverts:])


def test_seed_geometry_closed_form():
    m = mesh(0, 0)
    v = m.vert_x.T
    # icosahedron: poles and two rings at latitude +-atan(1/2), 72 deg apart
    lat = np.arctan(0.5)
    np.testing.assert_allclose(v[0], [0, 0, 1], atol=1e-16)
    np.testing.assert_allclose(v[11], [0, 0, -1], atol=1e-16)
    for k in range(5):
        lon_u = np.deg2rad(36 + 72 * k)
        lon_l = np.deg2rad(72 * k)
        np.testing.assert_allclose(
            v[1 + k], [np.cos(lat) * np.cos(lon_u), np.cos(lat) * np.sin(lon_u),
                       np.sin(lat)], atol=2e-16)
        np.testing.assert_allclose(
            v[6 + k], [np.cos(lat) * np.cos(lon_l), np.cos(lat) * np.sin(lon_l),
                       -np.sin(lat)], atol=2e-16)
    # face centres are the normalised vertex means
    for f in range(20):
        c = v[m.face_verts[f]].mean(0)
        np.testing.assert_allclose(m.face_x[:, f], c / np.linalg.norm(c), atol=3e-16)
    mc = mesh(1, 0)
    np.testing.assert_allclose(np.abs(mc.vert_x), 1 / np.sqrt(3), atol=1e-16)


@pytest.mark.parametrize("seed", [0, 1])
def test_refinement_structure(seed):
    m = mesh(seed, 4)
    # divided faces precede the leaves; leaves are the last n_leaves faces
    assert np.all(m.face_mask[m.leaf_slice] == 0)
    assert np.all(m.face_mask[:m.n_faces - m.n_leaves] == 1)
    assert np.all(m.face_area[m.face_mask == 1] == 0)
    assert np.all(m.face_area[m.face_mask == 0] > 0)
    # unit sphere
    np.testing.assert_allclose(np.linalg.norm(m.vert_x, axis=0), 1, atol=3e-16)
    np.testing.assert_allclose(np.linalg.norm(m.face_x, axis=0), 1, atol=3e-16)
    # children of face f are 4 consecutive faces whose parent is f
    par = m.face_parent
    nseed = 20 if seed == 0 else 6
    assert np.all(par[:nseed] == -1)
    kids = par[nseed:].reshape(-1, 4)
    assert np.all(kids == kids[:, :1])
    # leaf centres equal the normalised mean of their corner vertices
    leaf = np.arange(m.n_faces)[m.leaf_slice][:200]
    c = m.vert_x[:, m.face_verts[leaf]].mean(axis=2)
    c /= np.linalg.norm(c, axis=0)
    np.testing.assert_allclose(m.face_x[:, leaf], c, atol=3e-16)
    # 4^k consecutive leaves are the leaves of one level-(L-k) face, so
    # packing tiles are spatially compact (the mixed-precision kernel's
    # tile-relative offsets rely on it)
    m6 = mesh(seed, 6)
    lx = m6.face_x[:, m6.leaf_slice]
    par = m6.face_parent[m6.leaf_slice]
    assert np.all(par.reshape(-1, 4) == par.reshape(-1, 4)[:, :1])
    for t0 in range(0, lx.shape[1], 256 * 37):
        t = lx[:, t0:t0 + 256]
        assert np.max(np.linalg.norm(t - t[:, :1], axis=0)) < 0.6


# ----------------------------------------------------------------- ABI
def _header_symbols():
    syms = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            syms |= set(re.findall(r"\b(lpm_[a-z0-9_]+)\s*\(", txt))
    return syms


def test_library_exports_every_declared_symbol():
    L = _abi.lib()
    syms = _header_symbols()
    assert len(syms) > 10
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding covers every declared entry point
    assert set(_abi.exported_symbols()) == syms


def test_abi_version_matches_header():
    txt = open(os.path.join(ROOT, "include", "lpm_mi355x.h")).read()
    v = int(re.search(r"#define LPM_ABI_VERSION (\d+)", txt).group(1))
    assert _abi.lib().lpm_abi_version() == v


def test_constants_match_header():
    txt = open(os.path.join(ROOT, "include", "lpm_mi355x.h")).read()
    for name in ["LPM_FLAG_CHECK_FINITE", "LPM_FLAG_SKIP_MASKED_TARGETS",
                 "LPM_RK4_REFERENCE_COMPAT", "LPM_RK4_RENORMALIZE",
                 "LPM_RK4_USE_GRAPH", "LPM_SEED_ICOS_TRI",
                 "LPM_SEED_CUBED_SPHERE", "LPM_KERNEL_BVE_VEL_F64",
                 "LPM_KERNEL_BVE_VEL_MIXED"]:
        m = re.search(r"#define %s (0x[0-9a-fA-F]+u?|\d+)" % name, txt)
        val = int(m.group(1).rstrip("u"), 0)
        assert getattr(_abi, name) == val, name
    for name, val in [("LPM_OK", 0), ("LPM_ERR_ARG", -1), ("LPM_ERR_HIP", -2),
                      ("LPM_ERR_NONFINITE", -3)]:
        assert re.search(r"%s = %d" % (name, val), txt)


def test_mesh_entry_point_errors():
    L = _abi.lib()
    h = ctypes.c_void_p()
    assert L.lpm_mesh_create(7, 2, ctypes.byref(h)) == _abi.LPM_ERR_ARG
    assert b"seed" in L.lpm_last_error(None)
    assert L.lpm_mesh_create(0, -1, ctypes.byref(h)) == _abi.LPM_ERR_ARG
    assert L.lpm_mesh_create(0, 2, None) == _abi.LPM_ERR_ARG


def test_null_context_is_rejected():
    L = _abi.lib()
    rc = L.lpm_bve_velocity(None, None, 1, 1, 0, None, 1, 1, None, None, None,
                            0, 0, 0.0, 0, 0, None, 1, 1, None)
    assert rc == _abi.LPM_ERR_ARG
    assert L.lpm_timing_enable(None, 1) == _abi.LPM_ERR_ARG


def test_mixed_stepper_without_renormalize_warns():
    """Mixed sums assume positions on the unit sphere: a mixed stepper that
    does not renormalise warns; FP64 or renormalised mixed steppers do not."""
    import warnings
    from lpm_amd.incompressible2d import Incompressible2DRK2
    with pytest.warns(UserWarning, match="unit sphere"):
        Incompressible2DRK2(0.01, None, precision="mixed")
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        Incompressible2DRK2(0.01, None, precision="mixed", renormalize=True)
        Incompressible2DRK2(0.01, None)
