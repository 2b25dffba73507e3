"""Adaptive refinement of the host quadtree mesh (PolyMesh2d::
divide_flagged_faces) and the tree invariants it must keep."""
import numpy as np
import pytest

from lpm_amd.mesh import PolyMesh2d, SphereMesh

FIELDS = ("vert_x", "face_x", "face_area", "face_mask", "face_parent",
          "face_verts", "face_level")


@pytest.mark.parametrize("seed", ["icos", "cube"])
@pytest.mark.parametrize("depth", [0, 1, 2])
def test_dividing_every_leaf_is_one_more_uniform_level(seed, depth):
    # divide_flagged_faces walks the faces in index order, as tree_init does
    # level by level, so flagging every leaf must reproduce depth + 1 exactly
    pm = PolyMesh2d(seed, depth, amr_limit=1)
    m = pm.sphere_mesh()
    n, lim = pm.divide_flagged_faces(m.face_mask == 0)
    assert (n, lim) == (m.n_leaves, False)
    a, b = pm.sphere_mesh(), SphereMesh.build(seed, depth + 1)
    for k in FIELDS:
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    assert (a.n_leaves, a.n_edges) == (b.n_leaves, b.n_edges)


def _check_tree(m):
    kids = np.zeros(m.n_faces, int)
    has_parent = m.face_parent >= 0
    np.add.at(kids, m.face_parent[has_parent], 1)
    # a face is masked exactly when it has been divided (4 children)
    assert set(np.unique(kids)) <= {0, 4}
    np.testing.assert_array_equal(m.face_mask != 0, kids == 4)
    assert m.n_leaves == int((m.face_mask == 0).sum())
    np.testing.assert_array_equal(m.face_level[has_parent],
                                  m.face_level[m.face_parent[has_parent]] + 1)
    leaf = m.face_mask == 0
    # leaves tile the sphere; divided faces carry zero area
    assert abs(m.face_area[leaf].sum() - 4 * np.pi) < 1e-12 * 4 * np.pi
    assert (m.face_area[~leaf] == 0).all()
    np.testing.assert_allclose(np.linalg.norm(m.vert_x, axis=0), 1, atol=1e-15)


@pytest.mark.parametrize("seed", ["icos", "cube"])
def test_local_refinement_keeps_tree_invariants_and_level_limit(seed):
    depth, limit = 2, 2
    pm = PolyMesh2d(seed, depth, amr_limit=limit)
    counts = []
    for rnd in range(4):
        m = pm.sphere_mesh()
        flags = (m.face_mask == 0) & (m.face_x[2] > 0.6)
        n0 = m.n_leaves
        n, lim = pm.divide_flagged_faces(flags)
        counts.append((n, lim))
        a = pm.sphere_mesh()
        _check_tree(a)
        within = flags & (m.face_level <= depth + limit)
        assert n == int(within.sum())
        assert lim == bool((flags & ~within).any())
        assert a.n_leaves == n0 + 3 * n
        # faces present before keep their arrays (append-only growth)
        np.testing.assert_array_equal(a.face_x[:, :m.n_faces], m.face_x)
        np.testing.assert_array_equal(a.vert_x[:, :m.n_verts], m.vert_x)
    # a face may be divided while its level is <= depth + amr_limit = 4, so
    # rounds 1-3 refine the cap to levels 3, 4, 5 and round 4 meets the limit
    assert all(c[0] > 0 and not c[1] for c in counts[:3])
    assert counts[3][1]
    m = pm.sphere_mesh()
    assert m.face_level.max() == depth + limit + 1


def test_shared_edges_are_divided_once():
    # two neighbouring faces divided one after the other share the midpoint
    # vertex of their common edge: vertex count grows by 3 + 2, not 3 + 3
    pm = PolyMesh2d("icos", 0, amr_limit=3)
    m = pm.sphere_mesh()
    fv = m.face_verts
    a = 0
    b = next(j for j in range(1, m.n_faces)
             if len(set(fv[a]) & set(fv[j])) == 2)
    flags = np.zeros(m.n_faces, np.uint8)
    flags[[a, b]] = 1
    pm.divide_flagged_faces(flags)
    assert pm.sizes()[0] == m.n_verts + 5


def test_already_divided_and_allocation_limit():
    pm = PolyMesh2d("cube", 1, amr_limit=3, max_faces=40)
    m = pm.sphere_mesh()
    # the reference counts a flagged, already divided face but leaves it
    flags = np.zeros(m.n_faces, np.uint8)
    flags[0] = 1  # seed face 0 is divided at depth 1
    assert pm.divide_flagged_faces(flags) == (1, False)
    assert pm.sizes() == (m.n_verts, m.n_faces, m.n_leaves, m.n_edges)
    # room for (40 - 30) / 4 = 2 divisions: 3 flagged faces divide nothing
    leaves = np.flatnonzero(m.face_mask == 0)
    flags[:] = 0
    flags[leaves[:3]] = 1
    assert pm.divide_flagged_faces(flags)[0] == -1
    assert pm.sizes()[1] == m.n_faces
    flags[leaves[2]] = 0
    assert pm.divide_flagged_faces(flags) == (2, False)
    assert pm.sizes()[1] == m.n_faces + 8


def test_argument_errors():
    from lpm_amd import _abi
    pm = PolyMesh2d("icos", 1)
    with pytest.raises(_abi.LpmError):
        pm.divide_flagged_faces(np.zeros(3, np.uint8))
