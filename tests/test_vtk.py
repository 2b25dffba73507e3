"""The .vtp writer and reader (host I/O, no GPU)."""
import base64
import struct
import zlib

import numpy as np
import pytest

from lpm_amd.mesh import SphereMesh
from lpm_amd.vtk import VtkPolymeshInterface, diff_vtp, read_vtp


@pytest.fixture(scope="module", params=[(0, 2), (1, 2)], ids=["tri", "quad"])
def mesh(request):
    return SphereMesh.build(*request.param)


def _fill(vtk, mesh, rng):
    vtk.add_scalar_point_data(rng.normal(size=mesh.n_verts), "s")
    vtk.add_vector_point_data(rng.normal(size=(3, mesh.n_verts)), "v")
    vtk.add_scalar_cell_data(rng.normal(size=mesh.n_faces), "s")
    vtk.add_vector_cell_data(rng.normal(size=(mesh.n_faces, 3)), "v")


@pytest.mark.parametrize("encoding", ["binary", "ascii"])
def test_round_trip_is_exact(tmp_path, mesh, encoding):
    vtk = VtkPolymeshInterface(mesh)
    _fill(vtk, mesh, np.random.default_rng(1))
    f = str(tmp_path / f"m.{encoding}.vtp")
    vtk.write(f, encoding)
    r = read_vtp(f)
    np.testing.assert_array_equal(r["points"], mesh.vert_x.T)
    leaf = mesh.face_mask == 0
    np.testing.assert_array_equal(r["cell_data"]["area"], mesh.face_area[leaf])
    np.testing.assert_array_equal(r["cell_data"]["lag_crds"], mesh.face_x.T[leaf])
    np.testing.assert_array_equal(r["point_data"]["lag_crds"], mesh.vert_x.T)
    for kind in ("point_data", "cell_data"):
        for k in ("s", "v"):
            np.testing.assert_array_equal(r[kind][k], vtk.__dict__[kind][k])
    assert len(r["offsets"]) == mesh.n_leaves
    d = diff_vtp(f, f)
    assert all(v == (0.0, 0.0) for v in d.values())


def test_leaf_polygons_tile_the_sphere(mesh):
    # every leaf polygon is counter-clockwise seen from outside, and the leaf
    # polygons form a closed surface of genus 0: V - E + F = 2 and every edge
    # is shared by exactly two polygons
    vtk = VtkPolymeshInterface(mesh)
    P = vtk.polys
    X = mesh.vert_x.T
    a, b, c = X[P[:, 0]], X[P[:, 1]], X[P[:, 2]]
    assert (np.einsum("ij,ij->i", np.cross(b - a, c - a), a) > 0).all()
    k = P.shape[1]
    edges = np.sort(np.stack([P, np.roll(P, -1, axis=1)], -1).reshape(-1, 2), 1)
    uniq, cnt = np.unique(edges, axis=0, return_counts=True)
    assert (cnt == 2).all()
    used = np.unique(P)
    assert len(used) - len(uniq) + len(P) == 2
    assert len(used) == mesh.n_verts  # uniform refinement: every vertex used
    # leaf areas sum to the sphere's area
    assert abs(vtk.cell_data["area"].sum() - 4 * np.pi) < 1e-12 * 4 * np.pi
    assert k in (3, 4)


def test_height_field_scales_points(mesh):
    h = np.linspace(0, 0.1, mesh.n_verts)
    vtk = VtkPolymeshInterface(mesh, height_field=h)
    np.testing.assert_allclose(vtk.points, mesh.vert_x.T * (1 + h)[:, None],
                               rtol=0, atol=0)


def test_diff_reports_relative_errors(tmp_path, mesh):
    rng = np.random.default_rng(2)
    a = VtkPolymeshInterface(mesh)
    s = rng.normal(size=mesh.n_verts)
    a.add_scalar_point_data(s, "s")
    b = VtkPolymeshInterface(mesh)
    b.add_scalar_point_data(s * (1 + 1e-6), "s")
    fa, fb = str(tmp_path / "a.vtp"), str(tmp_path / "b.vtp")
    a.write(fa)
    b.write(fb)
    d = diff_vtp(fb, fa)
    assert d["point_data/s"][0] == pytest.approx(1e-6, rel=1e-6)
    assert d["point_data/s"][1] == pytest.approx(1e-6, rel=1e-3)
    assert d["points"] == (0.0, 0.0)


def _b64(b):
    return base64.b64encode(b).decode()


def test_reader_decodes_the_other_vtk_encodings(tmp_path):
    # one triangle with a Float64 cell array, written by hand in the four
    # encodings VTK's XML writer produces: inline base64 with UInt32 header,
    # zlib-compressed inline, appended raw, appended base64 + compressed
    pts = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    conn = np.array([0, 1, 2], np.int64)
    offs = np.array([3], np.int64)
    val = np.array([0.125])

    def inline32(a):
        raw = a.tobytes()
        return _b64(struct.pack("<I", len(raw))) + _b64(raw)

    def inline_z(a):
        raw = a.tobytes()
        comp = zlib.compress(raw)
        hdr = struct.pack("<4Q", 1, len(raw), len(raw), len(comp))
        return _b64(hdr) + _b64(comp)

    def doc(head, arrays, appended=""):
        names = ["Points", "connectivity", "offsets", "q"]
        body = {n: a for n, a in zip(names, arrays)}
        return (f'<?xml version="1.0"?>\n<VTKFile type="PolyData" version="1.0" '
                f'byte_order="LittleEndian" {head}><PolyData>'
                f'<Piece NumberOfPoints="3" NumberOfPolys="1"><PointData/>'
                f'<CellData>{body["q"]}</CellData>'
                f'<Points>{body["Points"]}</Points>'
                f'<Polys>{body["connectivity"]}{body["offsets"]}</Polys>'
                f'</Piece></PolyData>{appended}</VTKFile>')

    def da(name, t, ncomp, fmt, text="", offset=None):
        off = f' offset="{offset}"' if offset is not None else ""
        return (f'<DataArray type="{t}" Name="{name}" NumberOfComponents="{ncomp}" '
                f'format="{fmt}"{off}>{text}</DataArray>')

    files = {}
    files["inline32"] = doc('header_type="UInt32"', [
        da("Points", "Float64", 3, "binary", inline32(pts)),
        da("connectivity", "Int64", 1, "binary", inline32(conn)),
        da("offsets", "Int64", 1, "binary", inline32(offs)),
        da("q", "Float64", 1, "binary", inline32(val))])
    files["inline_zlib"] = doc(
        'header_type="UInt64" compressor="vtkZLibDataCompressor"', [
            da("Points", "Float64", 3, "binary", inline_z(pts)),
            da("connectivity", "Int64", 1, "binary", inline_z(conn)),
            da("offsets", "Int64", 1, "binary", inline_z(offs)),
            da("q", "Float64", 1, "binary", inline_z(val))])
    # appended raw, UInt64 headers
    blobs, arrs, off = [], [], 0
    for (n, t, c), a in zip([("Points", "Float64", 3), ("connectivity", "Int64", 1),
                             ("offsets", "Int64", 1), ("q", "Float64", 1)],
                            [pts, conn, offs, val]):
        raw = a.tobytes()
        arrs.append(da(n, t, c, "appended", offset=off))
        blobs.append(struct.pack("<Q", len(raw)) + raw)
        off += 8 + len(raw)
    raw_doc = doc('header_type="UInt64"', arrs,
                  '<AppendedData encoding="raw">_@@@</AppendedData>')
    pre, post = raw_doc.split("@@@")
    files["appended_raw"] = pre.encode() + b"".join(blobs) + post.encode()
    # appended base64, compressed
    chunks, arrs, off = [], [], 0
    for (n, t, c), a in zip([("Points", "Float64", 3), ("connectivity", "Int64", 1),
                             ("offsets", "Int64", 1), ("q", "Float64", 1)],
                            [pts, conn, offs, val]):
        s = inline_z(a)
        arrs.append(da(n, t, c, "appended", offset=off))
        chunks.append(s)
        off += len(s)
    files["appended_b64_zlib"] = doc(
        'header_type="UInt64" compressor="vtkZLibDataCompressor"', arrs,
        '<AppendedData encoding="base64">_' + "".join(chunks) + "</AppendedData>")
    for name, content in files.items():
        p = tmp_path / f"{name}.vtp"
        p.write_bytes(content if isinstance(content, bytes) else content.encode())
        r = read_vtp(str(p))
        np.testing.assert_array_equal(r["points"], pts, err_msg=name)
        np.testing.assert_array_equal(r["connectivity"], conn, err_msg=name)
        np.testing.assert_array_equal(r["offsets"], offs, err_msg=name)
        np.testing.assert_array_equal(r["cell_data"]["q"], val, err_msg=name)


@pytest.mark.gpu
def test_bve_and_swe_field_sets(tmp_path):
    from lpm_amd import galleries as G
    from lpm_amd.bve import BVESphere
    from lpm_amd.swe import SWE
    from lpm_amd.vtk import vtk_interface, vtk_mesh_interface
    sph = BVESphere(0, 2)
    sph.init_vorticity(lambda x: G.rh54_vorticity(x, 0.1, -0.1))
    sph.init_velocity()
    f = str(tmp_path / "bve.vtp")
    vtk_interface(sph).write(f)
    r = read_vtp(f)
    leaf = sph.mesh.face_mask == 0
    np.testing.assert_array_equal(r["point_data"]["velocity"],
                                  sph.velocity_verts.cpu().numpy().T)
    np.testing.assert_array_equal(r["cell_data"]["rel_vort"],
                                  sph.rel_vort_faces.cpu().numpy()[leaf])
    np.testing.assert_array_equal(r["points"], sph.vert_x.cpu().numpy().T)
    s = SWE(0, 2)
    s.init_surface(lambda x: G.tc2_state(x)["surface"])
    f = str(tmp_path / "swe.vtp")
    vtk_mesh_interface(s).write(f)
    r = read_vtp(f)
    np.testing.assert_array_equal(r["cell_data"]["surface_height"],
                                  s.surf_active.cpu().numpy()[leaf])
    assert "surface_laplaican" in r["cell_data"]
