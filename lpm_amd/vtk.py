"""VTK XML PolyData (.vtp) output and field diffing for sphere particle meshes.

Mirrors ``VtkPolymeshInterface<Seed>`` (src/vtk/lpm_vtk_io.hpp:35-80,
src/vtk/lpm_vtk_io_impl.hpp) and the field sets of ``vtk_interface(BVESphere)``
(src/lpm_bve_sphere_impl.hpp:106-125) and ``vtk_mesh_interface(SWE)``
(src/lpm_swe_impl.hpp:490-530), written without the VTK library:

* points are the vertices' physical coordinates, one tuple per vertex;
* one polygon per leaf face (faces that were not divided), its corners in the
  mesh's vertex order; cell arrays hold one tuple per leaf face, in face
  order;
* the constructor adds the cell array ``area`` and the point and cell arrays
  ``lag_crds`` (the particles' initial positions).

With a height field h the points are (1 + h) x.  (The reference's sphere
branch of that constructor computes these coordinates but never inserts
them, src/vtk/lpm_vtk_io_impl.hpp:64-72, so its file has no points; here
they are written.)

``read_vtp`` reads the XML PolyData encodings VTK writes (ascii, inline
base64, appended raw or base64, optionally zlib-compressed, UInt32 or UInt64
headers), and ``diff_vtp`` compares two files array by array, e.g. a file
from this package against one the reference wrote for the same mesh.
"""
from __future__ import annotations

import base64
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np

_VTK_TYPES = {
    "Float64": np.float64, "Float32": np.float32, "Int64": np.int64,
    "Int32": np.int32, "Int16": np.int16, "Int8": np.int8,
    "UInt64": np.uint64, "UInt32": np.uint32, "UInt16": np.uint16,
    "UInt8": np.uint8,
}
_NP_TO_VTK = {np.dtype(v).str[1:]: k for k, v in _VTK_TYPES.items()}


def _host(a) -> np.ndarray:
    """numpy copy of a numpy array or a (device) torch tensor."""
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def _rows(v, n: int) -> np.ndarray:
    """(n, k) rows of a vector field given as (n, k) or as (k, n) (the
    component-major layout of this package's device fields)."""
    v = _host(v)
    if v.ndim == 1:
        return v.reshape(n, 1)
    if v.shape[0] == n:
        return v
    if v.shape[1] == n:
        return v.T
    raise ValueError(f"vector field of shape {v.shape} does not match {n} tuples")


class VtkPolymeshInterface:
    """Points, leaf-face polygons and named point/cell arrays of one mesh.

    ``pm`` is a ``SphereMesh`` or a particle object holding one (``.mesh``)
    with current positions ``vert_x`` (BVESphere, SWE, Incompressible2D).
    """

    def __init__(self, pm, height_field=None):
        self._pm = pm
        mesh = getattr(pm, "mesh", pm)
        if getattr(mesh, "face_verts", None) is None:
            raise ValueError("the mesh has no face corner indices "
                             "(concatenated ensembles cannot be written)")
        self.mesh = mesh
        self.n_verts = mesh.n_verts
        self.leaf = np.flatnonzero(_host(mesh.face_mask) == 0)
        self.n_cells = len(self.leaf)
        self.point_data: dict[str, np.ndarray] = {}
        self.cell_data: dict[str, np.ndarray] = {}
        self._height = None if height_field is None else _host(height_field)
        self.points = self._make_points()
        self.polys = np.ascontiguousarray(
            _host(mesh.face_verts)[self.leaf].astype(np.int64))
        area = getattr(pm, "face_area", mesh.face_area)
        self.add_scalar_cell_data(area, "area")
        self.add_vector_point_data(mesh.vert_x, "lag_crds")
        self.add_vector_cell_data(mesh.face_x, "lag_crds")

    def _make_points(self) -> np.ndarray:
        x = _rows(getattr(self._pm, "vert_x", self.mesh.vert_x), self.n_verts)
        x = np.array(x, dtype=np.float64)
        if self._height is not None:
            x *= (1.0 + self._height.reshape(-1, 1))
        return x

    def update_positions(self):
        """Re-read the vertices' physical coordinates
        (src/vtk/lpm_vtk_io_impl.hpp:130-134)."""
        self.points = self._make_points()

    def add_scalar_point_data(self, s, name: str = ""):
        self.point_data[self._name(s, name)] = _rows(s, self.n_verts)[:, 0].copy()

    def add_vector_point_data(self, v, name: str = ""):
        self.point_data[self._name(v, name)] = np.array(_rows(v, self.n_verts))

    def add_scalar_cell_data(self, s, name: str = ""):
        a = _rows(s, self.mesh.n_faces)[:, 0]
        self.cell_data[self._name(s, name)] = a[self.leaf].copy()

    def add_vector_cell_data(self, v, name: str = ""):
        a = _rows(v, self.mesh.n_faces)
        self.cell_data[self._name(v, name)] = np.array(a[self.leaf])

    def add_tracers(self, point_tracers: dict, cell_tracers: dict):
        """Named tracers, one point and one cell array each
        (src/vtk/lpm_vtk_io_impl.hpp:100-109)."""
        if set(point_tracers) != set(cell_tracers):
            raise ValueError("point and cell tracers must have the same names")
        for k in point_tracers:
            self.add_scalar_point_data(point_tracers[k], k)
            self.add_scalar_cell_data(cell_tracers[k], k)

    @staticmethod
    def _name(a, name):
        if name:
            return name
        label = getattr(a, "label", None)
        if not label:
            raise ValueError("an array without a label needs a name")
        return label

    def write(self, filename: str, encoding: str = "binary"):
        """Write the .vtp file; ``encoding`` is "binary" (inline base64, a
        UInt64 byte-count header per array) or "ascii"."""
        if encoding not in ("binary", "ascii"):
            raise ValueError("encoding must be 'binary' or 'ascii'")
        offsets = np.arange(1, self.n_cells + 1, dtype=np.int64) * self.polys.shape[1]
        out = ['<?xml version="1.0"?>',
               '<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" '
               'header_type="UInt64">',
               "  <PolyData>",
               f'    <Piece NumberOfPoints="{self.n_verts}" NumberOfVerts="0" '
               f'NumberOfLines="0" NumberOfStrips="0" NumberOfPolys="{self.n_cells}">']
        out.append("      <PointData>")
        out += [_data_array(k, v, encoding) for k, v in self.point_data.items()]
        out.append("      </PointData>")
        out.append("      <CellData>")
        out += [_data_array(k, v, encoding) for k, v in self.cell_data.items()]
        out.append("      </CellData>")
        out.append("      <Points>")
        out.append(_data_array("Points", self.points, encoding))
        out.append("      </Points>")
        out.append("      <Polys>")
        out.append(_data_array("connectivity", self.polys.reshape(-1), encoding))
        out.append(_data_array("offsets", offsets, encoding))
        out.append("      </Polys>")
        out += ["    </Piece>", "  </PolyData>", "</VTKFile>", ""]
        with open(filename, "w") as f:
            f.write("\n".join(out))


def _data_array(name, a, encoding) -> str:
    a = np.ascontiguousarray(a)
    if a.dtype.byteorder == ">":
        a = a.astype(a.dtype.newbyteorder("<"))
    ncomp = 1 if a.ndim == 1 else a.shape[1]
    vt = _NP_TO_VTK[a.dtype.str[1:]]
    head = (f'        <DataArray type="{vt}" Name="{name}" '
            f'NumberOfComponents="{ncomp}" format="{encoding}">')
    if encoding == "ascii":
        flat = a.reshape(-1)
        if a.dtype.kind == "f":
            body = " ".join(repr(float(x)) for x in flat)
        else:
            body = " ".join(str(int(x)) for x in flat)
    else:
        raw = a.tobytes()
        # header and data are encoded as separate base64 blocks, as VTK does
        body = (base64.b64encode(struct.pack("<Q", len(raw))).decode() +
                base64.b64encode(raw).decode())
    return f"{head}{body}</DataArray>"


# --------------------------------------------------------------- reading --
def _decode_inline(text: str, dtype, header: str, compressed: bool) -> bytes:
    hsize = 8 if header == "UInt64" else 4
    hfmt = "<Q" if hsize == 8 else "<I"
    text = "".join(text.split())
    if compressed:
        # header [nblocks, block size, last block size, csize_0..] in its
        # own base64 block, then the compressed blocks
        first = base64.b64decode(text[: 4 * -(-3 * hsize // 3)])
        nblocks = struct.unpack(hfmt, first[:hsize])[0]
        hlen = (3 + nblocks) * hsize
        hchars = 4 * -(-hlen // 3)
        hdr = struct.unpack("<" + hfmt[1] * (3 + nblocks),
                            base64.b64decode(text[:hchars])[:hlen])
        data = base64.b64decode(text[hchars:])
        return _inflate(hdr, data, hsize)
    hchars = 4 * -(-hsize // 3)
    nbytes = struct.unpack(hfmt, base64.b64decode(text[:hchars])[:hsize])[0]
    rest = base64.b64decode(text[hchars:])
    if len(rest) >= nbytes:
        return rest[:nbytes]
    # header and data encoded as one base64 stream
    allb = base64.b64decode(text)
    return allb[hsize:hsize + nbytes]


def _inflate(hdr, data: bytes, hsize: int) -> bytes:
    nblocks = hdr[0]
    sizes = hdr[3:3 + nblocks]
    out, pos = [], 0
    for cs in sizes:
        out.append(zlib.decompress(data[pos:pos + cs]))
        pos += cs
    return b"".join(out)


def _read_appended(blob: bytes, offset: int, header: str, compressed: bool,
                   base64_enc: bool) -> bytes:
    hsize = 8 if header == "UInt64" else 4
    hfmt = "<Q" if hsize == 8 else "<I"
    if base64_enc:
        text = blob.decode("ascii")
        rest = text[offset:]
        if compressed:
            first = base64.b64decode(rest[: 4 * -(-hsize // 3)])
            nblocks = struct.unpack(hfmt, first[:hsize])[0]
            hlen = (3 + nblocks) * hsize
            hchars = 4 * -(-hlen // 3)
            hdr = struct.unpack("<" + hfmt[1] * (3 + nblocks),
                                base64.b64decode(rest[:hchars])[:hlen])
            csum = sum(hdr[3:])
            data = base64.b64decode(rest[hchars:hchars + 4 * -(-csum // 3)])
            return _inflate(hdr, data, hsize)
        hchars = 4 * -(-hsize // 3)
        nbytes = struct.unpack(hfmt, base64.b64decode(rest[:hchars])[:hsize])[0]
        return base64.b64decode(rest[hchars:hchars + 4 * -(-nbytes // 3)])[:nbytes]
    p = offset
    if compressed:
        nblocks = struct.unpack(hfmt, blob[p:p + hsize])[0]
        hdr = struct.unpack("<" + hfmt[1] * (3 + nblocks),
                            blob[p:p + (3 + nblocks) * hsize])
        p += (3 + nblocks) * hsize
        return _inflate(hdr, blob[p:p + sum(hdr[3:])], hsize)
    nbytes = struct.unpack(hfmt, blob[p:p + hsize])[0]
    return blob[p + hsize:p + hsize + nbytes]


def read_vtp(filename: str) -> dict:
    """Points, polygons and point/cell arrays of a .vtp file.

    Returns {"points": (n, 3), "connectivity", "offsets", "point_data":
    {name: array}, "cell_data": {name: array}}; vector arrays are (n, k).
    """
    with open(filename, "rb") as f:
        raw = f.read()
    blob = b""
    i = raw.find(b"<AppendedData")
    if i >= 0:
        j = raw.index(b">", i)
        k = raw.index(b"_", j)
        end = raw.rindex(b"</AppendedData>")
        blob = raw[k + 1:end].rstrip()
        app_enc = ET.fromstring(raw[i:j + 1].decode() + "</AppendedData>").get(
            "encoding", "raw")
        raw = raw[:k] + raw[end:]
    root = ET.fromstring(raw)
    header = root.get("header_type", "UInt32")
    compressed = root.get("compressor") is not None
    if root.get("byte_order", "LittleEndian") != "LittleEndian":
        raise ValueError("big-endian .vtp files are not supported")

    def arr(e):
        dt = np.dtype(_VTK_TYPES[e.get("type")])
        ncomp = int(e.get("NumberOfComponents", "1"))
        fmt = e.get("format", "ascii")
        if fmt == "ascii":
            a = np.array((e.text or "").split(), dtype=dt)
        elif fmt == "binary":
            a = np.frombuffer(_decode_inline(e.text or "", dt, header,
                                             compressed), dtype=dt)
        elif fmt == "appended":
            a = np.frombuffer(_read_appended(blob, int(e.get("offset")), header,
                                             compressed, app_enc == "base64"),
                              dtype=dt)
        else:
            raise ValueError(f"unknown DataArray format {fmt}")
        return a.reshape(-1, ncomp) if ncomp > 1 else a.copy()

    piece = root.find("PolyData/Piece")
    out = {"point_data": {}, "cell_data": {}}
    for e in piece.findall("PointData/DataArray"):
        out["point_data"][e.get("Name")] = arr(e)
    for e in piece.findall("CellData/DataArray"):
        out["cell_data"][e.get("Name")] = arr(e)
    out["points"] = arr(piece.find("Points/DataArray"))
    polys = {e.get("Name"): arr(e) for e in piece.findall("Polys/DataArray")}
    out["connectivity"] = polys.get("connectivity", np.zeros(0, np.int64))
    out["offsets"] = polys.get("offsets", np.zeros(0, np.int64))
    return out


def diff_vtp(a, b) -> dict:
    """Relative L2 and L-infinity differences (``a`` against ``b``) of the
    points and of every point/cell array present in both files; ``a`` and
    ``b`` are file names or ``read_vtp`` results.  Connectivity must agree
    exactly (same mesh)."""
    a = read_vtp(a) if isinstance(a, str) else a
    b = read_vtp(b) if isinstance(b, str) else b
    if not (np.array_equal(a["connectivity"], b["connectivity"]) and
            np.array_equal(a["offsets"], b["offsets"])):
        raise ValueError("the files' polygons differ: not the same mesh")

    def err(x, y):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        d = x - y
        ny = np.sqrt((y * y).sum())
        my = np.abs(y).max() if y.size else 0.0
        return (float(np.sqrt((d * d).sum()) / ny) if ny > 0 else float(np.sqrt((d * d).sum())),
                float(np.abs(d).max() / my) if my > 0 else float(np.abs(d).max(initial=0.0)))

    out = {"points": err(a["points"], b["points"])}
    for kind in ("point_data", "cell_data"):
        for k in sorted(set(a[kind]) & set(b[kind])):
            out[f"{kind}/{k}"] = err(a[kind][k], b[kind][k])
    return out


# ------------------------------------------------------ field collections --
def vtk_interface(bve) -> VtkPolymeshInterface:
    """The arrays of vtk_interface(BVESphere)
    (src/lpm_bve_sphere_impl.hpp:106-125)."""
    vtk = VtkPolymeshInterface(bve)
    vtk.add_scalar_point_data(bve.rel_vort_verts, "rel_vort")
    vtk.add_scalar_point_data(bve.abs_vort_verts, "abs_vort")
    vtk.add_scalar_point_data(bve.stream_fn_verts, "stream_fn")
    vtk.add_vector_point_data(bve.velocity_verts, "velocity")
    vtk.add_scalar_cell_data(bve.rel_vort_faces, "rel_vort")
    vtk.add_scalar_cell_data(bve.abs_vort_faces, "abs_vort")
    vtk.add_scalar_cell_data(bve.stream_fn_faces, "stream_fn")
    vtk.add_vector_cell_data(bve.velocity_faces, "velocity")
    return vtk


def vtk_mesh_interface(swe) -> VtkPolymeshInterface:
    """The arrays of vtk_mesh_interface(SWE) (src/lpm_swe_impl.hpp:490-530)
    that this package's SWE holds, under the reference's array names (the
    names are the reference's field labels, :62-92, including its spelling
    "surface_laplaican" of the face array).  Bottom height is zero."""
    vtk = VtkPolymeshInterface(swe)
    zv = np.zeros(swe.mesh.n_verts)
    zf = np.zeros(swe.mesh.n_faces)
    vtk.add_scalar_point_data(swe.rel_vort_passive, "relative_vorticity")
    vtk.add_scalar_point_data(swe.div_passive, "divergence")
    vtk.add_scalar_point_data(swe.surf_passive, "surface_height")
    vtk.add_scalar_point_data(swe.surf_lap_passive, "surface_laplacian")
    vtk.add_scalar_point_data(swe.depth_passive, "depth")
    vtk.add_scalar_point_data(swe.double_dot_passive, "double_dot")
    vtk.add_scalar_point_data(zv, "bottom_height")
    vtk.add_vector_point_data(swe.velocity_passive, "velocity")
    vtk.add_scalar_cell_data(swe.rel_vort_active, "relative_vorticity")
    vtk.add_scalar_cell_data(swe.div_active, "divergence")
    vtk.add_scalar_cell_data(swe.surf_active, "surface_height")
    vtk.add_scalar_cell_data(swe.depth_active, "depth")
    vtk.add_scalar_cell_data(swe.surf_lap_active, "surface_laplaican")
    vtk.add_scalar_cell_data(swe.double_dot_active, "double_dot")
    vtk.add_scalar_cell_data(zf, "bottom_height")
    vtk.add_vector_cell_data(swe.velocity_active, "velocity")
    vtk.add_scalar_cell_data(swe.mass_active, "mass")
    return vtk
