"""VTK output (ik_write_vtu, the role of Ikarus::Vtk::Writer, ikarus/io/vtkwriter.hh:37-146): the written XML
UnstructuredGrid parsed back and checked against the mesh, the bound displacement and the device's stresses.
VTK's corner order is checked geometrically: every hex face ring 0-1-2-3 / 4-5-6-7 runs along element edges (its
diagonals are longer than its sides) and the cell has positive orientation."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import ikarus_amd as ia
from ikarus_amd import _capi as capi
from ikarus_amd import problems as pb
from ikarus_amd import vtk

pytestmark = pytest.mark.gpu


def _read(path):
    piece = ET.parse(path).getroot().find("UnstructuredGrid/Piece")
    arrays = {}
    for sec in ("Points", "Cells", "PointData", "CellData"):
        node = piece.find(sec)
        for da in node.findall("DataArray"):
            comps = int(da.get("NumberOfComponents", "1"))
            vals = np.array(da.text.split(), dtype=np.float64)
            arrays[(sec, da.get("Name"))] = vals.reshape(-1, comps) if comps > 1 else vals
    return int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells")), arrays


def _ring_ok(P):
    """the 4 corners form a ring: sides shorter than the diagonals"""
    d = lambda a, b: np.linalg.norm(P[a] - P[b])
    return max(d(0, 1), d(1, 2), d(2, 3), d(3, 0)) < min(d(0, 2), d(1, 3))


@pytest.mark.parametrize("tag", [vtk.DataTag.asCellData, vtk.DataTag.asPointData])
def test_vtu_hex(tmp_path, tag):
    p = pb.hex_block((3, 2, 2), jitter=0.08, seed=5)
    print(f"case: vtu hex tag={tag}", flush=True)
    dev = ia.Device(p)
    u = 0.01 * np.random.default_rng(1).standard_normal(p.ndof) * (1 - p.fixed)
    dev.set_requirement(u, 1.0)
    w = vtk.Writer(dev)
    w.addInterpolation(u, "displacement")
    w.addInterpolation(np.arange(p.ne, dtype=float), "element_id", vtk.DataTag.asCellData)
    w.addResult(vtk.PK2, tag)
    w.addResult(vtk.CAUCHY, tag)
    path = w.write(str(tmp_path / "hex"))
    npts, ncells, a = _read(path)
    assert npts == len(p.X) and ncells == p.ne
    assert np.array_equal(a[("Points", "Points")], p.X)
    conn = a[("Cells", "connectivity")].astype(np.int64).reshape(-1, 8)
    assert np.array_equal(conn[:, [0, 1, 3, 2, 4, 5, 7, 6]], p.conn)
    assert np.all(a[("Cells", "types")] == 12)
    assert np.array_equal(a[("Cells", "offsets")], 8 * np.arange(1, p.ne + 1))
    for c in conn:
        P = p.X[c]
        assert _ring_ok(P[:4]) and _ring_ok(P[4:])
        assert np.dot(np.cross(P[1] - P[0], P[3] - P[0]), P[4] - P[0]) > 0
    assert np.array_equal(a[("PointData", "displacement")], u.reshape(-1, 3))
    assert np.array_equal(a[("CellData", "element_id")], np.arange(p.ne))
    sec = "CellData" if tag == vtk.DataTag.asCellData else "PointData"
    for kind, name in ((vtk.PK2, "PK2Stress"), (vtk.CAUCHY, "CauchyStress")):
        if tag == vtk.DataTag.asCellData:
            ref = dev.stress_at([0.5, 0.5, 0.5], kind)
        else:
            s, cnt = np.zeros((len(p.X), 6)), np.zeros(len(p.X))
            for corner in range(8):
                v = dev.stress_at([corner & 1, (corner >> 1) & 1, (corner >> 2) & 1], kind)
                np.add.at(s, p.conn[:, corner], v)
                np.add.at(cnt, p.conn[:, corner], 1.0)
            ref = s / cnt[:, None]
        assert np.allclose(a[(sec, name)], ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())
    dev.close()


def test_vtu_plane_strain(tmp_path):
    p = pb.cantilever2d(n=(4, 3), seed=2)
    print("case: vtu quad", flush=True)
    dev = ia.Device(p)
    u = 0.01 * np.random.default_rng(2).standard_normal(p.ndof) * (1 - p.fixed)
    dev.set_requirement(u, 1.0)
    w = vtk.Writer(dev)
    w.addInterpolation(u, "displacement")
    w.addResult(vtk.PK2, vtk.DataTag.asCellData)
    npts, ncells, a = _read(w.write(str(tmp_path / "quad.vtu")))
    assert npts == len(p.X) and ncells == p.ne
    assert np.array_equal(a[("Points", "Points")][:, :2], p.X) and np.all(a[("Points", "Points")][:, 2] == 0)
    conn = a[("Cells", "connectivity")].astype(np.int64).reshape(-1, 4)
    assert np.array_equal(conn[:, [0, 1, 3, 2]], p.conn)
    assert np.all(a[("Cells", "types")] == 9)
    for c in conn:
        assert _ring_ok(p.X[c])
    d = a[("PointData", "displacement")]
    assert np.array_equal(d[:, :2], u.reshape(-1, 2)) and np.all(d[:, 2] == 0)
    assert np.allclose(a[("CellData", "PK2Stress")], dev.stress_at([0.5, 0.5], 0), rtol=1e-14, atol=1e-12)
    dev.close()


def test_vtu_rejects_bad_fields(tmp_path):
    p = pb.hex_block((2, 2, 2), seed=1)
    print("case: vtu rejections", flush=True)
    dev = ia.Device(p)
    w = vtk.Writer(dev)
    with pytest.raises(capi.IkError):
        w.addInterpolation(np.zeros(7), "bad")
    w.addResult(vtk.PK2, vtk.DataTag.asCellData)
    with pytest.raises(capi.IkError) as ei:   # no requirement bound yet
        w.write(str(tmp_path / "x"))
    assert ei.value.code == capi.IK_ENOTBOUND
    dev.close()
