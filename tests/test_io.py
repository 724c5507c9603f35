"""Mesh input (UNV) and output (VTU), CPU only.

The universal-file reader is pinned by a write/read round trip of generated
meshes with boundary groups, and, where the reference tree is present in the
build container, by the counts of a mesh its cases ship (cylinder2D's
scaled3.unv: 7,050 quads, 7,240 nodes) and a particle step of the CPU oracle
on it.  VTU output is parsed back with an XML parser."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from pfem2_amd import io
from pfem2_amd import mesh as M

SHIPPED = "/root/reference/cylinder2D/scaled3.unv"


def _edges(m, verts):
    """Boundary edges of a structured channel whose both ends are in `verts`."""
    s = set(int(v) for v in verts)
    out = []
    for c in m.cells:
        for a, b in ((c[0], c[1]), (c[2], c[3]), (c[0], c[2]), (c[1], c[3])):
            if a in s and b in s:
                out.append((int(a), int(b)))
    return sorted(set(out))


def test_unv_round_trip(tmp_path):
    m = M.quad_2d(7, 5, (0.0, 0.0), (2.0, 1.0), jitter=0.2, seed=3)
    left = np.nonzero(np.abs(m.vertices[:, 0]) < 1e-12)[0]
    groups = {"inlet": _edges(m, left)}
    p = str(tmp_path / "m.unv")
    io.write_unv(p, m, groups)
    r = io.read_unv(p)
    assert r.n_cells == m.n_cells and r.n_vertices == m.n_vertices
    assert np.array_equal(r.vertices, m.vertices)
    assert np.array_equal(r.cells, m.cells)
    assert np.array_equal(r.boundary_vertices["inlet"], np.sort(left))


def test_vtu_parses_back(tmp_path):
    m = M.channel_2d(6, 3, 3.0, 1.0)
    V = M.poiseuille_2d(m, 1.0, 1.0, 0.0)
    p = str(tmp_path / "f.vtu")
    io.write_mesh_vtu(p, m, {"velocity": V})
    piece = ET.parse(p).getroot().find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == m.n_vertices and int(piece.get("NumberOfCells")) == m.n_cells
    arr = piece.find("PointData/DataArray")
    vals = np.array(arr.text.split(), dtype=float).reshape(-1, 2)
    assert np.array_equal(vals, V.T)
    parts = {"x": m.vertices[:5], "v": V.T[:5], "id": np.arange(1, 6, dtype=np.uint32),
             "cell": np.zeros(5, dtype=np.int32)}
    io.write_particles_vtu(str(tmp_path / "p.vtu"), parts)
    piece = ET.parse(str(tmp_path / "p.vtu")).getroot().find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfCells")) == 5


@pytest.mark.skipif(not os.path.exists(SHIPPED), reason="reference meshes are only in the build container")
def test_shipped_mesh_reads_and_steps():
    from oracle_binding import Oracle
    m = io.read_unv(SHIPPED)
    assert m.n_cells == 7050 and m.n_vertices == 7240
    P = m.vertices[m.cells]
    a = P[:, 1] - P[:, 0]
    b = P[:, 2] - P[:, 0]
    assert np.all(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0] > 0)
    assert M.quad_convex(m.vertices, m.cells).all()
    V = M.uniform_field(m, [1.0, 0.1])
    o = Oracle(m, 0.02, 3, 3, [2, 2])
    o.upload_fields(V, V)
    o.seed_particles()
    assert o.n_global_particles() == 4 * m.n_cells
    o.step()
    counts = o.cell_counts()
    assert counts.min() >= 4 and counts.sum() == o.n_global_particles()
