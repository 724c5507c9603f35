"""Mesh input and result output for drivers of the particle stage.

read_unv   I-DEAS universal file (the format of the meshes the reference's
           cases read through deal.II's GridIn::read_unv; its repairMesh tool
           edits the same sections, repairMesh/repairMesh.cc:22-124):
           dataset 2411 (nodes), 2412 (elements: 2-D linear quads, type 44,
           and boundary lines, type 11), 2467 (groups of elements or nodes).
write_unv  the same three datasets for a 2-D quad mesh (used to round-trip).
write_vtu  VTK XML unstructured grid (ASCII) of a mesh with nodal fields, or
           of particles as vertex cells with per-particle data, as the
           reference's drivers write with DataOut
           (TurekBenchmarkCuda/TurekBenchmark.cc:802-928).
"""
import numpy as np

from .mesh import Mesh, _orient_2d

# UNV quad nodes run around the cell; deal.II's order is lexicographic
_UNV_TO_DEALII = [0, 1, 3, 2]


def _datasets(lines):
    """(dataset id, body lines) of each '-1'-delimited block."""
    out = []
    i, n = 0, len(lines)
    while i < n:
        if lines[i].strip() != "-1":
            i += 1
            continue
        i += 1
        if i >= n:
            break
        did = int(lines[i].split()[0])
        j = i + 1
        while j < n and lines[j].strip() != "-1":
            j += 1
        out.append((did, lines[i + 1:j]))
        i = j + 1
    return out


def read_unv(path):
    """2-D quad mesh of a universal file.  Node labels may be sparse; the
    returned vertices follow the order of the node records.  Groups become
    `boundary_vertices[name]` (the vertices of the group's line elements or
    its nodes) and `cell_groups[name]` (indices of the group's quads)."""
    with open(path) as f:
        lines = f.read().splitlines()
    nodes, quads, lines_el = {}, [], {}
    quad_label = {}
    groups = {}
    for did, body in _datasets(lines):
        if did == 2411:
            for k in range(0, len(body) - 1, 2):
                label = int(body[k].split()[0])
                xyz = [float(t.replace("D", "E")) for t in body[k + 1].split()]
                nodes[label] = xyz
        elif did == 2412:
            k = 0
            while k < len(body):
                head = body[k].split()
                label, fe_type, n_nodes = int(head[0]), int(head[1]), int(head[5])
                k += 1
                if fe_type in (11, 21, 22, 23, 24):      # beam-type elements carry an orientation record
                    k += 1
                conn = []
                while len(conn) < n_nodes:
                    conn += [int(t) for t in body[k].split()]
                    k += 1
                if fe_type == 44:
                    quad_label[label] = len(quads)
                    quads.append(conn)
                elif fe_type == 11:
                    lines_el[label] = conn
        elif did == 2467:
            k = 0
            while k < len(body):
                head = body[k].split()
                n_ent = int(head[7])
                name = body[k + 1].strip()
                k += 2
                ents = []
                while len(ents) < 4 * n_ent:
                    ents += [int(t) for t in body[k].split()]
                    k += 1
                groups[name] = [(ents[4 * e], ents[4 * e + 1]) for e in range(n_ent)]
    labels = sorted(nodes, key=lambda l: l)
    index = {l: i for i, l in enumerate(labels)}
    vertices = np.array([nodes[l][:2] for l in labels], dtype=np.float64)
    cells = np.array([[index[q[u]] for u in _UNV_TO_DEALII] for q in quads], dtype=np.int32)
    cells = _orient_2d(vertices, cells).astype(np.int32)
    bv, cg = {}, {}
    for name, ents in groups.items():
        vs, cs = set(), []
        for etype, tag in ents:
            if etype == 8 and tag in lines_el:
                vs.update(index[n] for n in lines_el[tag])
            elif etype == 8 and tag in quad_label:
                cs.append(quad_label[tag])
            elif etype == 7 and tag in index:
                vs.add(index[tag])
        if vs:
            bv[name] = np.array(sorted(vs), dtype=np.int64)
        if cs:
            cg[name] = np.array(sorted(cs), dtype=np.int64)
    m = Mesh(2, vertices, cells, str(path), bv)
    m.cell_groups = cg
    return m


def write_unv(path, mesh, groups=None):
    """2411/2412/2467 datasets of a 2-D quad mesh; `groups` maps a name to
    boundary edges (pairs of vertex indices), written as line elements."""
    out = []
    out += ["    -1", "  2411"]
    for i, (x, y) in enumerate(mesh.vertices):
        out.append(f"{i + 1:10d}{1:10d}{1:10d}{11:10d}")
        out.append(f"{x:25.16E}{y:25.16E}{0.0:25.16E}")
    out += ["    -1", "    -1", "  2412"]
    label = 1
    group_elems = {}
    for name, edges in (groups or {}).items():
        group_elems[name] = []
        for a, b in edges:
            out.append(f"{label:10d}{11:10d}{2:10d}{1:10d}{7:10d}{2:10d}")
            out.append(f"{0:10d}{1:10d}{1:10d}")
            out.append(f"{a + 1:10d}{b + 1:10d}")
            group_elems[name].append(label)
            label += 1
    inv = [0, 1, 3, 2]
    for c in mesh.cells:
        out.append(f"{label:10d}{44:10d}{2:10d}{1:10d}{7:10d}{4:10d}")
        out.append("".join(f"{c[inv[u]] + 1:10d}" for u in range(4)))
        label += 1
    out += ["    -1"]
    if group_elems:
        out += ["    -1", "  2467"]
        for gi, (name, labels) in enumerate(group_elems.items()):
            out.append(f"{gi + 1:10d}{0:10d}{0:10d}{0:10d}{0:10d}{0:10d}{0:10d}{len(labels):10d}")
            out.append(name)
            row = []
            for l in labels:
                row.append(f"{8:10d}{l:10d}{0:10d}{0:10d}")
                if len(row) == 2:
                    out.append("".join(row))
                    row = []
            if row:
                out.append("".join(row))
        out += ["    -1"]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def _array(name, a, ncomp):
    flat = " ".join(repr(float(v)) if a.dtype.kind == "f" else str(int(v)) for v in np.asarray(a).ravel())
    typ = "Float64" if np.asarray(a).dtype.kind == "f" else "Int64"
    return (f'<DataArray type="{typ}" Name="{name}" NumberOfComponents="{ncomp}" format="ascii">'
            f"{flat}</DataArray>")


def write_vtu(path, points, cells=None, cell_type=None, point_data=None):
    """ASCII VTU.  points (n, dim); cells (m, k) vertex indices in VTK order
    (None: one vertex cell per point, as for particles); point_data maps a
    name to (n,) or (n, c) arrays."""
    pts = np.zeros((len(points), 3))
    pts[:, :points.shape[1]] = points
    if cells is None:
        cells = np.arange(len(points)).reshape(-1, 1)
        cell_type = 1                                   # VTK_VERTEX
    cells = np.asarray(cells)
    offsets = np.arange(1, len(cells) + 1) * cells.shape[1]
    parts = ['<?xml version="1.0"?>',
             '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">',
             "<UnstructuredGrid>",
             f'<Piece NumberOfPoints="{len(pts)}" NumberOfCells="{len(cells)}">',
             "<Points>", _array("Points", pts, 3), "</Points>",
             "<Cells>", _array("connectivity", cells.astype(np.int64), 1), _array("offsets", offsets, 1),
             _array("types", np.full(len(cells), cell_type, dtype=np.int64), 1), "</Cells>"]
    if point_data:
        parts.append("<PointData>")
        for name, a in point_data.items():
            a = np.asarray(a)
            parts.append(_array(name, a, 1 if a.ndim == 1 else a.shape[1]))
        parts.append("</PointData>")
    parts += ["</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
    with open(path, "w") as f:
        f.write("\n".join(parts) + "\n")


def write_mesh_vtu(path, mesh, fields=None):
    """Quad (2-D) or hexahedron (3-D) mesh with nodal fields given as
    (dim, n_dofs) arrays (component-major, as the C ABI's nodal vectors)."""
    if mesh.dim == 2:
        cells, vtk = mesh.cells[:, [0, 1, 3, 2]], 9            # VTK_QUAD
    else:
        cells, vtk = mesh.cells[:, [0, 1, 3, 2, 4, 5, 7, 6]], 12   # VTK_HEXAHEDRON
    pd = {k: (np.asarray(v).T if np.asarray(v).ndim == 2 else v) for k, v in (fields or {}).items()}
    write_vtu(path, mesh.vertices, cells, vtk, pd)


def write_particles_vtu(path, particles):
    """Particles as returned by ParticleStage.particles(): location, with
    velocity, id and cell as point data."""
    write_vtu(path, particles["x"], None, None,
              {"velocity": particles["v"], "id": particles["id"].astype(np.int64),
               "cell": particles["cell"].astype(np.int64)})
