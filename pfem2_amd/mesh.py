"""Synthetic meshes, velocity fields and particle layouts for the particle stage.

Meshes use deal.II's cell vertex order (lexicographic: local vertex k has
unit coordinates (bit0(k), bit1(k)[, bit2(k)])) with positive orientation, and
Q1 DoFs numbered like the vertices (cell_dofs == cell_vertices).  q2(mesh)
adds a degree-2 velocity DoF numbering on the same geometry (FE_Q<dim>(2):
3^dim DoFs per cell in deal.II's order) with its support points; the field
generators evaluate at `dof_points`, so they serve both degrees.

Generators:
  channel_2d        structured quads on [0,L]x[0,H], cells and vertices
                    lexicographic with x fastest
  cylinder_channel  block-structured quads around a cylinder in a channel
                    (O-grid ring graded toward the cylinder, stretched
                    downstream block), seeded vertex jitter so that cells are
                    general quadrilaterals, cells in Morton order
  pipe_3d           hexahedra of an nx x ny x nz box clipped to a circular
                    cross-section
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Mesh:
    dim: int
    vertices: np.ndarray            # (n_vertices, dim) float64
    cells: np.ndarray               # (n_cells, 2**dim) int32
    name: str = ""
    boundary_vertices: dict = field(default_factory=dict)   # tag -> vertex indices
    degree: int = 1                 # velocity field degree
    cell_dofs: np.ndarray = None    # degree 2: (n_cells, 3**dim) int32
    dof_xyz: np.ndarray = None      # degree 2: support points (n_dofs, dim)

    @property
    def n_cells(self):
        return self.cells.shape[0]

    @property
    def n_vertices(self):
        return self.vertices.shape[0]

    @property
    def n_dofs(self):
        return self.vertices.shape[0] if self.degree == 1 else self.dof_xyz.shape[0]

    @property
    def dofs_per_cell(self):
        return (1 << self.dim) if self.degree == 1 else 3 ** self.dim

    @property
    def dof_points(self):
        """Support points of the velocity DoFs (the vertices for degree 1)."""
        return self.vertices if self.degree == 1 else self.dof_xyz

    def cell_vertices_xyz(self):
        return self.vertices[self.cells]


# local vertex pairs of the lines and vertex quadruples of the faces, in
# deal.II's GeometryInfo numbering (2-D lines: x=0, x=1, y=0, y=1; 3-D lines:
# the four of face z=0, the four of z=1, then the four parallel to z; faces:
# x=0, x=1, y=0, y=1, z=0, z=1)
LINES = {2: [(0, 2), (1, 3), (0, 1), (2, 3)],
         3: [(0, 2), (1, 3), (0, 1), (2, 3), (4, 6), (5, 7), (4, 5), (6, 7), (0, 4), (1, 5), (2, 6), (3, 7)]}
FACES3 = [(0, 2, 4, 6), (1, 3, 5, 7), (0, 1, 4, 5), (2, 3, 6, 7), (0, 1, 2, 3), (4, 5, 6, 7)]


def q2(mesh):
    """The mesh with a degree-2 velocity DoF numbering (FE_Q<dim>(2) on the
    Q1 geometry).  Per cell 3^dim DoFs in deal.II's order: the vertices, one
    per line, (3-D) one per face, one interior.  Global numbering: vertex
    DoFs keep the vertex numbers, then the lines, faces and cell interiors
    (each in order of their sorted vertex keys / cells).  Support points are
    the vertex averages of the lines, faces and cells, which is where the
    multilinear map puts the unit-cell nodes."""
    dim, cells, X = mesh.dim, np.asarray(mesh.cells), np.asarray(mesh.vertices)
    nc, nv = cells.shape[0], X.shape[0]
    parts, points, base = [cells.astype(np.int64)], [X], nv
    groups = [LINES[dim]] + ([FACES3] if dim == 3 else [])
    bits = max(1, int(nv - 1).bit_length())
    assert 3 * bits <= 63, "too many vertices for the packed line/face keys"
    for grp in groups:
        keys = np.stack([np.sort(cells[:, list(g)], axis=1) for g in grp], axis=1)   # (nc, n_ent, |g|)
        flat = keys.reshape(-1, keys.shape[2]).astype(np.int64)
        # a line is its two vertices; a face of a conforming hex mesh is
        # fixed by its three smallest vertex numbers
        packed = flat[:, 0]
        for j in range(1, min(3, flat.shape[1])):
            packed = (packed << bits) | flat[:, j]
        _, first, inv = np.unique(packed, return_index=True, return_inverse=True)
        parts.append(base + inv.reshape(nc, len(grp)))
        points.append(X[flat[first]].mean(axis=1))
        base += len(first)
    parts.append(base + np.arange(nc)[:, None])
    points.append(X[cells].mean(axis=1))
    cd = np.concatenate(parts, axis=1).astype(np.int32)
    assert cd.shape[1] == 3 ** dim
    return Mesh(dim, X, cells, name=mesh.name + "-q2", boundary_vertices=mesh.boundary_vertices, degree=2,
                cell_dofs=cd, dof_xyz=np.concatenate(points, axis=0))


def _orient_2d(vertices, cells):
    """Swap local vertices 1 and 2 of negatively oriented quads."""
    P = vertices[cells]
    # Jacobian determinant at the cell centre
    a = 0.5 * ((P[:, 1] - P[:, 0]) + (P[:, 3] - P[:, 2]))
    b = 0.5 * ((P[:, 2] - P[:, 0]) + (P[:, 3] - P[:, 1]))
    det = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    neg = det < 0
    cells = cells.copy()
    cells[neg, 1], cells[neg, 2] = cells[neg, 2].copy(), cells[neg, 1].copy()
    return cells


def quad_convex(vertices, cells):
    """True where every corner of the quad turns left (CCW cycle 0,1,3,2)."""
    P = vertices[cells]
    cyc = [0, 1, 3, 2]
    ok = np.ones(len(cells), dtype=bool)
    for i in range(4):
        a, b, c = cyc[i], cyc[(i + 1) % 4], cyc[(i + 2) % 4]
        e0 = P[:, b] - P[:, a]
        e1 = P[:, c] - P[:, b]
        ok &= (e0[:, 0] * e1[:, 1] - e0[:, 1] * e1[:, 0]) > 0
    return ok


def channel_2d(nx=200, ny=20, L=10.0, H=1.0):
    xs = np.linspace(0.0, L, nx + 1)
    ys = np.linspace(0.0, H, ny + 1)
    X, Y = np.meshgrid(xs, ys)          # (ny+1, nx+1), x fastest
    vertices = np.stack([X.ravel(), Y.ravel()], axis=1)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    v0 = (j * (nx + 1) + i).ravel()
    cells = np.stack([v0, v0 + 1, v0 + nx + 1, v0 + nx + 2], axis=1).astype(np.int32)
    bv = {
        "bottom": np.arange(nx + 1),
        "top": ny * (nx + 1) + np.arange(nx + 1),
        "inlet": np.arange(ny + 1) * (nx + 1),
        "outlet": np.arange(ny + 1) * (nx + 1) + nx,
    }
    return Mesh(2, vertices, cells, f"channel_{nx}x{ny}", bv)


def _grade(n, ratio):
    """n+1 points on [0,1], geometric spacing with last/first cell size = ratio."""
    if abs(ratio - 1.0) < 1e-12:
        return np.linspace(0.0, 1.0, n + 1)
    q = ratio ** (1.0 / max(n - 1, 1))
    h = q ** np.arange(n)
    t = np.concatenate([[0.0], np.cumsum(h)])
    return t / t[-1]


def _block(x_of_st, s_pts, t_pts):
    """vertex grid of a block from a map (s, t) -> (x, y); returns (ns+1, nt+1, 2)."""
    S, T = np.meshgrid(s_pts, t_pts, indexing="ij")
    return x_of_st(S, T)


def _morton_order(points, bits=16):
    lo = points.min(axis=0)
    span = np.maximum(points.max(axis=0) - lo, 1e-300)
    q = np.clip(((points - lo) / span * (2 ** bits - 1)).astype(np.int64), 0, 2 ** bits - 1)
    code = np.zeros(len(points), dtype=np.int64)
    for b in range(bits):
        for d in range(points.shape[1]):
            code |= ((q[:, d] >> b) & 1) << (b * points.shape[1] + d)
    return np.argsort(code, kind="stable")


def cylinder_channel(scale=1.0, seed=0, jitter=0.12, length=2.2, height=0.41, center=(0.2, 0.2), radius=0.05,
                     box=0.1, morton=True):
    """Graded quad mesh of the Turek-Schaefer channel with a cylinder.

    scale=1 gives about 100k cells.  Interior vertices are displaced by a
    seeded uniform jitter of `jitter` x the local edge length.
    """
    cx, cy = center
    m = max(4, int(round(90 * scale)))          # cells along each side of the box
    r_cells = max(3, int(round(56 * scale)))    # radial cells in the O-grid
    n_left = max(2, int(round(45 * scale)))
    n_low = max(2, int(round(45 * scale)))
    n_up = max(2, int(round(50 * scale)))
    n_right = max(4, int(round(350 * scale)))
    x0, x1 = cx - box, cx + box
    y0, y1 = cy - box, cy + box
    blocks = []

    def lin(a, b, t):
        return a + (b - a) * t

    # O-grid: 4 blocks, s = tangential (0..1 along the box side), t = radial (0 circle .. 1 box)
    radial = _grade(r_cells, 6.0)
    tang = np.linspace(0.0, 1.0, m + 1)
    sides = [((x0, y0), (x1, y0)), ((x1, y0), (x1, y1)), ((x1, y1), (x0, y1)), ((x0, y1), (x0, y0))]
    for (ax, ay), (bx, by) in sides:
        def f(S, T, ax=ax, ay=ay, bx=bx, by=by):
            px = lin(ax, bx, S)
            py = lin(ay, by, S)
            ang = np.arctan2(py - cy, px - cx)
            qx = cx + radius * np.cos(ang)
            qy = cy + radius * np.sin(ang)
            return np.stack([lin(qx, px, T), lin(qy, py, T)], axis=-1)
        blocks.append(_block(f, tang, radial))
    # outer rectangles: columns [0,x0], [x0,x1], [x1,L]; rows [0,y0], [y0,y1], [y1,H]
    colx = [np.linspace(0.0, x0, n_left + 1), np.linspace(x0, x1, m + 1),
            x1 + (length - x1) * _grade(n_right, 8.0)]
    rowy = [np.linspace(0.0, y0, n_low + 1), np.linspace(y0, y1, m + 1), np.linspace(y1, height, n_up + 1)]
    for ci in range(3):
        for ri in range(3):
            if ci == 1 and ri == 1:
                continue
            xs, ys = colx[ci], rowy[ri]
            X, Y = np.meshgrid(xs, ys, indexing="ij")
            blocks.append(np.stack([X, Y], axis=-1))
    # merge vertices of all blocks
    allpts = []
    allcells = []
    base = 0
    for B in blocks:
        ns, nt = B.shape[0] - 1, B.shape[1] - 1
        allpts.append(B.reshape(-1, 2))
        I, J = np.meshgrid(np.arange(ns), np.arange(nt), indexing="ij")
        v0 = base + I * (nt + 1) + J
        c = np.stack([v0, v0 + (nt + 1), v0 + 1, v0 + (nt + 1) + 1], axis=-1).reshape(-1, 4)
        allcells.append(c)
        base += (ns + 1) * (nt + 1)
    pts = np.concatenate(allpts)
    cells = np.concatenate(allcells)
    key = np.round(pts * 1e9).astype(np.int64)
    _, first, inv = np.unique(key[:, 0] * 4_000_000_000_000 + key[:, 1], return_index=True, return_inverse=True)
    vertices = pts[first]
    cells = inv[cells]
    # drop unused vertices (none expected) and orient
    cells = _orient_2d(vertices, cells.astype(np.int64))
    # boundary tags before jitter
    tol = 1e-9
    on_wall = (np.abs(vertices[:, 1]) < tol) | (np.abs(vertices[:, 1] - height) < tol)
    on_in = np.abs(vertices[:, 0]) < tol
    on_out = np.abs(vertices[:, 0] - length) < tol
    on_cyl = np.abs(np.hypot(vertices[:, 0] - cx, vertices[:, 1] - cy) - radius) < 1e-7
    boundary = on_wall | on_in | on_out | on_cyl
    # seeded jitter of interior vertices, scaled by the shortest incident edge
    rng = np.random.default_rng(seed)
    if jitter > 0:
        P = vertices[cells]
        edges = [(0, 1), (1, 3), (3, 2), (2, 0)]
        hmin = np.full(len(vertices), np.inf)
        for a, b in edges:
            le = np.linalg.norm(P[:, a] - P[:, b], axis=1)
            np.minimum.at(hmin, cells[:, a], le)
            np.minimum.at(hmin, cells[:, b], le)
        disp = rng.uniform(-1.0, 1.0, size=vertices.shape) * (jitter * hmin)[:, None]
        disp[boundary] = 0.0
        vertices = vertices + disp
    if not quad_convex(vertices, cells).all():
        raise RuntimeError("cylinder_channel produced non-convex cells; lower the jitter")
    if morton:
        order = _morton_order(vertices[cells].mean(axis=1))
        cells = cells[order]
        # renumber vertices by first use
        flat = cells.ravel()
        _, first_use = np.unique(flat, return_index=True)
        vorder = np.argsort(first_use, kind="stable")
        uniq = np.unique(flat)[vorder]
        remap = np.empty(len(vertices), dtype=np.int64)
        remap[uniq] = np.arange(len(uniq))
        vertices = vertices[uniq]
        cells = remap[cells]
        boundary = boundary[uniq]
        on_wall, on_in, on_out, on_cyl = on_wall[uniq], on_in[uniq], on_out[uniq], on_cyl[uniq]
    bv = {
        "walls": np.nonzero(on_wall)[0],
        "inlet": np.nonzero(on_in)[0],
        "outlet": np.nonzero(on_out)[0],
        "cylinder": np.nonzero(on_cyl)[0],
    }
    return Mesh(2, np.ascontiguousarray(vertices), cells.astype(np.int32), f"cylinder_channel_{len(cells)}", bv)


def pipe_3d(nx=64, ny=64, nz=256, radius=0.5, length=4.0):
    """Hexahedral box mesh on [-R,R]^2 x [0,length] keeping cells whose centre lies in the circle."""
    xs = np.linspace(-radius, radius, nx + 1)
    ys = np.linspace(-radius, radius, ny + 1)
    zs = np.linspace(0.0, length, nz + 1)
    cxs = 0.5 * (xs[:-1] + xs[1:])
    cys = 0.5 * (ys[:-1] + ys[1:])
    CY, CX = np.meshgrid(cys, cxs, indexing="ij")          # (ny, nx)
    inside2d = (CX ** 2 + CY ** 2) <= radius ** 2
    jj, ii = np.nonzero(inside2d)                           # ascending j then i
    ncol = len(ii)
    k = np.repeat(np.arange(nz), ncol)
    i = np.tile(ii, nz)
    j = np.tile(jj, nz)

    def vid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    cells = np.stack([vid(i, j, k), vid(i + 1, j, k), vid(i, j + 1, k), vid(i + 1, j + 1, k),
                      vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i, j + 1, k + 1), vid(i + 1, j + 1, k + 1)],
                     axis=1)
    used = np.unique(cells)
    remap = np.full((nx + 1) * (ny + 1) * (nz + 1), -1, dtype=np.int64)
    remap[used] = np.arange(len(used))
    kk, rem = np.divmod(used, (nx + 1) * (ny + 1))
    jv, iv = np.divmod(rem, nx + 1)
    vertices = np.stack([xs[iv], ys[jv], zs[kk]], axis=1)
    cells = remap[cells].astype(np.int32)
    bv = {"inlet": np.nonzero(np.abs(vertices[:, 2]) < 1e-12)[0],
          "outlet": np.nonzero(np.abs(vertices[:, 2] - length) < 1e-12)[0]}
    return Mesh(3, vertices, cells, f"pipe_{nx}x{ny}x{nz}_{len(cells)}", bv)


def box_3d(nx=4, ny=4, nz=4, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), jitter=0.0, seed=0):
    """Structured hexahedra on a box; optional seeded jitter of interior vertices."""
    xs = np.linspace(lo[0], hi[0], nx + 1)
    ys = np.linspace(lo[1], hi[1], ny + 1)
    zs = np.linspace(lo[2], hi[2], nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    vertices = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()

    def vid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    cells = np.stack([vid(i, j, k), vid(i + 1, j, k), vid(i, j + 1, k), vid(i + 1, j + 1, k),
                      vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i, j + 1, k + 1), vid(i + 1, j + 1, k + 1)],
                     axis=1).astype(np.int32)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        h = np.array([(hi[d] - lo[d]) / n for d, n in enumerate((nx, ny, nz))])
        interior = np.all((vertices > np.array(lo) + 1e-12) & (vertices < np.array(hi) - 1e-12), axis=1)
        disp = rng.uniform(-1, 1, size=vertices.shape) * jitter * h
        vertices = vertices + np.where(interior[:, None], disp, 0.0)
    return Mesh(3, vertices, cells, f"box_{nx}x{ny}x{nz}")


def quad_2d(nx=4, ny=4, lo=(0.0, 0.0), hi=(1.0, 1.0), jitter=0.0, seed=0):
    """Structured quads on a rectangle; optional seeded jitter of interior vertices."""
    m = channel_2d(nx, ny, hi[0] - lo[0], hi[1] - lo[1])
    v = m.vertices + np.array(lo)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        h = np.array([(hi[0] - lo[0]) / nx, (hi[1] - lo[1]) / ny])
        interior = np.all((v > np.array(lo) + 1e-12) & (v < np.array(hi) - 1e-12), axis=1)
        v = v + np.where(interior[:, None], rng.uniform(-1, 1, size=v.shape) * jitter * h, 0.0)
    return Mesh(2, v, m.cells, f"quad_{nx}x{ny}", m.boundary_vertices)


# ---- velocity fields (nodal values, component-major (dim, n_dofs)) ---------

def fourier_perturbation(points, extent, seed, n_modes=8, kmax=4):
    """Sum of n_modes random Fourier modes per component, amplitudes N(0,1)/sqrt(n_modes)."""
    rng = np.random.default_rng(seed)
    dim = points.shape[1]
    out = np.zeros((dim, len(points)))
    for c in range(dim):
        for _ in range(n_modes):
            k = rng.integers(1, kmax + 1, size=dim)
            a = rng.normal() / np.sqrt(n_modes)
            ph = rng.uniform(0, 2 * np.pi)
            arg = 2 * np.pi * sum(k[d] * points[:, d] / extent[d] for d in range(dim)) + ph
            out[c] += a * np.cos(arg)
    return out


def poiseuille_2d(mesh, umax=1.0, H=1.0, eps=0.01, seed=0):
    x = mesh.dof_points
    V = np.zeros((2, mesh.n_dofs))
    V[0] = 4.0 * umax * x[:, 1] * (H - x[:, 1]) / H ** 2
    ext = x.max(axis=0) - x.min(axis=0)
    if eps:
        V += eps * umax * fourier_perturbation(x, ext, seed)
    return V


def cylinder_flow(mesh, U=1.0, center=(0.2, 0.2), radius=0.05, eps=0.01, seed=0):
    """Potential flow around a cylinder plus a seeded Fourier perturbation."""
    X = mesh.dof_points[:, 0] - center[0]
    Y = mesh.dof_points[:, 1] - center[1]
    r2 = np.maximum(X ** 2 + Y ** 2, radius ** 2)
    R2 = radius ** 2
    V = np.zeros((2, mesh.n_dofs))
    V[0] = U * (1.0 - R2 * (X ** 2 - Y ** 2) / r2 ** 2)
    V[1] = -U * 2.0 * R2 * X * Y / r2 ** 2
    if eps:
        ext = mesh.vertices.max(axis=0) - mesh.vertices.min(axis=0)
        V += eps * U * fourier_perturbation(mesh.dof_points, ext, seed)
    return V


def pipe_flow(mesh, umax=1.0, radius=0.5, eps=0.01, seed=0):
    x = mesh.dof_points
    r2 = x[:, 0] ** 2 + x[:, 1] ** 2
    V = np.zeros((3, mesh.n_dofs))
    V[2] = umax * np.clip(1.0 - r2 / radius ** 2, 0.0, None)
    if eps:
        ext = x.max(axis=0) - x.min(axis=0)
        V += eps * umax * fourier_perturbation(x, ext, seed)
    return V


def uniform_field(mesh, U):
    U = np.asarray(U, dtype=float)
    return np.repeat(U[:, None], mesh.n_dofs, axis=1)


# ---- particle layouts ------------------------------------------------------

def jittered_lattice(mesh, quantities, jitter=0.05, seed=0):
    """(cell, xi) of a lattice per cell with a seeded uniform jitter, cells
    ascending, first lattice direction outermost (the seeding order)."""
    q = list(quantities)[: mesh.dim]
    grids = np.meshgrid(*[(np.arange(n) + 0.5) / n for n in q], indexing="ij")
    lat = np.stack([g.ravel() for g in grids], axis=1)          # (prod q, dim)
    nper = lat.shape[0]
    cell = np.repeat(np.arange(mesh.n_cells, dtype=np.int32), nper)
    xi = np.tile(lat, (mesh.n_cells, 1))
    if jitter:
        rng = np.random.default_rng(seed)
        xi = xi + rng.uniform(-jitter, jitter, size=xi.shape)
    return cell, np.clip(xi, 0.0, 1.0)


def boundary_faces(mesh, tag=None):
    """(cell, face number) of the faces that belong to one cell only (deal.II
    face numbering: 2-D x=0, x=1, y=0, y=1; 3-D also z=0, z=1), in the
    reference's loop order (cells ascending, then faces); with `tag`, only
    the faces whose vertices all carry that boundary tag."""
    dim, cells = mesh.dim, np.asarray(mesh.cells)
    nc, nf = cells.shape[0], 2 * dim
    fv = [[k for k in range(1 << dim) if ((k >> (f >> 1)) & 1) == (f & 1)] for f in range(nf)]
    keys = np.stack([np.sort(cells[:, fv[f]], axis=1) for f in range(nf)], axis=1)   # (nc, nf, 2^(dim-1))
    flat = keys.reshape(nc * nf, -1)
    _, inv, cnt = np.unique(flat, axis=0, return_inverse=True, return_counts=True)
    bnd = (cnt[inv.reshape(-1)] == 1).reshape(nc, nf)
    if tag is not None:
        on = np.zeros(mesh.n_vertices, dtype=bool)
        on[np.asarray(mesh.boundary_vertices[tag])] = True
        bnd &= on[keys].all(axis=2)
    c, f = np.nonzero(bnd)
    return np.stack([c, f], axis=1).astype(np.int32)
