"""Mesh containers, synthesisers and the reference's text mesh format.

A mesh is what the reference reads from ``<base>.grid`` / ``<base>.bc``
(src/grid_procs.f90:65-164): node coordinates, triangles then quads as
counter-clockwise node lists, and boundaries, each a type name plus the list of
cells that carry one of its edges.  Indices here are 0-based.

Synthesisers:

* :func:`jittered_square` -- the isentropic-vortex square: ``nx*ny`` quads over
  ``[-5,5]^2``, each split by its bottom-left/top-right diagonal into two
  triangles, interior nodes displaced uniformly inside a disc of radius
  ``jitter*h``.  ``periodic=True`` identifies the seam nodes (builder-new
  periodic boundary); otherwise the whole outline is one ``dirichlet``
  boundary and the two corner quads whose lower/upper triangle would carry two
  boundary edges get the other diagonal, so every boundary cell has exactly one
  boundary edge (the reference pairs boundary cells and boundary edges one to
  one, src/residual.f90:115-125).
* :func:`naca_ogrid` -- O-grid around a symmetric NACA 4-digit airfoil.

Import: :func:`read_gambit` reads a Gambit neutral file (``.neu``) the way
utils/tools/read_gambit.f90 does; :func:`import_gambit` also writes the three
files that tool writes (``.grid``, ``.bc`` and a Tecplot ``.plt`` of the mesh).
"""
from __future__ import annotations

import dataclasses
import math
import os
import re
from typing import List, Optional, Sequence, Tuple

import numpy as np

BC_TYPES = {"freestream": 0, "slip_wall": 1, "dirichlet": 2, "solid_wall": 3}


@dataclasses.dataclass
class Mesh:
    xy: np.ndarray                      # (nnodes, 2) float64
    ntri: int
    nquad: int
    cell_nodes: np.ndarray              # ntri*3 + nquad*4 int32, 0-based, CCW
    boundaries: List[Tuple[str, np.ndarray]]
    canon: Optional[np.ndarray] = None  # periodic node identification (or None)
    period: Tuple[float, float] = (0.0, 0.0)

    @property
    def nnodes(self) -> int:
        return int(self.xy.shape[0])

    @property
    def ncells(self) -> int:
        return self.ntri + self.nquad

    def cells_as_array(self) -> np.ndarray:
        """(ncells, 4) node table, -1 in the 4th slot of triangles."""
        out = -np.ones((self.ncells, 4), np.int32)
        t = self.cell_nodes[: 3 * self.ntri].reshape(-1, 3)
        out[: self.ntri, :3] = t
        if self.nquad:
            out[self.ntri:, :] = self.cell_nodes[3 * self.ntri:].reshape(-1, 4)
        return out

    def bc_arrays(self):
        """(ncells per boundary, type ids, concatenated cells) as int32."""
        n = np.array([len(c) for _, c in self.boundaries], np.int32)
        t = np.array([BC_TYPES[name] for name, _ in self.boundaries], np.int32)
        cells = (np.concatenate([np.asarray(c, np.int32) for _, c in self.boundaries])
                 if self.boundaries else np.zeros(0, np.int32))
        return n, t, cells


def _tri_area(xy, tri):
    p1, p2, p3 = xy[tri[:, 0]], xy[tri[:, 1]], xy[tri[:, 2]]
    return 0.5 * ((p2[:, 0] - p1[:, 0]) * (p3[:, 1] - p1[:, 1]) - (p3[:, 0] - p1[:, 0]) * (p2[:, 1] - p1[:, 1]))


def jittered_square(nx: int, ny: Optional[int] = None, *, lo: float = -5.0, hi: float = 5.0,
                    jitter: float = 0.1, seed: int = 0, periodic: bool = False) -> Mesh:
    """Split-quad square with seeded jitter (see module docstring).

    Jitter sampling: for interior node k (row-major order), draw u1, u2 from
    ``numpy.random.Generator(PCG64(seed)).random((n_interior, 2))`` and move the
    node by ``jitter*h*sqrt(u1) * (cos 2 pi u2, sin 2 pi u2)``, h = (hi-lo)/nx.
    """
    ny = nx if ny is None else ny
    if nx < 3 or ny < 3:
        raise ValueError("jittered_square needs at least 3x3 quads")
    hx, hy = (hi - lo) / nx, (hi - lo) / ny
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    xy = np.stack([lo + ii.ravel() * hx, lo + jj.ravel() * hy], axis=1).astype(np.float64)
    interior = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < ny)).ravel()
    idx = np.nonzero(interior)[0]
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.random((idx.size, 2))
    r = jitter * min(hx, hy) * np.sqrt(u[:, 0])
    th = 2.0 * math.pi * u[:, 1]
    xy[idx, 0] += r * np.cos(th)
    xy[idx, 1] += r * np.sin(th)

    qi, qj = np.meshgrid(np.arange(nx), np.arange(ny))
    qi, qj = qi.ravel(), qj.ravel()
    bl = qj * (nx + 1) + qi
    br, tl = bl + 1, bl + nx + 1
    tr = tl + 1
    flip = np.zeros(nx * ny, bool)
    if not periodic:
        flip[(qi == nx - 1) & (qj == 0)] = True
        flip[(qi == 0) & (qj == ny - 1)] = True
    t1 = np.where(flip[:, None], np.stack([bl, br, tl], 1), np.stack([bl, br, tr], 1))
    t2 = np.where(flip[:, None], np.stack([br, tr, tl], 1), np.stack([bl, tr, tl], 1))
    tri = np.empty((2 * nx * ny, 3), np.int64)
    tri[0::2] = t1
    tri[1::2] = t2
    assert np.all(_tri_area(xy, tri) > 0)
    cell_nodes = tri.astype(np.int32).ravel()

    if periodic:
        canon = ((jj % ny) * (nx + 1) + (ii % nx)).ravel().astype(np.int32)
        return Mesh(xy, 2 * nx * ny, 0, cell_nodes, [], canon=canon, period=(hi - lo, hi - lo))

    # cells with a boundary edge, ascending: an edge lies on the outline when
    # both of its end nodes sit on the same side of the square
    side = np.zeros((xy.shape[0], 4), bool)
    side[:, 0] = (ii == 0).ravel()
    side[:, 1] = (ii == nx).ravel()
    side[:, 2] = (jj == 0).ravel()
    side[:, 3] = (jj == ny).ravel()
    nbe = np.zeros(tri.shape[0], np.int64)
    for e in range(3):
        a, b = tri[:, e], tri[:, (e + 1) % 3]
        nbe += np.any(side[a] & side[b], axis=1)
    assert nbe.max() == 1, "boundary cell with two boundary edges"
    bcells = np.nonzero(nbe)[0].astype(np.int32)
    return Mesh(xy, 2 * nx * ny, 0, cell_nodes, [("dirichlet", bcells)])


def naca_thickness(x: np.ndarray, t: float, closed_te: bool = True) -> np.ndarray:
    """Half thickness of a symmetric NACA 00tt section (chord 1).

    The closed-trailing-edge variant uses -0.1036 for the x^4 coefficient so
    that y(1) = 0 (the open form uses -0.1015)."""
    a4 = -0.1036 if closed_te else -0.1015
    return 5.0 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 + a4 * x ** 4)


def _growth_ratio(first: float, n: int, total: float) -> float:
    """Geometric ratio r with first*(r^n - 1)/(r - 1) = total (bisection)."""
    lo, hi = 1.0 + 1e-12, 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        s = first * (mid ** n - 1.0) / (mid - 1.0)
        if s > total:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def naca_ogrid(n_around: int = 512, n_layers: int = 256, *, thickness: float = 0.12,
               first_spacing: float = 2e-3, far_radius: float = 50.0) -> Mesh:
    """O-grid around a symmetric NACA 4-digit airfoil with a closed trailing edge.

    ``n_around`` surface nodes at x = (1 + cos s)/2, s = 2 pi i / n (clustered
    at both edges), ordered trailing edge -> upper surface -> leading edge ->
    lower surface.  Node k of the ray from surface node i sits at distance
    d_k = first * (r^k - 1)/(r - 1) along a direction that turns from the
    surface normal (at the wall) to the ray from mid-chord (far away); r is
    solved so that d reaches ``far_radius``, and each ray's distances are
    scaled so its last node lies on the circle of that radius about mid-chord.
    Each quad is split into two counter-clockwise triangles; ``slip_wall`` lists
    the triangles on the airfoil, ``freestream`` those on the outer circle, one
    boundary edge each."""
    m = n_around
    s = 2.0 * math.pi * np.arange(m) / m
    xs = 0.5 * (1.0 + np.cos(s))
    ys = naca_thickness(xs, thickness) * np.where(s <= math.pi, 1.0, -1.0)
    ys[0] = 0.0
    if m % 2 == 0:
        ys[m // 2] = 0.0
    xc0 = 0.5
    ang = np.arctan2(ys, xs - xc0)
    # outward unit normals from central differences along the surface
    tx = np.roll(xs, -1) - np.roll(xs, 1)
    ty = np.roll(ys, -1) - np.roll(ys, 1)
    ln = np.hypot(tx, ty)
    nxs, nys = ty / ln, -tx / ln
    flip = (nxs * (xs - xc0) + nys * ys) < 0
    nxs[flip], nys[flip] = -nxs[flip], -nys[flip]
    rx, ry = np.cos(ang), np.sin(ang)
    cx, cy = xc0 + far_radius * rx, far_radius * ry
    span = np.hypot(cx - xs, cy - ys)
    r = _growth_ratio(first_spacing, n_layers, far_radius)
    d = first_spacing * (r ** np.arange(n_layers + 1) - 1.0) / (r - 1.0)
    xy = np.empty(((n_layers + 1) * m, 2))
    for k in range(n_layers + 1):
        beta = (d[k] / far_radius) ** 0.5
        dx = (1.0 - beta) * nxs + beta * rx
        dy = (1.0 - beta) * nys + beta * ry
        dn = np.hypot(dx, dy)
        dist = d[k] * span / far_radius
        if k == n_layers:
            xy[k * m:(k + 1) * m, 0], xy[k * m:(k + 1) * m, 1] = cx, cy
        else:
            xy[k * m:(k + 1) * m, 0] = xs + dist * dx / dn
            xy[k * m:(k + 1) * m, 1] = ys + dist * dy / dn
    kk, ii = np.meshgrid(np.arange(n_layers), np.arange(m), indexing="ij")
    kk, ii = kk.ravel(), ii.ravel()
    a = kk * m + ii
    b = kk * m + (ii + 1) % m
    c = (kk + 1) * m + (ii + 1) % m
    dd = (kk + 1) * m + ii
    # (a, b, c, d) runs clockwise (the fluid lies right of the surface
    # direction), so the triangles are (c, b, a) and (d, c, a)
    tri = np.empty((2 * a.size, 3), np.int64)
    tri[0::2] = np.stack([c, b, a], 1)
    tri[1::2] = np.stack([dd, c, a], 1)
    area = _tri_area(xy, tri)
    if not np.all(area > 0):
        raise RuntimeError(f"O-grid has {int(np.sum(area <= 0))} inverted cells")
    wall = np.arange(0, 2 * m, 2, dtype=np.int32)
    far = (2 * (n_layers - 1) * m + 1 + np.arange(0, 2 * m, 2)).astype(np.int32)
    return Mesh(xy, tri.shape[0], 0, tri.astype(np.int32).ravel(), [("slip_wall", wall), ("freestream", far)])


# ---------------------------------------------------------------------------
# reference text format: src/grid_procs.f90:65-164
# ---------------------------------------------------------------------------
def write_grid(mesh: Mesh, base: str) -> None:
    """Write ``base.grid`` and ``base.bc`` (1-based ids)."""
    cells = mesh.cells_as_array()
    with open(base + ".grid", "w") as f:
        f.write(" #nodes, #cells_tri, #cells_quad\n")
        f.write(f"{mesh.nnodes} {mesh.ntri} {mesh.nquad}\n")
        for x, y in mesh.xy:
            f.write(f"{float(x)!r} {float(y)!r}\n")
        for c in cells[: mesh.ntri]:
            f.write(f"{c[0] + 1} {c[1] + 1} {c[2] + 1}\n")
        for c in cells[mesh.ntri:]:
            f.write(f"{c[0] + 1} {c[1] + 1} {c[2] + 1} {c[3] + 1}\n")
    with open(base + ".bc", "w") as f:
        f.write(f"{len(mesh.boundaries)}\n")
        for name, c in mesh.boundaries:
            f.write(f"{len(c)} {name}\n")
        for _, c in mesh.boundaries:
            for v in c:
                f.write(f"{int(v) + 1}\n")


def read_grid(base: str) -> Mesh:
    """Read ``base.grid`` / ``base.bc`` written in the reference's format."""
    with open(base + ".grid") as f:
        f.readline()
        nn, nt, nq = (int(v) for v in f.readline().split()[:3])
        xy = np.array([[float(v.replace("d", "e").replace("D", "e")) for v in f.readline().split()[:2]]
                       for _ in range(nn)])
        tri = [[int(v) - 1 for v in f.readline().split()[:3]] for _ in range(nt)]
        quad = [[int(v) - 1 for v in f.readline().split()[:4]] for _ in range(nq)]
    cell_nodes = np.array([v for c in tri for v in c] + [v for c in quad for v in c], np.int32)
    bnds = []
    with open(base + ".bc") as f:
        nb = int(f.readline().split()[0])
        heads = []
        for _ in range(nb):
            parts = f.readline().split()
            heads.append((int(parts[0]), parts[1]))
        for n, name in heads:
            cells = np.array([int(f.readline().split()[0]) - 1 for _ in range(n)], np.int32)
            bnds.append((name, cells))
    return Mesh(xy, nt, nq, cell_nodes, bnds)


def _ints(line: str) -> List[int]:
    return [int(v) for v in line.split()]


def _float(tok: str) -> float:
    return float(tok.replace("d", "e").replace("D", "e"))


def read_gambit(path: str) -> Mesh:
    """Read a 2-D Gambit neutral file into a :class:`Mesh`.

    Follows utils/tools/read_gambit.f90:29-111 and its renumbering at :139-152:
    the counts line is the 7th line (nodes, cells, element groups, boundary
    sets); node coordinates are taken in file order; a cell is a triangle or a
    quadrilateral by its node count (third field of its record, :42-50); cells
    are renumbered triangles first, then quadrilaterals, each group in file
    order; every boundary set is a name followed by one record per boundary
    edge whose first field is the (file-numbered) cell that carries it, and
    those cell ids are mapped through the renumbering (:161-166).  Node lists
    are kept as written (Gambit writes them counter-clockwise).

    Where the tool would silently produce a broken mesh this raises
    ValueError: a cell with other than 3 or 4 nodes (the tool's cell-count
    check at :54-57), a node-based boundary set (type 0: its entries are node
    ids, which the tool would read as cell ids), or a boundary entry naming no
    cell.  The tool reads a boundary header at fixed columns (``a39,i1,4x,i4``,
    :101), which caps a set at 9999 entries; here the header is split on
    blanks, which reads the same fields on every file the tool reads.
    """
    with open(path) as f:
        lines = f.read().splitlines()
    if len(lines) < 7:
        raise ValueError(f"{path}: too short for a Gambit neutral file")
    counts = _ints(lines[6])
    if len(counts) < 4:
        raise ValueError(f"{path}: line 7 must hold NUMNP NELEM NGRPS NBSETS")
    nnodes, ncells, _, nbsets = counts[:4]

    def section(start: int, key: str) -> int:
        for i in range(start, len(lines)):
            if key in lines[i].upper():
                return i + 1
        raise ValueError(f"{path}: no '{key}' section")

    pos = section(7, "NODAL COORDINATES")
    xy = np.empty((nnodes, 2), np.float64)
    for i in range(nnodes):
        t = lines[pos + i].split()
        xy[i] = _float(t[1]), _float(t[2])
    pos = section(pos + nnodes, "ELEMENTS/CELLS")
    nv = np.empty(ncells, np.int64)
    conn = []
    for i in range(ncells):
        t = _ints(lines[pos + i])
        nv[i] = t[2]
        if t[2] not in (3, 4):
            raise ValueError(f"{path}: cell {t[0]} has {t[2]} nodes; only triangles and quadrilaterals")
        conn.append([v - 1 for v in t[3:3 + t[2]]])
        if len(conn[-1]) != t[2]:
            raise ValueError(f"{path}: cell {t[0]} lists {len(conn[-1])} of its {t[2]} nodes")
    tri = np.flatnonzero(nv == 3)
    quad = np.flatnonzero(nv == 4)
    new_id = np.empty(ncells, np.int64)
    new_id[tri] = np.arange(tri.size)
    new_id[quad] = tri.size + np.arange(quad.size)
    cell_nodes = np.array([v for i in tri for v in conn[i]] + [v for i in quad for v in conn[i]], np.int32)
    if cell_nodes.size and (cell_nodes.min() < 0 or cell_nodes.max() >= nnodes):
        raise ValueError(f"{path}: a cell names a node outside 1..{nnodes}")

    bnds = []
    pos += ncells
    for _ in range(nbsets):
        pos = section(pos, "BOUNDARY CONDITIONS")
        head = lines[pos].split()
        k = len(head)
        while k > 0 and re.fullmatch(r"[+-]?\d+", head[k - 1]):
            k -= 1
        if k == 0 or len(head) - k < 2:
            raise ValueError(f"{path}: malformed boundary header {lines[pos]!r}")
        name = " ".join(head[:k])
        itype, n = int(head[k]), int(head[k + 1])
        if itype != 1:
            raise ValueError(f"{path}: boundary '{name}' is node-based (type {itype}); only cell-based sets")
        ids = np.array([_ints(lines[pos + 1 + j])[0] for j in range(n)], np.int64)
        if ids.size and (ids.min() < 1 or ids.max() > ncells):
            raise ValueError(f"{path}: boundary '{name}' names a cell outside 1..{ncells}")
        bnds.append((name, new_id[ids - 1].astype(np.int32)))
        pos += 1 + n
    return Mesh(xy, int(tri.size), int(quad.size), cell_nodes, bnds)


def write_plt(mesh: Mesh, path: str) -> None:
    """Tecplot finite-element file of the mesh, as utils/tools/read_gambit.f90
    writes it (:170-196): a triangle zone, or, when the mesh has quadrilaterals,
    a quadrilateral zone in which triangles repeat their third node."""
    cells = mesh.cells_as_array()
    quad = mesh.nquad > 0
    with open(path, "w") as f:
        f.write(' TITLE = "Example: 2D Finite Element Data"\n')
        f.write(' VARIABLES = "X", "Y"\n')
        f.write(f"ZONE NODES={mesh.nnodes} ELEMENTS={mesh.ncells} DATAPACKING=POINT, "
                f"ZONETYPE={'FEQUADRILATERAL' if quad else 'FETRIANGLE'}\n")
        for x, y in mesh.xy:
            f.write(f"{float(x)!r} {float(y)!r}\n")
        for c in cells:
            if quad:
                d = c[3] if c[3] >= 0 else c[2]
                f.write(f"{c[0] + 1} {c[1] + 1} {c[2] + 1} {d + 1}\n")
            else:
                f.write(f"{c[0] + 1} {c[1] + 1} {c[2] + 1}\n")


def import_gambit(path: str) -> Mesh:
    """Convert ``<base>.neu`` to ``<base>.grid``, ``<base>.bc`` and ``<base>.plt``
    beside it, as utils/tools/read_gambit.f90 does; returns the mesh."""
    mesh = read_gambit(path)
    base = path[: path.rfind(".")] if "." in os.path.basename(path) else path
    write_grid(mesh, base)
    write_plt(mesh, base + ".plt")
    return mesh


def concat(meshes: Sequence[Mesh]) -> Tuple[Mesh, np.ndarray, np.ndarray]:
    """Concatenate independent meshes into one ragged batch.

    Returns the batch mesh and the per-instance cell and node offset tables
    (length n+1).  Triangles of all instances come first, then quads, so the
    batch keeps the reference's tri-before-quad numbering only when no
    instance has quads (the batched solver only uses triangle meshes)."""
    if any(m.nquad for m in meshes):
        raise ValueError("concat supports triangle meshes only")
    coff = np.zeros(len(meshes) + 1, np.int64)
    noff = np.zeros(len(meshes) + 1, np.int64)
    for i, m in enumerate(meshes):
        coff[i + 1] = coff[i] + m.ncells
        noff[i + 1] = noff[i] + m.nnodes
    xy = np.concatenate([m.xy for m in meshes])
    cn = np.concatenate([m.cell_nodes.astype(np.int64) + noff[i] for i, m in enumerate(meshes)]).astype(np.int32)
    names = []
    for m in meshes:
        for n, _ in m.boundaries:
            if n not in names:
                names.append(n)
    bnds = []
    for n in names:
        cells = [np.asarray(c, np.int64) + coff[i] for i, m in enumerate(meshes) for nm, c in m.boundaries if nm == n]
        bnds.append((n, np.concatenate(cells).astype(np.int32)))
    return Mesh(xy, int(coff[-1]), 0, cn, bnds), coff, noff


if __name__ == "__main__":
    import sys
    if len(sys.argv) != 2 or not sys.argv[1].endswith(".neu"):
        sys.exit("usage: python -m fvs2d_amd.meshgen <mesh>.neu  (writes <mesh>.grid, .bc, .plt)")
    m = import_gambit(sys.argv[1])
    print(f"nnodes: {m.nnodes}\nncells triangles: {m.ntri}\nncells quad: {m.nquad}\n"
          f"# boundaries: {len(m.boundaries)}")
    for i, (name, c) in enumerate(m.boundaries, 1):
        print(f"boundary({i}) {name}: {len(c)} cells")
