"""Gambit neutral-file import (utils/tools/read_gambit.f90).

The test meshes are written here in the neutral-file layout Gambit 2.4 uses
(control block, nodal coordinates, elements/cells, one element group, one
section per boundary set with its header at the columns the reference reads
with ``a39,i1,4x,i4``): a jittered unit square whose cells are a shuffled mix
of quadrilaterals and triangle pairs, so the tri-first renumbering and the
boundary-id remapping both have work to do.  Pins:

* the importer against the mesh the file was written from (a construction of
  this test, independent of the importer);
* against the reference's own converter built from its source
  (oracle/_ref/read_gambit), on its ``.grid``/``.bc``/``.plt`` output;
* the imported mesh through the library's host topology against the oracle's
  restatement of src/grid_procs.f90, as for the shipped meshes.
"""
import os
import subprocess

import numpy as np
import pytest

from fvs2d_amd import meshgen
from fvs2d_amd.solver import HostMesh
from oracle.oracle import GAMBIT_REF, Oracle


def mixed_square(n=6, seed=0):
    """Cells of an n x n jittered grid in a shuffled file order: each square
    is a quadrilateral or two triangles; corner squares are split through the
    corner node so every boundary cell carries exactly one boundary edge.
    Returns nodes, file-ordered cells (0-based node lists) and boundary sets
    as lists of (file cell index, local face)."""
    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    xy = np.column_stack([X.ravel(), Y.ravel()])
    inner = (X.ravel() > 0) & (X.ravel() < 1) & (Y.ravel() > 0) & (Y.ravel() < 1)
    xy[inner] += rng.uniform(-0.2, 0.2, (inner.sum(), 2)) / n
    nid = lambda i, j: i * (n + 1) + j  # noqa: E731
    cells = []
    for i in range(n):
        for j in range(n):
            bl, br, tr, tl = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            corner = (i in (0, n - 1)) and (j in (0, n - 1))
            if corner:
                if (i, j) in ((0, 0), (n - 1, n - 1)):
                    cells += [[bl, br, tr], [bl, tr, tl]]
                else:
                    cells += [[bl, br, tl], [br, tr, tl]]
            elif rng.random() < 0.5:
                cells.append([bl, br, tr, tl])
            elif rng.random() < 0.5:
                cells += [[bl, br, tr], [bl, tr, tl]]
            else:
                cells += [[bl, br, tl], [br, tr, tl]]
    order = rng.permutation(len(cells))
    cells = [cells[k] for k in order]
    # boundary edges: a cell face (k -> k+1) with both nodes on one side
    side = {"slip_wall": lambda p: abs(p[1]) < 1e-12,
            "freestream": lambda p: abs(p[1] - 1) < 1e-12 or abs(p[0]) < 1e-12 or abs(p[0] - 1) < 1e-12}
    sets = {name: [] for name in side}
    for c, nodes in enumerate(cells):
        for f in range(len(nodes)):
            a, b = xy[nodes[f]], xy[nodes[(f + 1) % len(nodes)]]
            for name, on in side.items():
                if on(a) and on(b) and (abs(a[0] - b[0]) < 1e-12 or abs(a[1] - b[1]) < 1e-12):
                    sets[name].append((c, f + 1))
                    break
    return xy, cells, sets


def write_neu(path, xy, cells, sets, node_based=()):
    L = ["        CONTROL INFO 2.4.6", "** GAMBIT NEUTRAL FILE", "mixed", "PROGRAM:                Gambit"
         "     VERSION:  2.4.6", "Dec 2026",
         "     NUMNP     NELEM     NGRPS    NBSETS     NDFCD     NDFVL",
         f"{len(xy):10d}{len(cells):10d}{1:10d}{len(sets):10d}{2:10d}{2:10d}", "ENDOFSECTION",
         "   NODAL COORDINATES 2.4.6"]
    L += [f"{i + 1:10d}{x:20.11e}{y:20.11e}" for i, (x, y) in enumerate(xy)]
    L += ["ENDOFSECTION", "      ELEMENTS/CELLS 2.4.6"]
    for c, nodes in enumerate(cells):
        L.append(f"{c + 1:8d} {3 if len(nodes) == 3 else 2:2d} {len(nodes):2d} "
                 + "".join(f"{v + 1:8d}" for v in nodes))
    L += ["ENDOFSECTION", "       ELEMENT GROUP 2.4.6",
          f"GROUP:          1 ELEMENTS:{len(cells):11d} MATERIAL:          2 NFLAGS:          1",
          "                           fluid", "       0"]
    ids = list(range(1, len(cells) + 1))
    L += ["".join(f"{v:8d}" for v in ids[k:k + 10]) for k in range(0, len(ids), 10)]
    L.append("ENDOFSECTION")
    for name, ent in sets.items():
        L += [" BOUNDARY CONDITIONS 2.4.6",
              f"{name:>32s}{0 if name in node_based else 1:8d}{len(ent):8d}{0:8d}{6:8d}"]
        L += [f"{c + 1:10d}{2 if c < len(cells) and len(cells[c]) == 4 else 3:5d}{f:5d}" for c, f in ent]
        L.append("ENDOFSECTION")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


def expected(xy, cells, sets):
    tri = [k for k, c in enumerate(cells) if len(c) == 3]
    quad = [k for k, c in enumerate(cells) if len(c) == 4]
    new = {k: i for i, k in enumerate(tri + quad)}
    cn = np.array([v for k in tri + quad for v in cells[k]], np.int32)
    bnds = [(name, np.array([new[c] for c, _ in ent], np.int32)) for name, ent in sets.items()]
    return len(tri), len(quad), cn, bnds


@pytest.mark.parametrize("seed", [0, 1])
def test_read_gambit_matches_construction(tmp_path, seed):
    xy, cells, sets = mixed_square(6, seed)
    p = str(tmp_path / "mixed.neu")
    write_neu(p, xy, cells, sets)
    m = meshgen.read_gambit(p)
    nt, nq, cn, bnds = expected(xy, cells, sets)
    assert (m.ntri, m.nquad) == (nt, nq) and nt > 0 and nq > 0
    np.testing.assert_allclose(m.xy, xy, rtol=1e-10, atol=1e-12)
    assert np.array_equal(m.cell_nodes, cn)
    assert [b[0] for b in m.boundaries] == [b[0] for b in bnds]
    for (_, a), (_, b) in zip(m.boundaries, bnds):
        assert np.array_equal(a, b)
    # every listed cell really carries its boundary edge after renumbering
    ca = m.cells_as_array()
    for name, ids in m.boundaries:
        for i in ids:
            nodes = ca[i][ca[i] >= 0]
            yv = m.xy[nodes, 1]
            assert (np.sum(np.abs(yv) < 1e-12) >= 2) if name == "slip_wall" else True


def test_import_writes_reference_format(tmp_path):
    xy, cells, sets = mixed_square(5, 3)
    p = str(tmp_path / "mesh.neu")
    write_neu(p, xy, cells, sets)
    m = meshgen.import_gambit(p)
    r = meshgen.read_grid(str(tmp_path / "mesh"))
    assert np.array_equal(r.xy, m.xy) and np.array_equal(r.cell_nodes, m.cell_nodes)
    assert (r.ntri, r.nquad) == (m.ntri, m.nquad)
    for (n1, a), (n2, b) in zip(r.boundaries, m.boundaries):
        assert n1 == n2 and np.array_equal(a, b)
    assert os.path.exists(tmp_path / "mesh.plt")


def _plt_numbers(path):
    lines = open(path).read().splitlines()
    head = lines[2]
    body = [[float(v.replace("D", "E")) for v in ln.split()] for ln in lines[3:] if ln.strip()]
    return head, body


@pytest.mark.skipif(not os.path.exists(GAMBIT_REF), reason="reference converter not built (needs amdflang)")
@pytest.mark.parametrize("seed", [0, 2])
def test_read_gambit_matches_reference_tool(tmp_path, seed):
    xy, cells, sets = mixed_square(7, seed)
    ref_dir, own_dir = tmp_path / "ref", tmp_path / "own"
    ref_dir.mkdir()
    own_dir.mkdir()
    write_neu(str(ref_dir / "mesh.neu"), xy, cells, sets)
    write_neu(str(own_dir / "mesh.neu"), xy, cells, sets)
    # the tool asks for the file name on stdin (list-directed: no slashes)
    subprocess.run([GAMBIT_REF], input="mesh.neu\n", text=True, cwd=ref_dir, check=True,
                   capture_output=True, timeout=60)
    ref = meshgen.read_grid(str(ref_dir / "mesh"))
    own = meshgen.import_gambit(str(own_dir / "mesh.neu"))
    assert (ref.ntri, ref.nquad) == (own.ntri, own.nquad)
    np.testing.assert_allclose(own.xy, ref.xy, rtol=1e-14, atol=1e-15)
    assert np.array_equal(own.cell_nodes, ref.cell_nodes)
    assert [b[0] for b in own.boundaries] == [b[0] for b in ref.boundaries]
    for (_, a), (_, b) in zip(own.boundaries, ref.boundaries):
        assert np.array_equal(a, b)
    h1, b1 = _plt_numbers(ref_dir / "mesh.plt")
    h2, b2 = _plt_numbers(own_dir / "mesh.plt")
    assert h1.split() == h2.split()
    assert len(b1) == len(b2)
    for r1, r2 in zip(b1, b2):
        np.testing.assert_allclose(r2, r1, rtol=1e-14, atol=1e-15)


def test_imported_mesh_topology_matches_oracle(tmp_path):
    xy, cells, sets = mixed_square(8, 4)
    p = str(tmp_path / "mesh.neu")
    write_neu(p, xy, cells, sets)
    m = meshgen.read_gambit(p)
    hm = HostMesh(m, grad_method=3, lsq_nn=1)
    orc = Oracle(m, grad_method=3, lsq_nn=1)
    e1, e2 = hm.edges(), orc.edges()
    for k in e1:
        assert np.array_equal(e1[k], e2[k]), k
    c1, c2 = hm.cells(), orc.cells()
    for k in ("x", "y", "vol", "nghbre", "edge"):
        assert np.array_equal(c1[k], c2[k]), k
    # closed-form check: the cell areas tile the unit square
    assert abs(np.sum(c1["vol"]) - 1.0) < 1e-13


def test_read_gambit_rejects_what_it_cannot_map(tmp_path):
    xy, cells, sets = mixed_square(4, 5)
    p = str(tmp_path / "node.neu")
    write_neu(p, xy, cells, sets, node_based=("slip_wall",))
    with pytest.raises(ValueError, match="node-based"):
        meshgen.read_gambit(p)
    bad = [c + [c[0]] if len(c) == 4 else c for c in cells]  # a 5-node cell
    p = str(tmp_path / "penta.neu")
    write_neu(p, xy, bad, sets)
    with pytest.raises(ValueError, match="5 nodes"):
        meshgen.read_gambit(p)
    p = str(tmp_path / "range.neu")
    write_neu(p, xy, cells, {"slip_wall": [(len(cells), 1)]})
    with pytest.raises(ValueError, match="outside"):
        meshgen.read_gambit(p)
