"""Host mesh preparation: the library's topology/stencils against the oracle's
literal restatement of src/grid_procs.f90 and src/gradient_lsq.f90 (integer
work bit-exact, geometry bitwise identical), plus invariants that pin the
conventions independently of either restatement.

Meshes: the seeded jittered squares (Dirichlet and periodic) and the two
meshes the reference ships (examples/isentropic_vortex: irregular triangles,
examples/naca0012_ogrid: quads), stored as fixtures in tests/golden/.
"""
import numpy as np
import pytest

from fvs2d_amd import meshgen
from fvs2d_amd.solver import HostMesh
from oracle.oracle import Oracle
from tests.golden_meshes import load_example

CASES = {
    "square_dirichlet": lambda: meshgen.jittered_square(12, seed=3),
    "square_periodic": lambda: meshgen.jittered_square(12, seed=3, periodic=True),
    "vortex_example": lambda: load_example("vortex_example"),
    "naca0012_example": lambda: load_example("naca0012_example"),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("nn", [1, 0])
def test_topology_matches_oracle(case, nn):
    m = CASES[case]()
    if m.canon is not None and nn == 0:
        pytest.skip("face-neighbour stencil without boundaries is the plain neighbour list")
    hm = HostMesh(m, grad_method=3, lsq_nn=nn)
    orc = Oracle(m, grad_method=3, lsq_nn=nn)
    e1, e2 = hm.edges(), orc.edges()
    for k in e1:
        assert np.array_equal(e1[k], e2[k]), k
    c1, c2 = hm.cells(), orc.cells()
    for k in ("x", "y", "vol", "nghbre", "edge"):
        assert np.array_equal(c1[k], c2[k]), k
    l1, l2 = hm.lists(), orc.lists()
    for k in l1:
        assert np.array_equal(l1[k], l2[k]), k
    s1, s2 = hm.lsq(), orc.lsq()
    for k in s1:
        assert np.array_equal(s1[k], s2[k]), k


@pytest.mark.parametrize("case", list(CASES))
def test_topology_invariants(case):
    """Conventions of src/grid_procs.f90:401-791 checked from the geometry:
    c1 < c2, n points from c1 to c2 (and out of the domain on boundaries),
    each interior edge is listed by exactly its two cells, every cell is closed
    (sum of signed n*A = 0), and the Green-theorem area equals the shoelace
    area (src/grid_procs.f90:824-840)."""
    m = CASES[case]()
    hm = HostMesh(m)
    e, c = hm.edges(), hm.cells()
    inter = e["c2"] >= 0
    assert np.all(e["c1"][inter] < e["c2"][inter])
    xr = c["x"][e["c2"][inter]] + e["sx"][inter]
    yr = c["y"][e["c2"][inter]] + e["sy"][inter]
    dot = (xr - c["x"][e["c1"][inter]]) * e["nx"][inter] + (yr - c["y"][e["c1"][inter]]) * e["ny"][inter]
    assert np.all(dot > 0)
    b = ~inter
    outward = (e["x"][b] - c["x"][e["c1"][b]]) * e["nx"][b] + (e["y"][b] - c["y"][e["c1"][b]]) * e["ny"][b]
    assert np.all(outward > 0)
    assert np.allclose(np.hypot(e["nx"], e["ny"]), 1.0, atol=1e-15)
    # each edge id appears in its cells' edge lists exactly as c1/c2 say
    N = hm.ncells
    cnt = np.zeros(hm.nedges, np.int64)
    sx = np.zeros(N)
    sy = np.zeros(N)
    green = np.zeros(N)
    for k in range(4):
        ek = c["edge"][:, k]
        ok = ek >= 0
        cells = np.nonzero(ok)[0]
        ed = ek[ok]
        cnt += np.bincount(ed, minlength=hm.nedges)
        assert np.all((e["c1"][ed] == cells) | (e["c2"][ed] == cells))
        sgn = np.where(e["c1"][ed] == cells, 1.0, -1.0)
        sx[cells] += sgn * e["nx"][ed] * e["area"][ed]
        sy[cells] += sgn * e["ny"][ed] * e["area"][ed]
        green[cells] += sgn * e["nx"][ed] * e["area"][ed] * (e["x"][ed] - np.where(sgn > 0, 0.0, e["sx"][ed]))
    assert np.array_equal(cnt, np.where(inter, 2, 1))
    scale = np.sqrt(c["vol"])
    assert np.max(np.abs(sx) / scale) < 1e-12 and np.max(np.abs(sy) / scale) < 1e-12
    assert np.allclose(green, c["vol"], rtol=1e-10, atol=1e-14)
    assert np.all(c["vol"] > 0)


def test_square_counts():
    """Closed-form counts of the split-quad square: 2n^2 cells, 3n^2+2n edges,
    4n boundary edges; Euler characteristic V - E + F = 1."""
    n = 10
    hm = HostMesh(meshgen.jittered_square(n))
    assert hm.ncells == 2 * n * n
    assert hm.nedges == 3 * n * n + 2 * n
    assert hm.nedges_bndr == 4 * n
    assert hm.nnodes - hm.nedges + hm.ncells == 1
    hp = HostMesh(meshgen.jittered_square(n, periodic=True))
    assert hp.nedges == 3 * n * n and hp.nedges_bndr == 0
    assert np.isclose(hm.cells()["vol"].sum(), 100.0, rtol=1e-13)
    assert np.isclose(hp.cells()["vol"].sum(), 100.0, rtol=1e-13)


def test_lsq_nn_is_vertex_neighbourhood():
    """src/gradient_lsq.f90:227-271: the node-neighbour stencil is every other
    cell sharing a vertex, ascending (brute force over the cell-node table)."""
    m = meshgen.jittered_square(6, seed=1)
    s = HostMesh(m).lsq()
    cn = m.cells_as_array()
    for i in range(m.ncells):
        vs = set(cn[i][cn[i] >= 0])
        nb = [j for j in range(m.ncells) if j != i and vs & set(cn[j][cn[j] >= 0])]
        assert list(s["cell"][s["ptr"][i]:s["ptr"][i + 1]]) == nb


def test_bad_meshes_rejected():
    from fvs2d_amd._lib import FvsError
    m = meshgen.jittered_square(4)
    cw = meshgen.Mesh(m.xy, m.ntri, 0, m.cell_nodes.reshape(-1, 3)[:, ::-1].ravel().copy(), m.boundaries)
    with pytest.raises(FvsError, match="counter-clockwise"):
        HostMesh(cw)
    short = meshgen.Mesh(m.xy, m.ntri, 0, m.cell_nodes, [("dirichlet", m.boundaries[0][1][:-1])])
    with pytest.raises(FvsError, match="boundary cells"):
        HostMesh(short)
    bad_type = meshgen.Mesh(m.xy, m.ntri, 0, m.cell_nodes, [("solid_wall", m.boundaries[0][1])])
    with pytest.raises(FvsError, match="not implemented"):
        HostMesh(bad_type)


@pytest.mark.parametrize("gm,nn,lim", [(3, 1, 1), (1, 1, 1), (2, 1, 0)])
def test_batch_preparation_equals_whole_concatenation(gm, nn, lim):
    """Preparing a batch instance by instance on several threads
    (fvs_mesh_create_batch) gives the topology, lists and stencils of
    preparing the whole concatenation at once, bit for bit."""
    from fvs2d_amd.solver import HostMesh
    meshes = [meshgen.naca_ogrid(32, 16, thickness=0.12), meshgen.jittered_square(6, seed=3),
              meshgen.naca_ogrid(40, 20, thickness=0.1), meshgen.naca_ogrid(24, 12, thickness=0.15)]
    big, off, _ = meshgen.concat(meshes)
    whole = HostMesh(big, grad_method=gm, lsq_nn=nn, limiter=lim)
    part = HostMesh.batch(meshes, grad_method=gm, lsq_nn=nn, limiter=lim, nthreads=3)
    for get in ("edges", "cells", "lists", "lsq"):
        if get == "lsq" and gm != 3 and lim == 0:
            continue
        a, b = getattr(whole, get)(), getattr(part, get)()
        for k in a:
            assert np.array_equal(a[k], b[k]), (get, k)


def test_batch_preparation_names_every_failing_instance():
    """A batch whose instances 1 and 3 each hold a clockwise cell fails with
    an error naming both instances, so a caller can drop them in one pass."""
    from fvs2d_amd.solver import HostMesh
    from fvs2d_amd._lib import FvsError
    meshes = [meshgen.naca_ogrid(16, 8, thickness=0.12) for _ in range(4)]
    for k in (1, 3):
        cn = meshes[k].cell_nodes.copy()
        cn[15], cn[16] = cn[16], cn[15]          # cell 5 runs clockwise
        meshes[k].cell_nodes = cn
    with pytest.raises(FvsError, match=r"instances \[1,3\] failed") as e:
        HostMesh.batch(meshes, nthreads=2)
    assert HostMesh.failed_instances(e.value) == [1, 3]
    HostMesh.batch([meshes[0], meshes[2]], nthreads=2)
