"""Device cell orders (fvs_mesh_cell_order, the `reorder` option), host only.

The order is a layout choice: every order gives bitwise the same states (the
GPU tests hold that).  Here the orders themselves are checked: they are
permutations, keep batch instances contiguous and in order, and recursive
coordinate bisection (reorder 2) has its defining structure -- every aligned
block of 2^k x 64 positions is cut at its middle along the longer side of its
centroids' bounding box, leaves hold ascending cell ids -- and gives more
compact blocks (smaller stencil halos) than the Morton order on an O-grid.
"""
import numpy as np
import pytest

from fvs2d_amd import meshgen
from fvs2d_amd.solver import FvsError, HostMesh


def halo_per_cell(perm, ptr, cell, bs):
    """Distinct outside cells read by the stencils of each aligned block of bs
    device positions, summed over blocks, per cell."""
    n = len(perm)
    iperm = np.empty(n, np.int64)
    iperm[perm] = np.arange(n)
    owner = np.repeat(np.arange(n), np.diff(ptr))
    blk = iperm[owner] // bs
    nb = iperm[cell]
    out = nb // bs != blk
    return len(np.unique(blk[out] * n + nb[out])) / n


def test_orders_are_permutations():
    m = meshgen.jittered_square(20, seed=1)
    h = HostMesh(m)
    assert np.array_equal(h.cell_order(0), np.arange(h.ncells))
    for r in (1, 2):
        p = h.cell_order(r)
        assert np.array_equal(np.sort(p), np.arange(h.ncells)), r
        assert np.array_equal(p, h.cell_order(r)), r  # deterministic


def test_bisection_structure():
    m = meshgen.naca_ogrid(64, 32, thickness=0.12)  # 4096 cells: every cut at a block middle
    h = HostMesh(m)
    c = h.cells()
    p = h.cell_order(2)
    x, y = c["x"][p], c["y"][p]
    n = len(p)
    assert n == 4096
    for b in range(0, n, 64):
        assert np.all(np.diff(p[b:b + 64]) > 0), b  # leaves in ascending id
    size = 128
    while size <= n:
        for b in range(0, n, size):
            s = slice(b, b + size)
            h2 = size // 2
            byx = np.ptp(x[s]) >= np.ptp(y[s])
            k = x if byx else y
            assert k[b:b + h2].max() <= k[b + h2:b + size].min(), (size, b, byx)
        size *= 2


def test_batch_instances_stay_contiguous():
    """Instances off the 64-cell alignment: each keeps its own positions, and
    cuts fall on global multiples of 64 (a 64-aligned window inside one
    instance is one leaf: ascending ids)."""
    meshes = [meshgen.naca_ogrid(16, 10, thickness=0.1), meshgen.jittered_square(9, seed=2),
              meshgen.jittered_square(3, seed=1),  # 18 cells: an instance smaller than one leaf
              meshgen.naca_ogrid(24, 12, thickness=0.15), meshgen.jittered_square(13, seed=5)]
    hb = HostMesh.batch(meshes)
    from fvs2d_amd.meshgen import concat
    _, off, _ = concat(meshes)
    assert any(o % 64 for o in off[1:-1])
    for r in (1, 2):
        p = hb.cell_order(r, off)
        for k in range(len(meshes)):
            assert np.array_equal(np.sort(p[off[k]:off[k + 1]]), np.arange(off[k], off[k + 1])), (r, k)
    p = hb.cell_order(2, off)
    for b in range(0, len(p), 64):
        for k in range(len(meshes)):
            lo, hi = max(b, off[k]), min(b + 64, off[k + 1])
            if hi > lo:
                assert np.all(np.diff(p[lo:hi]) > 0), (b, k)


def test_bisection_blocks_compact_on_ogrid():
    """Stencil halo of 256- and 128-cell blocks (the gradient kernel's LDS
    staging) on a graded O-grid: bisection below the Z-curve's."""
    m = meshgen.naca_ogrid(128, 64, thickness=0.12)
    h = HostMesh(m, grad_method=3, lsq_nn=1)
    l = h.lsq()
    for bs in (256, 128):
        hm = halo_per_cell(h.cell_order(1), l["ptr"], l["cell"], bs)
        hr = halo_per_cell(h.cell_order(2), l["ptr"], l["cell"], bs)
        print(f"O-grid 128x64, {bs}-cell blocks: halo per cell Morton {hm:.3f}, bisection {hr:.3f}")
        assert hr < 0.9 * hm, bs
        assert hr < halo_per_cell(h.cell_order(0), l["ptr"], l["cell"], bs)


def test_cell_order_errors():
    h = HostMesh(meshgen.jittered_square(6, seed=0))
    with pytest.raises(FvsError, match="reorder"):
        h.cell_order(3)
    with pytest.raises(FvsError, match="cell_offsets"):
        h.cell_order(2, [0, 5])
    with pytest.raises(FvsError, match="cell_offsets"):
        h.cell_order(2, [0, 0, h.ncells])
