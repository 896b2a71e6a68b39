"""fvs_time_integration_noalloc: the variant of fvs_time_integration that
allocates nothing (its per-step outputs are staged in a fixed device array of
the library).  Through the C-ABI, on two identically prepared contexts, it
gives the same state, residual norms and vortex errors as
fvs_time_integration, bit for bit (same kernels, same order), on two calls in
a row, the second with fewer steps; a step count beyond its staging is
rejected with an error before anything runs."""
import ctypes as C

import numpy as np
import pytest

from fvs2d_amd import _lib, meshgen
from fvs2d_amd.solver import Solver
from tests.common import vortex_scheme

pytestmark = pytest.mark.gpu


def _run(L, s, fn, nsteps, with_verr):
    norms = np.zeros((nsteps, 4))
    verr = np.zeros((nsteps, 15)) if with_verr else None
    rc = fn(s._h, 0.0, nsteps, _lib.ptr(norms), _lib.ptr(verr))
    assert rc == 0, L.fvs_last_error()
    return norms, verr


@pytest.mark.parametrize("case", ["dirichlet_vortex", "periodic_graph"])
def test_noalloc_equals_time_integration(case):
    L = _lib.load()
    periodic = case == "periodic_graph"
    m = meshgen.jittered_square(24, periodic=periodic, seed=2)
    sch = vortex_scheme(dt=0.01, reorder=1)
    a, b = Solver(m, sch), Solver(m, sch)
    if periodic:
        a.set_graph(True)  # the allocating entry point may replay a graph; the variant never does
    for s in (a, b):
        s.initialize_solution(1, 0.0)
    for nsteps in (37, 11):
        na, va = _run(L, a, L.fvs_time_integration, nsteps, not periodic)
        nb, vb = _run(L, b, L.fvs_time_integration_noalloc, nsteps, not periodic)
        assert np.array_equal(na, nb)
        if not periodic:
            assert np.array_equal(va, vb)
        assert np.array_equal(a.cvar, b.cvar)
    assert np.all(np.isfinite(na)) and np.any(na != 0)


def test_noalloc_rejects_too_many_steps():
    L = _lib.load()
    m = meshgen.jittered_square(8, periodic=True, seed=0)
    s = Solver(m, vortex_scheme(dt=0.01))
    s.initialize_solution(1, 0.0)
    before = s.cvar
    assert L.fvs_time_integration_noalloc(s._h, 0.0, 4097, None, None) == -1
    assert b"4096" in L.fvs_last_error()
    assert np.array_equal(s.cvar, before)
