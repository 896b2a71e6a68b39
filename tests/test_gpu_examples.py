"""The reference's own example cases on the GPU.

Its NACA 0012 O-grid (65,536 quadrilaterals) with its steady transonic deck,
and its vortex square (7,226 triangles) with its unsteady deck and vortex
file.  Meshes are stored as arrays and decks as text in tests/golden/.  Every
other GPU test runs on generated triangle meshes.  The O-grid is the only
all-quadrilateral mesh: four faces per cell in the ordered gather, four
vertices in the limiters, quadrilateral face lists in the fused kernels.

Bars as in the other parity tests: 1e-10 per cell after the stated steps,
or ten times the oracle's own spread from a start one ulp off where the
transonic start-up amplifies rounding beyond that; 1e-12 for one residual.
In both decks the free stream has no vertical velocity, so the vertical
momentum is rounding noise along the O-grid's symmetry line and 1e-11 or
less far from the vortex: the per-cell relative error takes as its floor
1e-3 of each variable's largest magnitude (there the bar is absolute, 1e-13
of that magnitude).
"""
import dataclasses
import os

import numpy as np
import pytest

from fvs2d_amd.solver import Solver, read_input, read_vortex
from oracle.oracle import Oracle
from tests.common import rel_err, scaled_err, verr_err
from tests.golden_meshes import HERE, load_example

pytestmark = pytest.mark.gpu

NACA_STEPS = 40


def deck(name):
    _, sch, run = read_input(os.path.join(HERE, name + ".fvs2d.input"))
    if sch.lvortex:
        sch = read_vortex(os.path.join(HERE, name + ".fvs2d.vortex"), sch)
    return sch, run


def sym_err(a, b):
    """Per-cell relative error with a floor of 1e-3 x max|b| per variable."""
    return rel_err(a, b, floor=1e-3 * np.max(np.abs(b), axis=0))


def oracle(m, sch):
    orc = Oracle(m, **sch.oracle_kwargs())
    orc.set_threads(8)  # bitwise the serial oracle
    return orc


@pytest.fixture(scope="module")
def naca():
    m = load_example("naca0012_example")
    sch, _ = deck("naca0012_example")
    orc = oracle(m, sch)
    orc.initialize(1, 0.0)
    norms = orc.time_integration(0.0, NACA_STEPS)
    pert = oracle(m, sch)
    pert.initialize(1, 0.0)
    cv = pert.cvar
    pert.cvar = cv * (1.0 + 2.2e-16 * np.random.default_rng(0).choice([-1.0, 1.0], cv.shape))
    pert.time_integration(0.0, NACA_STEPS)
    tol = max(1e-10, 10.0 * sym_err(pert.cvar, orc.cvar))
    return m, sch, orc, norms, tol


def test_naca0012_example_steady(naca):
    """The deck as given (steady SSP-RK(4,2), CFL 1.25, LSQ-nn, no limiter,
    M 0.8, 0 deg): the fused path in Morton, coordinate-bisection and the
    reference's cell order and its cell-by-cell gradient arrangement (all
    bitwise equal), and the face-buffer kernels, against the oracle; then the
    wall forces of the same state."""
    m, sch, orc, n_ref, tol = naca
    assert (m.ntri, m.nquad) == (0, 65536)
    out = {}
    for path, reorder in ((0, 1), (3, 1), (0, 0), (0, 2), (1, 0)):
        s = Solver(m, sch, kernel_path=path, reorder=reorder)
        s.initialize_solution(1, 0.0)
        n = s.time_integration(0.0, NACA_STEPS)
        out[path, reorder] = (s.cvar, n, s.wall_forces())
        s.close()
    u = out[0, 1][0]
    err = sym_err(u, orc.cvar)
    print(f"NACA 0012 example, {NACA_STEPS} steps: GPU vs oracle {err:.2e}, bar {tol:.2e}", flush=True)
    assert err <= tol
    assert np.array_equal(u, out[3, 1][0])
    assert np.array_equal(u, out[0, 0][0])
    assert np.array_equal(u, out[0, 2][0])
    assert sym_err(out[1, 0][0], orc.cvar) <= tol
    assert rel_err(out[0, 1][1], n_ref) <= 100 * tol
    orc.cvar2pvar()
    assert np.allclose(out[0, 1][2], orc.wall_forces(), rtol=tol, atol=1e-13)


def test_naca0012_example_ssp3(naca):
    """SSP-RK3 on the same O-grid and deck (CFL 0.3: the local step is four
    times the common convention), 20 steps: the stages on conservative states
    (default path) bitwise equal to the primitive-state arrangement
    (kernel_path 3), and within the bar of the oracle."""
    m, sch0, _, _, _ = naca
    sch = dataclasses.replace(sch0, rk_scheme=2, cfl=0.3)
    orc = oracle(m, sch)
    orc.initialize(1, 0.0)
    orc.time_integration(0.0, 20)
    out = []
    for path in (0, 3):
        s = Solver(m, sch, kernel_path=path, reorder=1)
        s.initialize_solution(1, 0.0)
        s.time_integration(0.0, 20)
        out.append(s.cvar)
        s.close()
    assert np.array_equal(out[0], out[1])
    assert sym_err(out[0], orc.cvar) <= 1e-10


@pytest.mark.parametrize("grad", [(3, 1), (3, 0), (1, 1), (2, 1)], ids=["lsq_nn", "lsq_fn", "ggcb", "ggnb"])
@pytest.mark.parametrize("limiter", [0, 1, 2])
def test_naca0012_example_residual(naca, grad, limiter):
    """One residual, gradient and limiter evaluation on the O-grid for every
    gradient and the Venkatakrishnan and Barth limiters, on the state the
    deck reaches after the steps above (a developed flow around the wall).

    Most of the grid is still exactly free stream.  There a Green-Gauss
    gradient is rounding noise (the closed-surface sum of n A), exactly zero
    on one side and 1e-17 on the other, and Barth's ratio (q_max - q) / b
    with b that noise flips between 0 and 1: limiter values are compared
    where the oracle's gradient is above rounding level, and must lie in
    [0, 1] everywhere.  The residual is unaffected (b multiplies the noise).
    Where b is small but not noise, Barth's ratio still amplifies rounding:
    the oracle's own limiter values move by up to ~4e-5 when its input state
    is moved by one ulp (Venkatakrishnan's by ~1e-13), so the bar on limiter
    values is max(1e-10, 10 x that spread), measured here on the oracle."""
    m, sch0, orc0, _, _ = naca
    cv = orc0.cvar
    kw = sch0.oracle_kwargs()
    kw.update(grad_method=grad[0], lsq_nn=grad[1], limiter=limiter)
    orc = Oracle(m, **kw)
    orc.cvar = cv
    orc.compute_residual(0.0)
    pert = Oracle(m, **kw)
    pert.cvar = cv * (1.0 + 2.2e-16 * np.random.default_rng(1).choice([-1.0, 1.0], cv.shape))
    pert.compute_residual(0.0)
    gpu = Solver(m, sch0, grad_method=grad[0], lsq_nn=grad[1], limiter=limiter, reorder=1)
    gpu.cvar = cv
    r, ws = gpu.compute_residual(0.0)
    assert np.all(np.isfinite(orc.resid))
    assert scaled_err(r, orc.resid) <= 1e-12
    assert rel_err(ws, orc.ws) <= 1e-13
    g = gpu.compute_gradient_cellcntr()
    assert scaled_err(g.reshape(2, -1).T, orc.grad.reshape(2, -1).T) <= 1e-12
    phi = gpu.compute_gradient_limiter()
    live = np.max(np.abs(orc.grad), axis=(0, 2)) > 1e-8
    assert live.sum() > 1000
    bar = max(1e-10, 10.0 * float(np.max(np.abs(pert.phi[live] - orc.phi[live]))))
    assert np.max(np.abs(phi[live] - orc.phi[live])) <= bar
    assert np.all((phi >= 0.0) & (phi <= 1.0))
    gpu.close()


def test_vortex_example_unsteady():
    """The vortex deck and vortex file as given (classic RK4, dt 0.01,
    LSQ-fn), 50 steps: states and the vortex error norms against the
    oracle."""
    m = load_example("vortex_example")
    sch, run = deck("vortex_example")
    assert (m.ntri, m.nquad) == (7226, 0) and sch.lvortex == 1
    steps = 50
    orc = oracle(m, sch)
    orc.initialize(1, 0.0)
    orc.time_integration(0.0, steps)
    s = Solver(m, sch, reorder=1)
    s.initialize_solution(1, 0.0)
    norms, verr = s.time_integration(0.0, steps, vortex_error=True)
    assert sym_err(s.cvar, orc.cvar) <= 1e-10
    t = steps * run["dt"]
    assert verr_err(s.error_isentropic_vortex(t), orc.vortex_error(t), 1.0) <= 1e-9
    assert verr_err(verr[-1], orc.vortex_error(t), 1.0) <= 1e-9
    s.close()
