"""Shared helpers for the tests: meshes, schemes and comparisons."""
import numpy as np

from fvs2d_amd import meshgen
from fvs2d_amd.solver import Scheme


def vortex_scheme(**kw):
    """Config-1 style isentropic vortex: rho=u=v=p=1, K=5, centre (0,0)."""
    d = dict(grad_method=3, lsq_nn=1, lsq_p=0.0, limiter=0, recon=2, kappa=0.0, gamma=1.4, mach=0.5,
             aoa_deg=0.0, lvortex=1, vortex_x0=0.0, vortex_y0=0.0, vortex_K=5.0,
             vortex_inf=(1.0, 1.0, 1.0, 1.0), rk_scheme=2, rk_order=4, steady=0, dt=0.01, cfl=0.5)
    d.update(kw)
    return Scheme(**d)


def rel_err(a, b, floor=1e-300):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))


def verr_err(got, ref, min_same_location=0.9):
    """Largest relative difference of the 13 error norms of vortex-error
    records (one record or one per step), after checking that the location
    of the largest density error (the last two values) is the same cell in at
    least ``min_same_location`` of the records (a near tie between two cells
    may resolve differently under rounding)."""
    got, ref = np.atleast_2d(got), np.atleast_2d(ref)
    same = np.all(got[:, 13:15] == ref[:, 13:15], axis=1)
    assert np.mean(same) >= min_same_location, (got[~same, 13:15], ref[~same, 13:15])
    return rel_err(got[:, :13], ref[:, :13])


def scaled_err(a, b):
    """max |a-b| / max|b| per column (for quantities that cross zero)."""
    a = np.asarray(a).reshape(b.shape[0], -1) if a.ndim > 1 else a[:, None]
    b = np.asarray(b).reshape(b.shape[0], -1) if b.ndim > 1 else b[:, None]
    s = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return float(np.max(np.abs(a - b).max(axis=0) / s))


def config1_dt(oracle):
    """dt = 0.5 * min_i V_i / (0.5 ws_i) on the initial state (the reference's
    local-step formula, src/runge_kutta.f90:433, minimised over cells)."""
    oracle.compute_residual(0.0)
    vol = oracle.cells()["vol"]
    return float(np.min(0.5 * vol / (0.5 * oracle.ws)))
