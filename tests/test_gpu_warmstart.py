"""Learned warm start on the GPU, at a small scale (32x16 O-grids).

* gain: a network trained on 48 converged solutions lowers the median
  iteration count of 8 held-out conditions against free-stream starts, with
  every held-out condition converging;
* fallback: a network whose output is not a physical state sends every
  instance back to the free stream, after which each one repeats its cold
  solve exactly (state, forces, iterations + the one check step).
"""
import numpy as np
import pytest

from fvs2d_amd import warmstart
from fvs2d_amd.mlp import WarmStartNet

pytestmark = pytest.mark.gpu

KW = dict(cfl=1.25, max_steps=40000, drop=1e-8, chunk=250, freeze_after=20000, freeze_drop=None,
          freeze_stall=(2000, 0.1))


def test_warm_start_lowers_iterations():
    out = warmstart.run(48, 8, (32, 16), seed=0, steps=1500, batch=16384, lr=1e-3, depth=5, diverge=1e3,
                        fallback=1.0, group=64, n_seq=0, log=lambda s: None, **KW)
    assert out["train_converged"] >= 40
    assert out["loss_first_last"][1] < 0.1 * out["loss_first_last"][0]
    assert out["warm"]["converged"] >= out["cold"]["converged"]
    assert out["iteration_ratio_median"] < 0.95
    ratios = [p["warm_first_ratio"] for p in out["per_instance"]]
    assert np.median(ratios) < 0.2        # the warm field starts far closer to convergence


def test_fallback_repeats_cold_solve():
    conds = warmstart.draw(4, 5, (32, 16))
    b, rc, _, _ = warmstart.cold_solve(conds, diverge=1e3, **KW)
    cv_cold = b.solver.cvar
    f_cold = b.solver.wall_forces_batch()
    net = WarmStartNet(depth=1, batch=4096, seed=0)      # untrained, scaled far out of range
    rw, _, fell = warmstart.warm_solve(b, net, np.full(4, 1e3), diverge=1e3, fallback=1.0, **KW)
    assert fell.all()
    ok = rc["iterations"] > 0
    assert ok.all()
    assert np.array_equal(rw["iterations"], rc["iterations"] + 1)
    assert np.array_equal(b.solver.cvar, cv_cold)
    assert np.array_equal(b.solver.wall_forces_batch(), f_cold)
