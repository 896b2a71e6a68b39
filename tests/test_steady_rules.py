"""The host-side rules of the batched steady solve (BatchSolver.steady_solve_batch)
on a scripted stand-in for the device: each instance's density-residual norm is
a given function of its own step count and of whether its limiter is frozen.

* freezing on stall, then convergence;
* the give-up rule: an instance whose residual, after its limiter froze, falls
  by less than 1/factor over the window is given up, while one that keeps
  converging slowly is not;
* queued (continuous) batching gives every instance the iterations, freezing
  step and outcome it has when all instances step at once.
"""
import numpy as np

from fvs2d_amd.solver import BatchSolver


def converging(s, frozen):           # 1e-8 drop at step 18,421
    return np.exp(-s / 1000.0)


def stalling(s, frozen):             # flat at 1e-3 before and after freezing
    return max(np.exp(-s / 500.0), 1e-3) * (1.0 + 0.01 * np.sin(s))


class Slow:
    """Flat at 1e-2 until frozen, then falls by 0.4x every 8,000 steps."""

    def __init__(self):
        self.t0 = None

    def __call__(self, s, frozen):
        if not frozen:
            self.t0 = None
            return 1e-2 * (1.0 + 0.01 * np.sin(s))
        if self.t0 is None:
            self.t0 = s
        return 1e-2 * 0.4 ** ((s - self.t0) / 8000.0)


class Scripted(BatchSolver):
    """BatchSolver whose device calls are replaced by scripted norms."""

    def __init__(self, scripts):
        self.scripts = scripts
        self.ninst = len(scripts)
        self.active = np.ones(self.ninst, np.int32)
        self.steps = np.zeros(self.ninst, np.int64)
        self.frozen = np.zeros(self.ninst, np.int32)

    def set_active(self, flags):
        self.active = np.ascontiguousarray(flags, np.int32)

    def freeze_limiter_batch(self, flags):
        self.frozen = np.array(flags, np.int32)

    def freeze_limiter(self, frozen):
        self.frozen[:] = int(frozen)

    def batch_time_integration(self, t1, nsteps):
        out = np.zeros((nsteps, self.ninst, 4))
        for k in np.nonzero(self.active)[0]:
            for j in range(nsteps):
                self.steps[k] += 1
                out[j, k, :] = self.scripts[k](int(self.steps[k]), bool(self.frozen[k]))
        return out

    def __del__(self):
        pass


RULES = dict(drop=1e-8, chunk=250, freeze_stall=(2000, 0.1), r_ref=np.ones(3))


def test_give_up_rule():
    s = Scripted([converging, stalling, Slow()])
    r = s.steady_solve_batch(200000, give_up=(8000, 0.5), **RULES)
    assert r["iterations"][0] == 18421 and not r["gave_up"][0]
    assert r["iterations"][1] == -1 and r["gave_up"][1]
    assert r["frozen_at"][1] > 0
    # given up within two windows of freezing, far from the 200,000-step cap
    assert s.steps[1] <= r["frozen_at"][1] + 2 * 8000
    assert r["iterations"][2] > 0 and not r["gave_up"][2]
    assert r["frozen_at"][2] > 0


def test_without_give_up_the_stalled_instance_runs_to_the_cap():
    s = Scripted([converging, stalling, Slow()])
    r = s.steady_solve_batch(60000, **RULES)
    assert r["iterations"][1] == -1 and not r["gave_up"][1]
    assert s.steps[1] == 60000


def test_queued_batch_matches_all_active():
    a = Scripted([converging, stalling, Slow()])
    ra = a.steady_solve_batch(200000, give_up=(8000, 0.5), **RULES)
    q = Scripted([converging, stalling, Slow()])
    rq = q.steady_solve_batch(200000, give_up=(8000, 0.5), max_active=1, **RULES)
    for key in ("iterations", "frozen_at", "gave_up", "diverged"):
        assert np.array_equal(ra[key], rq[key]), key
    assert np.array_equal(a.steps, q.steps)
    # one at a time: the run lasts the sum of the instances' own steps
    assert rq["steps_run"] == int(q.steps.sum())
