"""Learned warm start of batched steady airfoil solves.

The reference starts every steady run from the free stream
(src/initialize.f90:36-43); nothing in it is learned.  This module adds the
warm start on top of the batched FP64 solver:

1. conditions are drawn from a seed (Mach U[0.5, 0.85], angle of attack
   U[-2, 4] deg, thickness U[0.10, 0.15]) on one O-grid resolution;
2. the batched solver converges them from the free stream to a 1e-8 residual
   drop; every cell of every converged instance becomes a training row
   (cell record -> scaled departure from the free stream);
3. the network (include/fvs2d_mlp.h) is trained on the device;
4. held-out conditions are solved twice in batches, cold (free stream) and
   warm (network output), to the same absolute tolerance: 1e-8 times the
   cold start's first-step density residual.  A warm instance falls back to
   the free stream if its field is not physical (density or pressure not
   positive), if its first-step residual exceeds ``fallback`` times the cold
   one, or if it diverges later (non-finite or above ``diverge`` times the
   cold first-step residual).

Run as ``python -m fvs2d_amd.warmstart`` (one JSON line at the end, progress
lines before it).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
from scipy.spatial import cKDTree

from . import cases, meshgen
from .mlp import WarmStartNet
from .solver import BatchSolver, Solver

GAMMA = 1.4


def cell_centres(mesh: meshgen.Mesh) -> np.ndarray:
    tri = mesh.cell_nodes[: 3 * mesh.ntri].reshape(-1, 3)
    return mesh.xy[tri].mean(axis=1)


def wall_distance(mesh: meshgen.Mesh, n_around: int, pts: np.ndarray) -> np.ndarray:
    """Distance from pts to the closed surface polyline of an O-grid (its
    first n_around nodes): nearest surface node, then exact distance to the
    two segments at that node and at its two neighbours."""
    surf = mesh.xy[:n_around]
    _, nn = cKDTree(surf).query(pts, k=3)
    best = np.full(pts.shape[0], np.inf)
    for c in range(3):
        for off in (-1, 0):
            i0 = (nn[:, c] + off) % n_around
            i1 = (i0 + 1) % n_around
            a, b = surf[i0], surf[i1]
            ab = b - a
            t = np.clip(np.einsum("ij,ij->i", pts - a, ab) / np.einsum("ij,ij->i", ab, ab), 0.0, 1.0)
            d = np.hypot(*(pts - (a + t[:, None] * ab)).T)
            best = np.minimum(best, d)
    return best


def freestream(mach: float, aoa: float, gamma: float = GAMMA) -> np.ndarray:
    """Conservative free stream (src/initialize.f90:36-43: rho = 1, p = 1/gamma)."""
    u, v = mach * np.cos(np.radians(aoa)), mach * np.sin(np.radians(aoa))
    return np.array([1.0, u, v, (1.0 / gamma) / (gamma - 1.0) + 0.5 * (u * u + v * v)])


def records(mesh: meshgen.Mesh, n_around: int, cond: dict) -> np.ndarray:
    """Cell records (x, y, wall distance, Mach, AoA, thickness) as float32."""
    c = cell_centres(mesh)
    d = wall_distance(mesh, n_around, c)
    n = c.shape[0]
    return np.column_stack([c, d, np.full(n, cond["mach"]), np.full(n, cond["aoa"]),
                            np.full(n, cond["thickness"])]).astype(np.float32)


def draw(n: int, seed: int, res):
    """Seeded conditions on one resolution (the condition ranges of
    cases.draw_instances)."""
    return cases.draw_instances(n, seed, (tuple(res),))


def _physical(U: np.ndarray, gamma: float = GAMMA) -> bool:
    rho = U[:, 0]
    p = (gamma - 1.0) * (U[:, 3] - 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2) / np.maximum(rho, 1e-300))
    return bool(np.all(np.isfinite(U)) and np.all(rho > 0.0) and np.all(p > 0.0))


class Batch:
    """Instances of one batched solve with their meshes and records."""

    def __init__(self, conds, cfl: float, graph: bool = True, limiter: int = 1):
        self.conds = conds
        self.meshes = [meshgen.naca_ogrid(c["n_around"], c["n_layers"], thickness=c["thickness"]) for c in conds]
        _, sch = cases.naca(16, 8, cfl=cfl, limiter=limiter)
        self.solver = BatchSolver(self.meshes, [(c["mach"], c["aoa"]) for c in conds], sch)
        self.solver.set_graph(graph)
        self.off = self.solver.cell_offsets
        self.solver.initialize_freestream()
        self.r_ref = self.solver.reference_norms()

    def records(self) -> np.ndarray:
        return np.concatenate([records(m, c["n_around"], c) for m, c in zip(self.meshes, self.conds)])

    def reset(self, k: int):
        cv = self.solver.cvar
        cv[self.off[k]:self.off[k + 1]] = freestream(self.conds[k]["mach"], self.conds[k]["aoa"])
        self.solver.cvar = cv
        return True

    def solve(self, max_steps, drop, chunk, freeze_after, freeze_drop, freeze_stall, diverge, max_active=None,
              give_up=None):
        return self.solver.steady_solve_batch(max_steps, drop=drop, chunk=chunk, freeze_after=freeze_after,
                                              freeze_drop=freeze_drop, r_ref=self.r_ref, diverge=diverge,
                                              on_diverge=self.reset, freeze_stall=freeze_stall,
                                              max_active=max_active, give_up=give_up)


def cold_solve(conds, cfl, max_steps, drop, chunk, freeze_after, freeze_drop, freeze_stall, diverge, limiter=1,
               max_active=None, give_up=None):
    t0 = time.perf_counter()
    b = Batch(conds, cfl, limiter=limiter)
    setup = time.perf_counter() - t0
    b.solver.initialize_freestream()
    t0 = time.perf_counter()
    res = b.solve(max_steps, drop, chunk, freeze_after, freeze_drop, freeze_stall, diverge, max_active, give_up)
    sec = time.perf_counter() - t0
    return b, res, sec, setup


def warm_solve(b: Batch, net: WarmStartNet, scale, cfl, max_steps, drop, chunk, freeze_after, freeze_drop,
               freeze_stall, diverge, fallback: float, limiter=1, max_active=None, give_up=None):
    """Warm start of the instances of b (its solver is reused).  Returns the
    solve result with one extra step counted for the first-step check (and
    res["first_ratio"], that step's norm over the cold first-step norm), the
    wall time (prediction included) and the per-instance fallback flags."""
    s = b.solver
    n = len(b.conds)
    t0 = time.perf_counter()
    rec = b.records()
    out = net.predict(rec).astype(np.float64) * scale
    cv = np.empty((s.ncells, 4))
    fell = np.zeros(n, bool)
    for k, c in enumerate(b.conds):
        u_inf = freestream(c["mach"], c["aoa"])
        U = u_inf + out[b.off[k]:b.off[k + 1]]
        if not _physical(U):
            U = np.broadcast_to(u_inf, U.shape)
            fell[k] = True
        cv[b.off[k]:b.off[k + 1]] = U
    s.cvar = cv
    first = s.batch_time_integration(0.0, 1)[0, :, 0]
    fell |= ~np.isfinite(first) | (first > fallback * b.r_ref)
    if np.any(fell):
        # fallen-back instances restart from the exact free stream, so they
        # repeat their cold solve (one step later)
        cv = s.cvar
        for k in np.nonzero(fell)[0]:
            cv[b.off[k]:b.off[k + 1]] = freestream(b.conds[k]["mach"], b.conds[k]["aoa"])
        s.cvar = cv
    res = b.solve(max_steps, drop, chunk, freeze_after, freeze_drop, freeze_stall, diverge, max_active, give_up)
    sec = time.perf_counter() - t0
    res["first_ratio"] = first / b.r_ref
    it = res["iterations"]
    res["iterations"] = np.where(it > 0, it + 1, it)
    return res, sec, fell


def write_warm_start(net: WarmStartNet, scale, mesh: meshgen.Mesh, n_around: int, cond: dict, base: str = "cont"):
    """Warm start of one O-grid instance as the reference's restart files
    (``<base>.cd`` / ``.s8``), so its Fortran driver can start from it with
    ntstart > 1 (src/initialize.f90:56-77).  Returns the state (ncells, 4)."""
    from . import restart
    U = freestream(cond["mach"], cond["aoa"]) + net.predict(records(mesh, n_around, cond)).astype(np.float64) * scale
    if not _physical(U):
        U = np.broadcast_to(freestream(cond["mach"], cond["aoa"]), U.shape).copy()
    restart.write_save(base, U, mesh.nnodes, 0)
    return U


def plain_launch_path(cond, cfl, limiter=1, steps=300):
    """The faster kernel arrangement for one instance solved on its own: the
    per-stage gradient + fused flux kernels (kernel_path 0) or the whole stage
    in one kernel (2), timed over a few hundred steps of a free-stream start."""
    best, path_best = None, 0
    for path in (0, 2):
        m, sch = cases.naca(cond["n_around"], cond["n_layers"], mach=cond["mach"], aoa=cond["aoa"],
                            thickness=cond["thickness"], cfl=cfl, limiter=limiter, kernel_path=path)
        s = Solver(m, sch)
        s.initialize_solution(1, 0.0)
        s.time_integration(0.0, 20)
        s.time_integration(0.0, steps)
        ms = s.last_elapsed_ms
        s.close()
        if best is None or ms < best:
            best, path_best = ms, path
    return path_best


def sequential_cold(conds, cfl, max_steps, drop, chunk, freeze_after, freeze_drop, freeze_stall, limiter=1,
                    max_active=None, give_up=None):
    """Plain (one instance at a time) cold solves: their wall time and count.
    Each runs in the faster kernel arrangement for its resolution (picked
    once per resolution, outside the timed solves)."""
    paths = {}
    for c in conds:
        key = (c["n_around"], c["n_layers"])
        if key not in paths:
            paths[key] = plain_launch_path(c, cfl, limiter)
    t0 = time.perf_counter()
    ok = 0
    its = []
    for c in conds:
        m, sch = cases.naca(c["n_around"], c["n_layers"], mach=c["mach"], aoa=c["aoa"], thickness=c["thickness"],
                            cfl=cfl, limiter=limiter, kernel_path=paths[(c["n_around"], c["n_layers"])])
        s = Solver(m, sch)
        s.set_graph(True)
        s.initialize_solution(1, 0.0)
        r = s.steady_solve(max_steps, drop=drop, chunk=chunk, freeze_after=freeze_after, freeze_at=freeze_drop,
                           freeze_stall=freeze_stall, give_up=give_up)
        ok += r["iterations"] is not None
        its.append(r["iterations"] or -1)
        s.close()
    return time.perf_counter() - t0, ok, its


def make_dataset(conds, group, cfl, max_steps, drop, chunk, freeze_after, freeze_drop, freeze_stall, diverge,
                 limiter=1, log=print, give_up=None):
    """Cold batched solves in groups; rows of the converged instances."""
    raws, dus, its, secs, n_ok = [], [], [], 0.0, 0
    for g0 in range(0, len(conds), group):
        part = conds[g0:g0 + group]
        b, res, sec, _ = cold_solve(part, cfl, max_steps, drop, chunk, freeze_after, freeze_drop, freeze_stall,
                                    diverge, limiter, give_up=give_up)
        secs += sec
        cv = b.solver.cvar
        rec = b.records()
        for k, c in enumerate(part):
            if res["iterations"][k] > 0:
                sl = slice(b.off[k], b.off[k + 1])
                raws.append(rec[sl])
                dus.append(cv[sl] - freestream(c["mach"], c["aoa"]))
                its.append(int(res["iterations"][k]))
                n_ok += 1
        log(f"dataset: {g0 + len(part)}/{len(conds)} instances solved, {n_ok} converged, {secs:.1f} s")
        b.solver.close()
    return np.concatenate(raws), np.concatenate(dus), its, secs


def run(n_train, n_test, res, seed, steps, batch, lr, depth, cfl, max_steps, drop, chunk, freeze_after, freeze_drop,
        freeze_stall, diverge, fallback, group, n_seq, limiter=1, log=print, test_res=None, stall_scale=False,
        max_active=None, give_up=None):
    """Train on n_train converged solves at resolution ``res``, then solve
    n_test held-out conditions at ``test_res`` (default ``res``; the network's
    inputs do not depend on the grid) cold and warm.  With ``stall_scale``
    the stall and give-up windows scale with n_around / 64 at each
    resolution (cases.stall_rule_for)."""
    test_res = tuple(res if test_res is None else test_res)
    kw = dict(cfl=cfl, max_steps=max_steps, drop=drop, chunk=chunk, freeze_after=freeze_after,
              freeze_drop=freeze_drop, freeze_stall=freeze_stall, limiter=limiter, give_up=give_up)
    out = dict(resolution=list(res), test_resolution=list(test_res), n_train=n_train, n_test=n_test, seed=seed,
               cfl=cfl, drop=drop, freeze_after=freeze_after, freeze_drop=freeze_drop, freeze_stall=freeze_stall,
               give_up=give_up, fallback=fallback, depth=depth, batch=batch, lr=lr, steps=steps, limiter=limiter)
    train_c = draw(n_train, seed, res)
    test_c = draw(n_test, seed + 1_000_003, test_res)

    def scaled(rule, r):
        return None if rule is None or not stall_scale else (int(rule[0] * r[0] / 64), rule[1])
    kw_train = dict(kw)
    kw_test = dict(kw, max_active=max_active)
    if stall_scale:
        kw_train.update(freeze_stall=scaled(freeze_stall, res), give_up=scaled(give_up, res))
        kw_test.update(freeze_stall=scaled(freeze_stall, test_res), give_up=scaled(give_up, test_res))
    raw, du, its, gen_s = make_dataset(train_c, group, diverge=diverge, log=log, **kw_train)
    out["train_converged"] = len(its)
    out["train_rows"] = int(raw.shape[0])
    out["train_solve_seconds"] = gen_s
    out["train_iterations_p50"] = float(np.percentile(its, 50))
    scale = du.std(axis=0)
    out["target_scale"] = scale.tolist()
    net = WarmStartNet(depth=depth, batch=batch, lr=lr, seed=seed)
    net.set_data(raw, (du / scale).astype(np.float32))
    t0 = time.perf_counter()
    loss = []
    for s0 in range(0, steps, 500):
        loss.append(net.train(min(500, steps - s0)))
        log(f"train: step {s0 + len(loss[-1])}/{steps}, loss {loss[-1][-1]:.4g}")
    loss = np.concatenate(loss)
    out["train_seconds"] = time.perf_counter() - t0
    out["train_ms_per_step"] = net.last_elapsed_ms / min(500, steps)
    out["loss_first_last"] = [float(loss[:10].mean()), float(loss[-50:].mean())]
    del raw, du
    # held-out conditions: cold, then warm on the same batch solver
    b, rc, cold_s, setup = cold_solve(test_c, diverge=diverge, **kw_test)
    cv_cold = b.solver.cvar
    f_cold = b.solver.wall_forces_batch()
    log(f"test cold: {int((rc['iterations'] > 0).sum())}/{n_test} converged in {cold_s:.2f} s")
    rw, warm_s, fell = warm_solve(b, net, scale, diverge=diverge, fallback=fallback, **kw_test)
    cv_warm = b.solver.cvar
    f_warm = b.solver.wall_forces_batch()
    log(f"test warm: {int((rw['iterations'] > 0).sum())}/{n_test} converged in {warm_s:.2f} s, "
        f"{int(fell.sum())} fell back")
    ic, iw = rc["iterations"], rw["iterations"]
    both = (ic > 0) & (iw > 0)
    dstate = []
    for k in range(n_test):
        sl = slice(b.off[k], b.off[k + 1])
        dstate.append(float(np.max(np.abs(cv_warm[sl] - cv_cold[sl])) / np.max(np.abs(cv_cold[sl]))))
    dstate = np.array(dstate)
    out.update(
        cold=dict(seconds=cold_s, converged=int((ic > 0).sum()), diverged=int(rc["diverged"].sum()),
                  gave_up=int(rc["gave_up"].sum()),
                  solves_per_hour=3600.0 * (ic > 0).sum() / cold_s,
                  iterations_p50=float(np.percentile(ic[ic > 0], 50)) if np.any(ic > 0) else None,
                  iterations_p99=float(np.percentile(ic[ic > 0], 99)) if np.any(ic > 0) else None),
        warm=dict(seconds=warm_s, converged=int((iw > 0).sum()), diverged=int(rw["diverged"].sum()),
                  gave_up=int(rw["gave_up"].sum()),
                  fell_back=int(fell.sum()), restarted=int(rw["restarted"].sum()),
                  solves_per_hour=3600.0 * (iw > 0).sum() / warm_s,
                  iterations_p50=float(np.percentile(iw[iw > 0], 50)) if np.any(iw > 0) else None,
                  iterations_p99=float(np.percentile(iw[iw > 0], 99)) if np.any(iw > 0) else None),
        iteration_ratio_median=float(np.median(iw[both] / ic[both])) if np.any(both) else None,
        cl_abs_diff_max=float(np.max(np.abs(f_warm[both, 0] - f_cold[both, 0]))) if np.any(both) else None,
        cd_abs_diff_max=float(np.max(np.abs(f_warm[both, 1] - f_cold[both, 1]))) if np.any(both) else None,
        cl_abs_diff_p50=float(np.median(np.abs(f_warm[both, 0] - f_cold[both, 0]))) if np.any(both) else None,
        cd_abs_diff_p50=float(np.median(np.abs(f_warm[both, 1] - f_cold[both, 1]))) if np.any(both) else None,
        state_rel_diff_max=float(np.max(dstate[both])) if np.any(both) else None,
        # instances converged both ways whose warm C_L and C_D are within 1e-5 of the cold ones
        forces_within_1e5=(float(np.mean((np.abs(f_warm[both, 0] - f_cold[both, 0]) < 1e-5) &
                                         (np.abs(f_warm[both, 1] - f_cold[both, 1]) < 1e-5)))
                           if np.any(both) else None),
        setup_seconds=setup,
    )
    out["per_instance"] = [dict(mach=c["mach"], aoa=c["aoa"], thickness=c["thickness"], cold=int(ic[k]),
                                warm=int(iw[k]), fell_back=bool(fell[k]), warm_first_ratio=float(rw["first_ratio"][k]),
                                cold_frozen_at=int(rc["frozen_at"][k]), warm_frozen_at=int(rw["frozen_at"][k]),
                                cold_gave_up=bool(rc["gave_up"][k]), warm_gave_up=bool(rw["gave_up"][k]),
                                cl_cold=float(f_cold[k, 0]),
                                cl_warm=float(f_warm[k, 0]), cd_cold=float(f_cold[k, 1]),
                                cd_warm=float(f_warm[k, 1]), state_rel_diff=float(dstate[k]))
                           for k, c in enumerate(test_c)]
    b.solver.close()
    if n_seq:
        sq, ok, its_s = sequential_cold(test_c[:n_seq], **kw_test)
        out["plain_cold"] = dict(instances=n_seq, seconds=sq, converged=ok, solves_per_hour=3600.0 * ok / sq)
        out["warm_batched_vs_plain_cold"] = out["warm"]["solves_per_hour"] / out["plain_cold"]["solves_per_hour"]
        log(f"plain cold: {ok}/{n_seq} in {sq:.2f} s")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m fvs2d_amd.warmstart")
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=256)
    ap.add_argument("--res", type=str, default="64x32")
    ap.add_argument("--test-res", type=str, default=None, help="resolution of the held-out solves (default --res)")
    ap.add_argument("--stall-scale", type=int, default=0, help="scale the stall window with n_around / 64")
    ap.add_argument("--max-active", type=int, default=0,
                    help="held-out instances stepping at once, the next starting as one finishes (0 = all)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--depth", type=int, default=5)
    ap.add_argument("--cfl", type=float, default=1.25)
    ap.add_argument("--max-steps", type=int, default=60000)
    ap.add_argument("--drop", type=float, default=1e-8)
    ap.add_argument("--chunk", type=int, default=250)
    ap.add_argument("--freeze-after", type=int, default=20000)
    ap.add_argument("--freeze-drop", type=float, default=None)
    ap.add_argument("--freeze-stall", type=str, default="2000,0.1",
                    help="window,factor: freeze an instance's limiter when its residual fell by less than "
                         "1/factor over the last window steps ('none' to disable)")
    ap.add_argument("--give-up", type=str, default="none",
                    help="window,factor: give an instance up once, after its limiter was frozen, its residual "
                         "fell by less than 1/factor over the last window steps ('none' to disable)")
    ap.add_argument("--diverge", type=float, default=1e3)
    ap.add_argument("--fallback", type=float, default=1.0)
    ap.add_argument("--group", type=int, default=256)
    ap.add_argument("--seq", type=int, default=16)
    ap.add_argument("--limiter", type=int, default=1, help="1 Venkatakrishnan (K = 5), 0 none")
    a = ap.parse_args(argv)
    res = tuple(int(v) for v in a.res.split("x"))
    stall = None if a.freeze_stall == "none" else (int(a.freeze_stall.split(",")[0]),
                                                    float(a.freeze_stall.split(",")[1]))
    out = run(a.train, a.test, res, a.seed, a.steps, a.batch, a.lr, a.depth, a.cfl, a.max_steps, a.drop, a.chunk,
              a.freeze_after, a.freeze_drop, stall, a.diverge, a.fallback, a.group, a.seq, a.limiter,
              log=lambda s: print(s, flush=True),
              test_res=None if a.test_res is None else tuple(int(v) for v in a.test_res.split("x")),
              stall_scale=bool(a.stall_scale), max_active=a.max_active or None,
              give_up=None if a.give_up == "none" else (int(a.give_up.split(",")[0]), float(a.give_up.split(",")[1])))
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
