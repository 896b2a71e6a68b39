"""Throughput of the fvs2d residual / time-step path on one MI355X.

Workload (default): the isentropic vortex on the periodic jittered square of
1024 x 1024 quads split into 2,097,152 triangles, FP64, Roe flux, unlimited
node-neighbour LSQ gradients, upwind-2nd reconstruction, SSP-RK3 at CFL 0.5
(dt = 0.5 * min_i V_i / (0.5 ws_i) on the initial state).  A "step" is one
SSP-RK3 time step (three residual evaluations and stage updates) over the whole
mesh; value = cells * stages * steps / seconds (cell-stage updates/s).

Prints ONE JSON line.  Extra fields: the per-kernel breakdown, the roofline
object of the dominant kernel (algorithmic bytes per launch / mean HIP-event
duration of that kernel on its stream; HBM traffic per launch from the
committed rocprofv3 counter summary of the same workload, when there is one),
the other gradient / limiter variants of the same mesh, the CPU baseline (the
C oracle on all host cores, and on one, on a few steps of the same mesh), and
two objects for the steady part of the metric:

* "steady": steady solves/hr to a 1e-8 density-residual drop on NACA O-grids
  (128x64 quads -> 16,384 triangles, the smallest resolution of the batch
  workload; Venkatakrishnan, SSP-RK(4,2), CFL 1.25): a warm-start network
  trained on batched cold solves of seeded conditions, then held-out
  conditions solved batched cold, batched warm (32 stepping at once, the next
  starting as one finishes) and one at a time cold (plain);
* "mlp": training-step rate of the warm-start network (68 -> 5 x 256 -> 4,
  65,536-row batches) with its per-kernel matrix-core rate against the bf16
  dense peak.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 matrix cores
TRAFFIC = os.path.join(ROOT, "profiles", "traffic_vortex1024_fused.json")


def model_bytes_per_cell_stage(grad_method: int, limiter: int) -> float:
    """Compulsory-traffic model per cell-stage for SSP-RK3 on triangle meshes
    (S 96 + G 160 + M 108 + stencil + L): LSQ-nn 604 B, GGCB 440 B, GGNB 508 B,
    +40 B with a limiter on LSQ, +88 B on GG (node-neighbour ids)."""
    stencil = {3: 240.0, 1: 76.0, 2: 144.0}[grad_method]
    lim = 0.0 if limiter == 0 else (40.0 if grad_method == 3 else 88.0)
    return 96.0 + 160.0 + 108.0 + stencil + lim


def build_case(n: int, periodic: bool = True):
    from fvs2d_amd import meshgen
    return meshgen.jittered_square(n, periodic=periodic, seed=0)


def scheme_for(grad_method=3, limiter=0, dt=0.01, reorder=1, path=0):
    from fvs2d_amd.solver import Scheme
    return Scheme(grad_method=grad_method, lsq_nn=1, lsq_p=0.0, limiter=limiter, recon=2, gamma=1.4, mach=0.5,
                  lvortex=1, vortex_x0=0.0, vortex_y0=0.0, vortex_K=5.0, vortex_inf=(1.0, 1.0, 1.0, 1.0),
                  rk_scheme=2, dt=dt, cfl=0.5, reorder=reorder, kernel_path=path)


def timed_run(solver, steps: int, warmup: int, init: bool = True):
    """Warmup steps, then the timed steps.  With init=False the run starts
    from the state on the device (the autotuner leaves the initial state
    restored there), so the device is not idle while the host builds and
    uploads an initial state right before the measurement."""
    if init:
        solver.initialize_solution(1, 0.0)
    if warmup:
        solver.time_integration(0.0, warmup)
    t0 = time.perf_counter()
    solver.time_integration(0.0, steps)  # synchronises the stream before returning
    wall = time.perf_counter() - t0
    return wall, solver.last_elapsed_ms / 1e3


def cpu_baseline(mesh, dt: float, steps: int, threads: int = 1):
    from oracle.oracle import Oracle
    sch = scheme_for(dt=dt)
    orc = Oracle(mesh, **sch.oracle_kwargs())
    orc.set_threads(threads)
    orc.initialize(1, 0.0)
    t0 = time.perf_counter()
    orc.time_integration(0.0, steps)
    sec = time.perf_counter() - t0
    return mesh.ncells * 3 * steps / sec, sec


def measured_traffic(n: int, reorder: int, path: int, kernel: str, fused_choice: int):
    """HBM bytes per launch of the dominant kernel from the committed counter
    summary, if it was taken on this workload with the same fused-kernel
    launch choice (block size and workgroup numbering)."""
    if not os.path.exists(TRAFFIC):
        return None
    with open(TRAFFIC) as f:
        entries = json.load(f)
    for t in entries if isinstance(entries, list) else [entries]:
        if (t.get("n"), t.get("reorder"), t.get("path"), t.get("kernel_class"), t.get("fused_choice")) == \
                (n, reorder, path, kernel, fused_choice):
            return t.get("bytes_per_launch")
    return None


def mlp_bench(steps: int):
    """Training-step rate of the warm-start network on synthetic rows."""
    from fvs2d_amd.mlp import WarmStartNet
    rng = np.random.default_rng(0)
    n = 1 << 20
    r = np.exp(rng.uniform(np.log(0.01), np.log(50), n))
    th = rng.uniform(-np.pi, np.pi, n)
    raw = np.stack([0.5 + r * np.cos(th), r * np.sin(th), rng.uniform(0, 50, n), rng.uniform(0.5, 0.85, n),
                    rng.uniform(-2, 4, n), rng.uniform(0.1, 0.15, n)], 1).astype(np.float32)
    tgt = np.stack([np.sin(2 * th) * np.exp(-r), np.cos(th) / (1 + r), raw[:, 3] - 0.675, raw[:, 4] / 4],
                   1).astype(np.float32)
    M, depth, width = 65536, 5, 256
    net = WarmStartNet(depth=depth, batch=M, seed=0)
    net.set_data(raw, tgt)
    net.train(5)
    loss = net.train(steps)
    ms_step = net.last_elapsed_ms / steps
    kt = net.kernel_times(10)
    flops_step = sum(v[1] * v[2] for v in kt.values())
    # algorithmic HBM bytes per launch of the matrix-core kernels (bfloat16
    # activations, each streamed once; weights are KBs): the forward layer reads
    # its input (68 features padded to 128 columns for the first layer) and
    # writes Z and GELU(Z); backward-data reads dZ and Z_prev and writes
    # dZ_prev; the weight gradient reads its input and dZ and writes float32
    # partials of 64 row slices
    kin = [128] + [width] * (depth - 1)
    slices = 64
    hbm = {"gemm_forward": sum(M * k * 2 + 2 * M * width * 2 for k in kin) / depth,
           "gemm_backward_data": 3 * M * width * 2,
           "gemm_weight_grad": sum(M * k * 2 + M * width * 2 + slices * (k + 1) * width * 4 for k in kin) / depth}
    gemm = {k: v for k, v in kt.items() if k.startswith("gemm")}
    dom = max(gemm, key=lambda k: gemm[k][0] * gemm[k][2])
    ms, fl, _ = kt[dom]
    # the binding ceiling of that kernel: HBM time of its bytes or matrix-core
    # time of its flops, whichever is longer
    t_hbm = hbm[dom] / (HBM_PEAK_GBS * 1e9)
    t_mma = fl / (BF16_PEAK_TFLOPS * 1e12)
    if t_hbm >= t_mma:
        ach = hbm[dom] / (ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "kernel": dom,
                "mfma_tflops": round(fl / (ms * 1e-3) / 1e12, 1)}
    else:
        ach = fl / (ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(ach, 1), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / BF16_PEAK_TFLOPS, 4), "kernel": dom}
    return {"network": "68 -> 5 x 256 (GELU) -> 4, bf16 operands / fp32 accumulate, Adam",
            "batch_rows": M, "params": net.nparams, "steps": steps, "ms_per_step": round(ms_step, 4),
            "rows_per_s": round(M / (ms_step * 1e-3), 1),
            "mfma_tflops_per_step": round(flops_step / (ms_step * 1e-3) / 1e12, 1),
            "loss_first_last": [float(loss[0]), float(loss[-1])],
            "kernels": {k: {"ms_per_launch": round(v[0], 5), "launches_per_step": v[2],
                            "TFLOPs": round(v[1] / (v[0] * 1e-3) / 1e12, 1) if v[0] > 0 and v[1] > 0 else None,
                            "alg_bytes_per_launch": int(hbm[k]) if k in hbm else None,
                            "GBps": round(hbm[k] / (v[0] * 1e-3) / 1e9, 1) if k in hbm and v[0] > 0 else None}
                        for k, v in kt.items()},
            "roofline": roof}


def steady_bench(n_train: int, n_test: int, n_seq: int, train_res=(128, 64), test_res=(128, 64), max_active=None,
                 give_up=(8000, 0.5)):
    """Warm versus cold steady solves/hr (see the module docstring)."""
    from fvs2d_amd import warmstart
    out = warmstart.run(n_train, n_test, train_res, seed=0, steps=16000, batch=65536, lr=1e-3, depth=5, cfl=1.25,
                        max_steps=80000, drop=1e-8, chunk=250, freeze_after=30000, freeze_drop=None,
                        freeze_stall=(2000, 0.1), diverge=1e3, fallback=1.0, group=256, n_seq=n_seq,
                        log=lambda s: print(f"[bench] steady: {s}", file=sys.stderr, flush=True),
                        test_res=test_res, stall_scale=True, max_active=max_active, give_up=give_up)
    keep = ("n_train", "n_test", "train_converged", "train_solve_seconds", "cold", "warm", "plain_cold",
            "iteration_ratio_median",
            "cl_abs_diff_p50", "cd_abs_diff_p50", "forces_within_1e5", "warm_batched_vs_plain_cold")
    d = {k: out[k] for k in keep if k in out}
    tr, te = train_res, test_res
    d["workload"] = (f"batched steady solves of a seeded draw of {n_test} NACA 4-digit O-grids at {te[0]}x{te[1]} "
                     f"quads -> {2 * te[0] * te[1]} triangles each (the smallest resolution of the batch workload), "
                     "M U[0.5,0.85], AoA U[-2,4] deg, t U[0.10,0.15], Venkatakrishnan K=5 (frozen on stall), "
                     "SSP-RK(4,2) CFL 1.25 (reference convention), to 1e-8 x the cold first-step density residual; "
                     f"an instance is given up when, after its limiter froze, its residual fell by less than "
                     f"{1 / give_up[1]:g}x over {give_up[0]} x n_around/64 steps (cold and warm alike); "
                     f"warm start trained on {n_train} seeded solves of the same generator at {tr[0]}x{tr[1]} "
                     "(16,000 Adam steps of 65,536 rows); "
                     + (f"at most {max_active} instances stepping at once, the next one starting as one finishes"
                        if max_active else "all instances stepping from the start"))
    d["max_active"] = max_active
    d["warm_vs_cold_batched"] = round(out["warm"]["solves_per_hour"] / out["cold"]["solves_per_hour"], 3)
    return d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024, help="quads per side (1024 = 2,097,152 triangles)")
    ap.add_argument("--variants", type=int, default=1, help="also time GGCB/GGNB x limiter on/off")
    ap.add_argument("--cpu-steps", type=int, default=6, help="oracle steps for the CPU baseline (0 = skip)")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--reorder", type=int, default=1, help="device cell order: 1 Morton, 2 coordinate bisection, 0 reference")
    ap.add_argument("--path", type=int, default=0, help="0 fused LDS kernels, 1 face-flux buffer kernels")
    ap.add_argument("--mlp-steps", type=int, default=200, help="warm-start network training steps (0 = skip)")
    ap.add_argument("--steady", type=int, default=1, help="warm vs cold steady solves/hr (0 = skip)")
    ap.add_argument("--steady-train", type=int, default=256)
    ap.add_argument("--steady-test", type=int, default=64)
    ap.add_argument("--steady-seq", type=int, default=8)
    ap.add_argument("--steady-active", type=int, default=32,
                    help="instances stepping at once in the batched steady solves (0 = all)")
    ap.add_argument("--choices", type=str, default=None,
                    help="launch choices gradient,flux,update,fused (the autotuner's encoding) applied instead "
                         "of autotuning, so that a profiler sees only the timed configuration")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only stream this many MiB through a plain read-once/write-once kernel (x1.0 "
                         "multiply), 3 times, and print its rate: the calibration of the HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this benchmark runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "decomposition is a design, not an implementation", file=sys.stderr)
        sys.exit(2)
    if args.copy_mib:
        import torch
        n = args.copy_mib * (1 << 20) // 8
        src = torch.full((n,), 1.5, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / 3
        print(json.dumps({"copy_mib_each_way": args.copy_mib, "ms": round(sec * 1e3, 4),
                          "TBps_read_plus_write": round(2 * n * 8 / sec / 1e12, 3)}))
        return

    from fvs2d_amd.solver import Solver

    t_setup = time.perf_counter()
    mesh = build_case(args.n)
    N = mesh.ncells
    # dt from the reference's local-step formula on the initial state
    s0 = Solver(mesh, scheme_for(dt=1.0, reorder=args.reorder, path=args.path))
    s0.initialize_solution(1, 0.0)
    dt = float(np.min(s0.compute_local_time()))
    s0.set_timestep(dt)
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] mesh {N} cells, dt={dt:.6e}, setup {setup_s:.1f}s", file=sys.stderr, flush=True)

    s0.initialize_solution(1, 0.0)
    if args.choices:
        tuned = dict(zip(("gradient", "flux", "update", "fused"), (int(v) for v in args.choices.split(","))))
        s0.set_launch_choices(tuned)
        tuned_ms = None
        print(f"[bench] launch choices: {tuned}", file=sys.stderr, flush=True)
    else:
        # 30-step trials: the two flux block sizes differ by about 1.5% of a step
        tuned, tuned_ms = s0.autotune(30)
        print(f"[bench] autotune: {tuned} ({tuned_ms:.4f} ms/step)", file=sys.stderr, flush=True)
    wall, dev = timed_run(s0, args.steps, args.warmup, init=False)
    stages = 3
    value = N * stages * args.steps / wall
    print(f"[bench] LSQ-nn: {args.steps} steps in {wall:.3f}s wall ({dev:.3f}s events) -> {value / 1e9:.3f} G "
          "cell-stage/s", file=sys.stderr, flush=True)

    # per-kernel breakdown with HIP events on the solver's stream
    kt = s0.kernel_times(0.0, args.profile_steps)
    per_step = {k: v[0] * v[2] / args.profile_steps for k, v in kt.items()}
    dom = max(per_step, key=per_step.get)
    ms, by, _ = kt[dom]
    achieved = by / (ms * 1e-3) / 1e9
    kernels = {k: {"ms_per_launch": round(v[0], 5), "launches_per_step": v[2] / args.profile_steps,
                   "alg_bytes_per_launch": int(v[1]),
                   "GBps": round(v[1] / (v[0] * 1e-3) / 1e9, 1) if v[0] > 0 else None} for k, v in kt.items()}
    print(f"[bench] kernels: {json.dumps(kernels)}", file=sys.stderr, flush=True)
    model_b = model_bytes_per_cell_stage(3, 0)

    variants = {}
    if args.variants:
        for gm, name in ((1, "ggcb"), (2, "ggnb"), (3, "lsq_nn")):
            for lim in (0, 1):
                if gm == 3 and lim == 0:
                    variants["lsq_nn"] = round(value, 1)
                    continue
                sv = Solver(mesh, scheme_for(grad_method=gm, limiter=lim, dt=dt, reorder=args.reorder,
                                             path=args.path))
                sv.initialize_solution(1, 0.0)
                sv.autotune(5)
                w, _ = timed_run(sv, max(5, args.steps // 4), 1, init=False)
                key = name + ("_venk" if lim else "")
                variants[key] = round(N * stages * max(5, args.steps // 4) / w, 1)
                sv.close()
                print(f"[bench] {key}: {variants[key] / 1e9:.3f} G cell-stage/s", file=sys.stderr, flush=True)

    cpu = None
    if args.cpu_steps > 0:
        # the oracle on one core, then on every core this process may use
        # (OpenMP over cells, bit-identical to the serial run)
        rate1, sec1 = cpu_baseline(mesh, dt, args.cpu_steps)
        print(f"[bench] cpu oracle, 1 thread: {rate1 / 1e6:.3f} M cell-stage/s ({sec1:.1f}s)", file=sys.stderr,
              flush=True)
        threads = max(1, int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1))
        cpu = {"value": round(rate1, 1), "unit": "cell-stage updates/s", "cores": 1, "kind": "port",
               "sample": f"C oracle (gcc -O2, serial), {args.cpu_steps} SSP-RK3 steps of the same "
                         f"{N}-cell mesh and scheme, {sec1:.1f} s"}
        if threads > 1:
            steps_n = max(args.cpu_steps, int(0.8 * args.cpu_steps * threads))  # a sample of several seconds
            rate, sec = cpu_baseline(mesh, dt, steps_n, threads)
            print(f"[bench] cpu oracle, {threads} threads: {rate / 1e6:.3f} M cell-stage/s ({sec:.1f}s)",
                  file=sys.stderr, flush=True)
            cpu = {"value": round(rate, 1), "unit": "cell-stage updates/s", "cores": threads, "kind": "port",
                   "sample": f"C oracle (gcc -O2, OpenMP over cells, {threads} threads), {steps_n} SSP-RK3 steps "
                             f"of the same {N}-cell mesh and scheme, {sec:.1f} s",
                   "single_core": {"value": round(rate1, 1), "sample": cpu["sample"]}}

    mlp = None
    if args.mlp_steps > 0:
        mlp = mlp_bench(args.mlp_steps)
        print(f"[bench] mlp: {mlp['ms_per_step']} ms/step, {mlp['mfma_tflops_per_step']} TFLOP/s",
              file=sys.stderr, flush=True)
    steady = None
    if args.steady:
        steady = steady_bench(args.steady_train, args.steady_test, args.steady_seq,
                              max_active=args.steady_active or None)

    traffic = measured_traffic(args.n, args.reorder, args.path, dom, tuned.get("fused"))
    out = {
        "metric": "cell-stage updates/s (FP64)",
        "value": round(value, 1),
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "device_ms_per_step": round(dev / args.steps * 1e3, 4),
        "autotune_ms_per_step": None if tuned_ms is None else round(tuned_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"isentropic vortex, periodic jittered square {args.n}x{args.n} quads -> {N} "
                               "triangles, Roe + unlimited LSQ-nn gradients, upwind-2nd, SSP-RK3, CFL 0.5",
                   "cell_order": ("reference", "morton", "bisection")[args.reorder],
                   "kernels": "fused" if args.path == 0 else "face-buffer",
                   "cells": N, "stages_per_step": stages, "dt": dt, "xcd_numbering": tuned},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom},
        "model_bytes_per_cell_stage": model_b,
        "effective_GBps_vs_model": round(value * model_b / 1e9, 1),
        "kernels": kernels,
        "variants_cell_stage_per_s": variants,
        "cpu_baseline": cpu,
        "steady": steady,
        "mlp": mlp,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
