"""Throughput benchmark of the fused stream-collide step on one MI355X.

Workload: schwarzschild_1024sq_d64 (1024x1024 cells on [-8,8]^2, 64-direction
stencil, Schwarzschild M=1 with excision r<=2.1, geodesic tables by RK45,
eta=max(0,GRF), kappa0~U[0,2], kappa1=0.3 U[0,1], CFL 0.5), seeded synthetic
fields.  A step is one timestep (stream + moments + collide) over the grid.

Prints one JSON line: FP32-mixed cell*direction updates/s as the value, the
FP64 GPU path, their ratio and the FP32-vs-FP64 rel-L2 of I, E, Fx, Fy after
the timed steps, the HBM roofline of the fused kernel, the adjoint leg
(observer-ring gradients on Schwarzschild 256x256 x 32, adjoint/forward time
ratio, central-difference check), the batched leg (64 uneven problems in
graph-replayed grouped launches vs one after another) and the CPU baseline
(the oracle, OpenMP on the host cores, on a bounded number of steps).

Before each timed leg of the headline workload the GPU is brought to its
working clocks by CLOCK_SPINUP_S seconds of steps on a separate throwaway
context (256x256 x 32): after the host-side set-up the first milliseconds of
work otherwise run at lower clocks (steps 6..25 measured 0.306 ms/step
without and 0.271 with it, against 0.257 for steps 201..220).  The timed
region itself is unchanged: W warm-up and then exactly K steps of the
workload from I = 0.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = "schwarzschild_1024sq_d64"
HBM_PEAK = 8.0e12
CLOCK_SPINUP_S = 0.3
_spinner = None


def clock_spinup(rt, seconds=CLOCK_SPINUP_S):
    """Steps a throwaway 256^2 x 32 context for `seconds` (untimed; no
    state of the benchmarked contexts is touched), so the timed region that
    follows starts at the GPU's working clocks."""
    global _spinner
    if _spinner is None:
        _spinner, _ = rt.workloads.make_context("schwarzschild_256sq_d32", rt.FP32_MIXED)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        _spinner.step(100, want_moments=False)
        _spinner.synchronize()


def algorithmic_bytes_per_update(nDir, b_i, b_f, curved=True, nCoeff=5):
    """2 b_I (read + write I) + (6 fields * b_f + geometry) / nDir, geometry =
    alpha + 5 series * nCoeff coefficients in the kernel's precision for
    curved spacetimes: the FP64 step reads the FP64 tables, the FP32-mixed
    kernel an FP32 copy of them in relative form (S - 1, departure offsets in
    cells), so b_geo is 8 (1 + 25) = 208 B or 4 (1 + 25) = 104 B per cell.
    Derived per-(cell, direction) tables are not algorithmic: the FP32
    kernel's 4-byte angular record per update (which stands in for the two
    velocity series) counts against the achieved fraction, in `traffic`."""
    b_geo = b_f * (1 + 5 * nCoeff) if curved else 0
    return 2 * b_i + (6 * b_f + b_geo) / nDir


def profiled_traffic_bytes():
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    summary (FETCH_SIZE and WRITE_SIZE runs scaled by a plain-copy
    calibration), or None if the summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_fused_f32_d64_schwarzschild_1024sq.md")
    if not os.path.exists(path):
        return None
    import re
    m = re.search(r"= \*\*([0-9.]+) MB\*\*", open(path).read())
    return float(m.group(1)) * 1e6 if m else None


def rel_l2(a, b):
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def cpu_model():
    """Model name of the host CPU (lscpu's "Model name", read from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(rt, name, budget_s=12.0, max_steps=40):
    """The oracle's -O3 -ffast-math OpenMP build on the host cores, same
    workload, as many steps as fit in the time budget (at least 2)."""
    from oracle import oracle as O
    spec = rt.workloads.WORKLOADS[name]
    nx_, ny_, start, end = spec["grid"]
    og = O.make_grid(nx_, ny_, start, end, spec["cfl"])
    gs = rt.GridSpec(nx_, ny_, start, end, spec["cfl"])
    kind = rt.workloads.spacetime_kind(spec)
    r2 = rt.workloads.excision_r2(spec)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    L = O.lib(fast=True)
    L.or_set_threads(threads)
    M = O.metric(og, kind)
    tabs = O.geodesic_tables(og, M, r2, spec.get("tol", 1e-4)) if kind else None
    f = rt.workloads.fields(name, gs)
    P = O.Problem(og, spec["nDir"], 2 if kind else 0, M, bh_r2=r2, tables=tabs, fields=f, fast=True)
    I = np.zeros(og.N * spec["nDir"])
    t0 = time.perf_counter()
    I = P.run(I, 1)[0]
    t1 = time.perf_counter()
    per = t1 - t0
    n = int(max(1, min(max_steps - 1, budget_s / max(per, 1e-6))))
    t2 = time.perf_counter()
    P.run(I, n)
    t3 = time.perf_counter()
    steps = n
    elapsed = t3 - t2
    act = int(np.count_nonzero(_active(gs, r2)))
    return {"value": act * spec["nDir"] * steps / elapsed, "unit": "cell*dir updates/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{name}, {steps} timed steps after 1 warm-up step, oracle -O3 -ffast-math OpenMP "
                      f"({threads} threads), {elapsed:.1f} s"}


def _active(gs, r2):
    x, y = gs.coords()
    i = np.arange(gs.N) % gs.nx
    j = np.arange(gs.N) // gs.nx
    inter = (i > 0) & (i < gs.nx - 1) & (j > 0) & (j < gs.ny - 1)
    return inter & ~((r2 >= 0) & (x * x + y * y <= r2))


def adjoint_leg(rt, steps=500, every=50, n_fd=16, h=1e-3):
    """Gradients of an observer-ring loss: Schwarzschild 256x256 on [-4,4]^2,
    32 directions, 500 steps with checkpoints every 50, kappa0 ~ U[0,1],
    eta = max(0.05, 0.5 + 0.25 GRF); target = ring samples of a run with
    eta* = max(0, GRF) (second seed).
    FP32-mixed forward and adjoint timed; FP64 adjoint and central
    differences (h = 1e-3, FP64 forward, 16 seeded active cells) as checks."""
    name = "schwarzschild_256sq_d32"
    th = 2 * np.pi * (np.arange(128) + 0.5) / 128
    px, py = 3.8 * np.cos(th), 3.8 * np.sin(th)
    out = {"workload": name, "steps": steps, "checkpoint_every": every, "ring_points": 128, "ring_radius": 3.8}
    grads = {}
    for prec in (rt.FP32_MIXED, rt.FP64):
        ctx, meta = rt.workloads.make_context(name, prec)
        gs = meta["grid"]
        f = rt.workloads.fields(name, gs)
        ftarget = dict(f)
        ftarget["eta"] = rt.workloads.target_eta(name, gs)
        zero = np.zeros(gs.N * ctx.nDir)
        ctx.set_fields(**ftarget)
        ctx.upload_I(zero)
        ctx.step(steps, want_moments=False)
        target = ctx.sample_outgoing(px, py)
        ctx.set_fields(**f)
        ctx.upload_I(zero)
        ctx.step(2, want_moments=False)  # warm-up
        ctx.upload_I(zero)
        ctx.synchronize()
        ms_fwd = ctx.time_steps(steps)
        ctx.upload_I(zero)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.forward_checkpointed(steps, every)
        t1 = time.perf_counter()
        obs = ctx.sample_outgoing(px, py)
        loss = 0.5 * float(np.sum((obs - target) ** 2))
        t2 = time.perf_counter()
        ge, gk = ctx.adjoint(px, py, obs - target)
        t3 = time.perf_counter()
        key = "fp32" if prec == rt.FP32_MIXED else "fp64"
        grads[key] = (ge, gk)
        out[key] = {"loss": loss, "forward_ms": ms_fwd, "forward_checkpointed_ms": 1e3 * (t1 - t0),
                    "adjoint_ms": 1e3 * (t3 - t2), "adjoint_over_forward": (t3 - t2) * 1e3 / ms_fwd}
        if prec == rt.FP64:
            # central differences of the FP64 forward on seeded active cells
            x, y = gs.coords()
            i = np.arange(gs.N) % gs.nx
            j = np.arange(gs.N) // gs.nx
            ok = (i > 1) & (i < gs.nx - 2) & (j > 1) & (j < gs.ny - 2) & (x * x + y * y > 2.1 ** 2 + 0.5)
            rng = np.random.default_rng(1004)
            cells = rng.choice(np.nonzero(ok)[0], n_fd, replace=False)
            fd_err = {}
            for hh in (h, 1e-5):
                for fname, grad in (("eta", ge), ("kappa0", gk)):
                    fd = []
                    for c in cells:
                        vals = []
                        for sgn in (1, -1):
                            ff = dict(f)
                            arr = ff[fname].copy()
                            arr[c] += sgn * hh
                            ff[fname] = arr
                            ctx.set_fields(**ff)
                            ctx.upload_I(zero)
                            ctx.step(steps, want_moments=False)
                            o = ctx.sample_outgoing(px, py)
                            vals.append(0.5 * float(np.sum((o - target) ** 2)))
                        fd.append((vals[0] - vals[1]) / (2 * hh))
                    fd = np.array(fd)
                    fd_err[f"{fname}_h{hh:g}"] = float(np.linalg.norm(grad[cells] - fd) / np.linalg.norm(fd))
            ctx.set_fields(**f)
            out["fd_rel_err_fp64"] = fd_err
            out["fd_cells"] = int(n_fd)
        ctx.close()
    out["rel_l2_grad_fp32_vs_fp64"] = {"eta": rel_l2(grads["fp32"][0], grads["fp64"][0]),
                                       "kappa0": rel_l2(grads["fp32"][1], grads["fp64"][1])}
    out["adjoint_over_forward"] = out["fp32"]["adjoint_over_forward"]
    return out


def batch_leg(rt, steps=200, n=64):
    """64 independent FP32-mixed problems (nx = ny log-uniform in [64, 512],
    nDir from {16, 24, 32, 48, 64}, Minkowski or Schwarzschild by a seeded coin),
    200 steps each, graph-replayed: one persistent launch draining a (problem,
    tile) queue per step, or one grouped launch per class, as the autotuner
    finds faster (with its rows per tile) against the 64 problems stepped one
    after another with their own launches.  Both runs must agree bit for bit."""
    specs = rt.workloads.batch_specs(n)
    A, B, upd = [], [], 0
    for sp in specs:
        a, grid = rt.workloads.make_batch_member(sp, rt.FP32_MIXED)
        bb, _ = rt.workloads.make_batch_member(sp, rt.FP32_MIXED)
        A.append(a)
        B.append(bb)
        upd += int(np.count_nonzero(_active(grid, sp["excision"] ** 2 if sp["spacetime"] == "schwarzschild" else -1.0))) * sp["nDir"]
    batch = rt.Batch(A)
    t0 = time.perf_counter()
    chosen = batch.autotune(2)
    t_tune = time.perf_counter() - t0
    ms_batch = batch.time_steps(steps)
    # sequential: each problem's own launches, one problem after another
    for c in B:
        c.step(1, want_moments=False)  # state convention set-up outside the timing
    zero_I = None
    ms_seq = 0.0
    for c in B:
        ms_seq += c.time_steps(steps - 1)
    # compare: A took `steps` steps, B took 1 + (steps - 1)
    same = all(np.array_equal(a.download_I(), bb.download_I()) for a, bb in zip(A, B))
    out = {"problems": n, "steps": steps, "classes": int(batch.groups()),
           "updates_per_step": upd, "batched_ms": ms_batch, "sequential_ms_for_steps_minus_1": ms_seq,
           "batched_updates_per_s": upd * steps / (ms_batch * 1e-3),
           "sequential_updates_per_s": upd * (steps - 1) / (ms_seq * 1e-3),
           "speedup": (upd * steps / ms_batch) / (upd * (steps - 1) / ms_seq),
           "autotune_s": t_tune, "rows_per_workgroup_chosen": [int(h) for h in chosen],
           "fp32_launch": "persistent queue" if batch.uses_queue() else "one launch per class",
           "bitwise_equal_to_sequential": bool(same)}
    batch.close()
    for c in A + B:
        c.close()
    return out


def configs_leg(rt):
    """The other single-problem configs at their full step counts from I = 0:
    FP32-mixed and FP64 throughput, and FP32-mixed against FP64 (rel-L2 of I
    and of the moments E, Fx, Fy of the final state)."""
    out = {}
    for name in ("minkowski_100sq_d20", "kerrschild_200sq_d32", "schwarzschild_256sq_d32"):
        spec = rt.workloads.WORKLOADS[name]
        K, D = spec["steps"], spec["nDir"]
        w, phi, cx, cy = rt.build_stencil(D, tables=False)
        res, finals = {"steps": K, "nDir": D}, {}
        for prec, tag in ((rt.FP32_MIXED, "fp32_mixed"), (rt.FP64, "fp64")):
            ctx, meta = rt.workloads.make_context(name, prec)
            gs = meta["grid"]
            act = int(np.count_nonzero(_active(gs, meta["excision_r2"])))
            ms = ctx.time_steps(K)
            res[tag] = {"value": act * D * K / (ms * 1e-3), "ms_per_step": ms / K, "fused": ctx.uses_fused()}
            I = ctx.download_I().reshape(-1, D)
            finals[tag] = (I, I @ w, I @ (w * cx), I @ (w * cy))
            ctx.close()
        a, b = finals["fp32_mixed"], finals["fp64"]
        res["rel_l2_fp32_vs_fp64"] = {k: rel_l2(x, y) for k, x, y in zip(("I", "E", "Fx", "Fy"), a, b)}
        res["fp32_over_fp64"] = res["fp32_mixed"]["value"] / res["fp64"]["value"]
        out[name] = res
    return out


def adaptive_cpu_baseline(rt, name, budget_s=8.0, max_steps=10):
    """The oracle's -O3 -ffast-math OpenMP build running the adaptive workload
    from I = 0 on the host cores: one warm-up step, then as many steps as fit
    in the budget."""
    from oracle import oracle as O
    spec = rt.workloads.WORKLOADS[name]
    nx_, ny_, start, end = spec["grid"]
    og = O.make_grid(nx_, ny_, start, end, spec["cfl"])
    gs = rt.GridSpec(nx_, ny_, start, end, spec["cfl"])
    kind = rt.workloads.spacetime_kind(spec)
    r2 = rt.workloads.excision_r2(spec)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    L = O.lib(fast=True)
    L.or_set_threads(threads)
    M = O.metric(og, kind)
    tabs = O.geodesic_tables(og, M, r2, spec.get("tol", 1e-4)) if kind else None
    f = rt.workloads.fields(name, gs)
    P = O.Problem(og, spec["nDir"], 3 if kind else 1, M, bh_r2=r2, tables=tabs, fields=f, fast=True,
                  nGhost=spec["nGhost"])
    t0 = time.perf_counter()
    I, mom, rot, rn = P.run_adaptive(np.zeros(og.N * spec["nDir"]), 1)[:4]
    per = time.perf_counter() - t0
    n = int(max(1, min(max_steps, budget_s / max(per, 1e-6))))
    t1 = time.perf_counter()
    P.run_adaptive(I, n, mom=mom, rot=rot, rotNew=rn)
    elapsed = time.perf_counter() - t1
    act = int(np.count_nonzero(_active(gs, r2)))
    return {"value": act * spec["nDir"] * n / elapsed, "unit": "cell*dir updates/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{name}, {n} timed steps after 1 warm-up step, oracle -O3 -ffast-math OpenMP "
                      f"({threads} threads), {elapsed:.1f} s"}


def adaptive_leg(rt, steps=40, warmup=3, cpu=True):
    """CurvedAdaptive streaming of the 1024^2 Schwarzschild problem with
    Stencil(48,16) (64 directions): rotation update, rotated stream, moments
    and collide per step, FP32-mixed and FP64, both from I = 0 with all
    angles 0; FP32-mixed is compared with FP64 after the timed steps."""
    name = "schwarzschild_1024sq_adaptive_48_16"
    spec = rt.workloads.WORKLOADS[name]
    out = {"workload": name, "stencil": f"Stencil({spec['nOrder']},{spec['nGhost']})", "steps": steps}
    states = {}
    for prec, tag in ((rt.FP32_MIXED, "fp32_mixed"), (rt.FP64, "fp64")):
        ctx, meta = rt.workloads.make_context(name, prec)
        gs = meta["grid"]
        act = int(np.count_nonzero(_active(gs, meta["excision_r2"])))
        zero_i, zero_n = np.zeros(gs.N * spec["nDir"]), np.zeros(gs.N)
        ctx.step(warmup)
        ctx.upload_I(zero_i)
        ctx.set_rotation(zero_n, zero_n)
        ctx.set_moments([zero_n] * 6)
        ctx.synchronize()
        ms = ctx.time_steps(steps)
        upd = act * spec["nDir"] * steps
        b_i = 4 if prec == rt.FP32_MIXED else 8
        # algorithmic bytes: I in and out, the fields, the tables and both angle buffers
        B = algorithmic_bytes_per_update(spec["nDir"], b_i, b_i) + 2 * 8 / spec["nDir"]
        rate = upd / (ms * 1e-3)
        out[tag] = {"value": rate, "unit": "cell*dir updates/s", "ms_per_step": ms / steps,
                    "hbm_frac": B * rate / HBM_PEAK, "bytes_per_update": B}
        states[tag] = ctx.download_I()
        rot = ctx.rotation()[0]
        out[tag]["turned_cells"] = int(np.count_nonzero(rot))
        ctx.close()
    out["fp32_over_fp64"] = out["fp32_mixed"]["value"] / out["fp64"]["value"]
    out["rel_l2_fp32_vs_fp64"] = rel_l2(states["fp32_mixed"], states["fp64"])
    if cpu:
        out["cpu_baseline"] = adaptive_cpu_baseline(rt, name)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (HBM counter calibration)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-problems leg")
    ap.add_argument("--no-adaptive", action="store_true", help="skip the adaptive-streaming leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the other single-problem configs")
    ap.add_argument("--only-adaptive", action="store_true", help="run only the adaptive-streaming leg")
    ap.add_argument("--only-adjoint", action="store_true", help="run only the adjoint leg")
    ap.add_argument("--only-batch", action="store_true", help="run only the batched-problems leg")
    ap.add_argument("--ref-steps", type=int, default=2000,
                    help="steps from I = 0 of the comparison with the FP64 reference-order kernels "
                         "(default: the workload's full 2000; 0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(the domain-decomposition design is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.rand(n, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 4 * n}))
        return
    rt = importlib.import_module("2dradiation_amd")
    rt.lib()
    if args.only_adaptive:
        print(json.dumps({"adaptive": adaptive_leg(rt, cpu=not args.no_cpu)}), flush=True)
        return
    if args.only_adjoint:
        print(json.dumps({"adjoint": adjoint_leg(rt)}), flush=True)
        return
    if args.only_batch:
        print(json.dumps({"batch": batch_leg(rt)}), flush=True)
        return
    spec = rt.workloads.WORKLOADS[WORKLOAD]
    K, W = args.steps, args.warmup

    print(f"[bench] setting up {WORKLOAD} FP32-mixed", file=sys.stderr, flush=True)
    c32, meta = rt.workloads.make_context(WORKLOAD, rt.FP32_MIXED)
    gs = meta["grid"]
    act = int(np.count_nonzero(_active(gs, meta["excision_r2"])))
    upd_per_step = act * spec["nDir"]
    assert c32.uses_fused(), "fused path not selected"
    zero = np.zeros(gs.N * spec["nDir"])

    # W untimed steps from I = 0, then K timed steps continuing from there
    # (no upload or state conversion between them)
    clock_spinup(rt)
    c32.step(W, want_moments=False)
    c32.synchronize()
    c32.reset_iteration_stats()
    c32.synchronize()
    print(f"[bench] timing {K} FP32-mixed steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ms32 = c32.time_steps(K)
    wall32 = time.perf_counter() - t0
    rate32 = upd_per_step * K / (ms32 * 1e-3)
    it_max, it_sum = c32.iteration_stats()

    result = {
        "metric": "cell*dir updates/s FP32-mixed vs FP64 at rel-L2<=1e-5; adjoint/forward ratio",
        "value": rate32, "unit": "cell*dir updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": ms32 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "dtype_detail": "f32 intensities, gather arithmetic and Fourier series (from an f32 copy "
                                        "of the f64 geodesic tables in relative form); f64 moment sums "
                                        "(uncompensated) and lambda iteration",
        "data": "synthetic (seeded fields)",
        "config": {"workload": WORKLOAD, "grid": "1026x1026 stored (1024x1024 interior) on [-8,8]^2",
                   "nDir": spec["nDir"], "spacetime": "Schwarzschild M=1, excision r<=2.1",
                   "cfl": spec["cfl"], "steps_per_run": K, "active_cells": act, "parallelism": "single GPU",
                   "clock_spinup_s": CLOCK_SPINUP_S},
        "lambda_iterations": {"max": it_max, "mean": it_sum / max(1, act * K), "steps": f"{W + 1}..{W + K}"},
        "wall_s_fp32": wall32,
    }

    B32 = algorithmic_bytes_per_update(spec["nDir"], 4, 4)
    per_launch_ms = ms32 / K
    achieved = B32 * upd_per_step / (per_launch_ms * 1e-3)
    traffic = profiled_traffic_bytes()
    result["roofline"] = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                          "frac": achieved / HBM_PEAK, "traffic": traffic,
                          "traffic_over_algorithmic": traffic / (B32 * upd_per_step) if traffic else None,
                          "bytes_per_update": B32, "kernel": "k_fused<float,64,curved>"}
    # the reference's own figure: stored cells x steps / s (src/Radiation.cpp:880)
    result["mlups_reference_convention"] = gs.N * K / (ms32 * 1e-3) / 1e6

    if not args.no_fp64:
        print(f"[bench] FP64 leg: {K} steps", file=sys.stderr, flush=True)
        c64, _ = rt.workloads.make_context(WORKLOAD, rt.FP64)
        clock_spinup(rt)
        c64.step(W, want_moments=False)
        c64.synchronize()
        ms64 = c64.time_steps(K)
        rate64 = upd_per_step * K / (ms64 * 1e-3)
        # quality: both runs started from I = 0 and took W + K steps
        I32, I64 = c32.download_I(), c64.download_I()
        # moments of the final state (one more step with moments requested on copies is not
        # needed: compare the state's moments computed on the host from I)
        w, phi, cx, cy = rt.build_stencil(spec["nDir"], tables=False)
        D = spec["nDir"]

        def mom(I):
            I2 = I.reshape(-1, D)
            return I2 @ w, I2 @ (w * cx), I2 @ (w * cy)

        E32, Fx32, Fy32 = mom(I32)
        E64, Fx64, Fy64 = mom(I64)
        B64 = algorithmic_bytes_per_update(spec["nDir"], 8, 8)
        result["fp64"] = {"value": rate64, "ms_per_step": ms64 / K,
                          "hbm_frac": B64 * upd_per_step / (ms64 / K * 1e-3) / HBM_PEAK}
        result["fp32_over_fp64"] = rate32 / rate64
        result["rel_l2_fp32_vs_fp64"] = {"steps": W + K, "I": rel_l2(I32, I64), "E": rel_l2(E32, E64),
                                         "Fx": rel_l2(Fx32, Fx64), "Fy": rel_l2(Fy32, Fy64)}
        c64.close()
        if args.ref_steps > 0:
            # FP64 reference-order kernels (the oracle's operation order, one
            # method at a time) as the FP64 truth; the product paths compared
            # with it after ref_steps steps from I = 0
            R = args.ref_steps
            print(f"[bench] FP64 reference-order leg: {R} steps", file=sys.stderr, flush=True)
            cr, _ = rt.workloads.make_context(WORKLOAD, rt.FP64)
            t0 = time.perf_counter()
            for k in range(R):
                cr.stream()
                cr.compute_moments()
                cr.collide()
                if (k + 1) % 500 == 0:
                    cr.synchronize()
                    print(f"[bench]   reference order: {k + 1} steps", file=sys.stderr, flush=True)
            cr.synchronize()
            tref = time.perf_counter() - t0
            Iref = cr.download_I()
            cr.close()
            a32, _ = rt.workloads.make_context(WORKLOAD, rt.FP32_MIXED)
            a32.step(R, want_moments=False)
            a64, _ = rt.workloads.make_context(WORKLOAD, rt.FP64)
            a64.step(R, want_moments=False)
            J32, J64 = a32.download_I(), a64.download_I()
            a32.close()
            a64.close()
            Er, Fxr, Fyr = mom(Iref)
            Ea, Fxa, Fya = mom(J32)
            Eb, Fxb, Fyb = mom(J64)
            result["rel_l2_vs_reference_order_fp64"] = {
                "steps": R, "ms_per_step_reference_order": 1e3 * tref / R,
                "fp32_mixed": {"I": rel_l2(J32, Iref), "E": rel_l2(Ea, Er), "Fx": rel_l2(Fxa, Fxr), "Fy": rel_l2(Fya, Fyr)},
                "fp64_step": {"I": rel_l2(J64, Iref), "E": rel_l2(Eb, Er), "Fx": rel_l2(Fxb, Fxr), "Fy": rel_l2(Fyb, Fyr)}}
    c32.close()

    if not args.no_adjoint:
        print("[bench] adjoint leg", file=sys.stderr, flush=True)
        result["adjoint"] = adjoint_leg(rt)
    if not args.no_batch:
        print("[bench] batch leg", file=sys.stderr, flush=True)
        result["batch"] = batch_leg(rt)
    if not args.no_adaptive:
        print("[bench] adaptive-streaming leg", file=sys.stderr, flush=True)
        result["adaptive"] = adaptive_leg(rt, cpu=not args.no_cpu)
    if not args.no_configs:
        print("[bench] other configs leg", file=sys.stderr, flush=True)
        result["configs"] = configs_leg(rt)
    if not args.no_cpu:
        print("[bench] CPU baseline leg", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(rt, WORKLOAD)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
