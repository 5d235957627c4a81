#!/usr/bin/env python3
"""Benchmark: cell-updates/s per RK stage of the fused FP32 MUSCL-HLLC-RK3 stage.

Workload (one step = one SSP-RK3 step = 3 fused stage launches): 2-D Euler,
MUSCL with the monotonized-central limiter + HLLC (Thornber) + SSP-RK3 on a
periodic 4096^2 grid, gamma = 1.4, CFL 0.4 with the dt reduction fused into
stage 3; initial state = isentropic vortex (beta = 5) plus 8 seeded Fourier
modes (amplitude 0.02).  Synthetic data, generated on the host in FP64.

Prints ONE JSON line (rank 0):
  value     = N_cells * 3 * steps / wall time of the timed steps (state resident in HBM)
  roofline  = algorithmic HBM bytes per stage launch / mean stage-kernel time (HIP events
              on the solver's stream) vs the MI355X HBM peak (8 TB/s)
  cpu_baseline = the CPU FP64 oracle (the reference path restated in C) on a bounded sample
              of the same workload on this host's cores (its threaded step, bitwise the serial
              one), with the single-thread figure beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "cell-updates/s per RK stage, % HBM peak; time-to-L1-error learned vs MUSCL"
# algorithmic bytes per cell per RK3 step (4 stored components):
#   stage 1 reads U, writes U1; stages 2 and 3 read U^k and U^0, write U^{k+1}
BYTES_PER_CELL_STEP = {"fp32": 4 * 4 * (2 + 3 + 3), "fp64": 8 * 4 * (2 + 3 + 3)}


def cpu_baseline(n_cpu: int = 1024, steps: int = 6, steps_threaded: int = 30):
    """The oracle (-O3 FP64, the reference path restated) on an n_cpu^2 sample of the same
    workload, on this host: on all the cores it may use (its threaded step, bitwise the serial
    one; OMP_NUM_THREADS or the visible count) and on one thread."""
    from oracle import orc
    from dev_field_efvm_amd import cases
    st = cases.vortex_modes_state(n_cpu, n_cpu, seed=1)
    threads = int(os.environ.get("OMP_NUM_THREADS", 0) or os.cpu_count() or 1)
    out = {}
    for label, th, k in (("all", threads, steps_threaded), ("one", 1, steps)):
        sim = orc.Sim(n_cpu, n_cpu, 1, hi=(10.0, 10.0, 10.0 / n_cpu), flux="hllc", muscl=True, limiter="mc",
                      cfl=0.4, bcs=["periodic"] * 4 + ["slip", "slip"])
        sim.set_interior(st.reshape(1, n_cpu, n_cpu, 5))
        used = sim.set_threads(th)
        sim.step(1)                      # warm-up (page-in)
        t0 = time.perf_counter()
        sim.step(k)
        dt = time.perf_counter() - t0
        out[label] = (n_cpu * n_cpu * 3 * k / dt, used, k, dt)
        del sim
    va, ca, ka, da = out["all"]
    v1, _, k1, d1 = out["one"]
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), model)
    except OSError:
        pass
    return {"value": va, "unit": "cell-updates/s", "cores": ca, "kind": "port",
            "single_thread_value": v1,
            "sample": f"oracle/efvm_oracle.c FP64 MUSCL-MC+HLLC+RK3, periodic {n_cpu}^2 vortex+modes: {ka} steps on "
                      f"{ca} threads ({da:.1f} s) and {k1} steps on 1 thread ({d1:.1f} s); {os.cpu_count()} host "
                      f"cores visible, {model}"}


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 10) -> float:
    """Achievable HBM bandwidth on this device: a 1 GiB device-to-device copy (read + write
    bytes / time, HIP events on torch's stream), best of `reps`."""
    import torch
    a = torch.ones(nbytes // 4, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    b.copy_(a)
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    del a, b
    torch.cuda.empty_cache()
    return round(2 * nbytes / (best * 1e-3) / 1e9, 1)


def load_traffic(n: int, prec: str, name: str = None):
    """HBM bytes per stage launch from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", name or f"stage_{prec}_{n}_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        return d.get("hbm_bytes_per_step", d.get("hbm_bytes_per_launch"))
    except Exception:
        return None


def bench_batch(args):
    """Batched ensemble: 512 independent 2-D Euler cases (64^2..1024^2 log-uniform, gamma in
    [1.3, 1.67], periodic vortices and four-quadrant Riemann problems), MUSCL-HLLC-RK3 FP32,
    per-case dt, one launch per stage with a wave per (case, strip), 100 steps replayed from a
    hipGraph, launch shapes from the autotuner."""
    import torch
    import dev_field_efvm_amd as efvm
    from dev_field_efvm_amd import cases
    from dev_field_efvm_amd.autotune import tune_batch
    specs, builders = cases.ensemble(n_cases=args.cases, seed=2024)
    b = efvm.Batch(specs, muscl=True, limiter="mc", cfl=0.4)
    for i, make in enumerate(builders):
        b.upload(i, make())
    sides = [s["nx"] for s in specs]
    total = sum(s["nx"] * s["ny"] for s in specs)
    t0 = time.perf_counter()
    rows, blocks, tuned_ms = tune_batch(b, sides, steps=4, use_cache=not args.retune)
    t_tune = time.perf_counter() - t0
    b.step(args.warmup)
    b.synchronize()
    b.graph_capture(2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps // 2):
        b.graph_launch()
    b.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    b.step(2)
    ev_ms = b.last_elapsed_ms()
    times = [b.clock(i)["time"] for i in (0, 1)]
    value = total * 3 * args.steps / wall
    # per RK stage: a third of a step (the step's launches: stage 1 of every case, stages 2 and
    # 3 of the cases that are not fused, the fused launch of those that are)
    mean_launch_s = ev_ms * 1e-3 / 6
    bytes_per_launch = BYTES_PER_CELL_STEP["fp32"] / 3.0 * total
    achieved = bytes_per_launch / mean_launch_s / 1e9
    fz = b.fused()
    return {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded vortices and four-quadrant Riemann problems, host FP64)",
        "config": {"workload": f"batched ensemble of {args.cases} 2-D Euler cases, MUSCL(MC)+HLLC+SSP-RK3, per-case dt, "
                               "one launch per stage (a wave per case strip), hipGraph", "total_cells": total,
                   "sides": [min(sides), max(sides)], "tuned_rows": {str(k): v for k, v in rows.items()},
                   "tuned_blocks": blocks, "tune_s": round(t_tune, 2), "sim_time_case0_1": times,
                   "fused_cases": fz[2], "fused_rows": fz[1] if fz[0] else 0},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": load_traffic(len(sides), "fp32", f"batch_fp32_{len(sides)}_pmc.json"),
                     "kernel": "stage_f32_batch_pair_kernel + stage_f32_batch_kernel + stage_f32_batch_pair23_kernel",
                     "mean_launch_us": round(mean_launch_s * 1e6, 2), "algorithmic_bytes_per_launch": bytes_per_launch,
                     "per": "RK stage (a third of a step)"},
    }


def bench_cylinder(args):
    """Compressible Navier-Stokes past a cylinder with the ghost-cell IBM: channel 22D x 4.1D,
    80 cells per D (1760 x 328), Re = 20 and Mach 0.1 on the mean inflow, MUSCL(MC)+HLLC +
    viscous + SSP-RK3, run until the marker drag plateaus; Cd/Cl in FP64."""
    import torch
    from dev_field_efvm_amd.cylinder import CylinderCase, run_to_plateau
    case = CylinderCase(nx=args.cyl_nx, ny=args.cyl_ny, D=0.1, mach_mean=0.1, muscl=True, limiter="mc",
                        precision=args.precision, cfl=0.4)
    torch.cuda.synchronize()
    r = run_to_plateau(case, max_steps=args.cyl_max_steps)
    cells = case.nx * case.ny
    value = cells * 3 * r["steps"] / (r["gpu_ms"] * 1e-3)
    return {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": r["steps"], "warmup": 0,
        "ms_per_step": r["gpu_ms"] / r["steps"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic (parabolic channel inflow, cylinder DFG 2D-1 geometry)",
        "config": {"workload": f"cylinder IBM, {case.nx}x{case.ny} (80/D), Re 20, Mach 0.1, MUSCL(MC)+HLLC+viscous+RK3, "
                               "run to drag plateau", "cd": r["cd"], "cl": r["cl"], "cd_ptp": r["ptp"],
                   "sim_time": r["time"], "wall_s": round(r["wall_s"], 2), "gpu_s": round(r["gpu_ms"] * 1e-3, 3),
                   "history": [(h["step"], round(h["cd"], 5), round(h["ptp"], 5)) for h in r["history"][-5:]]},
    }


TRUTH_PATH = os.path.join(ROOT, "tests", "golden", "learned_truth_seed101_512.npz")


def learned_comparison(seed: int = 101, n: int = 512, T: float = 1.0, truth_rho=None):
    """The second half of the metric: time to the plain-MUSCL 2048^2 density L1 error.

    Learned n^2 (network on every face side, and the bypass mode that evaluates the gate only
    where the open-gate face value is near the TVD bound) against plain MUSCL at n^2 and 4n^2,
    on held-out seed `seed` to exactly t = T; density L1 against an FP64 MUSCL run 16x finer
    box-filtered to n^2 (the committed tests/golden/learned_truth_seed101_512.npz unless
    `truth_rho` is given).  GPU time: HIP events around hipGraph replays of each whole run."""
    import numpy as np
    from dev_field_efvm_amd import learned
    from dev_field_efvm_amd.efvm import Mlp
    if truth_rho is None:
        z = np.load(TRUTH_PATH)
        if int(z["seed"]) != seed or z["rho"].shape != (n, n) or float(z["T"]) != T:
            raise ValueError("the committed truth is for seed 101, 512^2, t = 1")
        truth_rho, truth_src = z["rho"].astype(np.float64), f"committed FP64 MUSCL {int(z['n_truth'])}^2"
    else:
        truth_src = f"FP64 MUSCL {16 * n}^2 run here"
    net = Mlp(learned.load_weights())
    runs = {}
    for name, nn, m, byp in (("learned", n, net, False), ("learned_bypass", n, net, True), ("muscl", n, None, False),
                             ("muscl_fine", 4 * n, None, False)):
        r = learned.run_case(seed, nn, T, m, bypass=byp)
        runs[name] = dict(l1=float(np.mean(np.abs(learned.box(r["rho"], nn // n) - truth_rho))), steps=r["steps"],
                          gpu_ms=round(r["gpu_ms"], 3), faces=r["faces"], fallback_faces=r["fallback"],
                          network_face_sides=r["gated_sides"])
    fine = runs["muscl_fine"]
    out = {"seed": seed, "n": n, "T": T, "truth": truth_src}
    for name in ("learned", "learned_bypass"):
        out[name] = {"time_ratio_vs_muscl_fine": round(runs[name]["gpu_ms"] / fine["gpu_ms"], 4),
                     "reaches_muscl_fine_l1": runs[name]["l1"] <= fine["l1"], "target_met": (
                         runs[name]["gpu_ms"] / fine["gpu_ms"] <= 0.1 and runs[name]["l1"] <= fine["l1"])}
    out["l1_gain_vs_muscl_same_grid"] = round(runs["muscl"]["l1"] / runs["learned"]["l1"], 2)
    out["runs"] = runs
    return out


def bench_learned(args):
    """Learned reconstruction on the periodic [0,10]^2 box (8 Fourier modes of amplitude 0.05
    plus 1-4 isentropic vortices, held-out seed): learned 512^2 against plain MUSCL at 512^2
    and 2048^2, density L1 at t = 1 against an FP64 MUSCL run 16x finer (8192^2, about three
    minutes), all box-filtered to 512^2.  value = GPU time of MUSCL 2048^2 / GPU time of learned 512^2 (HIP events around
    hipGraph replays), reported with whether the learned run reaches the 2048^2 error."""
    import numpy as np
    import torch
    from dev_field_efvm_amd import learned
    from dev_field_efvm_amd.efvm import Mlp
    torch.cuda.synchronize()
    net = Mlp(learned.load_weights())
    n, fine, truth_n, T, seed = args.learned_n, 4 * args.learned_n, 16 * args.learned_n, 1.0, args.learned_seed
    if seed == 101 and n == 512 and not args.learned_truth and os.path.exists(TRUTH_PATH):
        ref = np.load(TRUTH_PATH)["rho"].astype(np.float64)
    else:
        truth = learned.run_case(seed, truth_n, T, precision="fp64", timed=False,
                                 progress=lambda d, s: print(f"truth {truth_n}^2: {d}/{s} steps", file=sys.stderr, flush=True))
        ref = learned.box(truth["rho"], truth_n // n)
    runs = {}
    for name, nn, m, byp in (("learned", n, net, False), ("learned_bypass", n, net, True), ("muscl", n, None, False),
                             ("muscl_fine", fine, None, False)):
        r = learned.run_case(seed, nn, T, m, bypass=byp)
        runs[name] = dict(l1=float(np.mean(np.abs(learned.box(r["rho"], nn // n) - ref))), steps=r["steps"],
                          gpu_ms=round(r["gpu_ms"], 3), fallback_frac=(r["fallback"] / r["faces"]) if r["faces"] else 0.0,
                          gated_frac=(r["gated_sides"] / (2 * r["faces"])) if r["faces"] else 0.0)
    L, F, B = runs["learned"], runs["muscl_fine"], runs["learned_bypass"]
    speedup = F["gpu_ms"] / L["gpu_ms"]
    # matrix-core work of the network per learned stage: 4 face sides per cell, 20-64-64-4
    flops_stage = 4 * n * n * 2 * (20 * 64 + 64 * 64 + 64 * 4)
    stage_s = L["gpu_ms"] * 1e-3 / (3 * L["steps"])
    achieved = flops_stage / stage_s / 1e12
    return {
        "metric": METRIC, "value": speedup, "unit": f"x (GPU time MUSCL {fine}^2 / learned {n}^2 to equal t)",
        "n_gpus": 1, "steps": L["steps"], "warmup": 0, "ms_per_step": L["gpu_ms"] / L["steps"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (network) / f32 (flux, update)",
        "data": f"synthetic (seed {seed}: 8 Fourier modes amp 0.05 + isentropic vortices, held out of training)",
        "config": {"workload": f"learned reconstruction {n}^2 vs MUSCL {n}^2 / {fine}^2, t = {T}, truth FP64 MUSCL {truth_n}^2",
                   "l1_learned": L["l1"], "l1_muscl": runs["muscl"]["l1"], "l1_muscl_fine": F["l1"],
                   "reaches_fine_l1": L["l1"] <= F["l1"], "l1_gain_vs_muscl": runs["muscl"]["l1"] / L["l1"],
                   "gpu_time_ratio_learned_over_fine": L["gpu_ms"] / F["gpu_ms"],
                   "bypass": {"reaches_fine_l1": B["l1"] <= F["l1"], "gpu_time_ratio_over_fine": B["gpu_ms"] / F["gpu_ms"],
                              "note": "network gate evaluated only where the ungated face value is near the TVD bound"},
                   "runs": runs},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(achieved / 2500.0, 5), "traffic": None,
                     "kernel": "learned stage (face kernel: MLP + fluxes; update kernel), bf16 dense peak",
                     "mean_stage_us": round(stage_s * 1e6, 2), "network_flops_per_stage": flops_stage},
    }


def bench_learned_stage(args):
    """Per-stage GPU time of the learned reconstruction at n^2 (network on every face side; the
    bypass mode as the fused stage kernel and as its separate passes) next to plain MUSCL at
    n^2 and 4n^2, hipGraph replays of 10 steps (the short run to put under a kernel-trace
    profiler)."""
    import torch
    from dev_field_efvm_amd import learned
    from dev_field_efvm_amd.autotune import tune_context
    from dev_field_efvm_amd.efvm import Mlp
    torch.cuda.synchronize()
    net = Mlp(learned.load_weights())
    out = {}
    for name, n, use in (("learned", args.learned_n, 1), ("learned_bypass", args.learned_n, 2),
                         ("learned_bypass_separate", args.learned_n, 3), ("muscl", args.learned_n, 0),
                         ("muscl_fine", 4 * args.learned_n, 0)):
        c = learned.periodic_context(n)
        c.upload(learned.initial_state(args.learned_seed, n))
        if use:
            c.set_learned(net, bypass=use >= 2)
            c.set_learned_fused(use == 2)
        else:
            tune_context(c, steps=2, use_cache=False, store=False)
        c.step(2)
        c.graph_capture(10)
        ms = 0.0
        for _ in range(4):
            c.graph_launch()
            ms += c.last_elapsed_ms()
        out[name] = {"n": n, "us_per_stage": round(ms / 40 / 3 * 1e3, 2)}
        c.close()
    return {"metric": METRIC, "value": out["learned"]["us_per_stage"], "unit": "us per RK stage (learned, graph)",
            "n_gpus": 1, "steps": 40, "warmup": 2, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16 (network) / f32", "data": "synthetic", "config": {"workload": "learned stage timing", **out}}


def bench_fp32_error(args):
    """The FP32 production path against the FP64 path on the same workload (periodic n^2
    vortex + 8 modes, MUSCL(MC)+HLLC+RK3, CFL 0.4, `steps` steps): relative L-inf per stored
    component (normalised by the FP64 max |component|), density L1, final times."""
    import numpy as np
    import dev_field_efvm_amd as efvm
    from dev_field_efvm_amd import cases
    n = args.n
    st = cases.vortex_modes_state(n, n, seed=1)
    res = {}
    for prec in ("fp64", "fp32"):
        ctx = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, flux="hllc", muscl=True, limiter="mc",
                           precision=prec, cfl=0.4)
        ctx.set_boundaries(["periodic"] * 4)
        ctx.upload(st)
        ctx.step(args.steps)
        res[prec] = (ctx.download(), ctx.clock())
        ctx.close()
    a, b = res["fp32"][0], res["fp64"][0]
    rel = {name: float(np.max(np.abs(a[..., m] - b[..., m])) / np.max(np.abs(b[..., m])))
           for name, m in (("rho", 0), ("rho_u", 1), ("rho_v", 2), ("rho_E", 4))}
    l1 = float(np.mean(np.abs(a[..., 0] - b[..., 0])))
    return {"metric": "FP32 error against the FP64 path", "value": max(rel.values()), "unit": "relative L-inf",
            "n_gpus": 1, "steps": args.steps, "warmup": 0, "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 vs f64", "data": "synthetic (isentropic vortex + 8 seeded Fourier modes)",
            "config": {"workload": f"periodic {n}x{n} MUSCL(MC)+HLLC+SSP-RK3, CFL 0.4", "rel_linf": rel, "rho_l1": l1,
                       "time_fp32": res["fp32"][1]["time"], "time_fp64": res["fp64"][1]["time"]}}


def bench_sphere(args):
    """The reference's resolved sphere-drag validation on the 3-D context
    (examples/sphere_drag_validation, config.toml): 144 x 72 x 72 cells, D/dx = 12, Re 20,
    Mach 0.15, first-order HLLC + viscous + SSP-RK3, sponge, 2000 Fibonacci markers,
    9000 app.run() calls with Cd averaged over the last 20% (FP64 unless --precision fp32).
    value = interior cells x substages / GPU time of the bulk substages (HIP events),
    reported with Cd and its window peak-to-peak against the reference's 2.5402 / 2.7%."""
    import torch
    from dev_field_efvm_amd.sphere import SphereCase, run_reference
    prec = "fp64" if args.precision_sphere == "fp64" else "fp32"
    case = SphereCase(precision=prec)
    torch.cuda.synchronize()
    r = run_reference(case)
    ctx = r["ctx"]
    # GPU time per substage: graph replay of whole steps on the developed flow
    ctx.graph_capture(20)
    ctx.graph_launch()
    ctx.synchronize()
    ms = []
    for _ in range(5):
        ctx.graph_launch()
        ms.append(ctx.last_elapsed_ms())
    per_sub_s = min(ms) * 1e-3 / 60.0
    cells = case.nx * case.ny * case.nz
    # stage-kernel roofline: HBM bytes per raw cell-substage are U and U0 read and the new set
    # written, 5 components each; the kernel is bound by its FP64 arithmetic, not by these bytes
    raw = (case.nx + 4) * (case.ny + 4) * (case.nz + 4)
    esz = 8 if prec == "fp64" else 4
    stage_bytes = raw * 3 * 5 * esz
    ctx.step(1)
    ctx.synchronize()
    t_stage = _time_stage3(ctx)
    roof = {"bound": "hbm", "achieved": round(stage_bytes / t_stage / 1e9, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(stage_bytes / t_stage / 1e9 / PEAK_HBM_GBS, 4), "traffic": None,
            "kernel": "one 3-D substage, all its kernels (stage3_kernel is 78% of it, bound by FP64 arithmetic: six HLLC-Thornber faces per cell in the reference's form)",
            "mean_substage_us": round(t_stage * 1e6, 2), "algorithmic_bytes_per_launch": stage_bytes}
    out = {
        "metric": METRIC, "value": cells / per_sub_s, "unit": "cell-updates/s", "n_gpus": 1, "steps": 9000 // 3,
        "warmup": 0, "ms_per_step": 3e3 * per_sub_s, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else "f32",
        "data": "synthetic (uniform freestream past a fixed sphere)",
        "config": {"workload": f"sphere IBM 3-D, {case.nx}x{case.ny}x{case.nz} (D/dx 12), Re 20, Mach 0.15, "
                               "HLLC+viscous+RK3, 9000 substages", "cd": r["cd"], "cd_published": 2.5402,
                   "cd_ptp": r["ptp"], "cd_ptp_published": 0.027, "sim_time": r["time"],
                   "wall_s": round(r["wall_s"], 2), "us_per_substage": round(per_sub_s * 1e6, 2)},
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_sphere()
    return out


def _time_stage3(ctx, substages: int = 30):
    """Mean GPU time of one substage of the 3-D context (HIP events around whole steps, which
    also hold the fill, mirror, CFL and sponge kernels: an upper bound for the stage kernel)."""
    ctx.step(substages // 3)
    return ctx.last_elapsed_ms() * 1e-3 / substages


def cpu_baseline_sphere(n: int = 48, steps: int = 300):
    """The oracle (single thread, FP64) on the sphere case scaled to n x n/2 x n/2 cells."""
    from oracle import orc
    from dev_field_efvm_amd.sphere import SphereCase
    c = SphereCase(nx=n, ny=n // 2, nz=n // 2, markers=2000)
    fs = c.freestream()
    sim = orc.Sim(c.nx, c.ny, c.nz, lo=c.lo, hi=c.hi, flux="hllc", viscous=True, mu=c.mu, cfl=c.cfl,
                  bcs=["subsonic_inflow", "pressure_outlet", "symmetry", "symmetry", "symmetry", "symmetry"],
                  bc_params=[(c.rho, c.u_inf, 0.0, 0.0, c.t_inf), (c.p,), None, None, None, None],
                  spheres=[(0.0, 0.0, 0.0, 0.5 * c.D)], freestream=fs, sphere_markers=c.markers,
                  sponge_uniform=(c.sponge_x0, c.hi[0], c.sponge_sigma, fs))
    import numpy as np
    sim.set_interior(np.broadcast_to(fs, (c.nz, c.ny, c.nx, 5)).copy())
    sim.step(1)
    t0 = time.perf_counter()
    sim.step(steps)
    dt = time.perf_counter() - t0
    cells = c.nx * c.ny * c.nz
    return {"value": cells * 3 * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"oracle/efvm_oracle.c FP64 sphere case at {c.nx}x{c.ny}x{c.nz}, {steps} steps ({dt:.1f} s, 1 thread)"}


def bench_copy(args):
    """A 256 MiB device-to-device copy, three times: the known-size reference that scales the
    FETCH_SIZE / WRITE_SIZE counters of a PMC pass to bytes."""
    import torch
    n = 64 * 1024 * 1024
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"metric": "copy calibration", "value": 4 * n, "unit": "bytes per copy", "n_gpus": 1, "steps": 3,
            "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic", "config": {"workload": "256 MiB hipMemcpy device-to-device x3"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="single",
                    choices=["single", "batch", "cylinder", "learned", "learned-stage", "fp32-error", "copy",
                             "sphere"])
    ap.add_argument("--precision-sphere", default="fp64", choices=["fp32", "fp64"])
    ap.add_argument("--learned-n", type=int, default=512)
    ap.add_argument("--learned-seed", type=int, default=101)
    ap.add_argument("--cyl-nx", type=int, default=1760)
    ap.add_argument("--cyl-ny", type=int, default=328)
    ap.add_argument("--cyl-max-steps", type=int, default=200000)
    ap.add_argument("--cases", type=int, default=512)
    ap.add_argument("--retune", action="store_true")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--strip-rows", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--fused-rows", type=int, default=-1,
                    help="strip height of the fused stage-2+3 launch (0: separate stage launches; default: autotuner)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-MUSCL time-to-error comparison")
    ap.add_argument("--learned-truth", action="store_true", help="recompute the learned workload's FP64 truth")
    ap.add_argument("--cpu-n", type=int, default=1024)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1 (multi-GPU sharding is a design "
                                   "note, not a runnable path)"}))
        sys.exit(2)

    if args.workload == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "cylinder":
        print(json.dumps(bench_cylinder(args)), flush=True)
        return
    if args.workload == "learned":
        print(json.dumps(bench_learned(args)), flush=True)
        return
    if args.workload == "learned-stage":
        print(json.dumps(bench_learned_stage(args)), flush=True)
        return
    if args.workload == "copy":
        print(json.dumps(bench_copy(args)), flush=True)
        return
    if args.workload == "sphere":
        print(json.dumps(bench_sphere(args)), flush=True)
        return
    if args.workload == "fp32-error":
        print(json.dumps(bench_fp32_error(args)), flush=True)
        return

    import torch
    import dev_field_efvm_amd as efvm
    from dev_field_efvm_amd import cases

    n = args.n
    t_ic = time.perf_counter()
    st = cases.vortex_modes_state(n, n, seed=1)
    t_ic = time.perf_counter() - t_ic
    ctx = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, flux="hllc", muscl=True, limiter="mc",
                       integrator="rk3", precision=args.precision, cfl=0.4)
    ctx.set_boundaries(["periodic"] * 4)
    ctx.synchronize()
    t_up = time.perf_counter()
    ctx.upload(st)                   # host FP64 wire format -> device FP32 planes (synchronous)
    t_up = time.perf_counter() - t_up
    del st
    if args.strip_rows or args.waves:
        ctx.set_launch(args.strip_rows, args.waves)
        tuned = {"rows": args.strip_rows, "waves": args.waves, "source": "command line"}
    else:
        # launch shape from the autotuner (cached per architecture / precision / size class; it
        # times real steps and restores the state from a device snapshot afterwards)
        from dev_field_efvm_amd.autotune import tune_context
        r, w, _ = tune_context(ctx, steps=2, use_cache=not args.retune)
        tuned = {"rows": r, "waves": w, "source": "autotuner"}
    if args.fused_rows >= 0:
        ctx.set_fused_step(args.fused_rows > 0, args.fused_rows)
        tuned["fused_source"] = "command line"
    fz = ctx.fused_step()
    fused = fz[0] and fz[2]
    if fused:
        tuned["fused_rows"] = fz[1]

    # warm-up, then the timed region: K whole steps, bracketed by device synchronisation
    ctx.step(args.warmup)
    ctx.synchronize()
    # the timed steps replay one hipGraph of g steps (the largest of 20, 10, 4, 2 dividing K)
    g = next((k for k in (20, 10, 4, 2) if args.steps % k == 0), 0)
    use_graph = args.graph and g > 0
    if use_graph:
        ctx.graph_capture(g)
        ctx.graph_launch()
        ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if use_graph:
        for _ in range(args.steps // g):
            ctx.graph_launch()
        ctx.synchronize()
    else:
        ctx.step(args.steps)
        ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # kernel time: one eager timed pass of the same K steps with HIP events on the solver stream
    ctx.step(args.steps)
    ev_ms = ctx.last_elapsed_ms()
    clk = ctx.clock()

    t_down = time.perf_counter()
    ctx.download()                   # device -> host FP64 wire format (synchronous)
    t_down = time.perf_counter() - t_down
    copy_gbs = copy_bandwidth()
    cells = n * n
    value = cells * 3 * args.steps / wall
    if fused:
        # one step = stage 1 (stage_f32_pair_kernel) + stages 2 and 3 (stage_f32_pair23_kernel,
        # which also takes the CFL maximum): the roofline is taken over the step, the unit the
        # RK-stage metric counts three cell-stages in (128 B per cell algorithmic, 80 B moved)
        step_s = ev_ms * 1e-3 / args.steps
        bytes_per_step = BYTES_PER_CELL_STEP[args.precision] * cells
        achieved = bytes_per_step / step_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": load_traffic(n, args.precision, f"step_{args.precision}_{n}_pmc.json"),
                "traffic_per": "step (both stage launches)",
                "kernel": "stage_f32_pair_kernel (stage 1) + stage_f32_pair23_kernel (stages 2 and 3 fused)",
                "launches_per_step": 2, "mean_step_us": round(step_s * 1e6, 2),
                "mean_launch_us": round(step_s * 1e6 / 2, 2),
                "algorithmic_bytes_per_step": bytes_per_step, "measured_copy_GBps": copy_gbs}
    else:
        launches = 3 * args.steps
        mean_launch_s = ev_ms * 1e-3 / launches
        bytes_per_launch = BYTES_PER_CELL_STEP[args.precision] / 3.0 * cells
        achieved = bytes_per_launch / mean_launch_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": load_traffic(n, args.precision),
                "kernel": ctx.launch()[2], "mean_launch_us": round(mean_launch_s * 1e6, 2),
                "algorithmic_bytes_per_launch": bytes_per_launch, "measured_copy_GBps": copy_gbs}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic (isentropic vortex + 8 seeded Fourier modes, host FP64)",
        "config": {"workload": f"2-D Euler MUSCL(MC)+HLLC+SSP-RK3, periodic {n}x{n}, CFL 0.4, fused CFL "
                               f"(per-cell lambda in {'FP32' if args.precision == 'fp32' else 'FP64'}, maximum and dt in FP64)",
                   "grid": [n, n], "graph_steps": g if use_graph else 0, "sim_time": clk["time"], "launch": tuned,
                   "parallelism": "single GPU",
                   # the same K steps with one upload of the host state before and one download
                   # after (the boundary's host buffers, over PCIe); not the headline value
                   "pcie_inclusive": {"value": cells * 3 * args.steps / (wall + t_up + t_down),
                                      "upload_s": round(t_up, 4), "download_s": round(t_down, 4)}},
        "roofline": roof,
    }
    ctx.close()
    if not args.no_learned:
        out["learned"] = learned_comparison()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
