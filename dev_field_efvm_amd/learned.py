"""Learned face reconstruction: training data, training and evaluation.

The workflow of the learned part, all on the GPU through the C ABI:

1. ``make_training_set``: the solver's own MUSCL-HLLC-RK3 runs at fine
   resolution (default 2048^2, FP64) from seeded initial conditions
   (``cases.learned_state``: 8 Fourier modes of amplitude 0.05 plus 1-4
   isentropic vortices), snapshots box-filtered 4x on the device into
   training samples (features of both sides of every coarse face, targets from
   the fine solution's face line averages).
2. ``train``: Adam on minibatches with the library's FP32 forward/backward
   kernels.
3. ``evaluate``: held-out seeds at the coarse resolution with the network
   (bf16 matrix-core inference, per-face MUSCL fallback) against plain MUSCL at
   the coarse and at the fine resolution, density L1 measured against a much
   finer MUSCL run, all box-filtered to the coarse grid; GPU time of each run
   from HIP events around its stage kernels.

The reference has no learned part (it reconstructs with the limited MUSCL
slope of crates/cfd_solver/src/lib.rs:170-339); its MUSCL is what the network
replaces and what judges it.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from . import cases
from .efvm import Context, Mlp, NPARAM

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "weights", "learned_recon.npy")
L_BOX = 10.0
NIN, NH, NOUT = 20, 64, 4


B3_OPEN = 8.0     # initial gate bias: sigmoid(8) = 0.99966, the limiter starts open


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal hidden layers; output weights zero with the gates open (bias B3_OPEN); the
    learned correction zero (the untrained network is the unlimited central slope)."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0.0, math.sqrt(2.0 / NIN), (NH, NIN)).ravel(), np.zeros(NH),
             rng.normal(0.0, math.sqrt(2.0 / NH), (NH, NH)).ravel(), np.zeros(NH),
             np.zeros(NOUT * NH), np.full(NOUT, B3_OPEN), np.zeros(NOUT * NIN), np.zeros(NOUT)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == NPARAM
    return p


def load_weights(path: str = WEIGHTS) -> np.ndarray:
    p = np.load(path, allow_pickle=False)
    if p.shape != (NPARAM,):
        raise ValueError(f"{path}: expected {NPARAM} parameters")
    return p.astype(np.float32)


def periodic_context(n: int, precision: str = "fp32", muscl: bool = True, cfl: float = 0.4) -> Context:
    c = Context(n, n, hi=(L_BOX, L_BOX), dz=L_BOX / n, gamma=1.4, flux="hllc", muscl=muscl, limiter="mc",
                precision=precision, cfl=cfl, seam_mode=1)
    c.set_boundaries(["periodic"] * 4)
    return c


def initial_state(seed: int, n: int) -> np.ndarray:
    """Cell averages by sub x sub midpoint samples; the sub-cell spacing stays <= 10/4096
    (its O(h_sub^2) error is far below every run's discretisation error)."""
    return cases.learned_state(seed, n, L_BOX, sub=max(1, min(4, 4096 // n)))


def make_training_set(mlp: Mlp, seeds, n_fine: int = 2048, factor: int = 4, times=None, precision: str = "fp64",
                      verbose: bool = False):
    """Append the box-filtered snapshots of fine MUSCL runs to the network's sample store."""
    times = np.round(np.arange(0.1, 1.01, 0.1), 10) if times is None else times
    for s in seeds:
        c = periodic_context(n_fine, precision)
        c.upload(initial_state(s, n_fine))
        t0 = time.perf_counter()
        for T in times:
            c.advance_to(float(T))
            mlp.add_snapshot(c, factor)
        if verbose:
            print(f"  seed {s}: {len(times)} snapshots of {n_fine}^2 in {time.perf_counter() - t0:.1f} s "
                  f"({mlp.num_samples()} samples)", flush=True)
        c.close()


def advance_exact(c: Context, T: float, cfl: float = 0.4, safety: float = 0.9) -> int:
    """Advance from the context's current time to exactly T in equal CFL-limited steps."""
    c.set_time_step(cfl)
    t = c.clock()["time"]
    if T <= t:
        return 0
    n = int(math.ceil((T - t) / (safety * c.cfl_dt())))
    c.set_time_step(cfl, (T - t) / n)
    c.step(n)
    c.set_time_step(cfl)
    return n


def make_loop_set(mlp: Mlp, net: Mlp, seeds, n_fine: int = 2048, factor: int = 4, times=None, verbose: bool = False):
    """Samples of the states the network meets inside the solver: a coarse run with `net`
    attached and the fine MUSCL run advance side by side to the same exact times; the features
    come from the coarse run's own cells, the targets from the fine run's face line averages."""
    times = np.round(np.arange(0.1, 1.01, 0.1), 10) if times is None else times
    n_c = n_fine // factor
    for s in seeds:
        fine = periodic_context(n_fine, "fp64")
        fine.upload(initial_state(s, n_fine))
        coarse = periodic_context(n_c, "fp32")
        coarse.upload(initial_state(s, n_c))
        coarse.set_learned(net)
        t0 = time.perf_counter()
        for T in times:
            advance_exact(fine, float(T))
            advance_exact(coarse, float(T))
            mlp.add_snapshot(fine, factor, coarse=coarse)
        if verbose:
            print(f"  seed {s}: {len(times)} paired snapshots in {time.perf_counter() - t0:.1f} s "
                  f"({mlp.num_samples()} samples)", flush=True)
        fine.close()
        coarse.close()


def train(mlp: Mlp, steps: int = 20000, batch: int = 16384, lr0: float = 2e-3, lr1: float = 2e-5, seed: int = 1,
          chunk: int = 2000, train_skip: bool = False, verbose: bool = False):
    """Adam with a cosine schedule over `steps`, run in chunks (one schedule per chunk would
    restart the cosine, so the chunk boundaries interpolate the global schedule)."""
    losses = []
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        f0, f1 = done / max(steps - 1, 1), (done + k - 1) / max(steps - 1, 1)
        lr_a = lr1 + 0.5 * (lr0 - lr1) * (1 + math.cos(math.pi * f0))
        lr_b = lr1 + 0.5 * (lr0 - lr1) * (1 + math.cos(math.pi * f1))
        losses.append(mlp.train(k, batch, lr_a, lr_b, seed, train_skip))
        done += k
        if verbose:
            print(f"  step {done}: minibatch loss {losses[-1][-50:].mean():.3e}", flush=True)
    return np.concatenate(losses) if losses else np.zeros(0)


def box(a: np.ndarray, f: int) -> np.ndarray:
    """Box filter of an (n, n, ...) array by f."""
    n = a.shape[0]
    return a.reshape(n // f, f, n // f, f, *a.shape[2:]).mean(axis=(1, 3))


def exact_steps(c: Context, T: float, cfl: float = 0.4, safety: float = 0.9) -> int:
    """Switch the context to N equal steps that land exactly on T: dt = T / N with
    N = ceil(T / (safety * dt_cfl)) from the current state's CFL dt.  Runs compared at one
    final time must end there: stopping at the first t >= T leaves an overshoot of up to one
    dt that differs between resolutions and, at these error levels, would dominate the L1."""
    dt0 = c.cfl_dt()
    n = int(math.ceil(T / (safety * dt0)))
    c.set_time_step(cfl, T / n)
    return n


def timed_steps(c: Context, n: int, graph_steps: int = 20) -> float:
    """Advance n steps; GPU milliseconds from HIP events around the replayed hipGraphs of
    `graph_steps` steps (no launch gaps) plus the eager remainder."""
    ms = 0.0
    full, rest = divmod(n, graph_steps)
    if full:
        c.graph_capture(graph_steps)
        for _ in range(full):
            c.graph_launch()
            ms += c.last_elapsed_ms()
    if rest:
        c.step(rest)
        ms += c.last_elapsed_ms()
    return ms


def run_case(seed: int, n: int, T: float, mlp: Mlp | None = None, precision: str = "fp32", timed: bool = True,
             progress=None, chunk: int = 1000, bypass: bool = False):
    """One run to exactly T.  Untimed runs advance in chunks of `chunk` steps and report
    (steps done, steps) to `progress` after each, if given.  bypass: the network's gate only
    where the ungated face value is near the TVD bound."""
    c = periodic_context(n, precision)
    c.upload(initial_state(seed, n))
    if mlp is not None:
        c.set_learned(mlp, bypass=bypass)
    elif timed:
        # a timed MUSCL run gets its autotuned launch shape: the baseline is not handicapped
        from .autotune import tune_context
        tune_context(c, steps=2)
    steps = exact_steps(c, T)
    if timed:
        ms = timed_steps(c, steps)
    else:
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            c.step(k)
            c.synchronize()
            done += k
            if progress is not None:
                progress(done, steps)
        ms = float("nan")
    state = c.download()
    # the fixed dt must have stayed within the CFL limit of the final state too
    c.set_time_step(0.4)
    lam_dt = c.cfl_dt()
    if T / steps > 1.1 * lam_dt:
        raise RuntimeError(f"fixed dt {T / steps:.3e} exceeds the final CFL dt {lam_dt:.3e} by more than 10%")
    stats = c.learned_stats() if mlp is not None else (0, 0, 0)
    c.close()
    return dict(rho=state[..., 0], steps=steps, gpu_ms=ms, fallback=stats[0], faces=stats[1], gated_sides=stats[2])


def evaluate(mlp: Mlp, seed: int, n: int = 512, fine: int = 2048, truth_n: int = 8192, T: float = 1.0,
             truth_precision: str = "fp32", verbose: bool = False):
    """Density L1 (box-filtered to n) of learned-n, MUSCL-n and MUSCL-fine against MUSCL-truth_n."""
    t0 = time.perf_counter()
    truth = run_case(seed, truth_n, T, precision=truth_precision, timed=False)
    ref = box(truth["rho"], truth_n // n)
    out = dict(seed=seed, n=n, fine=fine, truth_n=truth_n, T=T)
    for name, nn, net in (("learned", n, mlp), ("muscl", n, None), ("muscl_fine", fine, None)):
        r = run_case(seed, nn, T, net)
        rho = box(r["rho"], nn // n)
        out[name] = dict(l1=float(np.mean(np.abs(rho - ref))), steps=r["steps"], gpu_ms=r["gpu_ms"],
                         fallback_frac=(r["fallback"] / r["faces"]) if r["faces"] else 0.0)
    out["truth_steps"] = truth["steps"]
    out["l1_gain_vs_muscl"] = out["muscl"]["l1"] / out["learned"]["l1"]
    out["time_ratio_vs_fine"] = out["learned"]["gpu_ms"] / out["muscl_fine"]["gpu_ms"]
    out["reaches_fine_l1"] = out["learned"]["l1"] <= out["muscl_fine"]["l1"]
    if verbose:
        print(json.dumps(out), flush=True)
    out["wall_s"] = time.perf_counter() - t0
    return out


def validation_error(p, seed: int, n: int, truth_rho: np.ndarray, T: float = 1.0) -> float:
    """Density L1 of a learned n^2 run (network p) against a finer run box-filtered to n."""
    r = run_case(seed, n, T, Mlp(p), timed=False)
    return float(np.mean(np.abs(r["rho"] - truth_rho)))


def train_default(train_seeds=range(1, 9), n_fine: int = 2048, factor: int = 4, steps: int = 20000, batch: int = 16384,
                  out: str = WEIGHTS, val_seed: int = 9, verbose: bool = True):
    """The committed network: 8 seeds of 2048^2 FP64 MUSCL, 10 snapshots each (t = 0.1 .. 1.0).

    1. the learned correction L x + bL in closed form (weighted normal equations on the
       device), kept exact for cubic profiles: the unconstrained optimum of the smooth
       training data amplifies grid-scale noise once it runs inside the solver;
    2. Adam on the limiter MLP (its output layer starts at zero with the gates open), the
       correction held fixed;
    3. model selection: the stage-1 and stage-2 networks each run the held-out seed
       `val_seed` at the coarse resolution, and the one with the smaller density L1 against
       an FP64 run 8x finer is kept (offline training loss and the error inside the solver
       need not agree)."""
    n_c = n_fine // factor
    mlp = Mlp(init_params(0))
    t0 = time.perf_counter()
    make_training_set(mlp, train_seeds, n_fine, factor, verbose=verbose)
    t1 = time.perf_counter()
    mlp.fit_linear(stride=4, constrained=True)
    linear_only = mlp.save()
    losses = train(mlp, steps, batch, train_skip=False, verbose=verbose)
    full = mlp.save()
    t2 = time.perf_counter()
    truth = run_case(val_seed, 8 * n_c, 1.0, precision="fp64", timed=False)
    truth_rho = box(truth["rho"], 8)
    e_lin = validation_error(linear_only, val_seed, n_c, truth_rho)
    e_full = validation_error(full, val_seed, n_c, truth_rho)
    p = full if e_full <= e_lin else linear_only
    if out:
        os.makedirs(os.path.dirname(out), exist_ok=True)
        np.save(out, p)
    info = dict(samples=mlp.num_samples(), data_s=t1 - t0, train_s=t2 - t1, final_loss=float(losses[-200:].mean()),
                first_loss=float(losses[:50].mean()), steps=steps, batch=batch, seeds=list(train_seeds), n_fine=n_fine,
                factor=factor, val_seed=val_seed, val_l1_correction_only=e_lin, val_l1_with_limiter=e_full,
                kept="with_limiter" if p is full else "correction_only")
    if verbose:
        print(json.dumps(info), flush=True)
    info["linear_params"] = linear_only
    return mlp, info


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="train / evaluate the learned face reconstruction")
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--seeds", default="1-8")
    ap.add_argument("--eval-seeds", default="101,102")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--fine", type=int, default=2048)
    ap.add_argument("--truth", type=int, default=8192)
    ap.add_argument("--T", type=float, default=1.0)
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args()
    if a.cmd == "train":
        lo, hi = (int(x) for x in a.seeds.split("-"))
        train_default(range(lo, hi + 1), a.fine, a.fine // a.n, a.steps, a.batch, a.out)
    else:
        net = Mlp(load_weights(a.out))
        for s in (int(x) for x in a.eval_seeds.split(",")):
            evaluate(net, s, a.n, a.fine, a.truth, a.T, verbose=True)
