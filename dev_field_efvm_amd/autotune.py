"""Launch-shape autotuner for the stage kernels.

The stage kernels are wave strips (64 or 128 columns x `strip_rows` rows per wave);
the shape trades halo re-reads (4 extra rows per strip) against parallelism
(strips per CU) and the tail of the last wave round.  The FP32 kernel runs a
strip's strip_rows + 1 row steps in groups of four, so the candidate heights
are one less than a multiple of four (no idle steps).  The tuner times real
steps of the caller's own context or batch with HIP events on its stream
(library-side `efvm_last_elapsed_ms`), restores the state from a device
snapshot afterwards, and caches the winner as JSON keyed by
architecture / precision / workload size.
"""
from __future__ import annotations

import json
import math
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE_PATH = os.environ.get("EFVM_TUNE_CACHE", os.path.join(HERE, "tuned.json"))

CTX_ROWS = (7, 15, 23, 27, 31, 39, 43, 47, 55, 63, 127, 255)
CTX_WAVES = (1, 2, 4)
# strip heights of the fused stage-2+3 launch (a strip of h rows runs h + 5 row steps of stage 2
# and h + 3 of stage 3, the loop in groups of four: h = 4k + 1 leaves no idle step)
FUSED_ROWS = (13, 21, 25, 29, 33, 37, 41, 45, 49, 57, 65, 73, 97)
BATCH_ROWS = (7, 15, 31, 63, 127)
FUSED_BATCH_ROWS = (13, 21, 29, 37, 45, 57)
BATCH_BLOCKS = (0, 2048, 4096, 8192)


def arch_name() -> str:
    """The device architecture from the library's own HIP runtime (torch's may not be the one
    the library links, and need not be initialised)."""
    from .efvm import device_arch
    return device_arch(0).split(":")[0]


def _load():
    try:
        return json.load(open(CACHE_PATH))
    except Exception:
        return {}


def _store(db):
    tmp = CACHE_PATH + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE_PATH)


def size_class(n: int) -> int:
    """Power-of-two bucket of a grid side."""
    return 1 << max(0, int(math.floor(math.log2(max(1, n)))))


def _time_ctx(ctx, steps):
    ctx.step(1)                     # compile/warm
    ctx.step(steps)
    return ctx.last_elapsed_ms() / steps


def tune_context(ctx, steps: int = 3, rows=CTX_ROWS, waves=CTX_WAVES, use_cache=True, verbose=False, store=True,
                 fused_rows=FUSED_ROWS):
    """Pick (strip_rows, waves_per_block) and, where both apply, the one- or two-column strip
    kernel for a context, then, where it applies, the fused stage-2+3 launch's strip height (or
    no fusion); returns (rows, waves, ms_per_step).  The two-column kernel halves the lane
    exchanges per cell but covers 124 columns per wave: at sizes just above a multiple of 124
    (512, 2048) its last strip is nearly empty, and the one-column kernel can win."""
    fits = ctx.launch()[3]          # 0: one-column only; 1: the periodic two-column form; 2: the general one
    fused_fits = ctx.fused_step()[2]
    key = f"{arch_name()}:{ctx.precision}:ctx:{size_class(ctx.nx)}x{size_class(ctx.ny)}:" + \
          ("general", "periodic-plain", "walls-plain")[fits]
    db = _load()
    if use_cache and key in db:
        e = db[key]
        ctx.set_pair_kernel(e.get("pair", 1))
        ctx.set_launch(e["rows"], e["waves"])
        if fused_fits:
            ctx.set_fused_step(e.get("fused", 0) > 0, e.get("fused", 0))
        return e["rows"], e["waves"], e["ms"]
    ctx.snapshot()
    if fused_fits:
        ctx.set_fused_step(False)
    best = None
    for pair in ((1, 0) if fits else (1,)):
        ctx.set_pair_kernel(pair)
        for r in rows:
            for w in waves:
                ctx.set_launch(r, w)
                ms = _time_ctx(ctx, steps)
                ctx.snapshot(restore=True)
                if verbose:
                    print(f"  pair={pair} rows={r:4d} waves={w}  {ms:.4f} ms/step", flush=True)
                if best is None or ms < best[2]:
                    best = (r, w, ms, pair)
    ctx.set_pair_kernel(best[3])
    ctx.set_launch(best[0], best[1])
    fused = 0
    if fused_fits:
        # stage 1 keeps the shape just chosen; the fused launch's strip height, or none
        for r in fused_rows:
            ctx.set_fused_step(True, r)
            ms = _time_ctx(ctx, steps)
            ctx.snapshot(restore=True)
            if verbose:
                print(f"  fused rows={r:4d}  {ms:.4f} ms/step", flush=True)
            if ms < best[2]:
                best, fused = (best[0], best[1], ms, best[3]), r
        ctx.set_fused_step(fused > 0, fused)
    if store:
        db[key] = {"rows": best[0], "waves": best[1], "ms": best[2], "pair": best[3], "fused": fused,
                   "when": time.time()}
        _store(db)
    return best[:3]


def _time_batch(b, steps, repeats=2):
    """Best of `repeats` timings of `steps` steps (HIP events), after one warm-up step: a
    single short timing of a 95 M-cell ensemble varies by a few percent from run to run."""
    b.step(1)
    best = None
    for _ in range(repeats):
        b.step(steps)
        ms = b.last_elapsed_ms() / steps
        best = ms if best is None else min(best, ms)
    return best


def tune_batch(batch, sides, steps: int = 4, rows=BATCH_ROWS, blocks=BATCH_BLOCKS, use_cache=True, verbose=False, store=True,
               key_extra=""):
    """Per size class strip rows, then the persistent grid size (coordinate descent).

    sides: per-case grid side (nx) used to form the size classes.
    Returns ({class: rows}, blocks, ms_per_step)."""
    classes = sorted(set(size_class(s) for s in sides))
    key = f"{arch_name()}:fp32:batch:{len(sides)}:{min(sides)}-{max(sides)}{key_extra}"
    db = _load()
    if use_cache and key in db:
        choice = {int(k): v for k, v in db[key]["rows"].items()}
        batch.set_pair_kernel(db[key].get("pair", 0))
        batch.set_fused(db[key].get("fused", 0) > 0, db[key].get("fused", 0))
        _apply_batch(batch, sides, choice)
        batch.set_grid(db[key]["blocks"], 4)
        return choice, db[key]["blocks"], db[key]["ms"]
    batch.snapshot()
    # the two-column kernel for the cases it takes, or the one-column kernel for all: both at
    # the default heights, the faster kept (then heights and grid are tuned for it)
    choice = {c: 31 for c in classes}
    _apply_batch(batch, sides, choice)
    nb = 0
    batch.set_grid(nb, 4)
    best_ms, pair = None, 0
    for p in (1, 0):
        batch.set_pair_kernel(p)
        ms = _time_batch(batch, steps)
        batch.snapshot(restore=True)
        if verbose:
            print(f"  pair={p}  {ms:.3f} ms/step", flush=True)
        if best_ms is None or ms < best_ms:
            best_ms, pair = ms, p
    batch.set_pair_kernel(pair)
    # one height for every class first, then class by class from the best uniform height
    for r in rows:
        if r == choice[classes[0]]:
            continue
        trial = {c: r for c in classes}
        _apply_batch(batch, sides, trial)
        ms = _time_batch(batch, steps)
        batch.snapshot(restore=True)
        if verbose:
            print(f"  all classes rows={r:4d}  {ms:.3f} ms/step", flush=True)
        if ms < best_ms:
            best_ms, choice = ms, trial
    _apply_batch(batch, sides, choice)
    for c in sorted(classes, reverse=True):          # largest classes dominate the time
        for r in rows:
            if r == choice[c]:
                continue
            trial = dict(choice)
            trial[c] = r
            _apply_batch(batch, sides, trial)
            ms = _time_batch(batch, steps)
            batch.snapshot(restore=True)
            if verbose:
                print(f"  class {c:5d} rows={r:4d}  {ms:.3f} ms/step", flush=True)
            if ms < best_ms:
                best_ms, choice = ms, trial
        _apply_batch(batch, sides, choice)
    for g in blocks:
        if g == nb:
            continue
        batch.set_grid(g, 4)
        ms = _time_batch(batch, steps)
        batch.snapshot(restore=True)
        if verbose:
            print(f"  blocks={g:5d}  {ms:.3f} ms/step", flush=True)
        if ms < best_ms:
            best_ms, nb = ms, g
    batch.set_grid(nb, 4)
    # the fused stage-2+3 launch of the periodic two-column cases: its strip height, or none
    fused = batch.fused()[1] if batch.fused()[2] > 0 else 0
    if pair and batch.fused()[2] > 0:
        for r in (0,) + FUSED_BATCH_ROWS:
            if r == fused:
                continue
            batch.set_fused(r > 0, r)
            ms = _time_batch(batch, steps)
            batch.snapshot(restore=True)
            if verbose:
                print(f"  fused rows={r:4d}  {ms:.3f} ms/step", flush=True)
            if ms < best_ms:
                best_ms, fused = ms, r
    batch.set_fused(fused > 0, fused)
    if store:
        db[key] = {"rows": {str(k): v for k, v in choice.items()}, "blocks": nb, "ms": best_ms, "pair": pair,
                   "fused": fused, "when": time.time()}
        _store(db)
    return choice, nb, best_ms


def _apply_batch(batch, sides, choice):
    batch.set_rows([choice[size_class(s)] for s in sides])
