"""Writes tests/golden/learned_truth_seed101_512.npz: the reference solution of the learned-
reconstruction evaluation, so that the benchmark can score the learned run without a long run.

Held-out seed 101 of the learned workload (periodic [0,10]^2, 8 Fourier modes of amplitude 0.05
plus 1-4 isentropic vortices), FP64 MUSCL(MC) + HLLC + SSP-RK3 at 8192^2, CFL 0.4, run in equal
steps to exactly t = 1, density box-filtered 16x to 512^2 and stored as float32 (its rounding,
6e-8 relative, is far below the 1e-5-level errors it judges).  Needs a GPU (about 3 minutes).

Usage: python tests/golden/make_learned_truth.py [out.npz]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from dev_field_efvm_amd import learned  # noqa: E402

SEED, N_TRUTH, N_EVAL, T = 101, 8192, 512, 1.0


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                             "learned_truth_seed101_512.npz")
    t0 = time.time()
    r = learned.run_case(SEED, N_TRUTH, T, precision="fp64", timed=False,
                         progress=lambda d, s: print(f"truth {N_TRUTH}^2: {d}/{s} steps {time.time() - t0:.0f}s", flush=True))
    rho = learned.box(r["rho"], N_TRUTH // N_EVAL).astype(np.float32)
    np.savez_compressed(out, rho=rho, seed=np.int64(SEED), n_truth=np.int64(N_TRUTH), T=np.float64(T),
                        steps=np.int64(r["steps"]))
    print("wrote", out, f"{os.path.getsize(out) / 1024:.0f} KiB, {r['steps']} steps, {time.time() - t0:.0f}s")


if __name__ == "__main__":
    main()
