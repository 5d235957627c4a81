"""Writes tests/golden/config2_fp64_4096.npz: the FP64 oracle's state after 200
steps of the 4096^2 periodic vortex-plus-modes workload, in compact form.

Workload: [0,10]^2 periodic, 4096 x 4096 cells, one slip cell in z (dz = 10/4096),
gamma 1.4, Shu vortex (beta 5, mean flow (1, 1)) plus 8 seeded Fourier modes
(seed 1, amplitude 0.02) on rho, u, v; MUSCL with the monotonized-central
limiter, HLLC with Thornber's correction, SSP-RK3, CFL 0.4, seam mode 0 (the
reference's periodic seam), 200 steps.

The oracle runs once, on the host's cores through orc_set_threads (bitwise the
single-threaded result; about 30 minutes on 8 cores and 27 GB of host memory).
The fixture holds:
  dt, time         per step (float64, compared bit for bit)
  sub              state at every 64th row and column (rows 3 + 64 a, columns 5 + 64 b),
                   components (rho, rho_u, rho_v, rho_E)
  row_sum, col_sum sums of each component over each row / each column (numpy float64 sums)
  sums             the oracle's cell-volume-weighted sums of the 5 components
  vmax, vmin       per-component maximum and minimum over all cells
  l1_rho_ic        mean |rho - rho_initial| (how far the flow moved)

Usage: python tests/golden/make_config2_fixture.py [N] [steps]   (defaults 4096 200)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from dev_field_efvm_amd import cases  # noqa: E402
from oracle import orc  # noqa: E402

COMP = (0, 1, 2, 4)
STRIDE, ROW0, COL0 = 64, 3, 5


def compact(state):
    """The fixture's view of an (ny, nx, 5) state."""
    s4 = state[..., list(COMP)]
    return dict(
        sub=np.ascontiguousarray(s4[ROW0::STRIDE, COL0::STRIDE, :]),
        row_sum=s4.sum(axis=1),
        col_sum=s4.sum(axis=0),
        vmax=s4.reshape(-1, 4).max(axis=0),
        vmin=s4.reshape(-1, 4).min(axis=0),
    )


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 4096
    steps = int(sys.argv[2]) if len(sys.argv) > 2 else 200
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "config2_fp64_4096.npz" if n == 4096 else f"config2_fp64_{n}.npz")
    st0 = cases.vortex_modes_state(n, n, seed=1, amplitude=0.02)
    sim = orc.Sim(n, n, 1, hi=(10.0, 10.0, 10.0 / n), gamma=1.4, flux="hllc", muscl=True, limiter="mc",
                  integrator="rk3", cfl=0.4, bcs=["periodic"] * 4 + ["slip", "slip"], seam_mode=0)
    sim.set_interior(st0.reshape(1, n, n, 5))
    sim.set_threads(int(os.environ.get("ORC_THREADS", os.cpu_count() or 1)))
    dts, times = np.zeros(steps), np.zeros(steps)
    t0 = time.time()
    for k in range(steps):
        sim.step(1)
        info = sim.info()
        dts[k], times[k] = info["dt"], info["time"]
        print(f"step {k + 1}/{steps} dt={info['dt']:.17g} t={info['time']:.17g} "
              f"elapsed={time.time() - t0:.0f}s", flush=True)
    fin = sim.interior()[0]
    fix = compact(fin)
    np.savez_compressed(out, n=np.int64(n), steps=np.int64(steps), dt=dts, time=times, sums=sim.sums(),
                        l1_rho_ic=np.float64(np.mean(np.abs(fin[..., 0] - st0[..., 0]))), **fix)
    print("wrote", out, f"{os.path.getsize(out) / 1024:.0f} KiB in {time.time() - t0:.0f}s")


if __name__ == "__main__":
    main()
