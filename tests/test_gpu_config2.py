"""The 4096^2 workload at its stated size: the FP64 GPU path over 200 steps against the
oracle's fixture (tests/golden/config2_fp64_4096.npz, written by
tests/golden/make_config2_fixture.py), and the FP32 production path's error against it.

Workload: periodic [0,10]^2, 4096 x 4096, isentropic vortex (beta 5, mean flow (1, 1)) plus
8 seeded Fourier modes (seed 1, amplitude 0.02); MUSCL(MC) + HLLC-Thornber + SSP-RK3,
CFL 0.4, seam mode 0 (the reference's periodic seam).
"""
import os
import struct

import numpy as np
import pytest

import dev_field_efvm_amd as efvm
from dev_field_efvm_amd import cases

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "config2_fp64_4096.npz")
COMP = (0, 1, 2, 4)
N = 4096
TOL64 = 1e-12
TOL32 = 1e-4            # the FP32 path's stated bound (relative L-inf per component)


@pytest.fixture(scope="module")
def fixture():
    if not os.path.exists(FIX):
        pytest.fail("missing tests/golden/config2_fp64_4096.npz (run tests/golden/make_config2_fixture.py)")
    return dict(np.load(FIX))


def run(precision, steps, per_step_clock=False):
    ctx = efvm.Context(N, N, hi=(10.0, 10.0), dz=10.0 / N, flux="hllc", muscl=True, limiter="mc",
                       integrator="rk3", precision=precision, cfl=0.4, seam_mode=0)
    ctx.set_boundaries(["periodic"] * 4)
    ctx.upload(cases.vortex_modes_state(N, N, seed=1, amplitude=0.02))
    clocks = []
    if per_step_clock:
        for _ in range(steps):
            ctx.step(1)
            c = ctx.clock()
            clocks.append((c["dt"], c["time"]))
    else:
        ctx.step(steps)
    st = ctx.download()
    sums = ctx.sums()
    clk = ctx.clock()
    ctx.close()
    return st, clocks, sums, clk


def errors(state, fx):
    """Relative L-inf of the strided subsample and of the per-row / per-column sums, per component,
    normalised by the fixture's max |component| (sums: by nx times it)."""
    s4 = state[..., list(COMP)]
    scale = np.maximum(np.abs(fx["vmax"]), np.abs(fx["vmin"]))
    sub = s4[3::64, 5::64, :]
    e_sub = np.max(np.abs(sub - fx["sub"]).reshape(-1, 4), axis=0) / scale
    e_row = np.max(np.abs(s4.sum(axis=1) - fx["row_sum"]), axis=0) / (N * scale)
    e_col = np.max(np.abs(s4.sum(axis=0) - fx["col_sum"]), axis=0) / (N * scale)
    e_ext = np.maximum(np.abs(s4.reshape(-1, 4).max(axis=0) - fx["vmax"]),
                       np.abs(s4.reshape(-1, 4).min(axis=0) - fx["vmin"])) / scale
    return e_sub, e_row, e_col, e_ext


def test_config2_fp64_200_steps_matches_the_oracle_fixture(fixture):
    st, clocks, sums, clk = run("fp64", 200, per_step_clock=True)
    dts = np.array([c[0] for c in clocks])
    times = np.array([c[1] for c in clocks])
    # the CFL reduction and the clock are FP64 and follow the oracle's order: bit for bit, every step
    assert np.array_equal(dts.view(np.int64), fixture["dt"].view(np.int64))
    assert np.array_equal(times.view(np.int64), fixture["time"].view(np.int64))
    assert clk["step"] == 200
    e_sub, e_row, e_col, e_ext = errors(st, fixture)
    print("FP64 4096^2 x 200 vs oracle: subsample", e_sub, "row sums", e_row, "column sums", e_col, "extrema", e_ext)
    assert np.max(e_sub) <= TOL64 and np.max(e_row) <= TOL64 and np.max(e_col) <= TOL64 and np.max(e_ext) <= TOL64
    # FP64 conservation sums (device reduction order vs the oracle's serial sum)
    ref = fixture["sums"][list(COMP)]
    print("sums rel diff", np.abs(sums[list(COMP)] - ref) / np.abs(ref))
    assert np.all(np.abs(sums[list(COMP)] - ref) <= 1e-12 * np.abs(ref))
    # the flow moved: the comparison is not of two copies of the initial state
    assert fixture["l1_rho_ic"] > 1e-4


def test_config2_fp32_error_against_the_fp64_fixture(fixture):
    """FP32 flux and update arithmetic (per-cell CFL lambda in FP32, its maximum in FP64): its error after
    the same 200 steps."""
    st, _, _, clk = run("fp32", 200)
    e_sub, e_row, e_col, e_ext = errors(st, fixture)
    rel_t = abs(clk["time"] - fixture["time"][-1]) / fixture["time"][-1]
    print("FP32 4096^2 x 200 vs FP64 oracle: subsample", e_sub, "row sums", e_row, "column sums", e_col,
          "extrema", e_ext, "time", rel_t)
    assert np.max(e_sub) <= TOL32 and np.max(e_ext) <= TOL32
    assert rel_t <= 1e-6
