"""Restart files on the GPU: the CFDRST01 checkpoint (examples/restart_determinism/main.rs:177-226)
and the bit-for-bit continuation the reference tests (crates/cfd_solver/tests/restart.rs:86-105)."""
import struct

import numpy as np
import pytest

import dev_field_efvm_amd as efvm
from dev_field_efvm_amd import cases
from oracle import orc

pytestmark = pytest.mark.gpu


def ctx(n, precision):
    c = efvm.Context(n, n, hi=(10.0, 10.0), flux="hllc", muscl=True, limiter="mc", integrator="rk3",
                     precision=precision, cfl=0.4, seam_mode=1)
    c.set_boundaries(["periodic"] * 4)
    c.upload(cases.vortex_modes_state(n, n, seed=3))
    return c


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_restored_run_continues_bit_for_bit(tmp_path, precision):
    n = 72
    ref = ctx(n, precision)
    ref.step(12)
    path = str(tmp_path / f"ck_{precision}.bin")
    ref.checkpoint_save(path)
    ref.step(18)
    restored = efvm.Context(n, n, hi=(10.0, 10.0), flux="hllc", muscl=True, limiter="mc", integrator="rk3",
                            precision=precision, cfl=0.4, seam_mode=1)
    restored.set_boundaries(["periodic"] * 4)
    restored.checkpoint_load(path)
    restored.step(18)
    a, b = ref.clock(), restored.clock()
    assert a["step"] == b["step"] == 30
    assert struct.pack("<d", a["time"]) == struct.pack("<d", b["time"])
    assert np.array_equal(ref.download(), restored.download())
    # the flow moved, so the comparison is not trivial
    assert not np.array_equal(ref.download(), cases.vortex_modes_state(n, n, seed=3))


def read_columns(path, n):
    data = open(path, "rb").read()
    assert data[:8] == b"CFDRST01"
    head = struct.unpack_from("<ddQB", data, 8)
    pos = 8 + 25
    nraw = (n + 4) * (n + 4) * 5
    cols = []
    for _ in range(3):
        (length,) = struct.unpack_from("<Q", data, pos)
        assert length == nraw
        cols.append(np.frombuffer(data, dtype="<f8", count=5 * nraw, offset=pos + 8).reshape(5, n + 4, n + 4, 5))
        pos += 8 + 40 * nraw
    assert pos == len(data)
    return head, cols


def test_checkpoint_layout(tmp_path):
    """u over the raw array with the device's interior; u0 is the reference's save_state copy
    (crates/cfd_solver/src/integrate.rs:76-78): the state before the completed step at a step
    boundary, U^n mid-step.  The 2-D device keeps no ghost cells and no rhs column, so those are
    written as zeros (the reference refills both before their next use)."""
    n = 24
    c = ctx(n, "fp64")
    c.step(2)
    before = c.download()
    c.step(1)
    path = str(tmp_path / "ck.bin")
    c.checkpoint_save(path)
    (dt, t, step, stage), (u, rhs, u0) = read_columns(path, n)
    clk = c.clock()
    assert (dt, t, step, stage) == (clk["dt"], clk["time"], 3, 0)
    interior = u[2, 2:n + 2, 2:n + 2]
    assert np.array_equal(interior, c.download())
    ghost = u.copy()
    ghost[2, 2:n + 2, 2:n + 2] = 0.0
    assert not ghost.any()                          # ghosts are written as zero
    assert not rhs.any()
    assert np.array_equal(u0[2, 2:n + 2, 2:n + 2], before)
    # against the oracle: its u0 column after the same three steps
    sim = orc.Sim(n, n, 1, hi=(10.0, 10.0, 10.0 / n), flux="hllc", muscl=True, limiter="mc", cfl=0.4, seam_mode=1,
                  bcs=["periodic"] * 4 + ["slip", "slip"])
    sim.set_interior(cases.vortex_modes_state(n, n, seed=3).reshape(1, n, n, 5))
    sim.step(3)
    ref_u0 = sim.raw_column("u0")[2, 2:n + 2, 2:n + 2]
    err = max(float(np.max(np.abs(u0[2, 2:n + 2, 2:n + 2, m] - ref_u0[..., m])) / np.max(np.abs(ref_u0[..., m])))
              for m in (0, 1, 2, 4))
    print(f"u0 column vs oracle rel L-inf {err:.2e}")
    assert err <= 1e-12
    # mid-step: after one substage the backup is U^n
    c.run()
    mid = str(tmp_path / "mid.bin")
    c.checkpoint_save(mid)
    (_, _, _, stage), (u_m, _, u0_m) = read_columns(mid, n)
    assert stage == 1 and np.array_equal(u0_m[2, 2:n + 2, 2:n + 2], interior)
    # load then save: the u0 column survives the round trip; the first step replaces it
    d = ctx(n, "fp64")
    d.checkpoint_load(path)
    again = str(tmp_path / "again.bin")
    d.checkpoint_save(again)
    assert open(again, "rb").read() == open(path, "rb").read()
    # a fresh state (no step yet) carries a zero backup, as a new CfdState does
    e = ctx(n, "fp64")
    fresh = str(tmp_path / "fresh.bin")
    e.checkpoint_save(fresh)
    assert not read_columns(fresh, n)[1][2].any()


def test_checkpoint_load_rejects_other_grids_and_mid_step_files(tmp_path):
    c = ctx(24, "fp64")
    c.step(2)
    path = str(tmp_path / "a.bin")
    c.checkpoint_save(path)
    other = ctx(32, "fp64")
    with pytest.raises(efvm.EfvmError) as e:
        other.checkpoint_load(path)
    assert e.value.status == 2                      # EFVM_ERR_CONFIG
    c.run()                                         # one substage: rk stage 1
    mid = str(tmp_path / "mid.bin")
    c.checkpoint_save(mid)
    assert open(mid, "rb").read()[32] == 1
    fresh = ctx(24, "fp64")
    with pytest.raises(efvm.EfvmError) as e:
        fresh.checkpoint_load(mid)
    assert e.value.status == 4                      # EFVM_ERR_UNSUPPORTED
    open(str(tmp_path / "bad.bin"), "wb").write(b"NOTACKPT" + bytes(40))
    with pytest.raises(efvm.EfvmError):
        fresh.checkpoint_load(str(tmp_path / "bad.bin"))
