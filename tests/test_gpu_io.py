"""Restart files on the GPU: the CFDRST01 checkpoint (examples/restart_determinism/main.rs:177-226)
and the bit-for-bit continuation the reference tests (crates/cfd_solver/tests/restart.rs:86-105)."""
import struct

import numpy as np
import pytest

import dev_field_efvm_amd as efvm
from dev_field_efvm_amd import cases

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


def test_checkpoint_layout(tmp_path):
    n = 24
    c = ctx(n, "fp64")
    c.step(3)
    path = str(tmp_path / "ck.bin")
    c.checkpoint_save(path)
    data = open(path, "rb").read()
    assert data[:8] == b"CFDRST01"
    dt, t, step, stage = struct.unpack_from("<ddQB", data, 8)
    clk = c.clock()
    assert (dt, t, step, stage) == (clk["dt"], clk["time"], 3, 0)
    pos = 8 + 25
    nraw = (n + 4) * (n + 4) * 5
    cols = []
    for _ in range(3):
        (length,) = struct.unpack_from("<Q", data, pos)
        assert length == nraw
        cols.append(np.frombuffer(data, dtype="<f8", count=5 * nraw, offset=pos + 8).reshape(5, n + 4, n + 4, 5))
        pos += 8 + 40 * nraw
    assert pos == len(data)
    u, rhs, u0 = cols
    interior = u[2, 2:n + 2, 2:n + 2]
    assert np.array_equal(interior, c.download())
    ghost = u.copy()
    ghost[2, 2:n + 2, 2:n + 2] = 0.0
    assert not ghost.any()                          # ghosts are written as zero
    assert not rhs.any()
    assert np.array_equal(u0[2, 2:n + 2, 2:n + 2], interior)   # at a step boundary the backup is the state


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
