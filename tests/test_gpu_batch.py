"""Batched ensembles: one persistent launch per RK stage over many independent cases.

Each case of a batch must equal the same case run alone through a single context
(same FP32 arithmetic, its own CFL-limited dt), bit for bit, and a sample is held
against the FP64 oracle to the FP32 tolerance.
"""
import numpy as np
import pytest

import dev_field_efvm_amd as efvm
from dev_field_efvm_amd import cases
from dev_field_efvm_amd.autotune import tune_batch
from oracle import orc

pytestmark = pytest.mark.gpu


def small_ensemble(n_cases=10, seed=7):
    specs, builders = cases.ensemble(n_cases=n_cases, seed=seed, nmin=40, nmax=200)
    return specs, [b() for b in builders]


def single_run(spec, st, steps):
    c = efvm.Context(spec["nx"], spec["ny"], hi=spec["hi"], dz=spec["hi"][0] / spec["nx"], gamma=spec["gamma"],
                     muscl=True, limiter="mc", precision="fp32")
    c.set_boundaries(spec["bcs"])
    c.upload(st)
    c.step(steps)
    return c.download(), c.clock()


def rel_linf(g, r):
    return max(float(np.max(np.abs(g[..., m] - r[..., m])) / np.max(np.abs(r[..., m]))) for m in (0, 1, 2, 4))


def test_batch_equals_single_contexts_bitwise():
    specs, states = small_ensemble()
    b = efvm.Batch(specs)
    for i, st in enumerate(states):
        b.upload(i, st)
    b.step(6)
    for i, (sp, st) in enumerate(zip(specs, states)):
        ref, clk = single_run(sp, st, 6)
        got = b.download(i)
        assert np.array_equal(got, ref), (i, sp["kind"], sp["nx"])
        bc = b.clock(i)
        assert bc["time"] == clk["time"] and bc["step"] == 6


def test_batch_two_column_kernel_is_bitwise_the_one_column_kernel():
    """The batch runs the cases the general two-column kernel takes (here every case: periodic
    vortices with even and odd sides, outflow Riemann problems) on it; forbidding it sends every
    case to the one-column kernel, with bitwise the same states and clocks."""
    specs, states = small_ensemble(n_cases=12, seed=5)
    assert any(sp["nx"] % 2 for sp in specs) and any(sp["nx"] % 2 == 0 for sp in specs)
    out = []
    for allow in (True, False):
        b = efvm.Batch(specs)
        b.set_pair_kernel(allow)
        for i, st in enumerate(states):
            b.upload(i, st)
        b.step(7)
        out.append(([b.download(i) for i in range(len(specs))], [b.clock(i) for i in range(len(specs))]))
    for i in range(len(specs)):
        assert np.array_equal(out[0][0][i], out[1][0][i]), (i, specs[i]["kind"], specs[i]["nx"])
        assert out[0][1][i] == out[1][1][i]


def test_batch_fused_stages_are_bitwise_the_separate_launches():
    """The periodic two-column cases (nx even, ny >= 8) run stages 2 and 3 as one fused launch,
    their U^n alternating between two plane sets by step parity.  Against fusion off: bitwise the
    same states and clocks after odd and even step counts, at several fused strip heights, with
    the setting switched mid-run at odd parity (the state moves to the set the new mode reads),
    through a snapshot save and restore, and from a graph replay."""
    specs, states = small_ensemble(n_cases=12, seed=5)

    def run(allow, rows, steps, switch_at=None):
        b = efvm.Batch(specs)
        b.set_fused(allow, rows)
        for i, st in enumerate(states):
            b.upload(i, st)
        if switch_at is None:
            b.step(steps)
        else:
            b.step(switch_at)
            b.set_fused(not allow, rows)
            b.step(steps - switch_at)
        return b, [b.download(i) for i in range(len(specs))], [b.clock(i) for i in range(len(specs))]

    for steps in (7, 8):
        b0, ref, ck = run(False, 0, steps)
        assert b0.fused()[2] == 0
        for rows, switch in ((37, None), (5, None), (13, 3), (57, None)):
            b, got, gk = run(True, rows, steps, switch)
            if switch is None:
                assert b.fused()[2] > 0
            for i in range(len(specs)):
                assert np.array_equal(got[i], ref[i]), (steps, rows, i, specs[i]["kind"], specs[i]["nx"], specs[i]["ny"])
                assert gk[i] == ck[i]
    # snapshot at odd parity, more steps, restore, the same steps again
    b = efvm.Batch(specs)
    for i, st in enumerate(states):
        b.upload(i, st)
    b.step(3)
    b.snapshot()
    b.step(4)
    first = [b.download(i) for i in range(len(specs))]
    b.snapshot(restore=True)
    b.step(4)
    for i in range(len(specs)):
        assert np.array_equal(b.download(i), first[i]), i
    # graph replay from odd parity
    g = efvm.Batch(specs)
    for i, st in enumerate(states):
        g.upload(i, st)
    g.step(3)
    g.graph_capture(2)
    g.graph_launch()
    g.graph_launch()
    g.synchronize()
    for i in range(len(specs)):
        assert np.array_equal(g.download(i), first[i]), i


def test_batch_per_case_dt_differs():
    specs, states = small_ensemble()
    b = efvm.Batch(specs)
    for i, st in enumerate(states):
        b.upload(i, st)
    b.step(2)
    dts = {round(b.clock(i)["dt"], 15) for i in range(len(specs))}
    assert len(dts) > len(specs) // 2


def test_batch_sample_against_fp64_oracle():
    specs, states = small_ensemble()
    b = efvm.Batch(specs)
    for i, st in enumerate(states):
        b.upload(i, st)
    b.step(5)
    for i in (0, 1):   # one vortex, one four-quadrant Riemann problem
        sp = specs[i]
        bcs = list(sp["bcs"]) + ["slip", "slip"]
        sim = orc.Sim(sp["nx"], sp["ny"], 1, hi=(sp["hi"][0], sp["hi"][1], sp["hi"][0] / sp["nx"]), gamma=sp["gamma"],
                      flux="hllc", muscl=True, limiter="mc", cfl=0.4, bcs=bcs)
        sim.set_interior(states[i].reshape(1, sp["ny"], sp["nx"], 5))
        sim.step(5)
        err = rel_linf(b.download(i), sim.interior()[0])
        print(f"case {i} ({sp['kind']}, {sp['nx']}^2): FP32 batch vs FP64 oracle rel L-inf {err:.2e}")
        # discontinuous Riemann data: the FP32/FP64 difference concentrates at the waves
        assert err <= (2e-3 if sp["kind"] == "riemann4" else 1e-4)


def test_batch_graph_replay_equals_eager_and_survives_retuning():
    specs, states = small_ensemble(n_cases=8, seed=11)
    eager, graph = efvm.Batch(specs), efvm.Batch(specs)
    for i, st in enumerate(states):
        eager.upload(i, st)
        graph.upload(i, st)
    sides = [s["nx"] for s in specs]
    tune_batch(graph, sides, steps=1, rows=(8, 32), blocks=(0, 64), use_cache=False, store=False)
    graph.graph_capture(2)
    for rep in range(3):
        eager.step(2)
        graph.graph_launch()
        graph.synchronize()
        for i in range(len(specs)):
            assert np.array_equal(eager.download(i), graph.download(i)), (rep, i)


def test_batch_rejects_bad_configs():
    spec = dict(nx=32, ny=32, hi=(1.0, 1.0), gamma=1.4, bcs=["periodic", "outflow", "outflow", "outflow"])
    with pytest.raises(efvm.EfvmError) as e:
        efvm.Batch([spec])
    assert e.value.status == 2


def test_batch_graph_replay_after_upload_and_odd_steps():
    """The batch graph replays from the current state: an upload is followed by a fresh CFL scan
    of every case before the replay; an odd number of eager steps (another CFL slot) is refused."""
    specs, states = small_ensemble(n_cases=6, seed=13)
    states2 = [1.1 * st for st in states]          # denser, same velocity, 10% higher pressure
    eager, graph = efvm.Batch(specs), efvm.Batch(specs)
    for i, st in enumerate(states):
        eager.upload(i, st)
        graph.upload(i, st)
    graph.graph_capture(2)
    graph.graph_launch()
    eager.step(2)
    for i, st in enumerate(states2):
        eager.upload(i, st)
        graph.upload(i, st)
    graph.graph_launch()
    eager.step(2)
    for i in range(len(specs)):
        assert np.array_equal(eager.download(i), graph.download(i)), i
        assert eager.clock(i)["dt"] == graph.clock(i)["dt"]
    graph.step(1)
    with pytest.raises(efvm.EfvmError):
        graph.graph_launch()
