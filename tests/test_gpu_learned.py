"""Learned face reconstruction on the GPU, against the CPU FP64 oracle (oracle/mlp_oracle.py).

* matrix-core inference == the oracle's bf16-quantised forward pass;
* the training kernels' loss and gradient == the oracle's FP64 backward pass
  (itself pinned by finite differences on the CPU), and one Adam step matches;
* training samples generated on the device from a fine state == the oracle's
  box filter, face line averages, features and targets;
* the face states of a learned stage == the oracle's (network + per-face MUSCL fallback);
* a learned run conserves mass, replays bit-identically from a hipGraph, and the
  trained network beats plain MUSCL at the same resolution.
"""
import os

import numpy as np
import pytest

import dev_field_efvm_amd as efvm
from dev_field_efvm_amd import cases, learned
from oracle import mlp_oracle as M
from oracle import orc

pytestmark = pytest.mark.gpu


def random_params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    d = M.unpack(M.init_params(seed))
    d["W3"] = rng.normal(0, scale, d["W3"].shape)
    d["b3"] = rng.normal(0, scale, 4)
    d["L"] = rng.normal(0, scale, d["L"].shape)
    d["bL"] = rng.normal(0, scale, 4)
    d["b1"] = rng.normal(0, 0.2, 64)
    d["b2"] = rng.normal(0, 0.2, 64)
    return M.pack(d).astype(np.float32)


def random_features(n, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (n, 20)).astype(np.float32)
    X[:, 8:12] = rng.uniform(0, 1, (n, 4))
    return X


def test_infer_matches_quantised_oracle():
    p = random_params(3)
    net = efvm.Mlp(p)
    X = random_features(5000, 1)                 # not a multiple of 64 or 256
    got = net.infer(X)
    ref = M.forward(p.astype(np.float64), X.astype(np.float64), quantized=True)
    full = M.forward(p.astype(np.float64), X.astype(np.float64), quantized=False)
    err = np.abs(got - ref)
    print(f"infer vs quantised oracle: max {err.max():.2e} mean {err.mean():.2e}; "
          f"bf16 effect (quantised vs FP64 network) mean {np.abs(ref - full).mean():.2e}")
    # FP32 accumulation order and the exp2/rcp GELU only (an activation that rounds to the
    # other bf16 neighbour moves the output by one bf16 step of one hidden unit)
    assert np.mean(err) < 2e-4
    assert np.quantile(err, 0.99) < 2e-3
    assert np.max(err) < 2e-2


def test_gradient_matches_oracle():
    p = random_params(4)
    net = efvm.Mlp(p)
    rng = np.random.default_rng(2)
    X = random_features(300, 2)                   # 5 blocks, the last one partial
    Y = rng.normal(0, 0.5, (300, 4)).astype(np.float32)
    loss, g = net.gradient(X, Y)
    lref, gref = M.loss_and_grad(p.astype(np.float64), X.astype(np.float64), Y.astype(np.float64))
    assert abs(loss - lref) <= 1e-5 * lref
    rel = np.abs(g - gref) / (np.abs(gref) + 1e-3 * np.abs(gref).max())
    print(f"gradient vs FP64 oracle: max rel {rel.max():.2e}")
    assert rel.max() < 1e-3


def test_linear_fit_matches_oracle():
    p = random_params(8)
    net = efvm.Mlp(p)
    rng = np.random.default_rng(9)
    X = random_features(3000, 9)
    X[:, 8:12] = rng.uniform(0.0, 0.2, (3000, 4)).astype(np.float32)
    Y = rng.normal(0, 0.3, (3000, 4)).astype(np.float32)
    net.add_samples(X, Y)
    sl = slice(M.OFF_L, M.NPARAM)
    for constrained in (False, True):
        net.load(p)
        net.fit_linear(stride=1, ridge=1e-9, constrained=constrained)
        got = net.save()
        ref = M.fit_linear(p.astype(np.float64), X.astype(np.float64), Y.astype(np.float64), ridge=1e-9,
                           constrained=constrained)
        err = np.max(np.abs(got[sl] - ref[sl]))
        print(f"linear fit (constrained={constrained}) vs oracle: max abs {err:.2e}")
        assert err < 1e-4 * max(1.0, np.abs(ref[sl]).max())
        assert np.array_equal(got[:M.OFF_L], p[:M.OFF_L])      # only the skip is fitted


def _mix64(z):
    m = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def test_adam_steps_match_oracle():
    p = random_params(5)
    net = efvm.Mlp(p)
    rng = np.random.default_rng(3)
    n = 1000
    X = random_features(n, 3)
    Y = rng.normal(0, 0.5, (n, 4)).astype(np.float32)
    net.add_samples(X, Y)
    batch, seed, lr = 256, 77, 1e-3
    losses = net.train(2, batch, lr, lr, seed)
    q = p.astype(np.float64)
    m = np.zeros_like(q)
    v = np.zeros_like(q)
    for t in (1, 2):
        idx = [_mix64(seed ^ _mix64((t * 0x100000001B3 + i) & ((1 << 64) - 1))) % n for i in range(batch)]
        l, g = M.loss_and_grad(q, X[idx].astype(np.float64), Y[idx].astype(np.float64))
        assert abs(losses[t - 1] - l) <= 1e-5 * l
        q, m, v = M.adam(q, g, m, v, t, lr)
    got = net.save()
    assert np.max(np.abs(got - q)) < 2e-6


def test_device_samples_match_oracle():
    n, f = 128, 4
    st = cases.learned_state(11, n)
    c = learned.periodic_context(n, "fp64")
    c.upload(st)
    c.step(3)
    U = c.download()
    net = efvm.Mlp(M.init_params(0).astype(np.float32))
    net.add_snapshot(c, f)
    assert net.num_samples() == 4 * (n // f) ** 2
    X, Y = net.get_samples(0, net.num_samples())
    Uo = np.stack([U[..., 0], U[..., 1], U[..., 2], U[..., 4]])
    cells, faces = M.coarse_grain(Uo, f)
    Xr, Yr = M.samples(cells, faces)
    assert np.max(np.abs(X - Xr)) < 1e-5
    assert np.max(np.abs(Y - Yr)) < 1e-4


def _oracle_faces(p, U, bypass=False):
    out, fb = M.face_states(p.astype(np.float64), U, quantized=True, bypass=bypass)
    return out, fb


def own4(p):
    """The correction restricted to each variable's own 4 neighbour cells (the cubic-exact
    fit's structure, which the face kernels evaluate as 4 taps instead of 20)."""
    p = p.copy()
    L = p[M.OFF_L:M.OFF_L + 80].reshape(4, 20)
    for v in range(4):
        for f in range(20):
            if not (f % 4 == v and f // 4 != 2):
                L[v, f] = 0.0
    return p


@pytest.mark.parametrize("bypass,sparse", [(False, False), (True, False), (False, True), (True, True)])
def test_learned_face_states_match_oracle(bypass, sparse):
    n = 96
    st = cases.learned_state(12, n)
    p = random_params(6, scale=0.1)               # rough network: a mix of accepted and fallen-back faces
    if sparse:
        p = own4(p)
    net = efvm.Mlp(p)
    c = learned.periodic_context(n, "fp32")
    c.upload(st)
    c.set_learned(net, bypass=bypass)
    F = c.learned_faces()                         # (axis, side, var, ny, nx)
    U = c.download()
    Uo = np.stack([U[..., 0], U[..., 1], U[..., 2], U[..., 4]]).astype(np.float64)
    out, fb = _oracle_faces(p, Uo, bypass)
    for axis in (0, 1):
        agree = ~fb[axis]
        for side, key in ((0, "L"), (1, "R")):
            ref = out[(axis, key)]
            got = F[axis, side].astype(np.float64)
            err = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
            # faces where GPU and oracle take the same branch agree to FP32 / bf16-boundary level
            frac_close = np.mean(err < 2e-4)
            print(f"axis {axis} side {side}: fallback frac {fb[axis].mean():.3f}, close {frac_close:.4f}, "
                  f"median err {np.median(err):.1e}")
            assert frac_close > 0.97
            assert np.median(err) < 1e-5
        assert 0.02 < fb[axis].mean() < 0.98
    # fallback faces are the solver's MUSCL exactly as in the oracle
    fbs, tot, nnet = c.learned_stats()
    assert tot == 0                               # learned_faces() does not count


@pytest.mark.parametrize("bypass,sparse", [(False, False), (True, False), (True, True)])
def test_learned_stage_fluxes_match_oracle(bypass, sparse):
    """One learned substage (forward-Euler stage of RK3, fixed dt) against the oracle's FP64
    Riemann fluxes of the face states the diagnostic face kernel returns: pins the stage
    kernels (fluxes, divergence, update) and, with bypass, that the light pass plus the
    deferred-face pass build the same face states as the single-pass kernel.  With the
    4-tap correction the light pass evaluates the open-gate value as a linear stencil (equal
    up to rounding), so a face at the deferral bound may be decided differently: a few
    cells are allowed to differ there."""
    n = 48
    st = cases.learned_state(5, n)
    p = random_params(6, scale=0.1)
    if sparse:
        p = own4(p)
        p[M.OFF_BL:M.OFF_BL + 4] = 0.0
    net = efvm.Mlp(p)
    c = learned.periodic_context(n, "fp32")
    c.upload(st)
    c.set_learned(net, bypass=bypass)
    F = c.learned_faces().astype(np.float64)      # (axis, side, var, ny, nx) primitives
    U = c.download().astype(np.float64)           # (ny, nx, 5)
    dt = 0.5 * c.cfl_dt()
    c.set_time_step(0.4, dt)
    c.run()                                       # stage 1 of RK3: U1 = U + dt L(U)
    U1 = c.download().astype(np.float64)
    h = 10.0 / n
    flux = np.zeros((2, 5, n, n))
    for axis in (0, 1):
        for j in range(n):
            for i in range(n):
                L = orc.prim_to_cons(np.array([F[axis, 0, 0, j, i], F[axis, 0, 1, j, i], F[axis, 0, 2, j, i], 0.0,
                                               F[axis, 0, 3, j, i], 1.0]), 1.4)
                R = orc.prim_to_cons(np.array([F[axis, 1, 0, j, i], F[axis, 1, 1, j, i], F[axis, 1, 2, j, i], 0.0,
                                               F[axis, 1, 3, j, i], 1.0]), 1.4)
                flux[axis, :, j, i] = orc.flux("hllc", L, R, 1.4, axis)
    fx, fy = np.moveaxis(flux[0], 0, -1), np.moveaxis(flux[1], 0, -1)
    div = (np.roll(fx, 1, axis=1) - fx) / h + (np.roll(fy, 1, axis=0) - fy) / h
    ref = U + dt * div
    cell_err = np.abs(U1 - ref)[..., [0, 1, 2, 4]].max(axis=-1) / np.abs(U).max()
    fb, tot, nnet = c.learned_stats()
    print("learned stage vs oracle fluxes: max rel err", cell_err.max(), "cells above 1e-5",
          int((cell_err >= 1e-5).sum()), "faces", tot, "gated sides", nnet)
    assert tot == 2 * n * n
    assert 0 < nnet < 2 * tot if bypass else nnet == 2 * tot
    if sparse:
        assert np.mean(cell_err < 1e-5) > 0.995
        assert cell_err.max() < 1e-2
    else:
        assert cell_err.max() < 1e-5


def trained_net():
    if not os.path.exists(learned.WEIGHTS):
        pytest.skip("no trained weights: run python -m dev_field_efvm_amd.learned train on the GPU")
    return efvm.Mlp(learned.load_weights())


def test_learned_run_conserves_and_replays_from_graph():
    net = trained_net()
    n = 128
    st = cases.learned_state(21, n)
    eager = learned.periodic_context(n)
    graph = learned.periodic_context(n)
    for c in (eager, graph):
        c.upload(st)
        c.set_learned(net)
    m0 = eager.download()[..., 0].sum(dtype=np.float64)
    graph.graph_capture(4)
    for rep in range(3):
        eager.step(4)
        graph.graph_launch()
        graph.synchronize()
        assert np.array_equal(eager.download(), graph.download()), rep
    m1 = eager.download()[..., 0].sum(dtype=np.float64)
    assert abs(m1 - m0) / m0 < 1e-6
    fb, tot, nnet = eager.learned_stats()
    assert tot == 12 * 3 * 2 * n * n               # 12 steps x 3 stages x 2 faces per cell
    assert nnet == 2 * tot                         # without bypass the network gates every face side
    assert fb < 0.05 * tot


def rough_own4(seed=6, scale=0.1):
    """A rough 4-tap network (no correction bias): many faces deferred to the network."""
    p = own4(random_params(seed, scale=scale))
    p[M.OFF_BL:M.OFF_BL + 4] = 0.0
    return p


def _bypass_pair(nx, ny, p, seed=7, fixed_dt=None):
    """Two contexts on one state and one 4-tap network in bypass mode: the fused stage
    kernel and the separate passes."""
    n = max(nx, ny)
    st = np.ascontiguousarray(cases.learned_state(seed, n, sub=1)[:ny, :nx])
    out = []
    for fused in (True, False):
        c = efvm.Context(nx, ny, hi=(10.0 * nx / n, 10.0 * ny / n), dz=10.0 / n, gamma=1.4, flux="hllc", muscl=True,
                         limiter="mc", precision="fp32", cfl=0.4, seam_mode=1, fixed_dt=fixed_dt)
        c.set_boundaries(["periodic"] * 4)
        c.upload(st)
        c.set_learned(efvm.Mlp(p), bypass=True)
        c.set_learned_fused(fused)
        assert c.learned_fused() == (fused, fused)
        out.append(c)
    return out


def _same(a, b, what):
    ua, ub = a.download(), b.download()
    assert np.array_equal(ua, ub), f"{what}: {np.count_nonzero(ua != ub)} values differ, max {np.abs(ua - ub).max():.3e}"
    ca, cb = a.clock(), b.clock()
    assert ca == cb, (what, ca, cb)


# (nx, ny): a tile narrower than the wave (columns wrap inside it), three tiles across with a
# ragged last one, a grid smaller than the stencil window in both axes (several periodic images
# per wave and per window); last tiles of one, two and three rows
@pytest.mark.parametrize("nx,ny", [(48, 48), (130, 61), (7, 9), (64, 41), (61, 130)])
@pytest.mark.parametrize("net", ["trained", "rough"])
def test_fused_bypass_stage_is_bitwise_the_separate_passes(nx, ny, net):
    """The fused bypass stage (faces, fluxes and update in one kernel, the cells next to a
    deferred face finished by the pending-cell pass) against the separate passes (light pass,
    deferred-face pass, flux-plane stage): states, clocks (the fused CFL) and face statistics
    bitwise equal after every substage of two RK3 steps and after two more whole steps.  The
    rough network defers many faces, so cells with one to four deferred faces, and deferred
    faces between two pending cells, all occur."""
    if net == "trained":
        if not os.path.exists(learned.WEIGHTS):
            pytest.skip("no trained weights")
        p = learned.load_weights()
    else:
        p = rough_own4()
    fz, sep = _bypass_pair(nx, ny, p)
    for k in range(6):
        fz.run()
        sep.run()
        _same(fz, sep, f"substage {k}")
    fz.step(2)
    sep.step(2)
    _same(fz, sep, "two steps")
    sf, ss = fz.learned_stats(), sep.learned_stats()
    print(f"{nx}x{ny} {net}: stats fused {sf} separate {ss}")
    assert sf == ss
    assert sf[1] == 4 * 3 * 2 * nx * ny            # every face once per substage
    if net == "rough":
        assert sf[2] > 0                           # faces were deferred to the network


def test_fused_bypass_stage_grid_stride_and_long_lists():
    """A grid whose tiles outnumber the fused kernel's workgroups (1024 x 1000: 6138 tiles on at
    most 2048 workgroups, each striding over several), with the rough network so that the
    pending-cell list holds thousands of cells: bitwise the separate passes."""
    fz, sep = _bypass_pair(1024, 1000, rough_own4(seed=4, scale=0.12), seed=13)
    for k in range(3):
        fz.run()
        sep.run()
        _same(fz, sep, f"substage {k}")
    sf, ss = fz.learned_stats(), sep.learned_stats()
    print("1024x1000 rough: stats", sf)
    assert sf == ss and sf[2] > 1000


def test_fused_bypass_stage_graph_replay_and_fixed_dt():
    """A graph of fused bypass steps replays bitwise the separate passes run eagerly, with the
    CFL dt and with a fixed dt (the fused CFL compiled out of the stage)."""
    p = rough_own4(seed=9, scale=0.08)
    for fixed in (None, 0.004):
        fz, sep = _bypass_pair(96, 80, p, seed=11, fixed_dt=fixed)
        fz.graph_capture(2)
        for rep in range(3):
            fz.graph_launch()
            fz.synchronize()
            sep.step(2)
            _same(fz, sep, f"replay {rep}, fixed dt {fixed}")
        assert fz.learned_stats() == sep.learned_stats()


def test_trained_network_beats_muscl():
    """What the learned reconstruction as a whole (the cubic-exact linear correction and the
    network's gate) gains over plain MUSCL: learned 128^2 vs MUSCL 128^2, both against MUSCL
    1024^2 box-filtered to 128^2, on a seed outside the training set.  The gate's own share is
    test_trained_gate_against_the_gate_fixed_open."""
    net = trained_net()
    out = learned.evaluate(net, seed=201, n=128, fine=512, truth_n=1024, T=0.5)
    print({k: out[k] for k in ("learned", "muscl", "muscl_fine", "l1_gain_vs_muscl")})
    assert out["learned"]["fallback_frac"] < 0.05
    assert out["l1_gain_vs_muscl"] > 2.0


def test_trained_gate_against_the_gate_fixed_open():
    """What the network itself gains over the method without it: the same learned correction with
    the network's gate fixed at 1 (output weights zero, output bias 40: sigmoid = 1 in FP32),
    both in network mode (no bypass), on a held-out seed at 256^2, density L1 at t = 1 against an
    FP64 MUSCL run 8x finer box-filtered to 256^2."""
    p = learned.load_weights()
    nin, nh, nout = learned.NIN, learned.NH, learned.NOUT
    o3 = nh * nin + nh + nh * nh + nh                  # W3 (nout x nh), then b3 (nout)
    open_gate = p.copy()
    open_gate[o3:o3 + nout * nh] = 0.0
    open_gate[o3 + nout * nh:o3 + nout * nh + nout] = 40.0
    seed, n, T = 103, 256, 1.0
    truth = learned.run_case(seed, 8 * n, T, precision="fp64", timed=False)
    ref = learned.box(truth["rho"], 8)
    errs = {}
    for name, params in (("trained gate", p), ("gate fixed at 1", open_gate)):
        r = learned.run_case(seed, n, T, efvm.Mlp(params), timed=False)
        errs[name] = float(np.mean(np.abs(r["rho"] - ref)))
        if name == "gate fixed at 1":
            assert r["gated_sides"] == 2 * r["faces"]     # the network still ran on every face side
    gain = errs["gate fixed at 1"] / errs["trained gate"]
    print(f"held-out seed {seed}, {n}^2: L1 {errs}, the network's gain x{gain:.3f}")
    assert errs["trained gate"] < errs["gate fixed at 1"]
