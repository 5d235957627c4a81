"""CPU checks that pin the learned-reconstruction oracle (oracle/mlp_oracle.py).

The reference has no learned part, so the oracle is pinned by what the
mathematics fixes: finite differences for the backward pass, closed forms for
the features, the targets and Adam's first step, mirror symmetry for the face
orientation, and hand-built stencils on both sides of each fallback bound.
"""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as M


def random_params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    p = M.init_params(seed)
    d = M.unpack(p)
    d["W3"] = rng.normal(0, scale, d["W3"].shape)
    d["b3"] = rng.normal(0, scale, 4)
    d["L"] = rng.normal(0, scale, d["L"].shape)
    d["bL"] = rng.normal(0, scale, 4)
    d["b1"] = rng.normal(0, 0.2, 64)
    d["b2"] = rng.normal(0, 0.2, 64)
    return M.pack(d)


def test_layout_size():
    assert M.NPARAM == 5848


def test_gelu_matches_erf_form_and_its_derivative():
    x = np.linspace(-6, 6, 241)
    exact = 0.5 * x * (1 + np.vectorize(math.erf)(x / math.sqrt(2)))
    assert np.max(np.abs(M.gelu(x) - exact)) < 1e-3           # tanh form vs erf form
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_grad(x))) < 1e-8


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = random_params(1)
    X = rng.uniform(-1, 1, (37, 20))
    Y = rng.normal(0, 0.5, (37, 4))
    loss, g = M.loss_and_grad(p, X, Y)
    assert np.isclose(loss, np.mean(M.sample_weight(X) * (M.forward(p, X) - Y) ** 2))
    idx = rng.choice(M.NPARAM, 300, replace=False)
    idx = np.concatenate([idx, [M.OFF_B1, M.OFF_B2 + 5, M.OFF_B3 + 2, M.OFF_BL + 1, M.OFF_L + 7, M.OFF_W3 + 11]])
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAM)
        e[i] = h
        fd = (M.loss_and_grad(p + e, X, Y)[0] - M.loss_and_grad(p - e, X, Y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_sample_weight_is_the_squared_face_value_scale():
    # x0 = sigma / (sigma + X0C scale): sigma / scale = 0.002 -> x0 = 0.002 / 0.102, weight (0.002 / 0.01)^2
    x0 = 0.002 / (0.002 + M.X0C)
    X = np.zeros((1, 20))
    X[0, 8:12] = [x0, 0.0, 0.5, 0.999]
    w = M.sample_weight(X)[0]
    assert np.allclose(w, [0.04, 0.0, 100.0, 100.0])


def test_linear_fit_solves_the_weighted_normal_equations():
    rng = np.random.default_rng(7)
    X = rng.uniform(-1, 1, (400, 20))
    X[:, 8:12] = rng.uniform(0.0, 0.2, (400, 4))
    Y = rng.normal(0, 0.3, (400, 4))
    d = M.unpack(random_params(3))
    d["W3"][:] = 0.0
    d["b3"][:] = 40.0                                    # gates fully open: phi = 1
    p = M.fit_linear(M.pack(d), X, Y, ridge=0.0)
    loss, g = M.loss_and_grad(p, X, Y)
    gl = np.concatenate([g[M.OFF_L:M.OFF_BL], g[M.OFF_BL:]])
    assert np.max(np.abs(gl)) < 1e-12 * max(1.0, loss) + 1e-14
    # and it beats the linear map it started from
    assert loss < M.loss_and_grad(M.pack(d), X, Y)[0]


def test_constrained_linear_fit_is_exact_for_cubic_profiles():
    rng = np.random.default_rng(8)
    X = rng.uniform(-1, 1, (300, 20))
    X[:, 8:12] = rng.uniform(0.0, 0.2, (300, 4))
    Y = rng.normal(0, 0.3, (300, 4))
    d = M.unpack(M.init_params(0))
    d["b3"][:] = 40.0                                    # gates fully open
    p = M.fit_linear(M.pack(d), X, Y, constrained=True)
    # cell averages of a cubic on unit cells centred at -2..2: the face value at +1/2 is exact
    coef = np.array([0.3, -1.1, 0.7, 0.9])               # c0 + c1 x + c2 x^2 + c3 x^3
    def cavg(k):
        P = np.polynomial.Polynomial(coef).integ()
        return P(k + 0.5) - P(k - 0.5)
    St = np.zeros((5, 4, 1))
    for k in range(5):
        St[k, :, 0] = 2.0 + 0.1 * cavg(k - 2)             # same shape on every variable, rho, p > 0
    Xs, sig, c = M.features(St)
    y = M.forward(p, Xs.reshape(1, 20))[0]
    face = c[0] + sig[0] * y
    exact = 2.0 + 0.1 * np.polynomial.Polynomial(coef)(0.5)
    assert np.allclose(face, exact, atol=1e-12)


def test_bf16_rounding_is_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7                                             # bf16 spacing at 1
    assert M.bf16(one + ulp / 2) == one                         # tie -> even (1.0)
    assert M.bf16(one + 3 * ulp / 2) == one + 2 * ulp           # tie -> even (1 + 2 ulp)
    assert M.bf16(one + 0.51 * ulp) == one + ulp
    assert M.bf16(-3.140625) == -3.140625                       # exactly representable


def test_adam_first_step_is_signed_lr():
    p = np.zeros(5)
    g = np.array([1e-3, -2.0, 5.0, -1e-4, 3.0])
    q, m, v = M.adam(p, g, np.zeros(5), np.zeros(5), 1, 0.01)
    assert np.allclose(q, -0.01 * np.sign(g), rtol=1e-4)


def smooth_field(n, seed=3, L=10.0):
    x = (np.arange(n) + 0.5) * L / n
    X, Yg = np.meshgrid(x, x)
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * np.pi, 8)
    rho = 1.0 + 0.1 * np.sin(2 * np.pi * X / L + ph[0]) * np.cos(2 * np.pi * 2 * Yg / L + ph[1])
    u = 0.5 + 0.2 * np.sin(2 * np.pi * (X + Yg) / L + ph[2])
    v = -0.3 + 0.1 * np.cos(2 * np.pi * 3 * X / L + ph[3])
    p = 1.0 + 0.05 * np.sin(2 * np.pi * Yg / L + ph[4])
    return np.stack([rho, u, v, p])


def test_minus_side_is_the_mirror_image_of_plus_side():
    W = smooth_field(32)
    St_minus = M.stencil(W, 0, -1)
    Wm = W[:, :, ::-1].copy()
    Wm[1] = -Wm[1]                                               # mirror x -> -x
    St_plus_mirror = M.stencil(Wm, 0, 1)[:, :, :, ::-1]
    assert np.allclose(St_minus, St_plus_mirror, atol=0, rtol=0)
    # y axis: u_n = v, u_t = u
    Sy = M.stencil(W, 1, 1)
    assert np.array_equal(Sy[2, 1], W[2]) and np.array_equal(Sy[2, 2], W[1])
    assert np.array_equal(Sy[3, 0], np.roll(W[0], -1, axis=0))


def test_features_of_a_linear_profile():
    # w_k = w0 + k s: sigma = 2|s|, x_k = k/2 sign(s) for k != 0
    St = np.zeros((5, 4, 1, 1))
    s = np.array([0.01, -0.02, 0.03, 0.04])
    base = np.array([1.0, 0.2, -0.1, 2.0])
    for k in range(5):
        St[k, :, 0, 0] = base + (k - 2) * s
    X, sig, c = M.features(St)
    X = X[0, 0]
    for k in range(5):
        if k == 2:
            continue
        assert np.allclose(X[4 * k:4 * k + 4], (k - 2) / 2 * np.sign(s))
    assert np.allclose(sig[0, 0], 2 * np.abs(s))
    scale = np.array([1.0, math.sqrt(1.4 * 2.0), math.sqrt(1.4 * 2.0), 2.0])
    assert np.allclose(X[8:12], sig[0, 0] / (sig[0, 0] + M.X0C * scale))
    # the central base is then exactly the face value of a linear profile: w0 + sigma * y = w0 + s / 2
    assert np.allclose(M.base(X[None])[0] * sig[0, 0], 0.5 * s)


def test_face_targets_converge_at_fourth_order():
    # f(x, y) = sin(2 pi x / L) (1 + 0.3 cos(2 pi y / L)); the east-face line average is known in closed form
    L = 10.0
    errs = []
    for nf in (64, 128):
        f = 4
        h = L / nf
        edges = np.arange(nf + 1) * h
        # exact cell averages
        cx = (np.cos(2 * np.pi * edges[:-1] / L) - np.cos(2 * np.pi * edges[1:] / L)) * L / (2 * np.pi) / h
        cy = 1 + 0.3 * (np.sin(2 * np.pi * edges[1:] / L) - np.sin(2 * np.pi * edges[:-1] / L)) * L / (2 * np.pi) / h
        U = np.ones((4, nf, nf))
        U[1] = np.outer(cy, cx)                                   # (ny, nx)
        cells, faces = M.coarse_grain(U, f)
        H = f * h
        ce = np.arange(nf // f + 1) * H
        xf = ce[1:]                                               # east faces of the coarse cells
        ly = 1 + 0.3 * (np.sin(2 * np.pi * ce[1:] / L) - np.sin(2 * np.pi * ce[:-1] / L)) * L / (2 * np.pi) / H
        exact = np.outer(ly, np.sin(2 * np.pi * xf / L))
        errs.append(np.max(np.abs(faces[0, 1] - exact)))
        assert np.allclose(cells[1], U[1].reshape(nf // f, f, nf // f, f).mean(axis=(1, 3)))
    order = math.log2(errs[0] / errs[1])
    assert order > 3.7, errs


def test_samples_targets_are_exact_face_offsets():
    U = M.cons(smooth_field(32))
    cells, faces = M.coarse_grain(U, 2)
    X, Y = M.samples(cells, faces)
    n = 16 * 16
    assert X.shape == (4 * n, 20) and Y.shape == (4 * n, 4)
    W = M.prim(cells)
    # sample 0 block: axis 0, side +: target rho offset * sigma + centre == face rho
    Xs, sig, c = M.features(M.stencil(W, 0, 1))
    rho_face = M.prim(faces[0])[0]
    assert np.allclose(c[..., 0] + sig[..., 0] * Y[:n, 0].reshape(16, 16), rho_face)
    # side - of axis 1 targets the north face of the cell below
    Xs, sig, c = M.features(M.stencil(W, 1, -1))
    rho_s = np.roll(M.prim(faces[1])[0], 1, axis=0)
    assert np.allclose(c[..., 0] + sig[..., 0] * Y[3 * n:, 0].reshape(16, 16), rho_s)


def test_side_ok_bounds():
    St = np.zeros((5, 4))
    for k in range(5):
        St[k] = [1.0 + 0.1 * k, 0.1 * k, 0.0, 1.0 + 0.1 * k]     # monotone: w0 = 1.2, w1 = 1.3
    St = St[:, :, None]
    face = St[2].copy()
    face[:, 0] = [1.25, 0.25, 0.0, 1.25]
    assert M.side_ok(St, face)[0]
    face[0, 0] = 1.3 + 1e-3                                       # curvature 0 -> strict bound (+ floor)
    assert not M.side_ok(St, face)[0]
    face[0, 0] = 1.25
    face[3, 0] = -0.1                                             # pressure positivity
    assert not M.side_ok(St, face)[0]
    # smooth maximum: curvatures both negative, overshoot allowed up to |curv|/2
    Sm = np.zeros((5, 4, 1))
    vals = np.array([0.0, 0.75, 1.0, 0.75, 0.0])                 # -(x^2)/4 shape, curvature -0.5
    for v in range(4):
        Sm[:, v, 0] = 2.0 + vals
    f = Sm[2].copy()
    f[:, 0] = 3.0 + 0.2                                           # max(w0, w1) = 3.0, allowance 0.25
    assert M.side_ok(Sm, f)[0]
    f[:, 0] = 3.0 + 0.26
    assert not M.side_ok(Sm, f)[0]


def test_face_states_fall_back_to_muscl_where_a_side_fails():
    rng = np.random.default_rng(5)
    U = M.cons(smooth_field(24))
    p = random_params(2, scale=0.1)                               # a rough network: many faces fail
    out, fb = M.face_states(p, U, quantized=False)
    assert fb[0].any() and (~fb[0]).any()
    up = np.roll(U, -1, 2)
    um = np.roll(U, 1, 2)
    s = M.mc(up - U, U - um)
    ml = M.prim(U + 0.5 * s)
    mr = np.roll(M.prim(U - 0.5 * s), -1, 2)
    assert np.allclose(out[(0, "L")][:, fb[0]], ml[:, fb[0]])
    assert np.allclose(out[(0, "R")][:, fb[0]], mr[:, fb[0]])
    # a planted mistake in the oracle (right side's stencil not mirrored) changes the
    # faces that pass, so the GPU comparison would catch a mis-oriented kernel
    good = out[(0, "R")][:, ~fb[0]]
    W = M.prim(U)
    bad_St = M.stencil(W, 0, 1)                                   # wrong orientation for the right side
    X, sig, c = M.features(bad_St)
    y = M.forward(p, X.reshape(-1, 20)).reshape(c.shape)
    fo = np.moveaxis(c + sig * y, -1, 0)
    assert not np.allclose(np.roll(fo, -1, 2)[:, ~fb[0]], good)
