"""CPU FP64 oracle of the learned face reconstruction (numpy).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this module, and only as the
checker.  The product (``dev_field_efvm_amd``) never imports it.

The reference has no learned part; this module is the specification the HIP
kernels are checked against:

* the network: the fixed central-slope offset plus a learned linear high-order
  correction (20 -> 4) gated per variable by sigmoid(MLP), the MLP 20 -> 64 -> 64 -> 4
  with tanh-form GELU: a learned limiter between the high-order map (gate open)
  and the unlimited central slope (gate closed);
* the stencil features: the 5-cell stencil of primitives (rho, u_n, u_t, p)
  oriented toward the face, normalised per variable by the stencil's largest
  deviation from the centre cell;
* the training targets: line averages of the fine solution over the coarse
  faces, from fourth-order face interpolation of the fine cell averages;
* the per-face fallback to the solver's plain MUSCL (conserved variables, MC
  limiter) when a side fails positivity or the curvature-relaxed TVD bound;
* the loss (sigma^2-weighted mean squared error of the normalised face offsets,
  i.e. the squared face-value error), its gradient, the closed-form weighted
  least-squares fit of the linear skip, and the Adam update.

How each part is pinned without a reference binary: the backward pass by
central finite differences of the forward pass; the features by their
symmetries (mirroring the field maps the + side onto the - side) and by
polynomial fields whose face values are known in closed form; the target
interpolation by exactness on polynomials of degree <= 3; the fallback test
by hand-built stencils on both sides of each bound.  All in FP64.
"""
from __future__ import annotations

import math

import numpy as np

NIN, NH, NOUT = 20, 64, 4
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NOUT * NH
OFF_L = OFF_B3 + NOUT
OFF_BL = OFF_L + NOUT * NIN
NPARAM = OFF_BL + NOUT

FLOOR = 1.0e-4       # sigma floor, relative to the centre cell's scale
X0C = 0.1            # centre feature: sigma / (sigma + X0C * scale)
GELU_C = math.sqrt(2.0 / math.pi)
GELU_A = 0.044715


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(NH, NIN), b1=p[OFF_B1:OFF_W2],
                W2=p[OFF_W2:OFF_B2].reshape(NH, NH), b2=p[OFF_B2:OFF_W3],
                W3=p[OFF_W3:OFF_B3].reshape(NOUT, NH), b3=p[OFF_B3:OFF_L],
                L=p[OFF_L:OFF_BL].reshape(NOUT, NIN), bL=p[OFF_BL:NPARAM])


def pack(d):
    return np.concatenate([d["W1"].ravel(), d["b1"], d["W2"].ravel(), d["b2"], d["W3"].ravel(), d["b3"],
                           d["L"].ravel(), d["bL"]])


B3_OPEN = 8.0        # initial gate bias: sigmoid(8) = 0.99966, the limiter starts open


def init_params(seed=0):
    """He-style init of the hidden layers; output weights zero with the gates open; skip zero."""
    rng = np.random.default_rng(seed)
    d = dict(W1=rng.normal(0, math.sqrt(2.0 / NIN), (NH, NIN)), b1=np.zeros(NH),
             W2=rng.normal(0, math.sqrt(2.0 / NH), (NH, NH)), b2=np.zeros(NH),
             W3=np.zeros((NOUT, NH)), b3=np.full(NOUT, B3_OPEN), L=np.zeros((NOUT, NIN)), bL=np.zeros(NOUT))
    return pack(d)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(GELU_C * (x + GELU_A * x ** 3)))


def gelu_grad(x):
    t = np.tanh(GELU_C * (x + GELU_A * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * x * x)


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def base(X):
    """Central-slope face offset in normalised units: (x_{+1} - x_{-1}) / 4 per variable."""
    return 0.25 * (X[..., 12:16] - X[..., 4:8])


def forward(p, X, quantized=False):
    """y = base(X) + phi * (L X + bL),  phi = sigmoid(W3 gelu(W2 gelu(W1 X + b1) + b2) + b3):
    the MLP gates (limits) the learned high-order correction per variable.

    quantized=True models the inference kernel: the MLP's weights (b1 included: it rides on
    a constant input of the bf16 layer), inputs and hidden activations rounded to bfloat16,
    sums exact; the skip, b2, b3 and the base in full precision."""
    d = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    q = bf16 if quantized else (lambda a: a)
    z1 = q(X) @ q(d["W1"]).T + q(d["b1"])
    h1 = q(gelu(z1))
    z2 = h1 @ q(d["W2"]).T + d["b2"]
    h2 = q(gelu(z2))
    phi = sigmoid(h2 @ q(d["W3"]).T + d["b3"])
    return base(X) + phi * (X @ d["L"].T + d["bL"])


WREF = 0.01
WMAX = 100.0


def sample_weight(X):
    """Per-variable loss weight min((sigma / (WREF scale))^2, WMAX), sigma / scale = X0C x0 / (1 - x0)
    from the centre features x0 = X[..., 8:12]: the squared face-value error in units of the
    centre cell's scale, capped so discontinuities do not dominate."""
    x0 = np.asarray(X, dtype=np.float64)[..., 8:12]
    r = X0C * x0 / (1.0 - x0)
    return np.minimum((r / WREF) ** 2, WMAX)


def loss_and_grad(p, X, Y):
    """Mean over samples and outputs of w (y - y*)^2 and its gradient w.r.t. the flat parameters."""
    d = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    n = len(X)
    z1 = X @ d["W1"].T + d["b1"]
    h1 = gelu(z1)
    z2 = h1 @ d["W2"].T + d["b2"]
    h2 = gelu(z2)
    phi = sigmoid(h2 @ d["W3"].T + d["b3"])
    s = X @ d["L"].T + d["bL"]
    y = base(X) + phi * s
    r = y - Y
    w = sample_weight(X)
    loss = float(np.mean(w * r * r))
    dy = 2.0 * w * r / (n * NOUT)
    dz3 = dy * s * phi * (1.0 - phi)
    g = dict(W3=dz3.T @ h2, b3=dz3.sum(0), L=(dy * phi).T @ X, bL=(dy * phi).sum(0))
    dz2 = (dz3 @ d["W3"]) * gelu_grad(z2)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ d["W2"]) * gelu_grad(z1)
    g["W1"] = dz1.T @ X
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


LIN5 = np.array([2.0, -13.0, 27.0, -3.0]) / 60.0      # fifth-order face value from cells -2, -1, +1, +2
NULL3 = np.array([1.0, -4.0, -4.0, 1.0])               # keeps exactness up to degree 3


def fit_linear(p, X, Y, ridge=1e-9, constrained=False):
    """Weighted least squares of the linear skip with the gates open (phi = 1): per variable, (L_v, bL_v) minimising
    sum_s w_sv ((L_v x_s + bL_v) - (y*_sv - base_sv))^2 with a ridge of ridge * trace / 21.

    constrained: the offsets stay exact for polynomial profiles up to degree 3 -- own-variable
    coefficients (LIN5 - base) + t NULL3 on cells -2, -1, +1, +2, all others 0 -- and only t is
    fitted, t = d.(b - G c0) / (d.G d + ridge trace)."""
    d = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    A = np.c_[X, np.ones(len(X))]
    R = np.asarray(Y, dtype=np.float64) - base(X)
    W = sample_weight(X)
    for v in range(NOUT):
        G = A.T @ (A * W[:, v:v + 1])
        b = A.T @ (W[:, v] * R[:, v])
        tr = np.trace(G)
        if constrained:
            idx = [v, 4 + v, 12 + v, 16 + v]
            c0 = np.zeros(NIN + 1)
            dv = np.zeros(NIN + 1)
            c0[idx] = LIN5 - np.array([0.0, -0.25, 0.25, 0.0])
            dv[idx] = NULL3
            t = dv @ (b - G @ c0) / (dv @ G @ dv + ridge * tr)
            c = c0 + t * dv
        else:
            G = G + ridge * tr / G.shape[0] * np.eye(G.shape[0])
            c = np.linalg.solve(G, b)
        d["L"][v] = c[:NIN]
        d["bL"][v] = c[NIN]
    return pack(d)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, Algorithm 1); t counts from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


# ---------------------------------------------------------------------------
# features and targets on periodic 2-D grids; arrays are (4, ny, nx)
# ---------------------------------------------------------------------------
def prim(U, gamma=1.4):
    r, ru, rv, re = U
    u = ru / r
    v = rv / r
    return np.stack([r, u, v, (gamma - 1.0) * (re - 0.5 * (ru * u + rv * v))])


def cons(W, gamma=1.4):
    r, u, v, p = W
    return np.stack([r, r * u, r * v, p / (gamma - 1.0) + 0.5 * r * (u * u + v * v)])


def stencil(W, axis, sgn):
    """Oriented 5-cell stencil (5, 4, ny, nx): k = -2..2 with k = +1 toward the face;
    variables (rho, u_n, u_t, p), u_n = sgn * (u for x faces, v for y faces)."""
    ax = 2 if axis == 0 else 1
    out = []
    for k in range(-2, 3):
        Wk = np.roll(W, -sgn * k, axis=ax)
        un = Wk[1] if axis == 0 else Wk[2]
        ut = Wk[2] if axis == 0 else Wk[1]
        out.append(np.stack([Wk[0], sgn * un, ut, Wk[3]]))
    return np.stack(out)


def features(St, gamma=1.4):
    """St (5, 4, ...) -> X (..., 20) with feature 4k+v, sigma (..., 4), centre (..., 4)."""
    c = St[2]
    a0 = np.sqrt(gamma * c[3] / c[0])
    scale = np.stack([c[0], a0, a0, c[3]])
    d = St - c[None]
    sig = np.maximum(np.abs(d).max(0), FLOOR * scale)
    X = d / sig[None]
    X[2] = sig / (sig + X0C * scale)
    X = np.moveaxis(X.reshape(20, *X.shape[2:]), 0, -1)
    return X, np.moveaxis(sig, 0, -1), np.moveaxis(c, 0, -1)


def coarse_grain(Uf, f):
    """Box-filtered coarse cells (4, n, n) and line-averaged conserved states on the coarse
    east (axis 0) and north (axis 1) faces, (2, 4, n, n).  A fine face value is the
    fourth-order interpolation (-u_{m-1} + 7 u_m + 7 u_{m+1} - u_{m+2}) / 12 of the fine cell
    averages; the coarse face value averages the f fine faces along it."""
    c, ny, nx = Uf.shape
    cells = Uf.reshape(c, ny // f, f, nx // f, f).mean(axis=(2, 4))
    faces = []
    for ax in (2, 1):
        sh = lambda k: np.roll(Uf, -k, ax)
        ff = (-sh(-1) + 7.0 * Uf + 7.0 * sh(1) - sh(2)) / 12.0
        if ax == 2:
            ff = ff[:, :, f - 1::f].reshape(c, ny // f, f, nx // f).mean(2)
        else:
            ff = ff[:, f - 1::f, :].reshape(c, ny // f, nx // f, f).mean(3)
        faces.append(ff)
    return cells, np.stack(faces)


def samples(cells, faces, gamma=1.4, clip=4.0):
    """Training samples of one coarse snapshot, ordered (axis, side, cell):
    X (4 n^2, 20) features, Y (4 n^2, 4) normalised target offsets (clipped to +-clip)."""
    W = prim(cells, gamma)
    Xs, Ys = [], []
    for axis in (0, 1):
        ax = 2 if axis == 0 else 1
        Fw = prim(faces[axis], gamma)
        for sgn in (1, -1):
            X, sig, c = features(stencil(W, axis, sgn), gamma)
            T = Fw if sgn == 1 else np.roll(Fw, 1, axis=ax)
            un = T[1] if axis == 0 else T[2]
            ut = T[2] if axis == 0 else T[1]
            To = np.moveaxis(np.stack([T[0], sgn * un, ut, T[3]]), 0, -1)
            Xs.append(X.reshape(-1, 20))
            Ys.append(np.clip((To - c) / sig, -clip, clip).reshape(-1, 4))
    return np.concatenate(Xs), np.concatenate(Ys)


def side_ok(St, face, gamma=1.4, relax=0.5):
    """Positivity (rho, p > 0) and the TVD bound relaxed at smooth extrema: every variable of the
    oriented face value lies in [min(w0, w1) - d, max(w0, w1) + d], d = |minmod(c0, c1)| / 2 +
    FLOOR * scale with the curvatures c0 = w1 - 2 w0 + w-1 and c1 = w2 - 2 w1 + w0 (the floor term
    is the feature normalisation's own resolution, so a near-uniform variable does not fail on
    offsets below it)."""
    w0, w1 = St[2], St[3]
    c = St[2]
    a0 = np.sqrt(gamma * c[3] / c[0])
    scale = np.stack([c[0], a0, a0, c[3]])
    c0 = St[3] - 2.0 * St[2] + St[1]
    c1 = St[4] - 2.0 * St[3] + St[2]
    d = relax * np.where(c0 * c1 <= 0.0, 0.0, np.minimum(np.abs(c0), np.abs(c1))) + FLOOR * scale
    ok = (face >= np.minimum(w0, w1) - d) & (face <= np.maximum(w0, w1) + d)
    return np.all(ok, axis=0) & (face[0] > 0.0) & (face[3] > 0.0)


def mc(a, b):
    s = a + b
    m = np.minimum(2.0 * np.minimum(np.abs(a), np.abs(b)), 0.5 * np.abs(s))
    return np.where(a * b > 0.0, np.copysign(m, s), 0.0)


def face_states(p, U, gamma=1.4, quantized=True, bypass=False):
    """Final face primitive states of the learned reconstruction on a periodic grid.

    bypass=True: the gate is evaluated only for face sides whose ungated value (phi = 1)
    fails the TVD bound with half its relaxation; the other sides keep phi = 1.

    Returns ({(axis, 'L'|'R'): (4, ny, nx) primitives of the east/north face of each cell},
    fallback mask per axis (ny, nx) = True where the face fell back to plain MUSCL)."""
    W = prim(U, gamma)
    out, fb = {}, {}
    for axis in (0, 1):
        ax = 2 if axis == 0 else 1
        sides, oks = [], []
        for sgn in (1, -1):
            St = stencil(W, axis, sgn)
            X, sig, c = features(St, gamma)
            y = forward(p, X.reshape(-1, 20), quantized).reshape(c.shape)
            if bypass:
                Xf = X.reshape(-1, 20)
                d = unpack(p)
                y1 = (base(Xf) + Xf @ d["L"].T + d["bL"]).reshape(c.shape)     # phi = 1
                open_ = side_ok(St, np.moveaxis(c + sig * y1, -1, 0), gamma, relax=0.25)
                y = np.where(open_[..., None], y1, y)
            fo = np.moveaxis(c + sig * y, -1, 0)
            ok = side_ok(St, fo, gamma)
            un, ut = sgn * fo[1], fo[2]
            Wf = np.stack([fo[0], un, ut, fo[3]]) if axis == 0 else np.stack([fo[0], ut, un, fo[3]])
            sides.append(Wf)
            oks.append(ok)
        left = sides[0]
        right = np.roll(sides[1], -1, ax)
        ok = oks[0] & np.roll(oks[1], -1, ax - 1)     # masks are (ny, nx)
        up = np.roll(U, -1, ax)
        um = np.roll(U, 1, ax)
        s = mc(up - U, U - um)
        ml = prim(U + 0.5 * s, gamma)
        mr = np.roll(prim(U - 0.5 * s, gamma), -1, ax)
        out[(axis, "L")] = np.where(ok[None], left, ml)
        out[(axis, "R")] = np.where(ok[None], right, mr)
        fb[axis] = ~ok
    return out, fb
