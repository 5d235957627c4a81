"""Initial conditions of the benchmark workloads (host side, FP64).

* ``vortex_state`` -- the Shu isentropic vortex of examples/isentropic_vortex/main.rs:97-125,
  with an optional periodic minimum-image distance for the periodic boxes.
* ``fourier_modes`` -- seeded smooth density/velocity perturbations: a sum of
  sinusoidal modes with integer wavenumbers on the periodic box.
* ``Rng`` -- a splitmix64 stream (the mixer of examples/restart_determinism/main.rs:345-350).

Arrays are (ny, nx, 5) in the reference's Component wire order
(rho, rho_u, rho_v, rho_w, rho_E).
"""
from __future__ import annotations

import math

import numpy as np

M64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


class Rng:
    """splitmix64 sequence: state advances by the golden gamma, output = mix(state)."""

    def __init__(self, seed: int):
        self.state = seed & M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo: int, hi: int) -> int:
        """Integer in [lo, hi] inclusive."""
        return lo + self.next_u64() % (hi - lo + 1)


def cell_centroids(nx, ny, lo=(0.0, 0.0), hi=(1.0, 1.0)):
    dx = (hi[0] - lo[0]) / nx
    dy = (hi[1] - lo[1]) / ny
    x = lo[0] + (np.arange(nx) + 0.5) * dx
    y = lo[1] + (np.arange(ny) + 0.5) * dy
    return np.meshgrid(x, y)


def prim_to_cons(gamma, rho, u, v, p):
    ke = 0.5 * rho * (u * u + v * v + 0.0 * 0.0)
    re = p / (gamma - 1.0) + ke
    return np.stack([rho, rho * u, rho * v, rho * 0.0, re], axis=-1)


def vortex_prim(x, y, t=0.0, L=10.0, beta=5.0, gamma=1.4, mean_u=1.0, mean_v=1.0, periodic=True,
                cx0=None, cy0=None):
    """(rho, u, v, p) of the translated Shu vortex (examples/isentropic_vortex/main.rs:97-125)."""
    cx = (0.5 * L if cx0 is None else cx0) + mean_u * t
    cy = (0.5 * L if cy0 is None else cy0) + mean_v * t
    dx = x - cx
    dy = y - cy
    if periodic:
        dx = dx - L * np.round(dx / L)
        dy = dy - L * np.round(dy / L)
    r2 = dx * dx + dy * dy
    f = np.exp(0.5 * (1.0 - r2))
    du = -beta / (2.0 * math.pi) * dy * f
    dv = beta / (2.0 * math.pi) * dx * f
    dtemp = -(gamma - 1.0) * beta * beta / (8.0 * gamma * (math.pi * math.pi)) * np.exp(1.0 - r2)
    temp = 1.0 + dtemp
    rho = temp ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    return rho, mean_u + du, mean_v + dv, p


def vortex_state(nx, ny, L=10.0, beta=5.0, gamma=1.4, mean_u=1.0, mean_v=1.0, periodic=True, t=0.0):
    x, y = cell_centroids(nx, ny, (0.0, 0.0), (L, L))
    rho, u, v, p = vortex_prim(x, y, t, L, beta, gamma, mean_u, mean_v, periodic)
    return prim_to_cons(gamma, rho, u, v, p)


def fourier_modes(seed: int, n_modes: int = 8, kmin: int = 1, kmax: int = 8):
    """n_modes x (kx, ky, phase_rho, phase_u, phase_v), drawn from a splitmix64 stream."""
    r = Rng(seed)
    modes = []
    for _ in range(n_modes):
        kx = r.randint(kmin, kmax)
        ky = r.randint(kmin, kmax)
        ph = [r.uniform(0.0, 2.0 * math.pi) for _ in range(3)]
        modes.append((kx, ky, *ph))
    return modes


def vortex_modes_state(nx, ny, seed=1, amplitude=0.02, L=10.0, beta=5.0, gamma=1.4, mean_u=1.0, mean_v=1.0,
                       n_modes=8):
    """Periodic vortex plus `n_modes` seeded Fourier modes on rho, u, v (pressure from the vortex)."""
    x, y = cell_centroids(nx, ny, (0.0, 0.0), (L, L))
    rho, u, v, p = vortex_prim(x, y, 0.0, L, beta, gamma, mean_u, mean_v, True)
    for kx, ky, pr, pu, pv in fourier_modes(seed, n_modes):
        arg = 2.0 * math.pi * (kx * x + ky * y) / L
        rho = rho + amplitude * np.sin(arg + pr)
        u = u + amplitude * np.sin(arg + pu)
        v = v + amplitude * np.sin(arg + pv)
    return prim_to_cons(gamma, rho, u, v, p)


def learned_vortices(seed: int, L: float = 10.0, min_sep: float = 3.0):
    """1..4 vortices (cx, cy, beta), beta ~ U[2, 6], centres uniform on the periodic box and at
    least `min_sep` apart (minimum-image distance), so superposed cores keep a positive temperature."""
    r = Rng(splitmix64(seed) ^ 0x6A09E667F3BCC909)
    nv = r.randint(1, 4)
    out = []
    while len(out) < nv:
        cx, cy, beta = r.uniform(0.0, L), r.uniform(0.0, L), r.uniform(2.0, 6.0)
        ok = True
        for ox, oy, _ in out:
            ddx = cx - ox - L * round((cx - ox) / L)
            ddy = cy - oy - L * round((cy - oy) / L)
            ok &= ddx * ddx + ddy * ddy >= min_sep * min_sep
        if ok:
            out.append((cx, cy, beta))
    return out


def learned_state(seed: int, n: int, L: float = 10.0, sub: int = 4, amplitude: float = 0.05, gamma: float = 1.4,
                  mean_u: float = 1.0, mean_v: float = 1.0):
    """Learned-reconstruction workload on the periodic [0,L]^2 box: mean flow (1, 1), rho = p = 1,
    1..4 superposed isentropic vortices (``learned_vortices``) and 8 seeded Fourier modes of
    amplitude `amplitude` on rho, u, v (pressure from the vortices).  Returned as cell averages
    (sub x sub midpoint samples per cell), so grids of different resolution discretise the same
    continuous field.  (n, n, 5) in the wire order."""
    vort = learned_vortices(seed, L)
    modes = fourier_modes(seed, 8)
    h = L / n
    acc = np.zeros((n, n, 5))
    g = gamma
    for a in range(sub):
        x1 = (np.arange(n) + (a + 0.5) / sub) * h
        for b in range(sub):
            y1 = (np.arange(n) + (b + 0.5) / sub) * h
            x, y = np.meshgrid(x1, y1)
            du = np.zeros_like(x)
            dv = np.zeros_like(x)
            dtemp = np.zeros_like(x)
            for cx, cy, beta in vort:
                ddx = x - cx
                ddx -= L * np.round(ddx / L)
                ddy = y - cy
                ddy -= L * np.round(ddy / L)
                r2 = ddx * ddx + ddy * ddy
                f = np.exp(0.5 * (1.0 - r2))
                du += -beta / (2.0 * math.pi) * ddy * f
                dv += beta / (2.0 * math.pi) * ddx * f
                dtemp += -(g - 1.0) * beta * beta / (8.0 * g * math.pi * math.pi) * np.exp(1.0 - r2)
            rho = (1.0 + dtemp) ** (1.0 / (g - 1.0))
            p = rho ** g
            u = mean_u + du
            v = mean_v + dv
            for kx, ky, pr, pu, pv in modes:
                arg = 2.0 * math.pi * (kx * x + ky * y) / L
                rho = rho + amplitude * np.sin(arg + pr)
                u = u + amplitude * np.sin(arg + pu)
                v = v + amplitude * np.sin(arg + pv)
            acc += prim_to_cons(g, rho, u, v, p)
    return acc / (sub * sub)


def density_l1(rho_num, rho_exact, cell_volume=1.0):
    """Volume-weighted L1 density error (examples/isentropic_vortex/main.rs:189-211)."""
    return float(np.sum(np.abs(rho_num - rho_exact) * cell_volume) / (rho_num.size * cell_volume))


def riemann4_state(nx, ny, quads, gamma, L=1.0):
    """Four-quadrant Riemann problem on [0,L]^2 split at (L/2, L/2); quads = 4 x (rho, u, v, p)
    for the quadrants (x>, y>), (x<, y>), (x<, y<), (x>, y<)."""
    x, y = cell_centroids(nx, ny, (0.0, 0.0), (L, L))
    q = np.where(x >= 0.5 * L, np.where(y >= 0.5 * L, 0, 3), np.where(y >= 0.5 * L, 1, 2))
    Q = np.asarray(quads, dtype=np.float64)
    rho, u, v, p = (Q[q, k] for k in range(4))
    return prim_to_cons(gamma, rho, u, v, p)


def ensemble(n_cases=512, seed=2024, nmin=64, nmax=1024):
    """The batched-ensemble workload: square grids with N = round(nmin * (nmax/nmin)^U) per side,
    gamma ~ U[1.3, 1.67]; even cases are periodic isentropic vortices (beta ~ U[2, 6], mean flow
    (1, 1), [0,10]^2), odd cases are four-quadrant Riemann problems on [0,1]^2 with
    rho ~ U[0.1, 1.5], p ~ U[0.03, 1.5], |u|, |v| <= 1 and zero-gradient outflow sides.
    Returns (case specs for Batch, initial-state builders)."""
    r = Rng(seed)
    specs, builders = [], []
    ratio = nmax / nmin
    for i in range(n_cases):
        n = int(round(nmin * ratio ** r.uniform()))
        gamma = r.uniform(1.3, 1.67)
        if i % 2 == 0:
            beta = r.uniform(2.0, 6.0)
            specs.append(dict(nx=n, ny=n, hi=(10.0, 10.0), gamma=gamma, bcs=["periodic"] * 4,
                              kind="vortex", beta=beta))
            builders.append(lambda n=n, beta=beta, gamma=gamma: vortex_state(n, n, beta=beta, gamma=gamma))
        else:
            quads = [(r.uniform(0.1, 1.5), r.uniform(-1.0, 1.0), r.uniform(-1.0, 1.0), r.uniform(0.03, 1.5))
                     for _ in range(4)]
            specs.append(dict(nx=n, ny=n, hi=(1.0, 1.0), gamma=gamma, bcs=["outflow"] * 4, kind="riemann4",
                              quads=quads))
            builders.append(lambda n=n, quads=quads, gamma=gamma: riemann4_state(n, n, quads, gamma))
    return specs, builders
