"""The fixed problem of the reference's restart and rank-count check
(examples/restart_determinism/main.rs) and its coordinate-keyed state digest.

The reference advances a smooth Euler problem for 10 steps and requires the digest of
the final state to equal a constant it holds (main.rs:79).  Each cell's five conserved
values enter the digest through their bit patterns, so the constant fixes every value
of the final state bit for bit.  This module restates the setup and the digest; the
tests run it on the oracle and on the 3-D context.
"""
import math

import numpy as np

# main.rs:77-79
STEPS_TOTAL = 10
STEPS_CHECKPOINT = 5
DIGEST_REF = (21_398_001_286.0, 21_099_935_807.0)

# main.rs:81-92: UniformMeshConfig { nx 16, ny 4, nz 2, ng 2, [0,1] x [0,0.25] x [0,0.125] }
NX, NY, NZ = 16, 4, 2
LO = (0.0, 0.0, 0.0)
HI = (1.0, 0.25, 0.125)
# main.rs:94-104: slip walls on the y and z sides; the x sides keep the registry default,
# SupersonicOutflow (crates/cfd_boundary/src/lib.rs:375-380)
BCS = ["outflow", "outflow", "slip", "slip", "slip", "slip"]
# main.rs:120-134: IdealGasPlugin (air), Rusanov flux, first order, forward Euler,
# cfl 0.2, fixed dt 2e-5, inviscid
SCHEME = dict(flux="rusanov", muscl=False, integrator="euler", viscous=False, cfl=0.2, fixed_dt=2.0e-5)
GAMMA, R_GAS = 1.4, 287.058

_M64 = (1 << 64) - 1
_LOW26 = (1 << 26) - 1


def initial_prim(x0: float, x1: float):
    """Primitive state (rho, u, v, w, p, T) of main.rs:106-115 at one centroid.  Python's
    math module calls the C library's exp, sin and cos, as the reference's f64 methods
    do; powi(2) is a product."""
    d0, d1 = x0 - 0.5, x1 - 0.125
    r2 = d0 * d0 + 16.0 * (d1 * d1)
    bump = math.exp(-80.0 * r2)
    rho = 1.0 + 0.04 * bump
    u = 9.0 + 0.7 * math.sin(2.0 * math.pi * x0)
    v = 0.15 * math.cos(4.0 * math.pi * x1)
    p = 100_000.0 * (1.0 + 0.03 * bump)
    return rho, u, v, 0.0, p, p / (rho * 287.058)


def initial_prims(cx, cy, nz):
    """initial_prim at every cell centroid, as an (nz, ny, nx, 6) array."""
    ny, nx = len(cy), len(cx)
    out = np.zeros((nz, ny, nx, 6))
    for j in range(ny):
        for i in range(nx):
            out[:, j, i] = initial_prim(float(cx[i]), float(cy[j]))
    return out


def splitmix64(x: int) -> int:
    """main.rs:340-345."""
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def state_digest(u, cx, cy, cz):
    """main.rs:306-338 over an (nz, ny, nx, 5) interior state.  Each key is built from the
    cell's integer position (floor of centroid over spacing), each component's bits are
    mixed in, and the low and the next 26 bits of every hash are summed.  The sums are
    integers below 2^53, so the order of summation does not matter."""
    nz, ny, nx = u.shape[:3]
    dx, dy, dz = (HI[0] - LO[0]) / NX, (HI[1] - LO[1]) / NY, (HI[2] - LO[2]) / NZ
    bits = np.ascontiguousarray(u, dtype=np.float64).view(np.uint64)
    lo = hi = 0
    for k in range(nz):
        iz = math.floor((float(cz[k]) - LO[2]) / dz)
        for j in range(ny):
            iy = math.floor((float(cy[j]) - LO[1]) / dy)
            for i in range(nx):
                ix = math.floor((float(cx[i]) - LO[0]) / dx)
                key = (ix << 32) ^ (iy << 16) ^ iz
                for c in range(5):
                    h = splitmix64(key ^ (c << 48) ^ int(bits[k, j, i, c]))
                    lo += h & _LOW26
                    hi += (h >> 26) & _LOW26
    return float(lo), float(hi)

