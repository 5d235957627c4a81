"""Flow past a circular cylinder in a channel (Schaefer-Turek DFG 2D-1, Re = 20) with
the ghost-cell immersed boundary, as set up by examples/cylinder_drag_validation
(main.rs:231-416, config.toml): channel 22 D x 4.1 D, cylinder of diameter D
centred 2 D from the inlet and 0.05 D below the centreline, parabolic velocity
inlet with pressure extrapolated, pressure outlet, no-slip walls, an outlet
sponge from x = 14 D, constant mu from Re on the mean inflow velocity, and the
drag/lift from the 64 analytic surface markers (FP64).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .efvm import Context
from . import cases

R_AIR = 287.058


@dataclass
class CylinderCase:
    nx: int = 440
    ny: int = 82
    D: float = 0.1
    rho: float = 1.0
    p: float = 1.0e5
    u_mean: float = 50.0            # reference case: M_max ~ 0.2
    mach_mean: float = None         # if set, u_mean = mach_mean * c_inf
    reynolds: float = 20.0
    cfl: float = 0.3
    muscl: bool = False
    limiter: str = "mc"
    precision: str = "fp64"
    sponge_sigma: float = 1.0e4

    def __post_init__(self):
        if self.mach_mean is not None:
            self.u_mean = self.mach_mean * math.sqrt(1.4 * self.p / self.rho)

    @property
    def length(self):
        return 22.0 * self.D

    @property
    def height(self):
        return 4.1 * self.D

    @property
    def lz(self):
        return 0.41 * self.D

    @property
    def u_max(self):
        return 1.5 * self.u_mean

    @property
    def mu(self):
        return self.rho * self.u_mean * self.D / self.reynolds

    @property
    def q_ref(self):
        """1/2 rho U^2 D L_z (examples/cylinder_drag_validation/main.rs:320)."""
        return 0.5 * self.rho * self.u_mean ** 2 * self.D * self.lz

    def freestream(self):
        """Uniform-mean freestream, IdealGas::prim_to_cons (main.rs:257-266)."""
        one = np.ones(1)
        return cases.prim_to_cons(1.4, self.rho * one, self.u_mean * one, 0.0 * one, self.p * one)[0]

    def initial_state(self):
        """Channel already carrying the parabolic profile (main.rs:268-272)."""
        x, y = cases.cell_centroids(self.nx, self.ny, (0.0, 0.0), (self.length, self.height))
        h = self.height
        uy = 4.0 * self.u_max * y * (h - y) / (h * h)
        return cases.prim_to_cons(1.4, self.rho + 0 * x, uy, 0 * x, self.p + 0 * x)

    def make_context(self) -> Context:
        c = Context(self.nx, self.ny, hi=(self.length, self.height), dz=self.lz, flux="hllc", muscl=self.muscl,
                    limiter=self.limiter, precision=self.precision, viscous=True, mu=self.mu, cfl=self.cfl)
        c.set_boundaries(["outflow", ("pressure_outlet", (self.p,)), "noslip", "noslip"])
        c.set_parabolic_inflow(self.u_max, self.height, self.rho)
        c.set_sponge(14.0 * self.D, self.length, self.sponge_sigma, self.u_max, self.height, self.rho, self.p)
        c.add_cylinder(2.0 * self.D, 2.0 * self.D, 0.5 * self.D, self.freestream())
        c.upload(self.initial_state())
        return c

    def coefficients(self, ctx: Context):
        f = ctx.marker_force()
        return (f[0] + f[3]) / self.q_ref, (f[1] + f[4]) / self.q_ref


def run_substeps_reference(case: CylinderCase, substeps: int, average_frac: float = 0.1):
    """The reference harness loop (main.rs:331-346): `substeps` app.run() calls, Cd/Cl sampled
    after every substage over the last `average_frac` of them."""
    ctx = case.make_context()
    avg_start = int((1.0 - average_frac) * substeps)
    full, rest = divmod(avg_start, 3)
    t0 = time.perf_counter()
    ctx.step(full)
    for _ in range(rest):
        ctx.run()
    cds, cls = [], []
    for _ in range(avg_start, substeps):
        ctx.run()
        cd, cl = case.coefficients(ctx)
        cds.append(cd)
        cls.append(cl)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    cd = float(np.mean(cds))
    return dict(cd=cd, cl=float(np.mean(cls)), ptp=float((max(cds) - min(cds)) / abs(cd)),
                time=ctx.clock()["time"], wall_s=wall, ctx=ctx)


def run_to_plateau(case: CylinderCase, chunk_steps: int = 2000, sample_every: int = 20, ptp_tol: float = 0.02,
                   drift_tol: float = 0.002, max_steps: int = 400000, min_steps: int = 6000, tune: bool = True):
    """Advance whole RK steps until the marker drag plateaus: over the last chunk the Cd samples
    have peak-to-peak <= ptp_tol of the mean (the reference's steady gate, config.toml:78) and
    the chunk mean moved by less than drift_tol from the previous chunk."""
    ctx = case.make_context()
    if tune:
        # small grids cannot fill 256 CUs with tall strips: let the tuner pick the strip height
        from .autotune import tune_context
        tune_context(ctx, steps=2)
    # whole chunks of `sample_every` steps replayed from a hipGraph (no per-kernel launch cost)
    ctx.graph_capture(sample_every)
    hist, prev_mean = [], None
    steps, t0 = 0, time.perf_counter()
    gpu_ms = 0.0
    while steps < max_steps:
        samples = []
        for _ in range(chunk_steps // sample_every):
            ctx.graph_launch()
            gpu_ms += ctx.last_elapsed_ms()
            steps += sample_every
            samples.append(case.coefficients(ctx))
        cd = np.array([s[0] for s in samples])
        cl = np.array([s[1] for s in samples])
        mean = float(cd.mean())
        hist.append(dict(step=steps, time=ctx.clock()["time"], cd=mean, cl=float(cl.mean()),
                         ptp=float((cd.max() - cd.min()) / abs(mean))))
        if (steps >= min_steps and prev_mean is not None and hist[-1]["ptp"] <= ptp_tol
                and abs(mean - prev_mean) <= drift_tol * abs(mean)):
            break
        prev_mean = mean
    return dict(cd=hist[-1]["cd"], cl=hist[-1]["cl"], ptp=hist[-1]["ptp"], steps=steps,
                time=ctx.clock()["time"], wall_s=time.perf_counter() - t0, gpu_ms=gpu_ms, history=hist, ctx=ctx)
