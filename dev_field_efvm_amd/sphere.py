"""Resolved drag of a fixed sphere at Re = 20 with the 3-D ghost-cell immersed boundary,
as set up by examples/sphere_drag_validation (main.rs:164-330, config.toml): box
12 D x 6 D x 6 D around a sphere of diameter D = 0.1 at the origin, 144 x 72 x 72
cells (D/dx = 12), uniform freestream rho = 1, p = 1e5, U = 56.1 (Mach 0.15),
subsonic velocity inlet, pressure outlet, symmetry on the four lateral sides,
constant mu = rho U D / Re, first-order HLLC + viscous + SSP-RK3 at CFL 0.3, a
sponge relaxing to the freestream over x in [0.5, 0.9], and the drag from 2000
spherical-Fibonacci surface markers (FP64):  Cd = F_x / (1/2 rho U^2 pi R^2).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from .efvm import Context3
from . import cases

R_AIR = 287.058


@dataclass
class SphereCase:
    nx: int = 144
    ny: int = 72
    nz: int = 72
    lo: tuple = (-0.3, -0.3, -0.3)
    hi: tuple = (0.9, 0.3, 0.3)
    D: float = 0.1
    rho: float = 1.0
    p: float = 1.0e5
    u_inf: float = 56.1
    reynolds: float = 20.0
    markers: int = 2000
    cfl: float = 0.3
    muscl: bool = False
    sponge_x0: float = 0.5
    sponge_sigma: float = 2.0e4
    precision: str = "fp64"

    @property
    def mu(self):
        return self.rho * self.u_inf * self.D / self.reynolds

    @property
    def t_inf(self):
        return self.p / (self.rho * R_AIR)

    @property
    def q_ref(self):
        """1/2 rho U^2 pi R^2 (main.rs:232)."""
        r = 0.5 * self.D
        return 0.5 * self.rho * self.u_inf * self.u_inf * math.pi * r * r

    def freestream(self):
        """IdealGas::prim_to_cons of (rho, U, 0, 0, p) (main.rs:189-191)."""
        one = np.ones(1)
        return np.asarray(cases.prim_to_cons(1.4, self.rho * one, self.u_inf * one, 0.0 * one, self.p * one)[0],
                          dtype=np.float64)

    def make_context(self) -> Context3:
        c = Context3(self.nx, self.ny, self.nz, lo=self.lo, hi=self.hi, flux="hllc", muscl=self.muscl,
                     precision=self.precision, viscous=True, mu=self.mu, cfl=self.cfl)
        c.set_boundaries([("subsonic_inflow", (self.rho, self.u_inf, 0.0, 0.0, self.t_inf)),
                          ("pressure_outlet", (self.p,)), "symmetry", "symmetry", "symmetry", "symmetry"])
        fs = self.freestream()
        c.set_sponge(self.sponge_x0, self.hi[0], self.sponge_sigma, fs)
        c.add_sphere((0.0, 0.0, 0.0), 0.5 * self.D, fs, markers=self.markers)
        c.upload(np.broadcast_to(fs, (self.nz, self.ny, self.nx, 5)).copy())   # CfdStatePlugin::new(init)
        return c

    def cd(self, ctx: Context3) -> float:
        f = ctx.marker_force()
        return (f[0] + f[3]) / self.q_ref


def run_reference(case: SphereCase, runs: int = 9000, average_frac: float = 0.2):
    """The reference harness loop (main.rs:253-283): `runs` app.run() calls (substages),
    Cd sampled after every one from int((1 - average_frac) runs) on, averaged; also the
    peak-to-peak over that window relative to the mean (the steady gate)."""
    ctx = case.make_context()
    avg_start = int((1.0 - average_frac) * runs)
    full, rest = divmod(avg_start, 3)
    t0 = time.perf_counter()
    ctx.step(full)
    for _ in range(rest):
        ctx.run()
    ctx.synchronize()
    t_bulk = time.perf_counter() - t0
    cds = []
    for _ in range(avg_start, runs):
        ctx.run()
        cds.append(case.cd(ctx))
    ctx.synchronize()
    wall = time.perf_counter() - t0
    cd = float(np.mean(cds))
    return dict(cd=cd, ptp=float((max(cds) - min(cds)) / abs(cd)), runs=runs, time=ctx.clock()["time"],
                wall_s=wall, bulk_s=t_bulk, bulk_substages=avg_start, ctx=ctx)
