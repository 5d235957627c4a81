"""Host-side mirror of the reference's plugin API for the FV update path.

A reference app is assembled from GRASS plugins (crates/cfd_solver/src/lib.rs:691-984,
crates/cfd_boundary/src/lib.rs:457-533, crates/cfd_ibm/src/lib.rs:621-680):

    app = App()
    app.add_plugins(FieldDefaultPlugins(mesh=UniformMeshConfig(nx=64, ny=64, nz=1, ng=2,
                                                               bounds_lo=(0, 0, 0), bounds_hi=(10, 10, 10 / 64))))
    app.add_plugins(CfdStatePlugin(init))
    app.add_plugins(IdealGasPlugin())
    app.add_plugins(StructuredBoundaryPlugin(BoundaryRegistry().with_axis(0, Periodic()).with_axis(2, SlipWall())))
    app.add_plugins(MusclPlugin.hllc().with_limiter(SlopeLimiter.MonotonizedCentral))
    app.add_plugins(IntegratorPlugin.rk3())
    app.add_plugins(SolverPlugin(SolverConfig(cfl=0.4)))
    app.prepare()
    app.run()          # one app.run() = one RK substage through every phase

Here the plugins only collect configuration; `prepare()` creates one GPU
context through the C ABI (include/efvm.h) and every `run()` enqueues the
fused stage kernels.  `GpuPlugin` selects precision (FP64 parity / FP32
production), the periodic-seam mode and the launch shape.  A one-cell-thick
mesh with slip or symmetry z sides runs on the 2-D context (efvm_*); any other
mesh (nz > 1, other z sides, a sphere) on the 3-D context (efvm3_*).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from enum import Enum
from typing import Callable, Optional, Sequence

import numpy as np

from .efvm import Context, Context3, EfvmError, ERR_UNSUPPORTED

R_AIR = 287.058


# ---------------------------------------------------------------- mesh / state
@dataclass
class UniformMeshConfig:
    nx: int
    ny: int
    nz: int = 1
    ng: int = 2
    bounds_lo: Sequence[float] = (0.0, 0.0, 0.0)
    bounds_hi: Sequence[float] = (1.0, 1.0, 1.0)
    # stretched y / z: ny+1 (nz+1) increasing cell edges from bounds_lo to bounds_hi, or None
    y_edges: Optional[Sequence[float]] = None
    z_edges: Optional[Sequence[float]] = None

    def centroids(self, axis: int) -> np.ndarray:
        """Interior cell centroids along an axis: edge midpoints on a stretched axis."""
        e = (None, self.y_edges, self.z_edges)[axis]
        if e is not None:
            e = np.asarray(e, dtype=np.float64)
            return 0.5 * (e[:-1] + e[1:])
        n = (self.nx, self.ny, self.nz)[axis]
        lo, hi = self.bounds_lo[axis], self.bounds_hi[axis]
        return lo + (np.arange(n) + 0.5) * ((hi - lo) / n)


@dataclass
class FieldDefaultPlugins:
    mesh: UniformMeshConfig


class CfdStatePlugin:
    """Initial condition: a callable init(x, y) -> (rho, rho_u, rho_v, rho_w, rho_E) over
    cell-centroid arrays (init(x, y, z) on a 3-D mesh), or an array of shape (ny, nx, 5)
    ((nz, ny, nx, 5)) (crates/cfd_solver/src/lib.rs:691-719)."""

    def __init__(self, init):
        self.init = init


# ---------------------------------------------------------------- gas
@dataclass
class IdealGas:
    gamma: float = 1.4
    r_gas: float = R_AIR

    @staticmethod
    def air() -> "IdealGas":
        return IdealGas(1.4, R_AIR)

    def prim_to_cons(self, rho, u, v, w, p):
        """IdealGas::prim_to_cons (crates/cfd_eos/src/lib.rs:55-59), vectorised."""
        ke = 0.5 * rho * (u * u + v * v + w * w)
        return np.stack(np.broadcast_arrays(rho, rho * u, rho * v, rho * w, p / (self.gamma - 1.0) + ke), axis=-1)


class Viscosity:
    """Viscosity::{Sutherland, Constant(mu)} (crates/cfd_eos/src/lib.rs:100-117)."""

    def __init__(self, mu: Optional[float] = None):
        self.mu = mu

    @staticmethod
    def sutherland() -> "Viscosity":
        return Viscosity(None)

    @staticmethod
    def constant(mu: float) -> "Viscosity":
        return Viscosity(mu)


class GasPlugin:
    def __init__(self, gas: IdealGas):
        self.gas = gas


class IdealGasPlugin(GasPlugin):
    def __init__(self):
        super().__init__(IdealGas.air())


# ---------------------------------------------------------------- flux / limiter / integrator
class SlopeLimiter(Enum):
    Minmod = "minmod"
    MonotonizedCentral = "mc"


class FluxPlugin:
    """First-order flux sweep (crates/cfd_solver/src/lib.rs:764-797)."""
    muscl = False

    def __init__(self, flux: str = "rusanov"):
        self.flux = flux
        self.limiter = SlopeLimiter.Minmod

    @classmethod
    def rusanov(cls):
        return cls("rusanov")

    @classmethod
    def hllc(cls):
        return cls("hllc")

    @classmethod
    def hllc_textbook(cls):
        return cls("hllc_textbook")


class MusclPlugin(FluxPlugin):
    """MUSCL reconstruction + flux (crates/cfd_solver/src/lib.rs:806-842); default limiter Minmod."""
    muscl = True

    def with_limiter(self, limiter: SlopeLimiter) -> "MusclPlugin":
        self.limiter = limiter
        return self


class IntegratorPlugin:
    def __init__(self, kind: str):
        self.kind = kind

    @classmethod
    def euler(cls):
        return cls("euler")

    @classmethod
    def rk3(cls):
        return cls("rk3")


@dataclass
class SolverConfig:
    """SolverConfig (crates/cfd_solver/src/lib.rs:39-58)."""
    cfl: float = 0.4
    fixed_dt: Optional[float] = None
    viscous: bool = False
    muscl: bool = False


class SolverPlugin:
    def __init__(self, config: SolverConfig = None):
        self.config = config or SolverConfig()


# ---------------------------------------------------------------- boundaries
class BoundarySide(Enum):
    XLo = 0
    XHi = 1
    YLo = 2
    YHi = 3
    ZLo = 4
    ZHi = 5

    def axis(self) -> int:
        return self.value // 2


class BoundaryCondition:
    kind = "outflow"
    params: tuple = ()

    def is_periodic(self) -> bool:
        return False


class SupersonicOutflow(BoundaryCondition):
    kind = "outflow"


class Periodic(BoundaryCondition):
    kind = "periodic"

    def is_periodic(self) -> bool:
        return True


class SlipWall(BoundaryCondition):
    kind = "slip"


class Symmetry(BoundaryCondition):
    kind = "symmetry"


class NoSlipWall(BoundaryCondition):
    kind = "noslip"


class SupersonicInflow(BoundaryCondition):
    kind = "supersonic_inflow"

    def __init__(self, rho, u, v, w, p):
        self.params = (rho, u, v, w, p)


class SubsonicInflow(BoundaryCondition):
    kind = "subsonic_inflow"

    def __init__(self, rho, u, v, w, t):
        self.params = (rho, u, v, w, t)


class PressureOutlet(BoundaryCondition):
    kind = "pressure_outlet"

    def __init__(self, p):
        self.params = (p,)


class FromOutside(BoundaryCondition):
    """crates/cfd_boundary/src/lib.rs:108-147: the ghost state another code supplies per
    boundary face through BoundaryInputs; faces without one copy the interior cell."""
    kind = "table"


class FnBoundary(BoundaryCondition):
    """crates/cfd_boundary/src/lib.rs:152-170 for conditions that depend on where the face is:
    fn(face_centroid) -> conserved state (wire order), evaluated once per face at prepare()
    into a ghost table.  (A condition that reads the interior state each phase has no
    device form here.)"""
    kind = "table"

    def __init__(self, fn):
        self.fn = fn


class BoundaryInputs:
    """crates/cfd_boundary/src/lib.rs:108-111: supplied ghost states, here per side as
    (states (n, 5), given (n,)), one per boundary face: n = ny for x sides and nx for y
    sides of a 2-D mesh; on a 3-D mesh ny nz (j + ny k), nx nz (i + nx k) and nx ny
    (i + nx j) for x, y and z sides."""

    def __init__(self):
        self.sides = {}

    def set(self, side: "BoundarySide", states, given) -> "BoundaryInputs":
        self.sides[side.value] = (np.asarray(states, dtype=np.float64), np.asarray(given, dtype=bool))
        return self


class BoundaryRegistry:
    """Per-side boundary conditions; every side defaults to SupersonicOutflow
    (crates/cfd_boundary/src/lib.rs:369-449)."""

    def __init__(self):
        self.patches = [SupersonicOutflow() for _ in range(6)]

    def set(self, side: BoundarySide, bc: BoundaryCondition) -> "BoundaryRegistry":
        self.patches[side.value] = bc
        return self

    def with_(self, side: BoundarySide, bc: BoundaryCondition) -> "BoundaryRegistry":
        return self.set(side, bc)

    def with_axis(self, axis: int, bc: BoundaryCondition) -> "BoundaryRegistry":
        self.patches[2 * axis] = bc
        self.patches[2 * axis + 1] = bc
        return self

    def get(self, side: BoundarySide) -> BoundaryCondition:
        return self.patches[side.value]

    def validate_periodic_pairs(self):
        for axis in range(3):
            lo, hi = self.patches[2 * axis], self.patches[2 * axis + 1]
            if lo.is_periodic() != hi.is_periodic():
                raise ValueError(f"periodic boundary on axis {axis} requires both sides; use with_axis(axis, Periodic())")


class BoundaryPlugin:
    def __init__(self, registry: BoundaryRegistry = None):
        self.registry = registry or BoundaryRegistry()


class StructuredBoundaryPlugin(BoundaryPlugin):
    def __init__(self, registry: BoundaryRegistry = None):
        super().__init__(registry)
        self.registry.validate_periodic_pairs()   # panics in the reference (cfd_boundary:516-518)


# ---------------------------------------------------------------- IBM / case systems
@dataclass
class Cylinder:
    center_xy: Sequence[float]
    radius: float


class GhostCellIbmPlugin:
    pass


class CylinderIbmPlugin:
    def __init__(self, cylinder: Cylinder, freestream):
        self.cylinders = [cylinder]
        self.freestream = np.asarray(freestream, dtype=np.float64)

    @classmethod
    def with_cylinders(cls, cylinders, freestream):
        p = cls(cylinders[0], freestream)
        p.cylinders = list(cylinders)
        return p


@dataclass
class SphereBody:
    """A fixed sphere (crates/cfd_ibm/src/sphere.rs:27-46, SphereBody::fixed)."""
    center: Sequence[float]
    radius: float

    @classmethod
    def fixed(cls, center, radius):
        return cls(tuple(center), radius)


class SphereIbmPlugin:
    """SphereIbmPlugin::new(bodies, freestream).with_markers(n) (sphere.rs:239-278)."""

    def __init__(self, bodies, freestream):
        self.bodies = list(bodies)
        self.freestream = np.asarray(freestream, dtype=np.float64)
        self.markers = 2000

    def with_markers(self, n_markers: int) -> "SphereIbmPlugin":
        self.markers = int(n_markers)
        return self


@dataclass
class Sponge:
    """The uniform-freestream outlet sponge of examples/sphere_drag_validation/main.rs:110-160:
    after each completed step u += min(sigma s^2 dt, 1) (target - u) for x >= x_start."""
    x_start: float
    x_end: float
    sigma: float
    target: Sequence[float]


@dataclass
class ParabolicInflow:
    """examples/cylinder_drag_validation/main.rs:146-177."""
    u_max: float
    height: float
    rho: float


@dataclass
class OutletSponge:
    """examples/cylinder_drag_validation/main.rs:179-229."""
    x_start: float
    x_end: float
    sigma: float
    u_max: float
    height: float
    rho: float
    p: float


@dataclass
class BodyForceX:
    """crates/cfd_solver/tests/poiseuille_channel_validation.rs:102-120."""
    gx: float


@dataclass
class OutputPlugin:
    """crates/cfd_output/src/lib.rs:126-160: one VTK file of the state when the App is
    finished (``App.finish()``, the reference's cleanup hook)."""
    path: str = "output.vtk"


@dataclass
class GpuPlugin:
    precision: str = "fp64"        # "fp64" (parity) or "fp32" (production)
    seam_mode: int = 0             # 0 reference, 1 conservative periodic seams
    strip_rows: int = 0            # 0: default / autotuned
    waves_per_block: int = 0


@dataclass
class SolverState:
    dt: float
    time: float
    step: int
    rk_stage: int


# ---------------------------------------------------------------- the App
class App:
    def __init__(self):
        self.plugins = []
        self.resources = {}
        self.ctx: Optional[Context] = None

    def add_plugins(self, p) -> "App":
        self.plugins.append(p)
        return self

    def add_resource(self, r) -> "App":
        self.resources[type(r)] = r
        return self

    def _find(self, cls):
        for p in self.plugins:
            if isinstance(p, cls):
                return p
        return None

    def prepare(self):
        mesh = self._find(FieldDefaultPlugins)
        if mesh is None:
            raise ValueError("mesh plugin missing (FieldDefaultPlugins)")
        m = mesh.mesh
        if m.ng != 2:
            raise EfvmError(ERR_UNSUPPORTED, "the GPU path keeps ng = 2 ghost layers")
        gasp = self._find(GasPlugin) or IdealGasPlugin()
        flux = self._find(FluxPlugin)
        if flux is None:
            raise ValueError("riemann_flux plugin missing (FluxPlugin / MusclPlugin)")
        integ = self._find(IntegratorPlugin) or IntegratorPlugin.euler()
        solver = self._find(SolverPlugin)
        if solver is None:
            raise ValueError("solver plugin missing (SolverPlugin)")
        bcp = self._find(BoundaryPlugin) or BoundaryPlugin()
        gpu = self._find(GpuPlugin) or GpuPlugin()
        visc = self.resources.get(Viscosity, Viscosity.sutherland())
        reg = bcp.registry
        reg.validate_periodic_pairs()
        self.mesh = m
        thin = m.nz == 1 and all(reg.get(s).kind in ("slip", "symmetry") for s in (BoundarySide.ZLo, BoundarySide.ZHi))
        stretched = m.y_edges is not None or m.z_edges is not None
        if not thin or stretched or self._find(SphereIbmPlugin) is not None or Sponge in self.resources:
            return self._prepare3(m, gasp, flux, integ, solver, reg, gpu, visc)
        lo, hi = m.bounds_lo, m.bounds_hi
        cfg = solver.config
        self.ctx = Context(m.nx, m.ny, lo=(lo[0], lo[1]), hi=(hi[0], hi[1]), dz=hi[2] - lo[2],
                           gamma=gasp.gas.gamma, r_gas=gasp.gas.r_gas, mu=visc.mu, flux=flux.flux,
                           muscl=flux.muscl, limiter=flux.limiter.value, integrator=integ.kind,
                           precision=gpu.precision, viscous=cfg.viscous, seam_mode=gpu.seam_mode,
                           cfl=cfg.cfl, fixed_dt=cfg.fixed_dt)
        if gpu.strip_rows or gpu.waves_per_block:
            self.ctx.set_launch(gpu.strip_rows, gpu.waves_per_block)
        self.ctx.set_boundaries([(reg.get(BoundarySide(s)).kind, reg.get(BoundarySide(s)).params) for s in range(4)])
        # ghost tables: FnBoundary from the face centroids, FromOutside from BoundaryInputs
        dx = (hi[0] - lo[0]) / m.nx
        dy = (hi[1] - lo[1]) / m.ny
        for s in range(4):
            bc = reg.get(BoundarySide(s))
            if isinstance(bc, FnBoundary):
                n = m.ny if s < 2 else m.nx
                if s < 2:
                    fc = [(lo[0] if s == 0 else hi[0], lo[1] + (k + 0.5) * dy, 0.5 * (lo[2] + hi[2])) for k in range(n)]
                else:
                    fc = [(lo[0] + (k + 0.5) * dx, lo[1] if s == 2 else hi[1], 0.5 * (lo[2] + hi[2])) for k in range(n)]
                self.ctx.set_ghost_table(s, np.array([bc.fn(c) for c in fc]), np.ones(n, bool))
        self.apply_boundary_inputs()
        inflow = self.resources.get(ParabolicInflow)
        if inflow:
            self.ctx.set_parabolic_inflow(inflow.u_max, inflow.height, inflow.rho)
        sp = self.resources.get(OutletSponge)
        if sp:
            self.ctx.set_sponge(sp.x_start, sp.x_end, sp.sigma, sp.u_max, sp.height, sp.rho, sp.p)
        bf = self.resources.get(BodyForceX)
        if bf:
            self.ctx.set_body_force(bf.gx)
        cyl = self._find(CylinderIbmPlugin)
        if cyl:
            for cy in cyl.cylinders:
                self.ctx.add_cylinder(cy.center_xy[0], cy.center_xy[1], cy.radius, cyl.freestream)
        st = self._find(CfdStatePlugin)
        if st is None:
            raise ValueError("CfdStatePlugin missing")
        init = st.init
        if callable(init):
            x, y = self.cell_centroids()
            u = np.asarray(init(x, y), dtype=np.float64)
        else:
            u = np.asarray(init, dtype=np.float64)
        self.ctx.upload(u.reshape(m.ny, m.nx, 5))
        return self

    def _prepare3(self, m, gasp, flux, integ, solver, reg, gpu, visc):
        """The 3-D context: six sides, bodies, the uniform sponge, body force."""
        cfg = solver.config
        if ParabolicInflow in self.resources or OutletSponge in self.resources:
            raise EfvmError(ERR_UNSUPPORTED, "the parabolic channel inflow and sponge are 2-D cases")
        self.ctx = Context3(m.nx, m.ny, m.nz, lo=tuple(m.bounds_lo), hi=tuple(m.bounds_hi), gamma=gasp.gas.gamma,
                            r_gas=gasp.gas.r_gas, mu=visc.mu, flux=flux.flux, muscl=flux.muscl,
                            limiter=flux.limiter.value, integrator=integ.kind, precision=gpu.precision,
                            viscous=cfg.viscous, seam_mode=gpu.seam_mode, cfl=cfg.cfl, fixed_dt=cfg.fixed_dt)
        self.ctx.set_boundaries([(reg.get(BoundarySide(s)).kind, reg.get(BoundarySide(s)).params) for s in range(6)])
        for axis, e in ((1, m.y_edges), (2, m.z_edges)):
            if e is not None:
                self.ctx.set_edges(axis, e)
        # ghost tables: FnBoundary from the face centroids, FromOutside from BoundaryInputs
        lo3, hi3 = tuple(m.bounds_lo), tuple(m.bounds_hi)
        for s in range(6):
            bc = reg.get(BoundarySide(s))
            if isinstance(bc, FnBoundary):
                ax = s // 2
                u, w = (1, 2) if ax == 0 else ((0, 2) if ax == 1 else (0, 1))   # in-plane axes, faster first
                cu, cw = m.centroids(u), m.centroids(w)
                fc = []
                for b in cw:
                    for a_ in cu:
                        c = [0.0, 0.0, 0.0]
                        c[ax] = lo3[ax] if s % 2 == 0 else hi3[ax]
                        c[u], c[w] = a_, b
                        fc.append(tuple(c))
                self.ctx.set_ghost_table(s, np.array([bc.fn(c) for c in fc]), np.ones(len(fc), bool))
        self.apply_boundary_inputs()
        sp = self.resources.get(Sponge)
        if sp:
            self.ctx.set_sponge(sp.x_start, sp.x_end, sp.sigma, sp.target)
        bf = self.resources.get(BodyForceX)
        if bf:
            self.ctx.set_body_force(bf.gx)
        sph = self._find(SphereIbmPlugin)
        cyl = self._find(CylinderIbmPlugin)
        if sph:
            for b in sph.bodies:
                self.ctx.add_sphere(b.center, b.radius, sph.freestream, markers=sph.markers)
        if cyl:
            for cy in cyl.cylinders:
                self.ctx.add_cylinder(cy.center_xy[0], cy.center_xy[1], cy.radius, cyl.freestream)
        st = self._find(CfdStatePlugin)
        if st is None:
            raise ValueError("CfdStatePlugin missing")
        init = st.init
        if callable(init):
            x, y, z = self.cell_centroids()
            u = np.asarray(init(x, y, z), dtype=np.float64)
        else:
            u = np.asarray(init, dtype=np.float64)
        self.ctx.upload(np.broadcast_to(u, (m.nz, m.ny, m.nx, 5)).copy() if u.shape == (5,)
                        else u.reshape(m.nz, m.ny, m.nx, 5))
        return self

    def apply_boundary_inputs(self):
        """Push the BoundaryInputs resource to the FromOutside sides (call again after changing it)."""
        inputs = self.resources.get(BoundaryInputs)
        if inputs is None:
            return
        for s, (states, given) in inputs.sides.items():
            self.ctx.set_ghost_table(s, states, given)

    def cell_centroids(self):
        """x, y (2-D context) or x, y, z (3-D context) centroid arrays, x fastest."""
        m = self._find(FieldDefaultPlugins).mesh
        c = [m.centroids(d) for d in range(3)]
        if isinstance(self.ctx, Context3):
            z, y, x = np.meshgrid(c[2], c[1], c[0], indexing="ij")
            return x, y, z
        return np.meshgrid(c[0], c[1])

    # one app.run() = one RK substage
    def run(self):
        self.ctx.run()

    def step(self, n: int = 1):
        self.ctx.step(n)

    def advance_to(self, t_final: float, max_steps: int = 10 ** 9) -> int:
        if isinstance(self.ctx, Context3):
            k = 0
            while self.ctx.clock()["time"] < t_final and k < max_steps:
                self.ctx.step(1)
                k += 1
            return k
        return self.ctx.advance_to(t_final, max_steps)

    def solver_state(self) -> SolverState:
        c = self.ctx.clock()
        return SolverState(c["dt"], c["time"], c["step"], c["stage"])

    def state(self) -> np.ndarray:
        """CfdState.u of the interior, shape (ny, nx, 5) ((nz, ny, nx, 5) in 3-D), Component wire order."""
        return self.ctx.download()

    def marker_force(self) -> np.ndarray:
        return self.ctx.marker_force()

    def wall_force(self) -> np.ndarray:
        """WallForce (pressure xyz, viscous xyz), the staircase estimate (cfd_ibm:309-362, 383-395)."""
        return self.ctx.wall_force()

    def conservation_sums(self) -> np.ndarray:
        return self.ctx.sums()

    # restart files (examples/restart_determinism/main.rs:177-226 write_checkpoint / read_checkpoint)
    def write_checkpoint(self, path: str):
        self.ctx.checkpoint_save(path)

    def read_checkpoint(self, path: str):
        self.ctx.checkpoint_load(path)

    def finish(self):
        """The cleanup hooks: OutputPlugin writes its VTK file."""
        out = self._find(OutputPlugin)
        if out is not None:
            self.ctx.write_vtk(out.path)
