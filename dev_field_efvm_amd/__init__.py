"""MI355X-native explicit compressible finite-volume update (HIP, gfx950).

The product is ``libefvm.so`` (C ABI in include/efvm.h); this package holds
its ctypes binding (``efvm``), a host-side mirror of the reference's plugin
API (``app``) and the benchmark initial conditions (``cases``).
"""
from .efvm import Context, Context3, Batch, Mlp, EfvmError, load_library, write_vtk, write_vtk3, LIB_PATH, EXPORTS, NPARAM  # noqa: F401
from .app import (  # noqa: F401
    App, UniformMeshConfig, FieldDefaultPlugins, CfdStatePlugin, IdealGas, IdealGasPlugin, GasPlugin,
    Viscosity, SlopeLimiter, FluxPlugin, MusclPlugin, IntegratorPlugin, SolverConfig, SolverPlugin,
    BoundarySide, SupersonicOutflow, Periodic, SlipWall, Symmetry, NoSlipWall, SupersonicInflow,
    SubsonicInflow, PressureOutlet, BoundaryRegistry, BoundaryPlugin, StructuredBoundaryPlugin,
    Cylinder, GhostCellIbmPlugin, CylinderIbmPlugin, ParabolicInflow, OutletSponge, BodyForceX,
    GpuPlugin, SolverState, OutputPlugin, FromOutside, FnBoundary, BoundaryInputs, SphereBody, SphereIbmPlugin,
    Sponge,
)
from . import cases  # noqa: F401
