"""GPU parity: the HIP path (through the C ABI) against the CPU FP64 oracle on the
same inputs.  FP64 mode must match to 1e-12 relative L-infinity (per component,
normalised by the oracle's max |component|); dt and the clock must agree
bit for bit.  FP32 mode is compared against FP64 with its tolerance stated.
"""
import math

import numpy as np
import pytest

import dev_field_efvm_amd as efvm
from dev_field_efvm_amd import cases
from oracle import orc

pytestmark = pytest.mark.gpu
TOL64 = 1e-12
R = orc.R_AIR


def rel_linf(gpu, ref):
    """max over components of max|gpu - ref| / max|ref| (rho_w skipped: identically 0)."""
    worst = 0.0
    for m in (0, 1, 2, 4):
        scale = max(np.max(np.abs(ref[..., m])), 1e-300)
        worst = max(worst, float(np.max(np.abs(gpu[..., m] - ref[..., m])) / scale))
    return worst


def make_pair(nx, ny, hi, bcs, flux="hllc", muscl=False, limiter="minmod", cfl=0.4, integrator="rk3",
              viscous=False, mu=None, seam_mode=0, precision="fp64", gamma=1.4, lo=(0.0, 0.0), body_gx=None,
              fixed_dt=None):
    """An oracle Sim and a GPU Context with identical configuration (z = one slip cell)."""
    dz = hi[2]
    sim = orc.Sim(nx, ny, 1, lo=(lo[0], lo[1], 0.0), hi=hi, gamma=gamma, flux=flux, muscl=muscl, limiter=limiter,
                  integrator=integrator, viscous=viscous, mu=mu, cfl=cfl, bcs=list(bcs) + ["slip", "slip"],
                  seam_mode=seam_mode, body_gx=body_gx, fixed_dt=fixed_dt)
    ctx = efvm.Context(nx, ny, lo=lo, hi=(hi[0], hi[1]), dz=dz, gamma=gamma, flux=flux, muscl=muscl,
                       limiter=limiter, integrator=integrator, precision=precision, viscous=viscous, mu=mu,
                       seam_mode=seam_mode, cfl=cfl, fixed_dt=fixed_dt)
    ctx.set_boundaries(list(bcs))
    if body_gx is not None:
        ctx.set_body_force(body_gx)
    return sim, ctx


def load_both(sim, ctx, state):
    sim.set_interior(state.reshape(1, ctx.ny, ctx.nx, 5))
    ctx.upload(state)


def gpu_state(ctx):
    return ctx.download()


def ora_state(sim):
    return sim.interior()[0]


# ------------------------------------------------------------------ config 1
def test_config1_vortex_first_order_periodic_fp64_matches_oracle_to_t2():
    """64x64 periodic vortex, 1st-order HLLC + SSP-RK3, CFL 0.4, FP64, to t = 2."""
    n = 64
    sim, ctx = make_pair(n, n, (10.0, 10.0, 10.0 / n), ["periodic"] * 4)
    load_both(sim, ctx, cases.vortex_state(n, n, mean_u=1.0, mean_v=1.0, periodic=True))
    steps = 0
    while sim.info()["time"] < 2.0:
        sim.step(1)
        ctx.step(1)
        steps += 1
        if steps % 20 == 0:
            co, cg = sim.info(), ctx.clock()
            assert cg["dt"] == co["dt"] and cg["time"] == co["time"], (steps, co, cg)
    co, cg = sim.info(), ctx.clock()
    assert cg["step"] == co["step"] and cg["time"] == co["time"] and cg["dt"] == co["dt"]
    err = rel_linf(gpu_state(ctx), ora_state(sim))
    print(f"config1 steps={steps} t={co['time']:.6f} rel_linf={err:.3e}")
    assert err <= TOL64


# ------------------------------------------------------------------ MUSCL variants
@pytest.mark.parametrize("seam_mode", [0, 1])
@pytest.mark.parametrize("limiter", ["mc", "minmod"])
def test_muscl_periodic_vortex_modes_fp64(seam_mode, limiter):
    """The config-2 physics (periodic vortex + 8 seeded modes, MUSCL + HLLC + RK3) at 128^2, 30 steps."""
    n = 128
    sim, ctx = make_pair(n, n, (10.0, 10.0, 10.0 / n), ["periodic"] * 4, muscl=True, limiter=limiter, seam_mode=seam_mode)
    load_both(sim, ctx, cases.vortex_modes_state(n, n, seed=1))
    sim.step(30)
    ctx.step(30)
    co, cg = sim.info(), ctx.clock()
    assert cg["time"] == co["time"]
    err = rel_linf(gpu_state(ctx), ora_state(sim))
    print(f"muscl {limiter} seam={seam_mode} rel_linf={err:.3e}")
    assert err <= TOL64


def test_muscl_outflow_vortex_reference_gate_setup_fp64():
    """The reference's MUSCL order-gate case (outflow x/y, CFL 0.3, mean (0.5,0.5)) at 40^2 to t = 1."""
    n = 40
    sim, ctx = make_pair(n, n, (10.0, 10.0, 10.0 / n), ["outflow"] * 4, muscl=True, limiter="mc", cfl=0.3)
    st = cases.vortex_state(n, n, mean_u=0.5, mean_v=0.5, periodic=False)
    load_both(sim, ctx, st)
    sim.advance_to(1.0)
    ctx.advance_to(1.0)
    assert ctx.clock()["time"] == sim.info()["time"]
    assert rel_linf(gpu_state(ctx), ora_state(sim)) <= TOL64


def test_muscl_order_gate_on_gpu():
    """crates/cfd_solver/tests/isentropic_vortex_muscl_convergence.rs:146-192 run on the GPU (FP64)."""
    ns, errs = [40, 80, 160], []
    for n in ns:
        ctx = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, muscl=True, limiter="mc", cfl=0.3)
        ctx.set_boundaries(["outflow"] * 4)
        ctx.upload(cases.vortex_state(n, n, mean_u=0.5, mean_v=0.5, periodic=False))
        ctx.advance_to(1.0)
        t = ctx.clock()["time"]
        rho = ctx.download()[..., 0]
        x, y = cases.cell_centroids(n, n, (0, 0), (10, 10))
        ex = cases.vortex_prim(x, y, t, mean_u=0.5, mean_v=0.5, periodic=False)[0]
        errs.append(float(np.mean(np.abs(rho - ex))))
    xs, ys = np.log(ns), np.log(errs)
    order = -np.polyfit(xs, ys, 1)[0]
    print("gpu muscl errs", errs, "order", order)
    assert order >= 1.8 and errs[-1] < 3e-4


@pytest.mark.parametrize("flux", ["rusanov", "hllc_textbook"])
def test_other_flux_schemes_fp64(flux):
    n = 48
    sim, ctx = make_pair(n, n, (10.0, 10.0, 10.0 / n), ["periodic"] * 4, flux=flux, muscl=True, limiter="mc")
    load_both(sim, ctx, cases.vortex_modes_state(n, n, seed=3))
    sim.step(20)
    ctx.step(20)
    assert rel_linf(gpu_state(ctx), ora_state(sim)) <= TOL64


def test_euler_integrator_and_ragged_sizes_fp64():
    """Euler substeps, a grid that is not a multiple of the strip shape, mixed BCs."""
    nx, ny = 61, 37
    sim, ctx = make_pair(nx, ny, (6.1, 3.7, 0.1), ["outflow", "pressure_outlet", "slip", "noslip"], muscl=True,
                         limiter="minmod", integrator="euler", cfl=0.3)
    ctx.set_boundaries(["outflow", ("pressure_outlet", (0.9,)), "slip", "noslip"])
    sim = orc.Sim(nx, ny, 1, hi=(6.1, 3.7, 0.1), flux="hllc", muscl=True, limiter="minmod", integrator="euler",
                  cfl=0.3, bcs=["outflow", "pressure_outlet", "slip", "noslip", "slip", "slip"],
                  bc_params=[None, (0.9,), None, None, None, None])
    st = cases.vortex_modes_state(nx, ny, seed=5, L=10.0)
    load_both(sim, ctx, st)
    sim.step(25)
    ctx.step(25)
    assert ctx.clock()["time"] == sim.info()["time"]
    assert rel_linf(gpu_state(ctx), ora_state(sim)) <= TOL64


def test_inflow_boundaries_fp64():
    nx, ny = 70, 20
    sim = orc.Sim(nx, ny, 1, hi=(1.0, 0.3, 0.01), flux="hllc", muscl=True, limiter="mc", cfl=0.4,
                  bcs=["supersonic_inflow", "outflow", "subsonic_inflow", "symmetry", "slip", "slip"],
                  bc_params=[(1.2, 500.0, 10.0, 0.0, 1.1e5), None, (1.0, 20.0, 0.0, 0.0, 300.0), None, None, None])
    ctx = efvm.Context(nx, ny, hi=(1.0, 0.3), dz=0.01, muscl=True, limiter="mc", cfl=0.4)
    ctx.set_boundaries([("supersonic_inflow", (1.2, 500.0, 10.0, 0.0, 1.1e5)), "outflow",
                        ("subsonic_inflow", (1.0, 20.0, 0.0, 0.0, 300.0)), "symmetry"])
    x, y = cases.cell_centroids(nx, ny, (0, 0), (1.0, 0.3))
    rho = 1.0 + 0.2 * np.sin(6 * x) * np.cos(9 * y)
    st = cases.prim_to_cons(1.4, rho, 100.0 + 0 * x, 5.0 * np.sin(3 * x), 1e5 + 1e3 * np.cos(4 * y))
    load_both(sim, ctx, st)
    sim.step(15)
    ctx.step(15)
    assert rel_linf(gpu_state(ctx), ora_state(sim)) <= TOL64


@pytest.mark.parametrize("muscl", [False, True])
def test_ghost_table_boundaries_fp64(muscl):
    """FromOutside / BoundaryInputs (crates/cfd_boundary/src/lib.rs:108-147) on the device: an x-lo
    inlet profile from the face centroids (boundary_face.rs:77-94, FnBoundary as a table) and an
    x-hi side with every other row supplied (:96-118, the rest copying the interior), against
    the oracle; then a new table between steps, replayed from a captured graph."""
    nx, ny = 40, 16
    sim, ctx = make_pair(nx, ny, (1.0, 1.0, 0.05), ["table", "table", "slip", "slip"], muscl=muscl, limiter="mc")
    x, y = cases.cell_centroids(nx, ny, (0, 0), (1.0, 1.0))
    p0 = 1e5
    st = cases.prim_to_cons(1.4, 1.0 + 0.1 * np.sin(5 * x), 30.0 + 0 * x, 0 * x, p0 + 0 * x)
    load_both(sim, ctx, st)
    yc = (np.arange(ny) + 0.5) / ny
    inlet = np.array([orc.prim_to_cons([1.1, 400.0 * yy * (1.0 - yy), 0.0, 0.0, p0, p0 / (1.1 * R)], 1.4) for yy in yc])
    given = np.arange(ny) % 2 == 0
    outside = np.array([orc.prim_to_cons([0.9 + 0.01 * j, -20.0, 3.0, 0.0, 0.95 * p0, 300.0], 1.4) for j in range(ny)])
    for obj in (sim, ctx):
        obj.set_ghost_table(0, inlet, np.ones(ny, bool))
        obj.set_ghost_table(1, outside, given)
    sim.step(20)
    ctx.step(20)
    assert rel_linf(gpu_state(ctx), ora_state(sim)) <= TOL64
    # a new outside state, picked up by the captured graph in place
    ctx.graph_capture(4)
    outside2 = outside.copy()
    outside2[:, 1] *= -1.0
    for obj in (sim, ctx):
        obj.set_ghost_table(1, outside2, ~given)
    sim.step(4)
    ctx.graph_launch()
    ctx.synchronize()
    assert rel_linf(gpu_state(ctx), ora_state(sim)) <= TOL64


def test_app_from_outside_and_fn_boundary():
    """The plugin mirror of boundary_face.rs: FnBoundary inlet at x-lo, FromOutside with
    BoundaryInputs at x-hi (rows 0 and 2 given), HLLC first order, Euler, dt 1e-9."""
    nx, ny = 6, 4
    p0 = 1e5
    air = lambda rho, u: orc.prim_to_cons([rho, u, 0.0, 0.0, p0, p0 / (rho * R)], 1.4)
    app = efvm.App()
    app.add_plugins(efvm.FieldDefaultPlugins(efvm.UniformMeshConfig(nx=nx, ny=ny, nz=1, ng=2, bounds_lo=(0, 0, 0),
                                                                    bounds_hi=(1.0, 1.0, 0.1))))
    app.add_plugins(efvm.CfdStatePlugin(np.tile(air(1.0, 0.0), (ny, nx, 1))))
    reg = (efvm.BoundaryRegistry()
           .with_(efvm.BoundarySide.XLo, efvm.FnBoundary(lambda c: air(1.0, 400.0 * c[1] * (1.0 - c[1]))))
           .with_(efvm.BoundarySide.XHi, efvm.FromOutside()).with_axis(1, efvm.SlipWall()).with_axis(2, efvm.SlipWall()))
    app.add_plugins(efvm.StructuredBoundaryPlugin(reg))
    app.add_plugins(efvm.FluxPlugin.hllc())
    app.add_plugins(efvm.IntegratorPlugin.euler())
    app.add_plugins(efvm.SolverPlugin(efvm.SolverConfig(cfl=0.4, fixed_dt=1e-9)))
    given = [j % 2 == 0 for j in range(ny)]
    outside = [air(1.0 + 0.1 * j, -50.0) for j in range(ny)]
    app.add_resource(efvm.BoundaryInputs().set(efvm.BoundarySide.XHi, outside, given))
    app.prepare()
    app.run()
    sim = orc.Sim(nx, ny, 1, hi=(1.0, 1.0, 0.1), flux="hllc", integrator="euler", fixed_dt=1e-9,
                  bcs=["table", "table", "slip", "slip", "slip", "slip"])
    sim.set_interior(np.tile(air(1.0, 0.0), (1, ny, nx, 1)))
    y = (np.arange(ny) + 0.5) / ny
    sim.set_ghost_table(0, [air(1.0, 400.0 * yy * (1.0 - yy)) for yy in y], [True] * ny)
    sim.set_ghost_table(1, outside, given)
    sim.run()
    assert rel_linf(app.state(), ora_state(sim)) <= TOL64
    assert not np.array_equal(app.state()[:, 0], app.state()[:, 2])   # the inlet did act on the first column


def test_sod_tube_ny1_fp64():
    """crates/cfd_solver/tests/sod_exact_convergence.rs setup (ny = 1) on the GPU."""
    n = 200
    sim, ctx = make_pair(n, 1, (1.0, 0.01, 0.01), ["outflow", "outflow", "slip", "slip"])
    left = orc.prim_to_cons([1.0, 0, 0, 0, 1e5, 1e5 / R], 1.4)
    right = orc.prim_to_cons([0.125, 0, 0, 0, 1e4, 1e4 / (0.125 * R)], 1.4)
    x = (np.arange(n) + 0.5) / n
    st = np.where((x < 0.5)[:, None], left[None, :], right[None, :]).reshape(1, n, 5)
    load_both(sim, ctx, st)
    sim.advance_to(6e-4)
    ctx.advance_to(6e-4)
    assert rel_linf(gpu_state(ctx), ora_state(sim)) <= TOL64


def test_poiseuille_viscous_noslip_body_force_fp64():
    """crates/cfd_solver/tests/poiseuille_channel_validation.rs setup, 300 RK3 steps."""
    p0 = R * 300.0
    sim, ctx = make_pair(3, 32, (1.0, 1.0, 1.0), ["outflow", "outflow", "noslip", "noslip"], viscous=True, mu=10.0,
                         cfl=0.3, body_gx=160.0)
    st = np.tile(orc.prim_to_cons([1.0, 0, 0, 0, p0, 300.0], 1.4), (32, 3, 1))
    load_both(sim, ctx, st)
    sim.step(300)
    ctx.step(300)
    assert ctx.clock()["time"] == sim.info()["time"]
    assert rel_linf(gpu_state(ctx), ora_state(sim)) <= TOL64


def test_viscous_sutherland_muscl_fp64():
    """Sutherland mu(T) through pow(): the device's pow rounds like the host's on these
    temperatures (measured: the states agree bit for bit after 20 steps), so the case is held
    to the common 1e-12 bound."""
    n = 40
    sim, ctx = make_pair(n, n, (1.0, 1.0, 0.025), ["noslip", "noslip", "noslip", "noslip"], viscous=True,
                         muscl=True, limiter="mc", cfl=0.3)
    x, y = cases.cell_centroids(n, n, (0, 0), (1, 1))
    st = cases.prim_to_cons(1.4, 1.0 + 0 * x, 30.0 * np.sin(np.pi * y), 10.0 * np.sin(np.pi * x), 1e5 + 0 * x)
    load_both(sim, ctx, st)
    sim.step(20)
    ctx.step(20)
    err = rel_linf(gpu_state(ctx), ora_state(sim))
    print(f"sutherland viscous MUSCL rel L-inf {err:.3e}")
    assert err <= TOL64


# ------------------------------------------------------------------ cylinder (ghost-cell IBM)
def cylinder_setup(nx, ny, precision="fp64", muscl=False):
    """examples/cylinder_drag_validation (config.toml at D/dx = 20: 440 x 82), case systems included."""
    D, H, L, lz = 0.1, 0.41, 2.2, 0.041
    rho0, p0, u_mean = 1.0, 1.0e5, 50.0
    mu = rho0 * u_mean * D / 20.0
    u_max = 1.5 * u_mean
    fs = orc.prim_to_cons([rho0, u_mean, 0, 0, p0, p0 / (rho0 * R)], 1.4)
    common = dict(flux="hllc", muscl=muscl, limiter="mc", cfl=0.3, viscous=True, mu=mu)
    sim = orc.Sim(nx, ny, 1, hi=(L, H, lz), bcs=["outflow", "pressure_outlet", "noslip", "noslip", "slip", "slip"],
                  bc_params=[None, (p0,), None, None, None, None], parabolic_inflow=(u_max, H, rho0),
                  sponge=(1.4, L, 1.0e4, u_max, H, rho0, p0), cylinders=[(0.2, 0.2, 0.5 * D)], freestream=fs, **common)
    sim.ic_channel(u_max, H, rho0, p0)
    ctx = efvm.Context(nx, ny, hi=(L, H), dz=lz, precision=precision, **common)
    ctx.set_boundaries(["outflow", ("pressure_outlet", (p0,)), "noslip", "noslip"])
    ctx.set_parabolic_inflow(u_max, H, rho0)
    ctx.set_sponge(1.4, L, 1.0e4, u_max, H, rho0, p0)
    ctx.add_cylinder(0.2, 0.2, 0.5 * D, fs)
    ctx.upload(sim.interior()[0])
    q = 0.5 * rho0 * u_mean * u_mean * D * lz
    return sim, ctx, q


def test_cylinder_ibm_first_order_fp64_matches_oracle():
    """Reference's own 440 x 82 first-order setup: 20 steps, state and marker forces."""
    sim, ctx, q = cylinder_setup(440, 82)
    for _ in range(20):
        sim.step(1)
        ctx.step(1)
    assert ctx.clock()["time"] == sim.info()["time"]
    assert rel_linf(gpu_state(ctx), ora_state(sim)) <= TOL64
    fo, fg = sim.marker_force(), ctx.marker_force()
    print("marker force oracle", fo, "gpu", fg, "Cd", (fo[0] + fo[3]) / q)
    assert np.allclose(fg, fo, rtol=1e-11, atol=1e-12 * np.max(np.abs(fo)))


def test_cylinder_ibm_muscl_fp64_matches_oracle():
    sim, ctx, q = cylinder_setup(220, 41, muscl=True)
    for _ in range(15):
        sim.step(1)
        ctx.step(1)
    assert rel_linf(gpu_state(ctx), ora_state(sim)) <= TOL64


def wall_force_scales(sim, gamma=1.4):
    """Bounds on the summed magnitudes of the staircase wall force's pressure and shear terms:
    the largest pressure (resp. mu |u| / (dn/2)) of the state times the solid-fluid face area."""
    st = sim.interior()[0]
    rho, mom, en = st[..., 0], st[..., 1:4], st[..., 4]
    p = (gamma - 1.0) * (en - 0.5 * (mom ** 2).sum(-1) / rho)
    speed = np.sqrt((mom ** 2).sum(-1)) / rho
    solid = sim.solid()[2, 2:-2, 2:-2].astype(bool)     # the one z layer of the raw mask, interior
    nx, ny = solid.shape[1], solid.shape[0]
    lx, ly, lz = sim.h
    faces_x = (solid[:, :-1] != solid[:, 1:]).sum()
    faces_y = (solid[:-1, :] != solid[1:, :]).sum()
    area = faces_x * ly * lz + faces_y * lx * lz + 2 * solid.sum() * lx * ly
    mu = sim.cfg.mu_const if sim.cfg.visc_model == 1 else 1.0e-4
    return np.abs(p).max() * area, mu * speed.max() / (0.5 * min(lx, ly, lz)) * area


def check_wall_force(sim, ctx, what, rtol=1e-11):
    fo, fg = sim.wall_force(), ctx.wall_force()
    sp, sv = wall_force_scales(sim)
    dp, dv = np.abs(fg[:3] - fo[:3]).max() / sp, np.abs(fg[3:] - fo[3:]).max() / sv
    print(f"{what}: wall force oracle {fo} gpu {fg} rel. to term scale: pressure {dp:.2e} shear {dv:.2e}")
    assert dp <= rtol and dv <= rtol, what
    return fo, fg


def test_cylinder_wall_force_fp64_matches_oracle():
    """Staircase wall force (crates/cfd_ibm/src/lib.rs:309-362) on the cylinder case, after whole
    steps and after each RK3 substage: the device sums the faces in the oracle's order, and its
    ghost cells (z slip images of the solid cells, the walls' boundary ghosts) carry the fills and
    updates of the substages since the step boundary, as the reference's raw array does.
    (Measured: bitwise equal to the oracle at every check.)"""
    sim, ctx, q = cylinder_setup(220, 41, muscl=True)
    assert np.array_equal(ctx.wall_force(), np.zeros(6)) and np.array_equal(sim.wall_force(), np.zeros(6))
    for k in range(6):
        sim.step(1)
        ctx.step(1)
        fo, _ = check_wall_force(sim, ctx, f"step {k + 1}")
    assert fo[0] > 0.0 and abs(fo[2]) == 0.0        # drag along the flow; the two z faces cancel exactly
    for s in range(3):
        sim.run()
        ctx.run()
        check_wall_force(sim, ctx, f"substage {s + 1}")


def test_cylinder_wall_force_euler_and_first_order_fp64():
    """Forward Euler keeps one substage per step: the ghosts are the fill of U^n."""
    sim, ctx, q = cylinder_setup(220, 41)
    for k in range(4):
        sim.step(1)
        ctx.step(1)
    check_wall_force(sim, ctx, "rk3 first order")
    D, H, L, lz = 0.1, 0.41, 2.2, 0.041
    rho0, p0, u_mean = 1.0, 1.0e5, 50.0
    fs = orc.prim_to_cons([rho0, u_mean, 0, 0, p0, p0 / (rho0 * R)], 1.4)
    common = dict(flux="rusanov", cfl=0.3, viscous=True, mu=0.1, integrator="euler")
    sim = orc.Sim(110, 21, 1, hi=(L, H, lz), bcs=["supersonic_inflow", "outflow", "slip", "slip", "slip", "slip"],
                  bc_params=[(rho0, u_mean, 0.0, 0.0, p0), None, None, None, None, None], cylinders=[(0.4, 0.2, 0.5 * D)],
                  freestream=fs, **common)
    sim.set_interior(np.broadcast_to(fs, (1, 21, 110, 5)).copy())
    ctx = efvm.Context(110, 21, hi=(L, H), dz=lz, **common)
    ctx.set_boundaries([("supersonic_inflow", (rho0, u_mean, 0.0, 0.0, p0)), "outflow", "slip", "slip"])
    ctx.add_cylinder(0.4, 0.2, 0.5 * D, fs)
    ctx.upload(sim.interior()[0])
    for k in range(5):
        sim.step(1)
        ctx.step(1)
        check_wall_force(sim, ctx, f"euler step {k + 1}")


def test_wall_force_of_bodies_cut_by_the_boundaries_fp64():
    """Two cylinders cut by the periodic x seam and the no-slip y-lo wall: solid cells whose
    neighbours are periodic images of solid cells and wall ghosts of themselves, both taken as
    the substages' fills saw them (before the IBM mirror overwrote the solid cells)."""
    rho0, p0, u0 = 1.0, 1.0e5, 40.0
    fs = orc.prim_to_cons([rho0, u0, 0, 0, p0, p0 / (rho0 * R)], 1.4)
    nx, ny, L, H, lz = 100, 50, 1.0, 0.5, 0.01
    cyl = [(0.0, 0.0, 0.08), (L, 0.0, 0.08), (0.5, 0.25, 0.06)]
    common = dict(flux="hllc", muscl=True, limiter="minmod", cfl=0.3, viscous=True, mu=0.02)
    sim = orc.Sim(nx, ny, 1, hi=(L, H, lz), bcs=["periodic", "periodic", "noslip", "slip", "slip", "slip"],
                  cylinders=cyl, freestream=fs, **common)
    sim.set_interior(np.broadcast_to(fs, (1, ny, nx, 5)).copy())
    ctx = efvm.Context(nx, ny, hi=(L, H), dz=lz, **common)
    ctx.set_boundaries(["periodic", "periodic", "noslip", "slip"])
    for c in cyl:
        ctx.add_cylinder(*c, fs)
    ctx.upload(sim.interior()[0])
    for k in range(4):
        sim.step(1)
        ctx.step(1)
        check_wall_force(sim, ctx, f"step {k + 1}")
    assert rel_linf(gpu_state(ctx), ora_state(sim)) <= TOL64
    for s in range(2):
        sim.run()
        ctx.run()
        check_wall_force(sim, ctx, f"substage {s + 1}")


def test_fp32_cylinder_wall_force_tracks_fp64():
    out = {}
    for prec in ("fp64", "fp32"):
        sim, ctx, q = cylinder_setup(220, 41, precision=prec)
        ctx.step(20)
        out[prec] = ctx.wall_force()
    sim.step(20)
    sp, sv = wall_force_scales(sim)
    d = np.abs(out["fp32"] - out["fp64"])
    print("fp32", out["fp32"], "fp64", out["fp64"], "rel. to term scale", d[:3].max() / sp, d[3:].max() / sv)
    assert d[:3].max() <= 1e-5 * sp and d[3:].max() <= 1e-5 * sv


# ------------------------------------------------------------------ the general two-column kernel
def rough_state(nx, ny, seed, hi=(1.0, 1.0)):
    """Positive density and pressure with seeded modes in all four components and a jump."""
    rng = np.random.default_rng(seed)
    y, x = np.meshgrid((np.arange(ny) + 0.5) / ny * hi[1], (np.arange(nx) + 0.5) / nx * hi[0], indexing="ij")
    def field(base, amp):
        f = np.full_like(x, base)
        for _ in range(3):
            k = rng.integers(1, 4, size=2)
            f += amp * np.sin(2 * np.pi * k[0] * x / hi[0] + rng.uniform(0, 6.3)) * np.cos(2 * np.pi * k[1] * y / hi[1])
        return f
    rho = field(1.0, 0.1) + 0.3 * (x > 0.6 * hi[0])
    u, v, p = field(0.3, 0.2), field(-0.2, 0.2), field(1.0, 0.1)
    return np.stack([rho, rho * u, rho * v, 0 * x, p / 0.4 + 0.5 * rho * (u * u + v * v)], axis=-1)


GENERAL_PAIR_CASES = [
    dict(n=(130, 37), bcs=["outflow"] * 4, flux="hllc", muscl=True, limiter="mc", integrator="rk3"),
    dict(n=(131, 40), bcs=["slip", "noslip", "symmetry", "outflow"], flux="rusanov", muscl=True, limiter="minmod",
         integrator="rk3"),
    dict(n=(128, 33), bcs=["periodic", "periodic", "slip", "noslip"], flux="hllc", muscl=True, limiter="mc",
         integrator="rk3", seam_mode=1),
    dict(n=(125, 36), bcs=["outflow", "noslip", "periodic", "periodic"], flux="hllc_textbook", muscl=False,
         integrator="rk3"),
    dict(n=(253, 21), bcs=["noslip", "outflow", "outflow", "slip"], flux="hllc", muscl=True, limiter="mc",
         integrator="euler"),
    dict(n=(5, 4), bcs=["outflow", "slip", "noslip", "symmetry"], flux="hllc", muscl=True, limiter="minmod",
         integrator="rk3"),
]


@pytest.mark.parametrize("cfg", GENERAL_PAIR_CASES, ids=lambda c: f"{c['n'][0]}x{c['n'][1]}-{'-'.join(c['bcs'])}")
def test_general_two_column_kernel_is_bitwise_the_one_column_kernel(cfg):
    """Outflow, slip, no-slip and symmetry sides, odd widths, a periodic axis beside walls, first
    order, Euler: the two-column kernel's ghost columns and rows (owner values reflected as each
    side's condition says), its edge columns' first-order boundary faces and its one-column
    stores of an odd row's last pair give bitwise the one-column kernel's states and clocks."""
    nx, ny = cfg["n"]
    kw = {k: v for k, v in cfg.items() if k not in ("n", "bcs")}
    out = []
    for allow in (True, False):
        c = efvm.Context(nx, ny, hi=(1.0, 1.0), dz=1.0 / nx, precision="fp32", cfl=0.4, **kw)
        c.set_boundaries(cfg["bcs"])
        c.set_pair_kernel(allow)
        kernel = c.launch()[2]
        assert ("pair" in kernel) == allow, kernel
        c.upload(rough_state(nx, ny, seed=nx + ny))
        c.step(9)
        c.run()                     # and one lone substage
        out.append((c.download(), c.clock()))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


# ------------------------------------------------------------------ FP32 production path
def test_fp32_error_against_fp64_config2_physics():
    """FP32 (flux/update arithmetic) vs the FP64 path on the config-2 physics, 256^2, 50 steps.
    Stated tolerance: relative L-inf <= 1e-4 of max|component|; density L1 <= 1e-5."""
    n = 256
    st = cases.vortex_modes_state(n, n, seed=1)
    out = {}
    for prec in ("fp64", "fp32"):
        ctx = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, muscl=True, limiter="mc", precision=prec)
        ctx.set_boundaries(["periodic"] * 4)
        ctx.upload(st)
        ctx.step(50)
        out[prec] = (ctx.download(), ctx.clock())
    e = rel_linf(out["fp32"][0], out["fp64"][0])
    l1 = float(np.mean(np.abs(out["fp32"][0][..., 0] - out["fp64"][0][..., 0])))
    print(f"fp32 vs fp64: rel_linf={e:.3e} rho_L1={l1:.3e} t32={out['fp32'][1]['time']} t64={out['fp64'][1]['time']}")
    assert e <= 1e-4 and l1 <= 1e-5
    assert abs(out["fp32"][1]["time"] / out["fp64"][1]["time"] - 1.0) < 1e-5


def test_fp32_first_order_and_rusanov_close_to_fp64():
    n = 96
    st = cases.vortex_state(n, n, mean_u=1.0, mean_v=1.0)
    for flux in ("hllc", "rusanov"):
        res = {}
        for prec in ("fp64", "fp32"):
            ctx = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, flux=flux, precision=prec)
            ctx.set_boundaries(["periodic"] * 4)
            ctx.upload(st)
            ctx.step(40)
            res[prec] = ctx.download()
        assert rel_linf(res["fp32"], res["fp64"]) <= 1e-4


def test_fp32_cylinder_runs_and_tracks_fp64():
    out = {}
    for prec in ("fp64", "fp32"):
        sim, ctx, q = cylinder_setup(220, 41, precision=prec)
        ctx.step(30)
        out[prec] = (ctx.download(), ctx.marker_force())
    assert rel_linf(out["fp32"][0], out["fp64"][0]) <= 1e-4
    assert abs(out["fp32"][1][0] - out["fp64"][1][0]) <= 1e-4 * abs(out["fp64"][1][0])


# ------------------------------------------------------------------ invariants and API behaviour
def test_periodic_mass_conservation_fp64():
    n = 128
    ctx = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, muscl=True, limiter="mc", seam_mode=1)
    ctx.set_boundaries(["periodic"] * 4)
    ctx.upload(cases.vortex_modes_state(n, n, seed=2))
    m0 = ctx.sums()[0]
    ctx.step(50)
    assert abs(ctx.sums()[0] - m0) / m0 < 1e-12


def test_graph_replay_equals_eager():
    n = 200
    st = cases.vortex_modes_state(n, n, seed=4)
    for prec in ("fp64", "fp32"):
        eager = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, muscl=True, limiter="mc", precision=prec)
        eager.set_boundaries(["periodic"] * 4)
        eager.upload(st)
        graph = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, muscl=True, limiter="mc", precision=prec)
        graph.set_boundaries(["periodic"] * 4)
        graph.upload(st)
        graph.graph_capture(4)
        for rep in range(3):
            eager.step(4)
            graph.graph_launch()
            graph.synchronize()
            a, b = eager.download(), graph.download()
            assert np.array_equal(a, b), (prec, rep)
            assert eager.clock()["time"] == graph.clock()["time"]


def test_graph_replay_after_upload_and_eager_work():
    """A captured graph replays from the state the context holds now: after an upload (the CFL
    maximum of the new state is computed first), after eager whole steps, and it refuses to start
    mid-step or from the other Euler plane set, where it would read the wrong planes or CFL slot."""
    n = 96
    st1, st2 = cases.vortex_modes_state(n, n, seed=11), cases.vortex_modes_state(n, n, seed=12)
    for prec in ("fp64", "fp32"):
        graph = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, muscl=True, limiter="mc", precision=prec)
        eager = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, muscl=True, limiter="mc", precision=prec)
        for c in (graph, eager):
            c.set_boundaries(["periodic"] * 4)
            c.upload(st1)
        graph.graph_capture(2)
        graph.graph_launch()
        eager.step(2)
        # a new state: the replayed first step must take its dt from this state
        graph.upload(st2)
        eager.upload(st2)
        graph.graph_launch()
        eager.step(2)
        assert np.array_equal(graph.download(), eager.download()), prec
        assert graph.clock()["dt"] == eager.clock()["dt"]
        # eager whole steps in between (an odd count moves RK3's CFL slot: refused, then an even count)
        graph.step(1)
        eager.step(1)
        with pytest.raises(efvm.EfvmError):
            graph.graph_launch()
        graph.step(1)
        eager.step(1)
        graph.graph_launch()
        eager.step(2)
        assert np.array_equal(graph.download(), eager.download()), prec
        # a lone substage leaves the context mid-step: refused
        graph.run()
        with pytest.raises(efvm.EfvmError):
            graph.graph_launch()
        graph.close()
        eager.close()
    # Euler: an odd number of eager steps leaves U^n on the other plane set
    e = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, integrator="euler", precision="fp64")
    e.set_boundaries(["periodic"] * 4)
    e.upload(st1)
    e.graph_capture(2)
    e.step(1)
    with pytest.raises(efvm.EfvmError):
        e.graph_launch()
    e.close()


def test_run_substages_equal_step():
    n = 64
    st = cases.vortex_modes_state(n, n, seed=6)
    a = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, muscl=True)
    b = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, muscl=True)
    for c in (a, b):
        c.set_boundaries(["periodic"] * 4)
        c.upload(st)
    a.step(3)
    for _ in range(9):
        b.run()
    assert np.array_equal(a.download(), b.download())
    assert b.clock()["step"] == 3


def test_launch_shapes_give_identical_results():
    n = 150
    st = cases.vortex_modes_state(n, n, seed=7)
    ref = None
    for rows, waves in [(32, 4), (8, 1), (64, 2), (200, 3)]:
        c = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, muscl=True, limiter="mc")
        c.set_boundaries(["periodic"] * 4)
        c.set_launch(rows, waves)
        c.upload(st)
        c.step(5)
        s = c.download()
        if ref is None:
            ref = s
        assert np.array_equal(s, ref), (rows, waves)


# grids for the fused stage-2+3 launch: strips holding the x and y seams, a strip holding several
# periodic images of each column (nx < 120), ragged last strips, ny at the minimum of 8
FUSED_GRIDS = [(130, 64, "hllc", "mc"), (40, 12, "hllc", "mc"), (250, 138, "hllc", "minmod"),
               (64, 64, "rusanov", "mc"), (96, 40, "hllc", None), (8, 8, "hllc", "mc"), (246, 9, "hllc_textbook", "mc")]


@pytest.mark.parametrize("seam_mode", [0, 1])
def test_fused_step_is_bitwise_the_substage_launches(seam_mode):
    """Whole FP32 RK3 steps on the periodic two-column path run stage 1 and one fused stage-2+3
    launch (U2 kept in registers).  State and clock must be bitwise those of the three substage
    launches (fusion off), at several fused strip heights, in both seam modes (in seam mode 0 the
    fused launch evaluates each seam face as each side's owner sees it)."""
    for nx, ny, flux, lim in FUSED_GRIDS:
        st = cases.vortex_modes_state(nx, ny, seed=3, L=10.0)
        ref = None
        for fused, rows in [(False, 0), (True, 73), (True, 5), (True, 9), (True, 1)]:
            c = efvm.Context(nx, ny, hi=(10.0, 10.0), dz=10.0 / nx, flux=flux, muscl=lim is not None,
                             limiter=lim or "minmod", precision="fp32", seam_mode=seam_mode)
            c.set_boundaries(["periodic"] * 4)
            c.set_fused_step(fused, rows)
            assert c.fused_step()[2], (nx, ny)
            c.upload(st)
            c.step(7)
            out = (c.download(), c.clock())
            if ref is None:
                ref = out
                continue
            assert np.array_equal(out[0], ref[0]), (nx, ny, flux, lim, rows)
            assert out[1] == ref[1], (nx, ny, flux, lim, rows)


def test_fused_step_with_fixed_dt_and_euler_fallback():
    """A fixed time step compiles the fused launch without its CFL epilogue: bitwise the
    substage launches, clock included.  Euler steps and FP64 contexts never take the fused
    launch (fits is false) and run as before."""
    n = 130
    st = cases.vortex_modes_state(n, 64, seed=9)
    out = []
    for fused in (False, True):
        c = efvm.Context(n, 64, hi=(10.0, 10.0), dz=10.0 / n, muscl=True, limiter="mc", precision="fp32",
                         fixed_dt=2.0e-3)
        c.set_boundaries(["periodic"] * 4)
        c.set_fused_step(fused, 21)
        c.upload(st)
        c.step(5)
        out.append((c.download(), c.clock()))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1] and abs(out[1][1]["time"] - 5 * 2.0e-3) < 1e-15
    for kw in (dict(integrator="euler", precision="fp32"), dict(precision="fp64")):
        c = efvm.Context(n, 64, hi=(10.0, 10.0), dz=10.0 / n, muscl=True, limiter="mc", **kw)
        c.set_boundaries(["periodic"] * 4)
        assert not c.fused_step()[2], kw


def test_fused_step_graph_replay_and_mid_step_entry():
    """Graph replay of fused steps equals eager fused steps (several replays with eager work
    between); a step entered mid-step (after one run() substage) runs the separate launches and
    lands on the same state as whole steps."""
    n = 136
    st = cases.vortex_modes_state(n, n, seed=5)

    def ctx():
        c = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, muscl=True, limiter="mc", precision="fp32")
        c.set_boundaries(["periodic"] * 4)
        c.upload(st)
        return c

    a = ctx()
    a.step(12)
    b = ctx()
    b.graph_capture(4)
    b.graph_launch()
    b.step(2)
    b.graph_launch()
    b.step(2)
    assert np.array_equal(a.download(), b.download())
    assert a.clock() == b.clock()
    c = ctx()
    c.step(3)
    c.run(); c.run(); c.run()     # step 4 as three run() substages
    c.run()                       # stage 1 of step 5
    c.step(1)                     # from mid-step: stages 2, 3 of step 5 and stage 1 of step 6
    c.run(); c.run()              # the rest of step 6
    c.step(6)                     # whole (fused) steps 7 .. 12
    assert np.array_equal(a.download(), c.download())
    assert a.clock() == c.clock()


def test_error_paths():
    c = efvm.Context(16, 16)
    with pytest.raises(efvm.EfvmError) as e:
        c.set_boundaries(["periodic", "outflow", "outflow", "outflow"])
    assert e.value.status == 2
    st = np.zeros((16, 16, 5))
    st[..., 0] = 1.0
    st[..., 4] = 2.5
    st[3, 3, 3] = 1.0
    with pytest.raises(efvm.EfvmError) as e:
        c.upload(st)
    assert e.value.status == 4
    with pytest.raises(ValueError):
        c.upload(np.zeros((10, 5)))


def test_nonphysical_state_is_flagged():
    """Failure detection: a negative density planted in the uploaded state trips the CFL scan."""
    n = 32
    c = efvm.Context(n, n, hi=(1.0, 1.0), dz=1.0 / n)
    st = cases.vortex_state(n, n, L=1.0, beta=0.0, mean_u=0.1, mean_v=0.0)
    st[5, 7, 0] = -1.0
    c.upload(st)
    c.step(1)
    with pytest.raises(efvm.EfvmError) as e:
        c.clock()
    assert e.value.status == 5


def test_app_plugin_api_matches_context():
    """The plugin-style mirror builds the same solver as the raw context."""
    n = 64
    st = cases.vortex_state(n, n, mean_u=1.0, mean_v=1.0)
    app = efvm.App()
    app.add_plugins(efvm.FieldDefaultPlugins(efvm.UniformMeshConfig(nx=n, ny=n, nz=1, ng=2, bounds_lo=(0, 0, 0),
                                                                    bounds_hi=(10.0, 10.0, 10.0 / n))))
    app.add_plugins(efvm.CfdStatePlugin(st))
    app.add_plugins(efvm.IdealGasPlugin())
    app.add_plugins(efvm.StructuredBoundaryPlugin(efvm.BoundaryRegistry().with_axis(0, efvm.Periodic())
                                                  .with_axis(1, efvm.Periodic()).with_axis(2, efvm.SlipWall())))
    app.add_plugins(efvm.MusclPlugin.hllc().with_limiter(efvm.SlopeLimiter.MonotonizedCentral))
    app.add_plugins(efvm.IntegratorPlugin.rk3())
    app.add_plugins(efvm.SolverPlugin(efvm.SolverConfig(cfl=0.4)))
    app.prepare()
    for _ in range(6):
        app.run()
    c = efvm.Context(n, n, hi=(10.0, 10.0), dz=10.0 / n, muscl=True, limiter="mc")
    c.set_boundaries(["periodic"] * 4)
    c.upload(st)
    c.step(2)
    assert np.array_equal(app.state(), c.download())
    assert app.solver_state().step == 2


def test_app_output_plugin_and_restart(tmp_path):
    """OutputPlugin writes the VTK file at finish(); write_checkpoint / read_checkpoint resume
    an App bit for bit (crates/cfd_solver/tests/restart.rs:86-105, with the plugin mirror)."""
    n = 48
    st = cases.vortex_state(n, n, mean_u=1.0, mean_v=1.0)

    def build(out):
        app = efvm.App()
        app.add_plugins(efvm.FieldDefaultPlugins(efvm.UniformMeshConfig(nx=n, ny=n, nz=1, ng=2, bounds_lo=(0, 0, 0),
                                                                        bounds_hi=(10.0, 10.0, 10.0 / n))))
        app.add_plugins(efvm.CfdStatePlugin(st))
        app.add_plugins(efvm.StructuredBoundaryPlugin(efvm.BoundaryRegistry().with_axis(0, efvm.Periodic())
                                                      .with_axis(1, efvm.Periodic()).with_axis(2, efvm.SlipWall())))
        app.add_plugins(efvm.MusclPlugin.hllc().with_limiter(efvm.SlopeLimiter.MonotonizedCentral))
        app.add_plugins(efvm.IntegratorPlugin.rk3())
        app.add_plugins(efvm.SolverPlugin(efvm.SolverConfig(cfl=0.4)))
        app.add_plugins(efvm.OutputPlugin(str(out)))
        return app.prepare()

    a = build(tmp_path / "a.vtk")
    a.step(5)
    a.write_checkpoint(str(tmp_path / "a.bin"))
    a.step(7)
    a.finish()
    b = build(tmp_path / "b.vtk")
    b.read_checkpoint(str(tmp_path / "a.bin"))
    b.step(7)
    b.finish()
    assert np.array_equal(a.state(), b.state())
    assert a.solver_state() == b.solver_state()
    text = open(tmp_path / "a.vtk").read()
    assert f"CELL_DATA {n * n}" in text and text == open(tmp_path / "b.vtk").read()
