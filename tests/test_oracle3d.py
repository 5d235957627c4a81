"""Pins the oracle's 3-D path (nz > 1) and its sphere body before the GPU is checked
against them.

The reference's own tests for these parts are the sphere geometry and Fibonacci-marker
tests (crates/cfd_ibm/src/sphere.rs:255-349), restated on the oracle below (the
moving-wall velocity of that test is left out: only fixed bodies are built).  The rest
is pinned by invariants that a wrong index, sign or transposed operand on the third
axis breaks:
  * a 2-D vortex extruded along z equals the nz = 1 run layer by layer, and the
    nz = 1 path is pinned by the published L1 figures (test_oracle.py);
  * the same vortex laid in the x-z and the y-z plane equals it after permuting the
    axes and the momentum components (z fluxes, z slopes, rho_w, z halo);
  * flow past a centred sphere stays symmetric under y -> -y, z -> -z and y <-> z
    (sphere mask, ghost list, image points, trilinear mirror).
"""
import math

import numpy as np
import pytest

from oracle import orc

R = orc.R_AIR


# ------------------------------------------------ crates/cfd_ibm/src/sphere.rs tests
def test_sphere_signed_distance_and_normal():
    """sphere.rs:259-280 (sphere_geometry_and_wall_velocity, fixed body)."""
    c, r = (1.0, 2.0, 3.0), 0.5
    assert abs(orc.sphere_signed_distance(c, r, (2.0, 2.0, 3.0)) - 0.5) < 1e-12
    assert abs(orc.sphere_signed_distance(c, r, (1.0, 2.0, 3.0)) + 0.5) < 1e-12
    n = orc.sphere_normal(c, r, (1.5, 2.0, 3.0))
    assert abs(n[0] - 1.0) < 1e-12 and abs(n[1]) < 1e-12
    # the centre itself has no direction: the reference returns +x
    assert list(orc.sphere_normal(c, r, c)) == [1.0, 0.0, 0.0]


def test_fibonacci_markers_sum_to_sphere_area_and_lie_on_surface():
    """sphere.rs:283-348: areas sum to 4 pi R^2, markers on the surface with unit radial
    normals, lattice centroid near the centre."""
    R0, n = 0.75, 4000
    m = orc.sphere_markers((0.0, 0.0, 0.0), R0, n)
    assert m.shape == (n, 7)
    assert abs(m[:, 6].sum() / (4.0 * math.pi * R0 * R0) - 1.0) < 1e-12
    r = np.linalg.norm(m[:, :3], axis=1)
    assert np.all(np.abs(r - R0) < 1e-9)
    assert np.all(np.abs(np.linalg.norm(m[:, 3:6], axis=1) - 1.0) < 1e-9)
    assert np.all(np.abs(m[:, :3].mean(axis=0)) < 1e-2)
    # equal-area lattice: z of marker m is 1 - 2 (m + 1/2) / n exactly (cell-centred in z)
    zr = 1.0 - 2.0 * (np.arange(n) + 0.5) / n
    assert np.allclose(m[:, 5], zr, rtol=0, atol=1e-15)


# ------------------------------------------------ third-axis invariants
N2 = 24
L2 = 10.0


def vortex_plane(n=N2):
    """2-D isentropic vortex (examples/isentropic_vortex/main.rs) on an n x n periodic grid:
    (n, n, 5) conserved, row = y, column = x."""
    h = L2 / n
    out = np.zeros((n, n, 5))
    for j in range(n):
        for i in range(n):
            out[j, i] = orc.vortex_point(L2, 5.0, 1.4, 0.5, 0.5, 0.0, (i + 0.5) * h, (j + 0.5) * h, True)
    return out


def run_sim(shape, hi, state, steps, muscl):
    nx, ny, nz = shape
    s = orc.Sim(nx, ny, nz, hi=hi, flux="hllc", muscl=muscl, limiter="mc", cfl=0.3, bcs=["periodic"] * 6)
    s.set_interior(state)
    s.step(steps)
    return s.interior()


@pytest.mark.parametrize("muscl", [False, True])
def test_extruded_and_permuted_vortex_match_the_plane_run(muscl):
    n, h, steps, layers = N2, L2 / N2, 6, 3
    V = vortex_plane(n)
    base = run_sim((n, n, 1), (L2, L2, h), V[None], steps, muscl)[0]          # (y, x, 5)
    scale = np.abs(base).max()
    # extruded along z: every layer equals the plane
    ext = run_sim((n, n, layers), (L2, L2, layers * h), np.repeat(V[None], layers, axis=0), steps, muscl)
    for k in range(layers):
        assert np.abs(ext[k] - base).max() < 1e-12 * scale
    # x-z plane: 2-D (x, y, u, v) -> 3-D (x, z, u, w), uniform along y
    S = np.zeros((n, layers, n, 5))
    for j in range(layers):
        S[:, j, :, 0] = V[:, :, 0]
        S[:, j, :, 1] = V[:, :, 1]
        S[:, j, :, 3] = V[:, :, 2]
        S[:, j, :, 4] = V[:, :, 4]
    xz = run_sim((n, layers, n), (L2, layers * h, L2), S, steps, muscl)
    for j in range(layers):
        back = np.stack([xz[:, j, :, 0], xz[:, j, :, 1], xz[:, j, :, 3], xz[:, j, :, 2], xz[:, j, :, 4]], axis=-1)
        assert np.abs(back - base).max() < 1e-12 * scale
        assert np.abs(xz[:, j, :, 2]).max() < 1e-12 * scale       # no momentum along the uniform axis
    # y-z plane: 2-D (x, y, u, v) -> 3-D (y, z, v, w), uniform along x
    S = np.zeros((n, n, layers, 5))
    for i in range(layers):
        S[:, :, i, 0] = V[:, :, 0]
        S[:, :, i, 2] = V[:, :, 1]
        S[:, :, i, 3] = V[:, :, 2]
        S[:, :, i, 4] = V[:, :, 4]
    yz = run_sim((layers, n, n), (layers * h, L2, L2), S, steps, muscl)
    for i in range(layers):
        back = np.stack([yz[:, :, i, 0], yz[:, :, i, 2], yz[:, :, i, 3], yz[:, :, i, 1], yz[:, :, i, 4]], axis=-1)
        assert np.abs(back - base).max() < 1e-12 * scale


def test_planted_z_mistake_breaks_the_permutation_invariant():
    """A swapped momentum component along z in the initial state (the kind of mistake a
    transposed rho_v / rho_w makes) is caught by the same comparison."""
    n, h, layers = 16, L2 / 16, 2
    V = vortex_plane(n)
    base = run_sim((n, n, 1), (L2, L2, h), V[None], 3, True)[0]
    S = np.zeros((n, layers, n, 5))
    for j in range(layers):
        S[:, j, :, 0] = V[:, :, 0]
        S[:, j, :, 1] = V[:, :, 1]
        S[:, j, :, 2] = V[:, :, 2]          # planted: v2d into rho_v instead of rho_w
        S[:, j, :, 4] = V[:, :, 4]
    xz = run_sim((n, layers, n), (L2, layers * h, L2), S, 3, True)
    back = np.stack([xz[:, 0, :, 0], xz[:, 0, :, 1], xz[:, 0, :, 3], xz[:, 0, :, 2], xz[:, 0, :, 4]], axis=-1)
    assert np.abs(back - base).max() > 1e-3 * np.abs(base).max()


def sphere_case(n=16, steps=4, viscous=True, muscl=False):
    rho, p, u = 1.0, 1.0e5, 56.1
    t = p / (rho * R)
    fs = orc.prim_to_cons([rho, u, 0.0, 0.0, p, t], 1.4)
    s = orc.Sim(n, n, n, lo=(-0.3, -0.3, -0.3), hi=(0.3, 0.3, 0.3), flux="hllc", muscl=muscl, viscous=viscous,
                mu=rho * u * 0.2 / 20.0, cfl=0.3,
                bcs=["subsonic_inflow", "pressure_outlet", "symmetry", "symmetry", "symmetry", "symmetry"],
                bc_params=[(rho, u, 0.0, 0.0, t), (p,), None, None, None, None],
                spheres=[(0.0, 0.0, 0.0, 0.1)], freestream=fs, sphere_markers=500)
    s.set_interior(np.broadcast_to(fs, (n, n, n, 5)).copy())
    s.step(steps)
    return s


def test_sphere_flow_keeps_the_mirror_symmetries():
    s = sphere_case()
    U = s.interior()                                   # (z, y, x, 5)
    scale = np.abs(U).max(axis=(0, 1, 2))
    # the body changed the flow
    assert np.abs(U[..., 1] - U[0, 0, 0, 1]).max() > 1e-3 * scale[1]
    flip_y = U[:, ::-1].copy(); flip_y[..., 2] *= -1
    flip_z = U[::-1].copy(); flip_z[..., 3] *= -1
    swap = np.transpose(U, (1, 0, 2, 3)).copy(); swap[..., [2, 3]] = swap[..., [3, 2]]
    for W in (flip_y, flip_z, swap):
        assert np.all(np.abs(W - U) <= 1e-9 * scale)
    # the solid mask is the set of cell centres inside the sphere
    solid = s.solid()[2:-2, 2:-2, 2:-2]
    h = 0.6 / 16
    c = -0.3 + (np.arange(16) + 0.5) * h
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    assert np.array_equal(solid.astype(bool), np.sqrt(X * X + Y * Y + Z * Z) - 0.1 < 0.0)
    # markers and drag: positive x force, transverse forces small next to it
    f = s.marker_force()
    fx, fy, fz = f[0] + f[3], f[1] + f[4], f[2] + f[5]
    assert fx > 0 and abs(fy) < 0.05 * fx and abs(fz) < 0.05 * fx


# ------------------------------------------------ FromOutside ghost tables in 3-D
def _table_plane(n, seed):
    """Per-face outside states of a 2-D side (n faces) and their given flags."""
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.2 * rng.random(n)
    u, v, p = 0.3 * rng.standard_normal(n), 0.3 * rng.standard_normal(n), 1.0 + 0.2 * rng.random(n)
    st = np.stack([rho, rho * u, rho * v, np.zeros(n), p / 0.4 + 0.5 * rho * (u * u + v * v)], axis=-1)
    return st, rng.random(n) < 0.6


def test_ghost_tables_index_the_faces_of_a_3d_side():
    """FromOutside tables on an x and a y side of a 2-D run (the nz = 1 layout, pinned by the
    reference's boundary-face tests in test_oracle.py) equal, layer by layer, the run
    extruded along z with each table repeated per layer (x side j + ny k, y side i + nx k),
    and the run laid in the x-z plane with the y-side table moved to the z side (i + nx j)."""
    n, h, steps, layers = 12, 1.0 / 12, 4, 3
    V = vortex_plane(n)
    tx, gx = _table_plane(n, 1)          # x-lo, one per row
    ty, gy = _table_plane(n, 2)          # y-hi, one per column
    kw = dict(flux="hllc", muscl=True, limiter="minmod", cfl=0.3)
    base_sim = orc.Sim(n, n, 1, hi=(L2, L2, h), bcs=["table", "outflow", "slip", "table", "slip", "slip"], **kw)
    base_sim.set_ghost_table(0, tx, gx)
    base_sim.set_ghost_table(3, ty, gy)
    base_sim.set_interior(V[None])
    base_sim.step(steps)
    base = base_sim.interior()[0]
    scale = np.abs(base).max()
    # extruded along z (periodic z): tables repeated per layer
    ext_sim = orc.Sim(n, n, layers, hi=(L2, L2, layers * h),
                      bcs=["table", "outflow", "slip", "table", "periodic", "periodic"], **kw)
    ext_sim.set_ghost_table(0, np.tile(tx, (layers, 1)), np.tile(gx, layers))
    ext_sim.set_ghost_table(3, np.tile(ty, (layers, 1)), np.tile(gy, layers))
    ext_sim.set_interior(np.repeat(V[None], layers, axis=0))
    ext_sim.step(steps)
    ext = ext_sim.interior()
    for k in range(layers):
        assert np.abs(ext[k] - base).max() < 1e-12 * scale
    # x-z plane: 2-D (x, y, u, v) -> 3-D (x, z, u, w), uniform (periodic) along y
    def perm(st):
        return st[:, [0, 1, 3, 2, 4]]
    S = np.zeros((n, layers, n, 5))
    for j in range(layers):
        S[:, j, :, :] = V[:, :, [0, 1, 3, 2, 4]]
    xz_sim = orc.Sim(n, layers, n, hi=(L2, layers * h, L2),
                     bcs=["table", "outflow", "periodic", "periodic", "slip", "table"], **kw)
    # x side: index j + ny k with k = the 2-D row; z side: index i + nx j
    xz_sim.set_ghost_table(0, np.repeat(perm(tx), layers, axis=0), np.repeat(gx, layers))
    xz_sim.set_ghost_table(5, np.tile(perm(ty), (layers, 1)), np.tile(gy, layers))
    xz_sim.set_interior(S)
    xz_sim.step(steps)
    xz = xz_sim.interior()
    for j in range(layers):
        back = xz[:, j, :, :][..., [0, 1, 3, 2, 4]]
        assert np.abs(back - base).max() < 1e-12 * scale
    # a transposed table (rows and layers swapped) is caught by the same comparison
    bad = orc.Sim(n, layers, n, hi=(L2, layers * h, L2), bcs=["table", "outflow", "periodic", "periodic", "slip", "table"], **kw)
    bad.set_ghost_table(0, np.tile(perm(tx), (layers, 1)), np.tile(gx, layers))
    bad.set_ghost_table(5, np.tile(perm(ty), (layers, 1)), np.tile(gy, layers))
    bad.set_interior(S)
    bad.step(steps)
    assert np.abs(bad.interior()[:, 0, :, :][..., [0, 1, 3, 2, 4]] - base).max() > 1e-6 * scale


# ------------------------------------------------ examples/restart_determinism/main.rs
def _restart_problem_sim():
    import restart_digest as rd
    s = orc.Sim(rd.NX, rd.NY, rd.NZ, lo=rd.LO, hi=rd.HI, gamma=rd.GAMMA, r_gas=rd.R_GAS, bcs=rd.BCS, **rd.SCHEME)
    cen = (s.centroids(0), s.centroids(1), s.centroids(2))
    prims = rd.initial_prims(cen[0], cen[1], rd.NZ)
    s.set_interior(np.apply_along_axis(lambda q: orc.prim_to_cons(q, rd.GAMMA), -1, prims))
    return s, cen


def test_fixed_problem_digest_is_the_references():
    """fixed_problem_digest_matches_reference (examples/restart_determinism/main.rs:235-240):
    10 forward-Euler Rusanov steps of the 16 x 4 x 2 problem give the digest the reference
    holds (main.rs:79).  The digest hashes the bits of every conserved value of every cell,
    so this pins the oracle's whole 3-D first-order step bit for bit: the face sweep order,
    the face areas over the cell volume, the slip-wall and outflow ghosts on all three
    axes, the fixed dt and the update."""
    import restart_digest as rd
    s, cen = _restart_problem_sim()
    start = s.interior()
    assert rd.state_digest(start, *cen) != rd.DIGEST_REF
    s.step(rd.STEPS_TOTAL)
    info = s.info()
    assert info["step"] == rd.STEPS_TOTAL and info["dt"] == 2.0e-5
    final = s.interior()
    assert rd.state_digest(final, *cen) == rd.DIGEST_REF
    # the digest moves with one unit in the last place of one value of one cell
    moved = final.copy()
    moved[1, 2, 7, 4] = np.nextafter(moved[1, 2, 7, 4], np.inf)
    assert rd.state_digest(moved, *cen) != rd.DIGEST_REF
    # and with the flux plugin swapped (HLLC for Rusanov)
    kw = dict(rd.SCHEME, flux="hllc")
    h = orc.Sim(rd.NX, rd.NY, rd.NZ, lo=rd.LO, hi=rd.HI, gamma=rd.GAMMA, r_gas=rd.R_GAS, bcs=rd.BCS, **kw)
    h.set_interior(start)
    h.step(rd.STEPS_TOTAL)
    assert rd.state_digest(h.interior(), *cen) != rd.DIGEST_REF


def test_restart_at_step_five_continues_to_the_references_digest():
    """restart_checkpoint_is_bitwise (main.rs:207-224) on the oracle: the raw state and the
    clock after 5 steps, loaded into a fresh simulation and advanced 5 more, equal the
    uninterrupted 10 steps bit for bit."""
    import restart_digest as rd
    a, cen = _restart_problem_sim()
    a.step(rd.STEPS_TOTAL)
    b, _ = _restart_problem_sim()
    b.step(rd.STEPS_CHECKPOINT)
    raw, clock = b.raw(), b.info()
    c, _ = _restart_problem_sim()
    c.set_raw(raw)
    c.set_clock(**clock)
    c.step(rd.STEPS_TOTAL - rd.STEPS_CHECKPOINT)
    assert np.array_equal(c.raw(), a.raw())
    assert c.info() == a.info()
    assert rd.state_digest(c.interior(), *cen) == rd.DIGEST_REF
