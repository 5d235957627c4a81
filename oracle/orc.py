"""ctypes wrapper of the CPU FP64 oracle (``oracle/efvm_oracle.c``).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this module, and only as the
checker.  The product (``dev_field_efvm_amd``) never imports it.

The configuration mirrors the reference's plugin choices:
``FluxPlugin::{rusanov,hllc}`` / ``MusclPlugin::hllc().with_limiter(..)``
(crates/cfd_solver/src/lib.rs:775-826), ``IntegratorPlugin::{euler,rk3}``
(:881-888), ``SolverConfig`` (:39-58), ``BoundaryRegistry`` (crates/cfd_boundary/src/lib.rs:369-449).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liborc.so")

FLUX = {"rusanov": 0, "hllc": 1, "hllc_textbook": 2}
LIMITER = {"minmod": 0, "mc": 1}
BC = {"outflow": 0, "periodic": 1, "slip": 2, "noslip": 3, "symmetry": 4,
      "supersonic_inflow": 5, "subsonic_inflow": 6, "pressure_outlet": 7, "table": 8}
R_AIR = 287.058


class OrcCfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("gamma", C.c_double), ("r_gas", C.c_double),
        ("flux", C.c_int), ("recon", C.c_int), ("limiter", C.c_int), ("integrator", C.c_int),
        ("viscous", C.c_int), ("visc_model", C.c_int),
        ("mu_const", C.c_double),
        ("cfl", C.c_double), ("fixed_dt", C.c_double),
        ("bc", C.c_int * 6),
        ("bc_par", (C.c_double * 5) * 6),
        ("seam_mode", C.c_int),
        ("parabolic_inflow", C.c_int), ("inflow_umax", C.c_double), ("inflow_h", C.c_double), ("inflow_rho", C.c_double),
        ("sponge", C.c_int), ("sp_x0", C.c_double), ("sp_x1", C.c_double), ("sp_sigma", C.c_double),
        ("sp_umax", C.c_double), ("sp_h", C.c_double), ("sp_rho", C.c_double), ("sp_p", C.c_double),
        ("body_force", C.c_int), ("body_gx", C.c_double),
        ("n_cyl", C.c_int), ("cyl", (C.c_double * 3) * 4),
        ("freestream", C.c_double * 5),
        ("n_sph", C.c_int), ("sph", (C.c_double * 4) * 4),
        ("sph_markers", C.c_int),
        ("sp_uniform", C.c_int), ("sp_target", C.c_double * 5),
    ]


def build() -> str:
    """Compile the oracle with its committed Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d = C.c_double
        pd = C.POINTER(C.c_double)
        L.orc_new.restype = C.c_void_p
        L.orc_new.argtypes = [C.POINTER(OrcCfg)]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_ntot.restype = C.c_long
        L.orc_ntot.argtypes = [C.c_void_p]
        L.orc_set_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_edges.restype = C.c_int
        L.orc_set_edges.argtypes = [C.c_void_p, C.c_int, pd, C.c_int]
        for n in ("orc_set_raw", "orc_get_raw", "orc_get_rhs_raw", "orc_get_u0_raw", "orc_set_interior", "orc_get_interior", "orc_get_rhs_interior",
                  "orc_marker_force", "orc_wall_force", "orc_sums"):
            getattr(L, n).argtypes = [C.c_void_p, pd]
        L.orc_info.argtypes = [C.c_void_p, pd, pd, C.POINTER(C.c_long), C.POINTER(C.c_int)]
        L.orc_set_clock.argtypes = [C.c_void_p, d, d, C.c_long, C.c_int]
        L.orc_set_ghost_table.restype = C.c_int
        L.orc_set_ghost_table.argtypes = [C.c_void_p, C.c_int, pd, C.c_long]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p]
        L.orc_advance_to.restype = C.c_long
        L.orc_advance_to.argtypes = [C.c_void_p, d, C.c_long]
        L.orc_advance_steps.argtypes = [C.c_void_p, C.c_long]
        L.orc_cfl_dt.restype = d
        L.orc_cfl_dt.argtypes = [C.c_void_p]
        L.orc_ic_vortex.argtypes = [C.c_void_p, d, d, d, d, d, C.c_int]
        L.orc_vortex_l1.restype = d
        L.orc_vortex_l1.argtypes = [C.c_void_p, d, d, d, d, d, C.c_int]
        L.orc_vortex_point.argtypes = [d, d, d, d, d, d, d, d, C.c_int, pd]
        L.orc_sphere_markers.argtypes = [pd, C.c_int, pd]
        L.orc_sphere_sd.restype = d
        L.orc_sphere_sd.argtypes = [pd, pd]
        L.orc_sphere_normal.argtypes = [pd, pd, pd]
        L.orc_ic_pulse.argtypes = [C.c_void_p, d, d, d, d, d, d, d]
        L.orc_pulse_l1.restype = d
        L.orc_pulse_l1.argtypes = [C.c_void_p, d, d, d, d, d, d, d]
        L.orc_ic_split_x.argtypes = [C.c_void_p, d, pd, pd]
        L.orc_ic_uniform.argtypes = [C.c_void_p, pd]
        L.orc_ic_channel.argtypes = [C.c_void_p, d, d, d, d, d]
        L.orc_flux.argtypes = [C.c_int, pd, pd, d, C.c_int, pd]
        L.orc_flux_normal_axis.argtypes = [C.c_int, pd, pd, d, C.c_int, C.c_int, pd]
        L.orc_physical_flux.argtypes = [pd, d, C.c_int, pd]
        L.orc_prim_to_cons.argtypes = [pd, d, pd]
        L.orc_cons_to_prim.argtypes = [pd, d, d, pd]
        L.orc_sound_speed.restype = d
        L.orc_sound_speed.argtypes = [pd, d]
        L.orc_sutherland.restype = d
        L.orc_sutherland.argtypes = [d]
        L.orc_bc_ghost.argtypes = [C.c_int, C.c_int, pd, pd, d, d, pd]
        L.orc_limiter.restype = d
        L.orc_limiter.argtypes = [C.c_int, d, d]
        L.orc_ibm_ghosts.restype = C.c_long
        L.orc_ibm_ghosts.argtypes = [C.c_void_p, C.POINTER(C.c_long), pd]
        L.orc_get_solid.argtypes = [C.c_void_p, C.POINTER(C.c_ubyte)]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _vec(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


# ---- scalar kernels (known-answer tests) ----
def flux(scheme: str, L, R, gamma: float, direction: int) -> np.ndarray:
    out = np.zeros(5)
    l, r = _vec(L), _vec(R)
    lib().orc_flux(FLUX[scheme], _p(l), _p(r), gamma, direction, _p(out))
    return out


def flux_normal_axis(scheme: str, owner, other, gamma: float, direction: int, sign: int) -> np.ndarray:
    out = np.zeros(5)
    o, t = _vec(owner), _vec(other)
    lib().orc_flux_normal_axis(FLUX[scheme], _p(o), _p(t), gamma, direction, sign, _p(out))
    return out


def physical_flux(U, gamma: float, direction: int) -> np.ndarray:
    out = np.zeros(5)
    u = _vec(U)
    lib().orc_physical_flux(_p(u), gamma, direction, _p(out))
    return out


def prim_to_cons(P, gamma: float) -> np.ndarray:
    out = np.zeros(5)
    p = _vec(P)
    lib().orc_prim_to_cons(_p(p), gamma, _p(out))
    return out


def cons_to_prim(U, gamma: float, r_gas: float = R_AIR) -> np.ndarray:
    out = np.zeros(6)
    u = _vec(U)
    lib().orc_cons_to_prim(_p(u), gamma, r_gas, _p(out))
    return out


def sound_speed(U, gamma: float) -> float:
    u = _vec(U)
    return lib().orc_sound_speed(_p(u), gamma)


def sutherland(t: float) -> float:
    return lib().orc_sutherland(t)


def bc_ghost(kind: str, side: int, interior, params=None, gamma: float = 1.4, r_gas: float = R_AIR) -> np.ndarray:
    """BoundaryCondition::ghost of a stock condition on box side `side` (0..5 = XLo..ZHi)."""
    out = np.zeros(5)
    par = _vec(list(params or []) + [0.0] * (5 - len(params or [])))
    u = _vec(interior)
    lib().orc_bc_ghost(BC[kind], side, _p(par), _p(u), gamma, r_gas, _p(out))
    return out


def limiter(name: str, a: float, b: float) -> float:
    return lib().orc_limiter(LIMITER[name], a, b)


def sphere_markers(center, radius, n) -> np.ndarray:
    """sphere_markers (crates/cfd_ibm/src/sphere.rs:213-237): rows of position, normal, area."""
    b = _vec(list(center) + [radius])
    out = np.zeros((n, 7))
    lib().orc_sphere_markers(_p(b), n, _p(out))
    return out


def sphere_signed_distance(center, radius, p) -> float:
    return lib().orc_sphere_sd(_p(_vec(list(center) + [radius])), _p(_vec(p)))


def sphere_normal(center, radius, p) -> np.ndarray:
    out = np.zeros(3)
    lib().orc_sphere_normal(_p(_vec(list(center) + [radius])), _p(_vec(p)), _p(out))
    return out


def vortex_point(L, beta, gamma, mean_u, mean_v, t, x, y, periodic=False) -> np.ndarray:
    out = np.zeros(5)
    lib().orc_vortex_point(L, beta, gamma, mean_u, mean_v, t, x, y, int(periodic), _p(out))
    return out


class Sim:
    """One reference App: a UniformMesh plus the solver/boundary/IBM plugins."""

    def __init__(self, nx, ny, nz=1, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), gamma=1.4, r_gas=R_AIR,
                 flux="hllc", muscl=False, limiter="minmod", integrator="rk3", viscous=False,
                 mu=None, cfl=0.4, fixed_dt=None, bcs=None, bc_params=None, seam_mode=0,
                 parabolic_inflow=None, sponge=None, body_gx=None, cylinders=(), freestream=None,
                 spheres=(), sphere_markers=2000, sponge_uniform=None):
        c = OrcCfg()
        c.nx, c.ny, c.nz = nx, ny, nz
        for d in range(3):
            c.lo[d] = lo[d]
            c.hi[d] = hi[d]
        c.gamma, c.r_gas = gamma, r_gas
        c.flux = FLUX[flux]
        c.recon = 1 if muscl else 0
        c.limiter = LIMITER[limiter]
        c.integrator = 1 if integrator == "rk3" else 0
        c.viscous = int(viscous)
        c.visc_model = 0 if mu is None else 1
        c.mu_const = 0.0 if mu is None else float(mu)
        c.cfl = cfl
        c.fixed_dt = -1.0 if fixed_dt is None else float(fixed_dt)
        bcs = bcs or ["outflow"] * 6
        for s in range(6):
            c.bc[s] = BC[bcs[s]]
            if bc_params and bc_params[s] is not None:
                for q, v in enumerate(bc_params[s]):
                    c.bc_par[s][q] = v
        c.seam_mode = seam_mode
        if parabolic_inflow:
            c.parabolic_inflow = 1
            c.inflow_umax, c.inflow_h, c.inflow_rho = parabolic_inflow
        if sponge:
            c.sponge = 1
            (c.sp_x0, c.sp_x1, c.sp_sigma, c.sp_umax, c.sp_h, c.sp_rho, c.sp_p) = sponge
        if body_gx is not None:
            c.body_force = 1
            c.body_gx = body_gx
        c.n_cyl = len(cylinders)
        for b, cy in enumerate(cylinders):
            for q in range(3):
                c.cyl[b][q] = cy[q]
        if freestream is not None:
            for q in range(5):
                c.freestream[q] = freestream[q]
        # SphereIbmPlugin bodies (cx, cy, cz, r); a simulation holds spheres or cylinders
        if spheres and cylinders:
            raise ValueError("spheres and cylinders in one simulation")
        c.n_sph = len(spheres)
        for b, sp in enumerate(spheres):
            for q in range(4):
                c.sph[b][q] = sp[q]
        c.sph_markers = int(sphere_markers)
        if sponge_uniform is not None:   # (x_start, x_end, sigma, target conserved state)
            x0, x1, sig, tgt = sponge_uniform
            c.sponge, c.sp_uniform = 1, 1
            c.sp_x0, c.sp_x1, c.sp_sigma = x0, x1, sig
            for q in range(5):
                c.sp_target[q] = tgt[q]
        self.cfg = c
        self.nx, self.ny, self.nz = nx, ny, nz
        self.h = tuple((hi[d] - lo[d]) / (nx, ny, nz)[d] for d in range(3))
        self.lo, self.hi = tuple(lo), tuple(hi)
        ptr = lib().orc_new(C.byref(c))
        if not ptr:
            raise ValueError("invalid periodic boundary configuration: a periodic axis needs both sides periodic")
        self.ptr = C.c_void_p(ptr)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_free(self.ptr)
            self.ptr = None

    def set_edges(self, axis: int, edges):
        """Stretch axis 1 (y) or 2 (z): its n+1 increasing cell-edge coordinates, lo to hi."""
        e = np.ascontiguousarray(edges, dtype=np.float64)
        if lib().orc_set_edges(self.ptr, int(axis), _p(e), int(e.size)) != 0:
            raise ValueError("edges: axis 1 or 2, n+1 increasing values from lo to hi, no immersed body")
        self.edges = getattr(self, "edges", {})
        self.edges[int(axis)] = e.copy()

    def centroids(self, axis: int) -> np.ndarray:
        """Interior cell centroids along an axis."""
        e = getattr(self, "edges", {}).get(axis)
        if e is not None:
            return 0.5 * (e[:-1] + e[1:])
        n = (self.nx, self.ny, self.nz)[axis]
        return self.lo[axis] + (np.arange(n) + 0.5) * self.h[axis]

    def set_threads(self, n: int) -> int:
        """Run steps on `n` OpenMP threads (bitwise the serial result); returns the count in use."""
        return lib().orc_set_threads(self.ptr, int(n))

    @property
    def ntot(self) -> int:
        return lib().orc_ntot(self.ptr)

    def interior(self) -> np.ndarray:
        out = np.zeros((self.nz, self.ny, self.nx, 5))
        lib().orc_get_interior(self.ptr, _p(out))
        return out

    def set_interior(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.nz, self.ny, self.nx, 5)
        lib().orc_set_interior(self.ptr, _p(a))

    def raw(self) -> np.ndarray:
        out = np.zeros((self.nz + 4, self.ny + 4, self.nx + 4, 5))
        lib().orc_get_raw(self.ptr, _p(out))
        return out

    def raw_column(self, name: str) -> np.ndarray:
        """The rhs or u0 column over the raw padded array (shape (nz+4, ny+4, nx+4, 5))."""
        out = np.zeros((self.nz + 4, self.ny + 4, self.nx + 4, 5))
        getattr(lib(), f"orc_get_{name}_raw")(self.ptr, _p(out))
        return out

    def set_raw(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == 5 * self.ntot
        lib().orc_set_raw(self.ptr, _p(a))

    def rhs(self) -> np.ndarray:
        out = np.zeros((self.nz, self.ny, self.nx, 5))
        lib().orc_get_rhs_interior(self.ptr, _p(out))
        return out

    def info(self):
        dt, t = C.c_double(), C.c_double()
        st, sg = C.c_long(), C.c_int()
        lib().orc_info(self.ptr, C.byref(dt), C.byref(t), C.byref(st), C.byref(sg))
        return dict(dt=dt.value, time=t.value, step=st.value, stage=sg.value)

    def set_ghost_table(self, side: int, states, given):
        """BoundaryInputs of a FromOutside ("table") side: states (n, 5) wire order, given (n,) bools."""
        st = np.asarray(states, dtype=np.float64).reshape(-1, 5)
        t = np.zeros((st.shape[0], 6))
        t[:, :5] = st
        t[:, 5] = np.asarray(given, dtype=bool)
        t = np.ascontiguousarray(t)
        if lib().orc_set_ghost_table(self.ptr, side, _p(t), st.shape[0]) != 0:
            raise ValueError("ghost table on a side whose kind is not 'table'")

    def set_clock(self, dt=0.0, time=0.0, step=0, stage=0):
        lib().orc_set_clock(self.ptr, dt, time, step, stage)

    def run(self) -> bool:
        """One ``app.run()``: one RK substage through every phase."""
        return bool(lib().orc_run(self.ptr))

    def step(self, n: int = 1):
        lib().orc_advance_steps(self.ptr, n)

    def advance_to(self, t_final: float, max_steps: int = 10 ** 9) -> int:
        return lib().orc_advance_to(self.ptr, t_final, max_steps)

    def cfl_dt(self) -> float:
        return lib().orc_cfl_dt(self.ptr)

    # ---- initial conditions (evaluated on every raw cell, ghosts included) ----
    def ic_vortex(self, L=10.0, beta=5.0, gamma=1.4, mean_u=0.5, mean_v=0.5, periodic=False):
        self._vortex = (L, beta, gamma, mean_u, mean_v, int(periodic))
        lib().orc_ic_vortex(self.ptr, L, beta, gamma, mean_u, mean_v, int(periodic))

    def vortex_l1(self) -> float:
        return lib().orc_vortex_l1(self.ptr, *self._vortex)

    def ic_pulse(self, L=1.0, center=0.85, sigma=0.08, amplitude=0.2, gamma=1.4, velocity=1.0, pressure=1.0):
        self._pulse = (L, center, sigma, amplitude, gamma, velocity, pressure)
        lib().orc_ic_pulse(self.ptr, *self._pulse)

    def pulse_l1(self) -> float:
        return lib().orc_pulse_l1(self.ptr, *self._pulse)

    def ic_split_x(self, x0, A, B):
        a, b = _vec(A), _vec(B)
        lib().orc_ic_split_x(self.ptr, x0, _p(a), _p(b))

    def ic_uniform(self, A):
        a = _vec(A)
        lib().orc_ic_uniform(self.ptr, _p(a))

    def ic_channel(self, u_max, height, rho0, p0, r_gas=R_AIR):
        lib().orc_ic_channel(self.ptr, u_max, height, rho0, p0, r_gas)

    def sums(self) -> np.ndarray:
        out = np.zeros(5)
        lib().orc_sums(self.ptr, _p(out))
        return out

    def mass(self) -> float:
        return float(self.sums()[0])

    def marker_force(self) -> np.ndarray:
        out = np.zeros(6)
        lib().orc_marker_force(self.ptr, _p(out))
        return out

    def wall_force(self) -> np.ndarray:
        out = np.zeros(6)
        lib().orc_wall_force(self.ptr, _p(out))
        return out

    def ibm_ghosts(self):
        n = lib().orc_ibm_ghosts(self.ptr, None, None)
        idx = np.zeros(n, dtype=np.int64)
        img = np.zeros((n, 3))
        lib().orc_ibm_ghosts(self.ptr, idx.ctypes.data_as(C.POINTER(C.c_long)), _p(img))
        return idx, img

    def solid(self) -> np.ndarray:
        out = np.zeros(self.ntot, dtype=np.uint8)
        lib().orc_get_solid(self.ptr, out.ctypes.data_as(C.POINTER(C.c_ubyte)))
        return out.reshape(self.nz + 4, self.ny + 4, self.nx + 4)
