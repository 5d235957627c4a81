"""ctypes binding of the C ABI in ``include/efvm.h`` (``libefvm.so``).

The library is loaded from this package directory.  There is no fallback: if
the HIP library is missing or no GPU is visible, constructing a context raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libefvm.so")

OK, ERR_ARG, ERR_CONFIG, ERR_HIP, ERR_UNSUPPORTED, ERR_NONPHYSICAL, ERR_NO_DEVICE = range(7)
STATUS_NAMES = {0: "EFVM_OK", 1: "EFVM_ERR_ARG", 2: "EFVM_ERR_CONFIG", 3: "EFVM_ERR_HIP",
                4: "EFVM_ERR_UNSUPPORTED", 5: "EFVM_ERR_NONPHYSICAL", 6: "EFVM_ERR_NO_DEVICE"}

FLUX = {"rusanov": 0, "hllc": 1, "hllc_textbook": 2}
LIMITER = {"minmod": 0, "mc": 1}
BC = {"outflow": 0, "periodic": 1, "slip": 2, "noslip": 3, "symmetry": 4,
      "supersonic_inflow": 5, "subsonic_inflow": 6, "pressure_outlet": 7, "table": 8}

# Every entry point include/efvm.h declares (checked by the CPU test suite).
EXPORTS = [
    "efvm_ctx_create", "efvm_ctx_destroy", "efvm_last_error", "efvm_set_launch", "efvm_get_launch", "efvm_set_pair_kernel", "efvm_set_fused_step", "efvm_get_fused_step", "efvm_device_arch", "efvm_set_boundaries",
    "efvm_set_parabolic_inflow", "efvm_set_sponge", "efvm_set_body_force", "efvm_set_time_step", "efvm_upload_state",
    "efvm_download_state", "efvm_get_clock", "efvm_set_clock", "efvm_cfl", "efvm_run", "efvm_step",
    "efvm_advance_to", "efvm_synchronize", "efvm_graph_capture", "efvm_graph_launch",
    "efvm_ibm_add_cylinder", "efvm_marker_force", "efvm_wall_force", "efvm_conservation_sums", "efvm_device_state",
    "efvm_last_elapsed_ms", "efvm_snapshot", "efvm_checkpoint_save", "efvm_checkpoint_load", "efvm_write_vtk",
    "efvm_set_ghost_table", "efvm_state_sums_noalloc",
]


def device_arch(device: int = 0) -> str:
    """gcnArchName of a HIP device as the library's runtime reports it ("gfx950:..."), or
    "unknown" without a device."""
    L = load_library()
    fn = L.efvm_device_arch
    fn.restype = C.c_int
    fn.argtypes = [C.c_int, C.c_char_p, C.c_size_t]
    buf = C.create_string_buffer(256)
    if fn(int(device), buf, 256) != 0:
        return "unknown"
    return buf.value.decode()


def state_sums_noalloc(planes, precision: str, n_cells: int, cell_volume: float, out5_ptr: int, stream: int = 0):
    """efvm_state_sums_noalloc: FP64 volume-weighted sums of four device planes (pointers, as
    efvm_device_state gives them) into 5 doubles at device address `out5_ptr`, enqueued on
    `stream` (a hipStream_t as an integer; 0 = the default stream).  No allocation, no
    synchronisation: the caller synchronises before reading the output."""
    L = load_library()
    fn = L.efvm_state_sums_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_longlong, C.c_double, C.c_void_p, C.c_void_p]
    arr = (C.c_void_p * 4)(*[int(p) for p in planes])
    st = fn(arr, {"fp64": 0, "fp32": 1}[precision], int(n_cells), float(cell_volume), C.c_void_p(int(out5_ptr)),
            C.c_void_p(int(stream)) if stream else None)
    if st != 0:
        raise EfvmError(st, "efvm_state_sums_noalloc failed")


def write_vtk(path: str, state, nx: int, ny: int, lo=(0.0, 0.0), hi=(1.0, 1.0), dz=None, gamma=1.4,
              r_gas=287.058):
    """VTK legacy RECTILINEAR_GRID file of a wire-order state (ny, nx, 5) on the host
    (crates/cfd_output/src/lib.rs write_vtk_to); needs no device."""
    L = load_library()
    dz = (hi[0] - lo[0]) / nx if dz is None else dz
    g = Grid(int(nx), int(ny), lo[0], lo[1], hi[0], hi[1], dz)
    gs = Gas(gamma, r_gas, 0, 0.0)
    u = np.ascontiguousarray(state, dtype=np.float64).reshape(-1)
    if u.size != 5 * nx * ny:
        raise ValueError("state must hold 5 values per cell")
    st = L.efvm_write_vtk(C.byref(g), C.byref(gs), u.ctypes.data_as(C.POINTER(C.c_double)), nx * ny,
                          os.fsencode(path))
    if st != OK:
        raise EfvmError(st, f"efvm_write_vtk({path}) failed")


class EfvmError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("lo_x", C.c_double), ("lo_y", C.c_double),
                ("hi_x", C.c_double), ("hi_y", C.c_double), ("dz", C.c_double)]


class Gas(C.Structure):
    _fields_ = [("gamma", C.c_double), ("r_gas", C.c_double), ("viscosity", C.c_int), ("mu", C.c_double)]


class Scheme(C.Structure):
    _fields_ = [("flux", C.c_int), ("recon", C.c_int), ("limiter", C.c_int), ("integrator", C.c_int),
                ("precision", C.c_int), ("viscous", C.c_int), ("seam_mode", C.c_int),
                ("cfl", C.c_double), ("fixed_dt", C.c_double)]


class Bc(C.Structure):
    _fields_ = [("kind", C.c_int), ("par", C.c_double * 5)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libefvm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library {path} is missing: build it with dev_field_efvm_amd/build.py")
    L = C.CDLL(path)
    vp, d, pd = C.c_void_p, C.c_double, C.POINTER(C.c_double)
    sig = {
        "efvm_ctx_create": [C.POINTER(Grid), C.POINTER(Gas), C.POINTER(Scheme), C.POINTER(vp)],
        "efvm_set_launch": [vp, C.c_int, C.c_int],
        "efvm_get_launch": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "efvm_set_pair_kernel": [vp, C.c_int],
        "efvm_set_fused_step": [vp, C.c_int, C.c_int],
        "efvm_get_fused_step": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "efvm_set_boundaries": [vp, C.POINTER(Bc)],
        "efvm_set_parabolic_inflow": [vp, d, d, d],
        "efvm_set_sponge": [vp, d, d, d, d, d, d, d],
        "efvm_set_body_force": [vp, d],
        "efvm_set_time_step": [vp, d, d],
        "efvm_upload_state": [vp, pd, C.c_size_t],
        "efvm_download_state": [vp, pd, C.c_size_t],
        "efvm_get_clock": [vp, pd, pd, C.POINTER(C.c_long), C.POINTER(C.c_int)],
        "efvm_set_clock": [vp, d, C.c_long],
        "efvm_cfl": [vp, pd],
        "efvm_run": [vp],
        "efvm_step": [vp, C.c_int],
        "efvm_advance_to": [vp, d, C.c_long, C.POINTER(C.c_long)],
        "efvm_synchronize": [vp],
        "efvm_graph_capture": [vp, C.c_int],
        "efvm_graph_launch": [vp],
        "efvm_ibm_add_cylinder": [vp, d, d, d, pd],
        "efvm_marker_force": [vp, pd],
        "efvm_wall_force": [vp, pd],
        "efvm_conservation_sums": [vp, pd],
        "efvm_device_state": [vp, C.POINTER(vp)],
        "efvm_last_elapsed_ms": [vp, pd],
        "efvm_snapshot": [vp, C.c_int],
        "efvm_set_ghost_table": [vp, C.c_int, pd, C.POINTER(C.c_ubyte), C.c_size_t],
        "efvm_checkpoint_save": [vp, C.c_char_p],
        "efvm_checkpoint_load": [vp, C.c_char_p],
        "efvm_write_vtk": [C.POINTER(Grid), C.POINTER(Gas), pd, C.c_size_t, C.c_char_p],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.efvm_ctx_destroy.argtypes = [vp]
    L.efvm_ctx_destroy.restype = None
    L.efvm_last_error.argtypes = [vp]
    L.efvm_last_error.restype = C.c_char_p
    _lib = L
    return L


def _pd(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """One solver context on the current HIP device (thin, 1:1 over the C ABI)."""

    def __init__(self, nx, ny, lo=(0.0, 0.0), hi=(1.0, 1.0), dz=None, gamma=1.4, r_gas=287.058, mu=None,
                 flux="hllc", muscl=False, limiter="minmod", integrator="rk3", precision="fp64",
                 viscous=False, seam_mode=0, cfl=0.4, fixed_dt=None):
        L = load_library()
        self._L = L
        self.nx, self.ny = int(nx), int(ny)
        dz = (hi[0] - lo[0]) / nx if dz is None else dz
        g = Grid(self.nx, self.ny, lo[0], lo[1], hi[0], hi[1], dz)
        gs = Gas(gamma, r_gas, 0 if mu is None else 1, 0.0 if mu is None else float(mu))
        sc = Scheme(FLUX[flux], 1 if muscl else 0, LIMITER[limiter], 1 if integrator == "rk3" else 0,
                    0 if precision == "fp64" else 1, int(viscous), int(seam_mode), cfl,
                    -1.0 if fixed_dt is None else float(fixed_dt))
        self.precision = precision
        self.grid, self.gas, self.scheme = g, gs, sc
        p = C.c_void_p()
        st = L.efvm_ctx_create(C.byref(g), C.byref(gs), C.byref(sc), C.byref(p))
        if st != OK:
            raise EfvmError(st, "efvm_ctx_create failed")
        self._p = p

    def close(self):
        if getattr(self, "_p", None):
            self._L.efvm_ctx_destroy(self._p)
            self._p = None

    def __del__(self):
        self.close()

    def _chk(self, st: int):
        if st != OK:
            raise EfvmError(st, self._L.efvm_last_error(self._p).decode())

    @property
    def ncell(self) -> int:
        return self.nx * self.ny

    def set_launch(self, strip_rows=0, waves_per_block=0):
        self._chk(self._L.efvm_set_launch(self._p, strip_rows, waves_per_block))

    KERNELS = ("stage_f32_kernel (one-column strips)", "stage_f32_pair_kernel (two-column strips)",
               "stage_f64_kernel", "learned face kernels")

    def launch(self):
        """(strip_rows, waves_per_block, kernel name, which two-column kernel applies: 0 none, 1 the
        periodic form, 2 the general form)."""
        r, w, k, f = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._chk(self._L.efvm_get_launch(self._p, C.byref(r), C.byref(w), C.byref(k), C.byref(f)))
        return r.value, w.value, self.KERNELS[k.value], f.value

    def set_pair_kernel(self, allow: bool):
        """Allow (default) or forbid the two-column periodic strip kernel."""
        self._chk(self._L.efvm_set_pair_kernel(self._p, int(bool(allow))))

    def set_fused_step(self, allow: bool = True, strip_rows: int = 0):
        """Allow (default) or forbid running RK3 stages 2 and 3 of whole steps as one fused launch,
        and set its strip height (0 keeps the current value)."""
        self._chk(self._L.efvm_set_fused_step(self._p, int(bool(allow)), strip_rows))

    def fused_step(self):
        """(allowed, strip_rows, fits): fits = the configuration admits the fused stage-2+3 launch
        (FP32 RK3, periodic two-column path, ny >= 8)."""
        a, r, f = C.c_int(), C.c_int(), C.c_int()
        self._chk(self._L.efvm_get_fused_step(self._p, C.byref(a), C.byref(r), C.byref(f)))
        return bool(a.value), r.value, bool(f.value)

    def set_boundaries(self, sides):
        """sides: 4 x (kind, params) for XLo, XHi, YLo, YHi."""
        arr = (Bc * 4)()
        for s, spec in enumerate(sides):
            kind, par = (spec, ()) if isinstance(spec, str) else spec
            arr[s].kind = BC[kind]
            for q, v in enumerate(par):
                arr[s].par[q] = v
        self._chk(self._L.efvm_set_boundaries(self._p, arr))

    def set_parabolic_inflow(self, u_max, height, rho):
        self._chk(self._L.efvm_set_parabolic_inflow(self._p, u_max, height, rho))

    def set_sponge(self, x0, x1, sigma, u_max, height, rho, p):
        self._chk(self._L.efvm_set_sponge(self._p, x0, x1, sigma, u_max, height, rho, p))

    def set_body_force(self, gx):
        self._chk(self._L.efvm_set_body_force(self._p, gx))

    def set_time_step(self, cfl=0.4, fixed_dt=None):
        self._chk(self._L.efvm_set_time_step(self._p, cfl, -1.0 if fixed_dt is None else float(fixed_dt)))

    def upload(self, cons: np.ndarray):
        a = np.ascontiguousarray(cons, dtype=np.float64).reshape(-1, 5)
        if a.shape[0] != self.ncell:
            raise ValueError("state has the wrong number of cells")
        self._chk(self._L.efvm_upload_state(self._p, _pd(a), a.shape[0]))

    def download(self) -> np.ndarray:
        out = np.empty((self.ny, self.nx, 5))
        self._chk(self._L.efvm_download_state(self._p, _pd(out), self.ncell))
        return out

    def clock(self):
        dt, t = C.c_double(), C.c_double()
        st, sg = C.c_long(), C.c_int()
        self._chk(self._L.efvm_get_clock(self._p, C.byref(dt), C.byref(t), C.byref(st), C.byref(sg)))
        return dict(dt=dt.value, time=t.value, step=st.value, stage=sg.value)

    def set_clock(self, time=0.0, step=0):
        self._chk(self._L.efvm_set_clock(self._p, time, step))

    def cfl_dt(self) -> float:
        d = C.c_double()
        self._chk(self._L.efvm_cfl(self._p, C.byref(d)))
        return d.value

    def run(self):
        self._chk(self._L.efvm_run(self._p))

    def step(self, n=1):
        self._chk(self._L.efvm_step(self._p, n))

    def advance_to(self, t_final, max_steps=10 ** 9) -> int:
        k = C.c_long()
        self._chk(self._L.efvm_advance_to(self._p, t_final, max_steps, C.byref(k)))
        return k.value

    def synchronize(self):
        self._chk(self._L.efvm_synchronize(self._p))

    def graph_capture(self, n_steps):
        self._chk(self._L.efvm_graph_capture(self._p, n_steps))

    def graph_launch(self):
        self._chk(self._L.efvm_graph_launch(self._p))

    def last_elapsed_ms(self) -> float:
        d = C.c_double()
        self._chk(self._L.efvm_last_elapsed_ms(self._p, C.byref(d)))
        return d.value

    def add_cylinder(self, cx, cy, r, freestream):
        fs = np.ascontiguousarray(freestream, dtype=np.float64)
        self._chk(self._L.efvm_ibm_add_cylinder(self._p, cx, cy, r, _pd(fs)))

    def marker_force(self) -> np.ndarray:
        out = np.zeros(6)
        self._chk(self._L.efvm_marker_force(self._p, _pd(out)))
        return out

    def wall_force(self) -> np.ndarray:
        """Staircase wall force (pressure xyz, viscous xyz) on the cylinders, FP64."""
        out = np.zeros(6)
        self._chk(self._L.efvm_wall_force(self._p, _pd(out)))
        return out

    def sums(self) -> np.ndarray:
        out = np.zeros(5)
        self._chk(self._L.efvm_conservation_sums(self._p, _pd(out)))
        return out

    def device_planes(self):
        arr = (C.c_void_p * 4)()
        self._chk(self._L.efvm_device_state(self._p, arr))
        return [arr[m] for m in range(4)]

    def snapshot(self, restore=False):
        self._chk(self._L.efvm_snapshot(self._p, 1 if restore else 0))

    def set_ghost_table(self, side: int, states, given):
        """BoundaryInputs of a FromOutside side (kind "table"): states (n, 5) in the wire order,
        given (n,) bools; faces not given copy the interior cell.  side 0..3 = XLo, XHi, YLo, YHi."""
        st = np.ascontiguousarray(np.asarray(states, dtype=np.float64).reshape(-1, 5))
        g = np.ascontiguousarray(np.asarray(given, dtype=np.uint8).reshape(-1))
        self._chk(self._L.efvm_set_ghost_table(self._p, int(side), _pd(st), g.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                               st.shape[0]))

    def checkpoint_save(self, path: str):
        """CFDRST01 restart file of the current state and clock (the reference's format)."""
        self._chk(self._L.efvm_checkpoint_save(self._p, os.fsencode(path)))

    def checkpoint_load(self, path: str):
        """Resume from a step-boundary CFDRST01 file of the same grid."""
        self._chk(self._L.efvm_checkpoint_load(self._p, os.fsencode(path)))

    def write_vtk(self, path: str):
        """VTK legacy output of the current state (OutputPlugin's file)."""
        u = np.ascontiguousarray(self.download(), dtype=np.float64)
        self._chk(self._L.efvm_write_vtk(C.byref(self.grid), C.byref(self.gas), _pd(u), self.nx * self.ny,
                                         os.fsencode(path)))

    def set_learned(self, mlp, bypass=False):
        """Use the network's current weights for every face state (None detaches).  bypass:
        evaluate the gate only where the ungated face value is near the TVD bound."""
        _bind_mlp(self._L)
        self._chk(self._L.efvm_set_learned(self._p, None if mlp is None else mlp._p))
        self._chk(self._L.efvm_set_learned_bypass(self._p, 1 if bypass else 0))

    def set_learned_fused(self, allow=True):
        """Bypass stages with the 4-tap correction as one fused stage kernel (faces, fluxes and
        update in one pass; cells next to a deferred face finished by a second pass), bitwise
        the separate passes."""
        _bind_mlp(self._L)
        self._chk(self._L.efvm_set_learned_fused(self._p, 1 if allow else 0))

    def learned_fused(self):
        """(allowed, whether the next stage takes the fused path)."""
        _bind_mlp(self._L)
        a, f = C.c_int(), C.c_int()
        self._chk(self._L.efvm_get_learned_fused(self._p, C.byref(a), C.byref(f)))
        return bool(a.value), bool(f.value)

    def learned_stats(self, reset=False):
        """(faces that fell back to MUSCL, faces evaluated, face sides the network gated)."""
        _bind_mlp(self._L)
        out = (C.c_ulonglong * 3)()
        self._chk(self._L.efvm_mlp_infer_stats(self._p, 1 if reset else 0, out))
        return int(out[0]), int(out[1]), int(out[2])

    def learned_faces(self):
        """Face primitives of the current state: (2 axes, 2 sides, 4 vars, ny, nx)."""
        _bind_mlp(self._L)
        out = np.empty((2, 2, 4, self.ny, self.nx), dtype=np.float32)
        self._chk(self._L.efvm_learned_faces(self._p, _pf(out), out.size))
        return out


class Case(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("lo_x", C.c_double), ("lo_y", C.c_double),
                ("hi_x", C.c_double), ("hi_y", C.c_double), ("dz", C.c_double), ("gamma", C.c_double),
                ("sides", Bc * 4)]


def _bind_batch(L):
    if getattr(L, "_batch_bound", False):
        return
    vp, d, pd = C.c_void_p, C.c_double, C.POINTER(C.c_double)
    sig = {
        "efvm_batch_create": [C.c_int, C.POINTER(Case), C.POINTER(Scheme), C.POINTER(vp)],
        "efvm_batch_set_launch": [vp, C.c_int, C.c_int],
        "efvm_batch_set_grid": [vp, C.c_int, C.c_int],
        "efvm_batch_upload": [vp, C.c_int, pd, C.c_size_t],
        "efvm_batch_download": [vp, C.c_int, pd, C.c_size_t],
        "efvm_batch_step": [vp, C.c_int],
        "efvm_batch_graph_capture": [vp, C.c_int],
        "efvm_batch_graph_launch": [vp],
        "efvm_batch_clock": [vp, C.c_int, pd, pd, C.POINTER(C.c_long)],
        "efvm_batch_synchronize": [vp],
        "efvm_batch_last_elapsed_ms": [vp, pd],
        "efvm_batch_snapshot": [vp, C.c_int],
        "efvm_batch_set_rows": [vp, C.POINTER(C.c_int)],
        "efvm_batch_set_pair_kernel": [vp, C.c_int],
        "efvm_batch_set_fused": [vp, C.c_int, C.c_int],
        "efvm_batch_get_fused": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.efvm_batch_destroy.argtypes = [vp]
    L.efvm_batch_destroy.restype = None
    L.efvm_batch_last_error.argtypes = [vp]
    L.efvm_batch_last_error.restype = C.c_char_p
    L._batch_bound = True


EXPORTS += [
    "efvm_batch_create", "efvm_batch_destroy", "efvm_batch_last_error", "efvm_batch_set_launch",
    "efvm_batch_set_grid", "efvm_batch_upload", "efvm_batch_download", "efvm_batch_step",
    "efvm_batch_graph_capture", "efvm_batch_graph_launch", "efvm_batch_clock", "efvm_batch_synchronize",
    "efvm_batch_last_elapsed_ms", "efvm_batch_snapshot", "efvm_batch_set_rows", "efvm_batch_set_pair_kernel", "efvm_batch_set_fused", "efvm_batch_get_fused",
]


class Batch:
    """A batched ensemble of independent 2-D cases (FP32, RK3) in one persistent launch per stage.

    cases: list of dicts {nx, ny, lo, hi, dz, gamma, bcs} with bcs as in Context.set_boundaries.
    """

    def __init__(self, cases, flux="hllc", muscl=True, limiter="mc", cfl=0.4, fixed_dt=None):
        L = load_library()
        _bind_batch(L)
        self._L = L
        arr = (Case * len(cases))()
        self.shapes = []
        for i, c in enumerate(cases):
            lo, hi = c.get("lo", (0.0, 0.0)), c["hi"]
            arr[i].nx, arr[i].ny = c["nx"], c["ny"]
            arr[i].lo_x, arr[i].lo_y, arr[i].hi_x, arr[i].hi_y = lo[0], lo[1], hi[0], hi[1]
            arr[i].dz = c.get("dz", (hi[0] - lo[0]) / c["nx"])
            arr[i].gamma = c.get("gamma", 1.4)
            for s, spec in enumerate(c["bcs"]):
                kind, par = (spec, ()) if isinstance(spec, str) else spec
                arr[i].sides[s].kind = BC[kind]
                for q, v in enumerate(par):
                    arr[i].sides[s].par[q] = v
            self.shapes.append((c["ny"], c["nx"]))
        sc = Scheme(FLUX[flux], 1 if muscl else 0, LIMITER[limiter], 1, 1, 0, 0, cfl,
                    -1.0 if fixed_dt is None else float(fixed_dt))
        p = C.c_void_p()
        st = L.efvm_batch_create(len(cases), arr, C.byref(sc), C.byref(p))
        if st != OK:
            raise EfvmError(st, "efvm_batch_create failed")
        self._p = p
        self.n = len(cases)

    def close(self):
        if getattr(self, "_p", None):
            self._L.efvm_batch_destroy(self._p)
            self._p = None

    def __del__(self):
        self.close()

    def _chk(self, st):
        if st != OK:
            raise EfvmError(st, self._L.efvm_batch_last_error(self._p).decode())

    def set_launch(self, strip_rows, case_id=-1):
        self._chk(self._L.efvm_batch_set_launch(self._p, case_id, strip_rows))

    def set_rows(self, rows):
        arr = (C.c_int * self.n)(*[int(r) for r in rows])
        self._chk(self._L.efvm_batch_set_rows(self._p, arr))

    def set_grid(self, blocks, waves_per_block=4):
        self._chk(self._L.efvm_batch_set_grid(self._p, blocks, waves_per_block))

    def set_pair_kernel(self, allow: bool):
        """Allow (default) or forbid the two-column strip kernel for the cases it takes."""
        self._chk(self._L.efvm_batch_set_pair_kernel(self._p, int(bool(allow))))

    def set_fused(self, allow: bool = True, strip_rows: int = 0):
        """Allow (default) or forbid the fused stage-2+3 launch for the periodic two-column cases,
        and set its strip height (0 keeps the current one)."""
        self._chk(self._L.efvm_batch_set_fused(self._p, int(bool(allow)), strip_rows))

    def fused(self):
        """(allowed, strip_rows, number of cases running fused)."""
        a, r, n = C.c_int(), C.c_int(), C.c_int()
        self._chk(self._L.efvm_batch_get_fused(self._p, C.byref(a), C.byref(r), C.byref(n)))
        return bool(a.value), r.value, n.value

    def upload(self, i, cons):
        a = np.ascontiguousarray(cons, dtype=np.float64).reshape(-1, 5)
        self._chk(self._L.efvm_batch_upload(self._p, i, _pd(a), a.shape[0]))

    def download(self, i):
        ny, nx = self.shapes[i]
        out = np.empty((ny, nx, 5))
        self._chk(self._L.efvm_batch_download(self._p, i, _pd(out), nx * ny))
        return out

    def step(self, n=1):
        self._chk(self._L.efvm_batch_step(self._p, n))

    def graph_capture(self, n_steps):
        self._chk(self._L.efvm_batch_graph_capture(self._p, n_steps))

    def graph_launch(self):
        self._chk(self._L.efvm_batch_graph_launch(self._p))

    def synchronize(self):
        self._chk(self._L.efvm_batch_synchronize(self._p))

    def clock(self, i):
        dt, t, s = C.c_double(), C.c_double(), C.c_long()
        self._chk(self._L.efvm_batch_clock(self._p, i, C.byref(dt), C.byref(t), C.byref(s)))
        return dict(dt=dt.value, time=t.value, step=s.value)

    def last_elapsed_ms(self):
        d = C.c_double()
        self._chk(self._L.efvm_batch_last_elapsed_ms(self._p, C.byref(d)))
        return d.value

    def snapshot(self, restore=False):
        self._chk(self._L.efvm_batch_snapshot(self._p, 1 if restore else 0))


# ---------------------------------------------------------------------------
# learned face reconstruction
# ---------------------------------------------------------------------------
NPARAM = 5848


def _bind_mlp(L):
    if getattr(L, "_mlp_bound", False):
        return
    vp, pf, pd, sz = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double), C.c_size_t
    sig = {
        "efvm_mlp_create": [C.POINTER(vp)],
        "efvm_mlp_load": [vp, pf, sz],
        "efvm_mlp_save": [vp, pf, sz],
        "efvm_mlp_add_snapshot": [vp, vp, C.c_int],
        "efvm_mlp_add_snapshot_pair": [vp, vp, vp, C.c_int],
        "efvm_mlp_add_samples": [vp, pf, pf, sz],
        "efvm_mlp_get_samples": [vp, sz, sz, pf, pf],
        "efvm_mlp_num_samples": [vp, C.POINTER(sz)],
        "efvm_mlp_clear_samples": [vp],
        "efvm_mlp_train_step": [vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_ulonglong, C.c_int, pd],
        "efvm_mlp_gradient": [vp, pf, pf, sz, pf, pd],
        "efvm_mlp_fit_linear": [vp, sz, C.c_double, C.c_int],
        "efvm_mlp_infer": [vp, pf, sz, pf],
        "efvm_set_learned": [vp, vp],
        "efvm_mlp_infer_stats": [vp, C.c_int, C.POINTER(C.c_ulonglong)],
        "efvm_set_learned_bypass": [vp, C.c_int],
        "efvm_set_learned_fused": [vp, C.c_int],
        "efvm_get_learned_fused": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "efvm_learned_faces": [vp, pf, sz],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.efvm_mlp_destroy.argtypes = [vp]
    L.efvm_mlp_destroy.restype = None
    L.efvm_mlp_last_error.argtypes = [vp]
    L.efvm_mlp_last_error.restype = C.c_char_p
    L._mlp_bound = True


EXPORTS += [
    "efvm_mlp_create", "efvm_mlp_destroy", "efvm_mlp_last_error", "efvm_mlp_load", "efvm_mlp_save",
    "efvm_mlp_add_snapshot", "efvm_mlp_add_snapshot_pair", "efvm_mlp_add_samples", "efvm_mlp_get_samples", "efvm_mlp_num_samples",
    "efvm_mlp_clear_samples", "efvm_mlp_train_step", "efvm_mlp_gradient", "efvm_mlp_fit_linear", "efvm_mlp_infer", "efvm_set_learned",
    "efvm_mlp_infer_stats", "efvm_learned_faces", "efvm_set_learned_bypass", "efvm_set_learned_fused",
    "efvm_get_learned_fused",
]


def _pf(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """The face-reconstruction network (device parameters, Adam state and training samples)."""

    def __init__(self, params=None):
        L = load_library()
        _bind_mlp(L)
        self._L = L
        p = C.c_void_p()
        st = L.efvm_mlp_create(C.byref(p))
        if st != OK:
            raise EfvmError(st, "efvm_mlp_create failed")
        self._p = p
        if params is not None:
            self.load(params)

    def close(self):
        if getattr(self, "_p", None):
            self._L.efvm_mlp_destroy(self._p)
            self._p = None

    def __del__(self):
        self.close()

    def _chk(self, st):
        if st != OK:
            raise EfvmError(st, self._L.efvm_mlp_last_error(self._p).decode())

    def load(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        self._chk(self._L.efvm_mlp_load(self._p, _pf(a), a.size))

    def save(self) -> np.ndarray:
        a = np.empty(NPARAM, dtype=np.float32)
        self._chk(self._L.efvm_mlp_save(self._p, _pf(a), a.size))
        return a

    def add_snapshot(self, fine: "Context", factor: int, coarse: "Context" = None):
        if coarse is None:
            self._chk(self._L.efvm_mlp_add_snapshot(self._p, fine._p, int(factor)))
        else:
            self._chk(self._L.efvm_mlp_add_snapshot_pair(self._p, fine._p, coarse._p, int(factor)))

    def add_samples(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        assert X.shape[1] == 20 and Y.shape[1] == 4 and len(X) == len(Y)
        self._chk(self._L.efvm_mlp_add_samples(self._p, _pf(X), _pf(Y), len(X)))

    def get_samples(self, first, n):
        X = np.empty((n, 20), dtype=np.float32)
        Y = np.empty((n, 4), dtype=np.float32)
        self._chk(self._L.efvm_mlp_get_samples(self._p, first, n, _pf(X), _pf(Y)))
        return X, Y

    def num_samples(self) -> int:
        n = C.c_size_t()
        self._chk(self._L.efvm_mlp_num_samples(self._p, C.byref(n)))
        return n.value

    def clear_samples(self):
        self._chk(self._L.efvm_mlp_clear_samples(self._p))

    def train(self, steps, batch=8192, lr0=2e-3, lr1=2e-5, seed=1, train_skip=True):
        losses = np.zeros(max(steps, 1))
        self._chk(self._L.efvm_mlp_train_step(self._p, int(steps), int(batch), lr0, lr1, int(seed),
                                              1 if train_skip else 0, _pd(losses)))
        return losses[:steps]

    def fit_linear(self, stride=1, ridge=1e-9, constrained=True):
        """Closed-form weighted least-squares fit of the linear skip to the held samples
        (constrained: exact for cubic profiles, one fitted coefficient per variable)."""
        self._chk(self._L.efvm_mlp_fit_linear(self._p, int(stride), float(ridge), 1 if constrained else 0))

    def gradient(self, X, Y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.empty(NPARAM, dtype=np.float32)
        loss = C.c_double()
        self._chk(self._L.efvm_mlp_gradient(self._p, _pf(X), _pf(Y), len(X), _pf(g), C.byref(loss)))
        return loss.value, g

    def infer(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty((len(X), 4), dtype=np.float32)
        self._chk(self._L.efvm_mlp_infer(self._p, _pf(X), len(X), _pf(Y)))
        return Y



# ---------------------------------------------------------------------------
# 3-D structured meshes (efvm3_*)
# ---------------------------------------------------------------------------
class Grid3(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("lo", C.c_double * 3), ("hi", C.c_double * 3)]


EXPORTS += [
    "efvm3_create", "efvm3_destroy", "efvm3_last_error", "efvm3_set_boundaries", "efvm3_set_time_step",
    "efvm3_set_body_force", "efvm3_set_sponge", "efvm3_upload_state", "efvm3_download_state", "efvm3_get_clock",
    "efvm3_set_clock", "efvm3_cfl", "efvm3_run", "efvm3_step", "efvm3_synchronize", "efvm3_graph_capture",
    "efvm3_graph_launch", "efvm3_ibm_add_sphere", "efvm3_ibm_add_cylinder", "efvm3_marker_force", "efvm3_wall_force",
    "efvm3_conservation_sums", "efvm3_last_elapsed_ms", "efvm3_checkpoint_save", "efvm3_checkpoint_load",
    "efvm3_write_vtk", "efvm3_set_ghost_table", "efvm3_set_edges", "efvm3_write_vtk_edges",
]


def _bind3(L):
    if getattr(L, "_bound3", False):
        return
    vp, d, pd = C.c_void_p, C.c_double, C.POINTER(C.c_double)
    sig = {
        "efvm3_create": [C.POINTER(Grid3), C.POINTER(Gas), C.POINTER(Scheme), C.POINTER(vp)],
        "efvm3_set_boundaries": [vp, C.POINTER(Bc)],
        "efvm3_set_ghost_table": [vp, C.c_int, pd, C.POINTER(C.c_ubyte), C.c_size_t],
        "efvm3_set_time_step": [vp, d, d],
        "efvm3_set_body_force": [vp, d],
        "efvm3_set_sponge": [vp, d, d, d, pd],
        "efvm3_upload_state": [vp, pd, C.c_size_t],
        "efvm3_download_state": [vp, pd, C.c_size_t],
        "efvm3_get_clock": [vp, pd, pd, C.POINTER(C.c_long), C.POINTER(C.c_int)],
        "efvm3_set_clock": [vp, d, C.c_long],
        "efvm3_cfl": [vp, pd],
        "efvm3_run": [vp],
        "efvm3_step": [vp, C.c_int],
        "efvm3_synchronize": [vp],
        "efvm3_graph_capture": [vp, C.c_int],
        "efvm3_graph_launch": [vp],
        "efvm3_ibm_add_sphere": [vp, d, d, d, d, pd, C.c_int],
        "efvm3_ibm_add_cylinder": [vp, d, d, d, pd],
        "efvm3_marker_force": [vp, pd],
        "efvm3_wall_force": [vp, pd],
        "efvm3_conservation_sums": [vp, pd],
        "efvm3_last_elapsed_ms": [vp, pd],
        "efvm3_checkpoint_save": [vp, C.c_char_p],
        "efvm3_checkpoint_load": [vp, C.c_char_p],
        "efvm3_write_vtk": [C.POINTER(Grid3), C.POINTER(Gas), pd, C.c_size_t, C.c_char_p],
        "efvm3_set_edges": [vp, C.c_int, pd, C.c_size_t],
        "efvm3_write_vtk_edges": [C.POINTER(Grid3), pd, pd, C.POINTER(Gas), pd, C.c_size_t, C.c_char_p],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.efvm3_destroy.argtypes = [vp]
    L.efvm3_destroy.restype = None
    L.efvm3_last_error.argtypes = [vp]
    L.efvm3_last_error.restype = C.c_char_p
    L._bound3 = True


class Context3:
    """A 3-D UniformMesh context (nx, ny, nz cells, five stored components), 1:1 over efvm3_*."""

    def __init__(self, nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), gamma=1.4, r_gas=287.058, mu=None,
                 flux="hllc", muscl=False, limiter="minmod", integrator="rk3", precision="fp64",
                 viscous=False, seam_mode=0, cfl=0.4, fixed_dt=None):
        L = load_library()
        _bind3(L)
        self._L = L
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        g = Grid3(self.nx, self.ny, self.nz, (C.c_double * 3)(*lo), (C.c_double * 3)(*hi))
        gs = Gas(gamma, r_gas, 0 if mu is None else 1, 0.0 if mu is None else float(mu))
        sc = Scheme(FLUX[flux], 1 if muscl else 0, LIMITER[limiter], 1 if integrator == "rk3" else 0,
                    0 if precision == "fp64" else 1, int(viscous), int(seam_mode), cfl,
                    -1.0 if fixed_dt is None else float(fixed_dt))
        self.precision = precision
        self.grid, self.gas, self.scheme = g, gs, sc
        self.lo, self.hi = tuple(lo), tuple(hi)
        self.edges = {}
        p = C.c_void_p()
        st = L.efvm3_create(C.byref(g), C.byref(gs), C.byref(sc), C.byref(p))
        if st != OK:
            raise EfvmError(st, "efvm3_create failed")
        self._p = p

    def close(self):
        if getattr(self, "_p", None):
            self._L.efvm3_destroy(self._p)
            self._p = None

    def __del__(self):
        self.close()

    def _chk(self, st: int):
        if st != OK:
            raise EfvmError(st, self._L.efvm3_last_error(self._p).decode())

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def nraw(self) -> int:
        return (self.nx + 4) * (self.ny + 4) * (self.nz + 4)

    def set_boundaries(self, sides):
        """sides: 6 x (kind, params) for XLo, XHi, YLo, YHi, ZLo, ZHi."""
        arr = (Bc * 6)()
        for s, spec in enumerate(sides):
            kind, par = (spec, ()) if isinstance(spec, str) else spec
            arr[s].kind = BC[kind]
            for q, v in enumerate(par):
                arr[s].par[q] = v
        self._chk(self._L.efvm3_set_boundaries(self._p, arr))

    def set_edges(self, axis: int, edges):
        """Stretch axis 1 (y) or 2 (z): its n+1 strictly increasing cell edges from lo to hi
        (the reference's UniformMeshConfig y_edges / z_edges)."""
        e = np.ascontiguousarray(edges, dtype=np.float64)
        self._chk(self._L.efvm3_set_edges(self._p, int(axis), _pd(e), e.size))
        self.edges[int(axis)] = e.copy()

    def set_ghost_table(self, side: int, states, given):
        """BoundaryInputs of a FromOutside side (kind "table"), side 0..5 = XLo .. ZHi: states
        (n, 5) in the wire order and given (n,) bools, one per boundary face with the
        lower-numbered in-plane axis fastest (x side j + ny k, y side i + nx k, z side
        i + nx j); faces not given copy the interior cell."""
        st = np.ascontiguousarray(np.asarray(states, dtype=np.float64).reshape(-1, 5))
        g = np.ascontiguousarray(np.asarray(given, dtype=np.uint8).reshape(-1))
        if g.shape[0] != st.shape[0]:
            raise ValueError("states and given differ in length")
        self._chk(self._L.efvm3_set_ghost_table(self._p, int(side), _pd(st), g.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                                st.shape[0]))

    def set_time_step(self, cfl=0.4, fixed_dt=None):
        self._chk(self._L.efvm3_set_time_step(self._p, cfl, -1.0 if fixed_dt is None else float(fixed_dt)))

    def set_body_force(self, gx):
        self._chk(self._L.efvm3_set_body_force(self._p, gx))

    def set_sponge(self, x0, x1, sigma, target):
        t = np.ascontiguousarray(target, dtype=np.float64)
        self._chk(self._L.efvm3_set_sponge(self._p, x0, x1, sigma, _pd(t)))

    def upload(self, cons: np.ndarray):
        """Interior (nz, ny, nx, 5) or the raw padded array (nz+4, ny+4, nx+4, 5)."""
        a = np.ascontiguousarray(cons, dtype=np.float64).reshape(-1, 5)
        if a.shape[0] not in (self.ncell, self.nraw):
            raise ValueError("state has the wrong number of cells")
        self._chk(self._L.efvm3_upload_state(self._p, _pd(a), a.shape[0]))

    def download(self, raw=False) -> np.ndarray:
        shape = (self.nz + 4, self.ny + 4, self.nx + 4, 5) if raw else (self.nz, self.ny, self.nx, 5)
        out = np.empty(shape)
        self._chk(self._L.efvm3_download_state(self._p, _pd(out), out.size // 5))
        return out

    def clock(self):
        dt, t = C.c_double(), C.c_double()
        st, sg = C.c_long(), C.c_int()
        self._chk(self._L.efvm3_get_clock(self._p, C.byref(dt), C.byref(t), C.byref(st), C.byref(sg)))
        return dict(dt=dt.value, time=t.value, step=st.value, stage=sg.value)

    def set_clock(self, time=0.0, step=0):
        self._chk(self._L.efvm3_set_clock(self._p, time, step))

    def cfl_dt(self) -> float:
        d = C.c_double()
        self._chk(self._L.efvm3_cfl(self._p, C.byref(d)))
        return d.value

    def run(self):
        self._chk(self._L.efvm3_run(self._p))

    def step(self, n=1):
        self._chk(self._L.efvm3_step(self._p, n))

    def synchronize(self):
        self._chk(self._L.efvm3_synchronize(self._p))

    def graph_capture(self, n_steps):
        self._chk(self._L.efvm3_graph_capture(self._p, n_steps))

    def graph_launch(self):
        self._chk(self._L.efvm3_graph_launch(self._p))

    def last_elapsed_ms(self) -> float:
        d = C.c_double()
        self._chk(self._L.efvm3_last_elapsed_ms(self._p, C.byref(d)))
        return d.value

    def add_sphere(self, center, r, freestream, markers=2000):
        fs = np.ascontiguousarray(freestream, dtype=np.float64)
        self._chk(self._L.efvm3_ibm_add_sphere(self._p, center[0], center[1], center[2], r, _pd(fs), int(markers)))

    def add_cylinder(self, cx, cy, r, freestream):
        fs = np.ascontiguousarray(freestream, dtype=np.float64)
        self._chk(self._L.efvm3_ibm_add_cylinder(self._p, cx, cy, r, _pd(fs)))

    def marker_force(self) -> np.ndarray:
        out = np.zeros(6)
        self._chk(self._L.efvm3_marker_force(self._p, _pd(out)))
        return out

    def wall_force(self) -> np.ndarray:
        """Staircase wall force (pressure xyz, viscous xyz) on the bodies, FP64."""
        out = np.zeros(6)
        self._chk(self._L.efvm3_wall_force(self._p, _pd(out)))
        return out

    def sums(self) -> np.ndarray:
        out = np.zeros(5)
        self._chk(self._L.efvm3_conservation_sums(self._p, _pd(out)))
        return out

    def checkpoint_save(self, path: str):
        """CFDRST01 restart file of the current state (raw padded array) and clock."""
        self._chk(self._L.efvm3_checkpoint_save(self._p, os.fsencode(path)))

    def checkpoint_load(self, path: str):
        """Resume from a step-boundary CFDRST01 file of the same grid."""
        self._chk(self._L.efvm3_checkpoint_load(self._p, os.fsencode(path)))

    def write_vtk(self, path: str):
        """VTK legacy output of the current state."""
        write_vtk3(path, self.download(), self.nx, self.ny, self.nz, self.lo, self.hi, self.gas.gamma, self.gas.r_gas,
                   y_edges=self.edges.get(1), z_edges=self.edges.get(2))


def write_vtk3(path: str, state, nx: int, ny: int, nz: int, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), gamma=1.4,
               r_gas=287.058, y_edges=None, z_edges=None):
    """VTK legacy RECTILINEAR_GRID file of a wire-order 3-D state (nz, ny, nx, 5) on the host;
    y_edges / z_edges: the cell edges of a stretched axis."""
    L = load_library()
    _bind3(L)
    g = Grid3(int(nx), int(ny), int(nz), (C.c_double * 3)(*lo), (C.c_double * 3)(*hi))
    gs = Gas(gamma, r_gas, 0, 0.0)
    u = np.ascontiguousarray(state, dtype=np.float64).reshape(-1)
    if u.size != 5 * nx * ny * nz:
        raise ValueError("state must hold 5 values per cell")
    pu = u.ctypes.data_as(C.POINTER(C.c_double))
    if y_edges is None and z_edges is None:
        st = L.efvm3_write_vtk(C.byref(g), C.byref(gs), pu, nx * ny * nz, os.fsencode(path))
    else:
        ey = None if y_edges is None else np.ascontiguousarray(y_edges, dtype=np.float64)
        ez = None if z_edges is None else np.ascontiguousarray(z_edges, dtype=np.float64)
        for e, n in ((ey, ny), (ez, nz)):
            if e is not None and e.size != n + 1:
                raise ValueError("edges: n + 1 values along the axis")
        st = L.efvm3_write_vtk_edges(C.byref(g), None if ey is None else _pd(ey), None if ez is None else _pd(ez),
                                     C.byref(gs), pu, nx * ny * nz, os.fsencode(path))
    if st != OK:
        raise EfvmError(st, f"efvm3_write_vtk({path}) failed")
