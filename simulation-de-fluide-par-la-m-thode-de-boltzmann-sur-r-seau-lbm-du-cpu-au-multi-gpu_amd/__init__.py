"""MI355X (gfx950) D2Q9 lattice-Boltzmann time step: Python host mirror of the C ABI.

The compute path is liblbm_hip.so (hand-written HIP kernels behind the C ABI declared in
include/lbm_hip.h); this module only binds it with ctypes and mirrors the reference solver
interface (SolverV3, src/v3_cuda_opt/solver_v3.hpp:27-57) with numpy arrays:

    s = Solver(nx, ny, tau, force_x, force_y, solid, precision="float")
    s.init_equilibrium(1.0, 0.0, 0.0); s.run(1000)
    rho, ux, uy = s.macroscopic(); S = s.populations()

Errors keep the reference's split: invalid arguments raise ValueError (std::invalid_argument
in the reference), runtime/HIP failures raise RuntimeError. There is no CPU fallback: if the
library is missing, importing the solver raises.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liblbm_hip.so"

LBM_OK, LBM_EINVAL, LBM_EHIP, LBM_ESTATE, LBM_ENOMEM = 0, -1, -2, -3, -4
FLOAT, DOUBLE, FP16 = 0, 1, 2
_PRECISION = {"float": FLOAT, "f32": FLOAT, "fp32": FLOAT, "double": DOUBLE, "f64": DOUBLE,
              "fp64": DOUBLE, "fp16": FP16, "half": FP16}

_lib = None
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class DomainDesc(C.Structure):
    """lbm_domain_desc (include/lbm_hip.h)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("tau", C.c_double), ("force_x", C.c_double),
                ("force_y", C.c_double), ("solid", C.POINTER(C.c_uint8)), ("rho0", C.c_double),
                ("ux0", C.c_double), ("uy0", C.c_double)]


def library_path() -> pathlib.Path:
    return LIB_PATH


def lib():
    """The loaded liblbm_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with python build.py "
                               "(there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    i, d, vp, sz = C.c_int, C.c_double, C.c_void_p, C.c_size_t
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    llp = C.POINTER(C.c_longlong)
    sig = {
        "lbm_last_error": ([], C.c_char_p),
        "lbm_device_info": ([C.c_char_p, sz], i),
        "lbm_create": ([i, i, d, d, d, _u8p, i, C.POINTER(vp)], i),
        "lbm_destroy": ([vp], i),
        "lbm_set_inflow_outflow": ([vp, d], i),
        "lbm_init_equilibrium": ([vp, d, d, d], i),
        "lbm_set_mask": ([vp, _u8p], i),
        "lbm_set_force": ([vp, d, d], i),
        "lbm_set_force_field": ([vp, vp], i),
        "lbm_set_tau": ([vp, d], i),
        "lbm_set_block": ([vp, i, i], i),
        "lbm_set_steps_per_launch": ([vp, i, i, i, i], i),
        "lbm_get_steps_per_launch": ([vp, ip, ip, ip, ip, ip], i),
        "lbm_get_block": ([vp, ip, ip], i),
        "lbm_set_kernel_options": ([vp, i, i, i], i),
        "lbm_get_kernel_options": ([vp, ip, ip, ip], i),
        "lbm_step": ([vp], i),
        "lbm_run": ([vp, i], i),
        "lbm_timed_run": ([vp, i, dp], i),
        "lbm_macroscopic": ([vp, vp, vp, vp], i),
        "lbm_macroscopic_noalloc": ([vp, vp, vp, vp], i),
        "lbm_vorticity": ([vp, vp, i], i),
        "lbm_total_mass": ([vp, dp], i),
        "lbm_momentum": ([vp, dp, dp], i),
        "lbm_populations": ([vp, vp], i),
        "lbm_populations_at": ([vp, i, i, vp], i),
        "lbm_set_populations": ([vp, vp], i),
        "lbm_get_state": ([vp, vp], i),
        "lbm_set_state": ([vp, vp], i),
        "lbm_threads": ([vp], i),
        "lbm_bench_info": ([vp, C.c_char_p, sz], i),
        "lbm_dims": ([vp, ip, ip, ip], i),
        "lbm_set_fp16_scale": ([vp, d], i),
        "lbm_autotune": ([vp, i, ip, ip, dp], i),
        "lbm_poiseuille_force": ([i, i, d, d], d),
        "lbm_momentum_exchange": ([vp, i, i, i, i, dp, dp], i),
        "lbm_copy_ceiling": ([sz, i, dp, dp, dp, dp], i),
        "lbm_strip_create": ([i, i, i, C.POINTER(vp)], i),
        "lbm_strip_create_ghost": ([i, i, i, i, C.POINTER(vp)], i),
        "lbm_strip_halo_values": ([vp, C.POINTER(sz)], i),
        "lbm_strip_compute_rows2": ([vp, i, i, i], i),
        "lbm_strip_compute_rows2_to": ([vp, i, i, i, vp, vp], i),
        "lbm_strip_set_two_step_tile": ([vp, i, i, i], i),
        "lbm_strip_destroy": ([vp], i),
        "lbm_strip_set_params": ([vp, d, d, d, i, d], i),
        "lbm_strip_upload_solid": ([vp, _u8p], i),
        "lbm_strip_upload_state": ([vp, vp], i),
        "lbm_strip_download_state": ([vp, vp], i),
        "lbm_strip_compute_rows": ([vp, i, i, i], i),
        "lbm_strip_pack_halo": ([vp, vp, vp], i),
        "lbm_strip_unpack_halo": ([vp, vp, vp], i),
        "lbm_strip_finish_step": ([vp], i),
        "lbm_strip_set_block": ([vp, i, i], i),
        "lbm_strip_timed_steps": ([vp, i, dp], i),
        "lbm_batch_create": ([i, C.POINTER(DomainDesc), i, C.POINTER(vp)], i),
        "lbm_batch_destroy": ([vp], i),
        "lbm_batch_set_graph_steps": ([vp, i], i),
        "lbm_batch_set_shape": ([vp, i, i, i], i),
        "lbm_batch_run": ([vp, i], i),
        "lbm_batch_timed_run": ([vp, i, dp], i),
        "lbm_batch_populations": ([vp, i, vp], i),
        "lbm_batch_total_cells": ([vp, llp], i),
        "lbm_batch_autotune": ([vp, i, ip, ip, ip, dp], i),
        "lbm_batch_set_order": ([vp, i, i], i),
        "lbm_batch_get_order": ([vp, ip, ip], i),
        "lbm_batch_set_steps_per_launch": ([vp, i, i, i, i], i),
        "lbm_batch_get_steps_per_launch": ([vp, ip, ip, ip, ip], i),
        "lbm_batch_two_step_layout": ([vp, ip, ip, ip, ip, dp], i),
        "lbm_batch_set_two_step_threads": ([vp, i], i),
        "lbm_adjoint_create": ([i, i, d, _u8p, i, i, i, i, i, i, i, i, C.POINTER(vp)], i),
        "lbm_adjoint_destroy": ([vp], i),
        "lbm_adjoint_set_force_field": ([vp, vp], i),
        "lbm_adjoint_set_tau": ([vp, d], i),
        "lbm_adjoint_forward": ([vp, dp, dp], i),
        "lbm_adjoint_gradient": ([vp, dp, vp, dp, dp, llp], i),
        "lbm_checkpoint_forward_steps": ([i, i, llp], i),
        "lbm_adjoint_set_schedule": ([vp, i], i),
        "lbm_adjoint_schedule_info": ([vp, ip, ip, ip, llp, llp], i),
        "lbm_checkpoint_schedule": ([i, i, i, ip, ip, llp, llp], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


def last_error() -> str:
    return lib().lbm_last_error().decode()


def check(rc: int) -> None:
    """Maps a C-ABI status to the reference's exception types."""
    if rc == LBM_OK:
        return
    msg = last_error()
    if rc == LBM_EINVAL:
        raise ValueError(msg)
    if rc == LBM_ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def precision_code(precision) -> int:
    if isinstance(precision, int):
        return precision
    try:
        return _PRECISION[precision]
    except KeyError:
        raise ValueError(f"unknown precision {precision!r} (float, double or fp16)") from None


def real_dtype(precision) -> type:
    return np.float64 if precision_code(precision) == DOUBLE else np.float32


def storage_dtype(precision) -> type:
    p = precision_code(precision)
    return np.float64 if p == DOUBLE else (np.uint16 if p == FP16 else np.float32)


def device_info() -> str:
    buf = C.create_string_buffer(512)
    check(lib().lbm_device_info(buf, len(buf)))
    return buf.value.decode()


def poiseuille_force(H: int, tau: float, u_max: float, precision="float") -> float:
    """Reference poiseuille_force (src/common/geometry.hpp:24-26) in the solver precision."""
    return lib().lbm_poiseuille_force(precision_code(precision), int(H), float(tau), float(u_max))


def copy_ceiling(nbytes: int = 1 << 32, reps: int = 20) -> dict:
    """HBM copy ceilings (GB/s, best of reps, two arrays of nbytes): the best grid-stride copy
    kernel, hipMemcpyDtoD, the reference's own copy configuration (bench/gpu_bandwidth.cu:
    17-22,66-67), and a nine-array copy (the stream structure of nine population planes)."""
    a, b, r, n = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    check(lib().lbm_copy_ceiling(int(nbytes), int(reps), C.byref(a), C.byref(b), C.byref(r),
                                 C.byref(n)))
    return {"copy_kernel": a.value, "memcpy": b.value, "reference_kernel": r.value,
            "nine_array_copy": n.value}


# ---- geometry (src/common/geometry.cpp) ------------------------------------------------------

def make_channel(nx: int, ny: int) -> np.ndarray:
    """Solid rows y = 0 and y = ny - 1, periodic in x (geometry.cpp:5-12); flat uint8 mask."""
    m = np.zeros((ny, nx), dtype=np.uint8)
    m[0, :] = 1
    m[ny - 1, :] = 1
    return m.reshape(-1)


def make_cylinder(nx: int, ny: int, xc: float, yc: float, D: float) -> np.ndarray:
    """Staircase disc (x - xc)^2 + (y - yc)^2 <= D^2/4, evaluated in double (geometry.cpp:14-23)."""
    y, x = np.mgrid[0:ny, 0:nx].astype(np.float64)
    dx, dy = x - xc, y - yc
    return (dx * dx + dy * dy <= 0.25 * D * D).astype(np.uint8).reshape(-1)


def seeded_block_position(seed: int, nx: int, ny: int, size: int = 8) -> tuple[int, int]:
    """Lower-left corner of the seeded square obstacle (the reference has no RNG; this is the
    build's seed -> position map): numpy default_rng(seed), x0 uniform in [0, nx - size], y0
    uniform in [1, ny - size - 1] (clear of the walls)."""
    rng = np.random.default_rng(seed)
    x0 = int(rng.integers(0, nx - size + 1))
    y0 = int(rng.integers(1, ny - size))
    return x0, y0


def seeded_block_mask(seed: int, nx: int, ny: int, size: int = 8) -> np.ndarray:
    """Walled channel plus a size x size solid block placed by `seed`."""
    m = make_channel(nx, ny).reshape(ny, nx)
    x0, y0 = seeded_block_position(seed, nx, ny, size)
    m[y0:y0 + size, x0:x0 + size] = 1
    return m.reshape(-1)


def add_disc(mask: np.ndarray, nx: int, ny: int, xc: float, yc: float, r: float) -> np.ndarray:
    """Adds a staircase disc of radius r to a flat mask (in place) and returns it."""
    return np.maximum(mask, make_cylinder(nx, ny, xc, yc, 2.0 * r), out=mask)


# ---- the plain solver ------------------------------------------------------------------------

class Solver:
    """SolverV3's interface on the MI355X step (src/v3_cuda_opt/solver_v3.hpp:27-57)."""

    def __init__(self, nx, ny, tau, force_x, force_y, solid, precision="float"):
        self.precision = precision_code(precision)
        self.dtype = real_dtype(self.precision)
        self._nx, self._ny = int(nx), int(ny)
        self._solid = np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
        if self._solid.size != self._nx * self._ny:
            raise ValueError("solid mask size does not match the grid")
        h = C.c_void_p()
        check(lib().lbm_create(self._nx, self._ny, float(tau), float(force_x), float(force_y),
                               self._solid, self.precision, C.byref(h)))
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.lbm_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # reference members
    def nx(self):
        return self._nx

    def ny(self):
        return self._ny

    def solid(self):
        return self._solid

    @property
    def N(self):
        return self._nx * self._ny

    def set_inflow_outflow(self, u_in):
        check(lib().lbm_set_inflow_outflow(self._h, float(u_in)))

    def init_equilibrium(self, rho0, ux0, uy0):
        check(lib().lbm_init_equilibrium(self._h, float(rho0), float(ux0), float(uy0)))

    def step(self, n: int | None = None):
        """One asynchronous step (reference step()); step(n) = run(n)."""
        if n is None:
            check(lib().lbm_step(self._h))
        else:
            self.run(n)

    def run(self, steps):
        check(lib().lbm_run(self._h, int(steps)))

    def timed_run(self, steps) -> float:
        sec = C.c_double()
        check(lib().lbm_timed_run(self._h, int(steps), C.byref(sec)))
        return sec.value

    def macroscopic(self):
        r, u, v = (np.empty(self.N, dtype=self.dtype) for _ in range(3))
        check(lib().lbm_macroscopic(self._h, r.ctypes.data, u.ctypes.data, v.ctypes.data))
        return r, u, v

    def macroscopic_noalloc(self):
        """macroscopic() without device allocation: the fields are staged in the library's
        fixed 12 MiB device array (3 * nx * ny reals must fit; larger grids raise ValueError)."""
        r, u, v = (np.empty(self.N, dtype=self.dtype) for _ in range(3))
        check(lib().lbm_macroscopic_noalloc(self._h, r.ctypes.data, u.ctypes.data,
                                            v.ctypes.data))
        return r, u, v

    def vorticity(self, periodic_x: int = -1) -> np.ndarray:
        """io.cpp:54-74 vorticity of the current velocity, on the device: periodic_x 1
        periodic, 0 one-sided at the open ends, -1 the solver's own x boundary."""
        w = np.empty(self.N, dtype=self.dtype)
        check(lib().lbm_vorticity(self._h, w.ctypes.data, int(periodic_x)))
        return w

    def total_mass(self) -> float:
        m = C.c_double()
        check(lib().lbm_total_mass(self._h, C.byref(m)))
        return m.value

    def momentum(self) -> tuple[float, float]:
        px, py = C.c_double(), C.c_double()
        check(lib().lbm_momentum(self._h, C.byref(px), C.byref(py)))
        return px.value, py.value

    def populations(self) -> np.ndarray:
        out = np.empty(9 * self.N, dtype=self.dtype)
        check(lib().lbm_populations(self._h, out.ctypes.data))
        return out

    def momentum_exchange(self, x0, y0, w, h) -> tuple[float, float]:
        """(drag, lift) on the solids bordering the fluid cells of a box (FP64)."""
        fx, fy = C.c_double(), C.c_double()
        check(lib().lbm_momentum_exchange(self._h, int(x0), int(y0), int(w), int(h), C.byref(fx),
                                          C.byref(fy)))
        return fx.value, fy.value

    def populations_at(self, x, y) -> np.ndarray:
        out = np.empty(9, dtype=self.dtype)
        check(lib().lbm_populations_at(self._h, int(x), int(y), out.ctypes.data))
        return out

    def set_populations(self, S):
        a = np.ascontiguousarray(S, dtype=self.dtype).reshape(-1)
        if a.size != 9 * self.N:
            raise ValueError("populations must hold 9 * nx * ny values")
        check(lib().lbm_set_populations(self._h, a.ctypes.data))

    def get_state(self) -> np.ndarray:
        out = np.empty(9 * self.N, dtype=storage_dtype(self.precision))
        check(lib().lbm_get_state(self._h, out.ctypes.data))
        return out

    def set_state(self, R):
        a = np.ascontiguousarray(R, dtype=storage_dtype(self.precision)).reshape(-1)
        if a.size != 9 * self.N:
            raise ValueError("state must hold 9 * nx * ny values")
        check(lib().lbm_set_state(self._h, a.ctypes.data))

    def threads(self) -> int:
        return lib().lbm_threads(self._h)

    def bench_info(self) -> str:
        buf = C.create_string_buffer(1024)
        check(lib().lbm_bench_info(self._h, buf, len(buf)))
        return buf.value.decode()

    # setters (the reference fixes these at construction)
    def set_mask(self, solid):
        m = np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
        if m.size != self.N:
            raise ValueError("solid mask size does not match the grid")
        check(lib().lbm_set_mask(self._h, m))
        self._solid = m

    def set_force(self, fx, fy):
        check(lib().lbm_set_force(self._h, float(fx), float(fy)))

    def set_force_field(self, fxy):
        if fxy is None:
            check(lib().lbm_set_force_field(self._h, None))
            return
        a = np.ascontiguousarray(fxy, dtype=self.dtype).reshape(-1)
        if a.size != 2 * self.N:
            raise ValueError("force field must hold 2 * nx * ny values (fx plane, fy plane)")
        self._force = a
        check(lib().lbm_set_force_field(self._h, a.ctypes.data))

    def set_tau(self, tau):
        check(lib().lbm_set_tau(self._h, float(tau)))

    def set_block(self, block_x, rows_per_thread):
        check(lib().lbm_set_block(self._h, int(block_x), int(rows_per_thread)))

    def set_steps_per_launch(self, steps_per_launch: int, block_x: int = 0, rows: int = 0,
                             tile_order: int = -1):
        """1: one step per launch; 2: run()/timed_run() advance pairs of steps in one launch of
        the two-step kernel where it applies (same bits). block_x: tile columns (0 keeps);
        rows: tile rows, -1 automatic (up to 64, filling the GPU evenly), 0 keeps; tile_order:
        0 round-robin, 1 XCD bands (default; FP16 storage: 2), 2 XCD rows, -1 keeps."""
        check(lib().lbm_set_steps_per_launch(self._h, int(steps_per_launch), int(block_x),
                                             int(rows), int(tile_order)))

    def steps_per_launch(self) -> dict:
        a, b, c, o, d = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().lbm_get_steps_per_launch(self._h, C.byref(a), C.byref(b), C.byref(c),
                                             C.byref(o), C.byref(d)))
        return {"steps_per_launch": a.value, "block_x": b.value, "rows": c.value,
                "tile_order": o.value, "in_use": bool(d.value)}

    def set_kernel_options(self, streaming_stores: bool, single_resource: bool,
                           tile_order: int = 0):
        """tile_order: 0 round-robin over the XCDs, 1 a band of tiles per XCD, 2 whole tile
        rows per XCD."""
        check(lib().lbm_set_kernel_options(self._h, int(bool(streaming_stores)),
                                           int(bool(single_resource)), int(tile_order)))

    def kernel_options(self) -> tuple[bool, bool, int]:
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().lbm_get_kernel_options(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return bool(a.value), bool(b.value), c.value

    def block(self) -> tuple[int, int]:
        bx, r = C.c_int(), C.c_int()
        check(lib().lbm_get_block(self._h, C.byref(bx), C.byref(r)))
        return bx.value, r.value

    def set_fp16_scale(self, scale):
        check(lib().lbm_set_fp16_scale(self._h, float(scale)))

    def autotune(self, steps=20):
        bx, r, m = C.c_int(), C.c_int(), C.c_double()
        check(lib().lbm_autotune(self._h, int(steps), C.byref(bx), C.byref(r), C.byref(m)))
        return bx.value, r.value, m.value


# ---- strip engine (one GPU's share of the v4 decomposition) ---------------------------------

HALO_UP = (2, 5, 6)    # c_y = +1: top owned row -> ghost row 0 of the strip above
HALO_DOWN = (4, 7, 8)  # c_y = -1: bottom owned row -> top ghost row of the strip below


class Strip:
    """StripEngine's interface (src/v4_mpi_cuda/engine.hpp:36-61) on the fused step: ny_local
    owned rows plus `ghost_rows` ghost rows below and above (1: the reference's strip; 2: one
    halo exchange per two fused steps, compute_rows2), SoA state of 9 * nx * (ny_local +
    2 * ghost_rows) reals. Owned rows are local rows ghost_rows .. ghost_rows + ny_local - 1."""

    def __init__(self, nx, ny_local, precision="float", ghost_rows=1):
        self.precision = precision_code(precision)
        self.dtype = real_dtype(self.precision)
        self.nx, self.ny_local, self.ghost = int(nx), int(ny_local), int(ghost_rows)
        h = C.c_void_p()
        check(lib().lbm_strip_create_ghost(self.nx, self.ny_local, self.precision, self.ghost,
                                           C.byref(h)))
        self._h = h
        n = C.c_size_t()
        check(lib().lbm_strip_halo_values(self._h, C.byref(n)))
        self.halo_values = n.value

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.lbm_strip_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    @property
    def rows(self):
        return self.ny_local + 2 * self.ghost

    @property
    def N(self):
        return self.nx * self.rows

    def set_params(self, tau, fx=0.0, fy=0.0, open_x=False, u_in=0.0):
        check(lib().lbm_strip_set_params(self._h, float(tau), float(fx), float(fy), int(open_x),
                                         float(u_in)))

    def upload_solid(self, solid_strip):
        m = np.ascontiguousarray(solid_strip, dtype=np.uint8).reshape(-1)
        if m.size != self.N:
            raise ValueError("strip mask must hold nx * (ny_local + 2 * ghost_rows) bytes")
        check(lib().lbm_strip_upload_solid(self._h, m))

    def upload_state(self, g):
        a = np.ascontiguousarray(g, dtype=self.dtype).reshape(-1)
        if a.size != 9 * self.N:
            raise ValueError("strip state must hold 9 * nx * (ny_local + 2 * ghost_rows) values")
        check(lib().lbm_strip_upload_state(self._h, a.ctypes.data))

    def download_state(self):
        out = np.empty(9 * self.N, dtype=self.dtype)
        check(lib().lbm_strip_download_state(self._h, out.ctypes.data))
        return out

    def compute_rows(self, first, last, stream=0):
        check(lib().lbm_strip_compute_rows(self._h, int(first), int(last), int(stream)))

    def compute_rows2(self, first, last, stream=0):
        """Two fused steps of owned local rows [first, last] (two ghost rows, periodic x)."""
        check(lib().lbm_strip_compute_rows2(self._h, int(first), int(last), int(stream)))

    def compute_rows2_to(self, first, last, stream, up: "Strip", down: "Strip"):
        """compute_rows2 that also stores the halo entries straight into the next-state ghost
        rows of the strips above (`up`) and below (`down`); no pack/unpack."""
        check(lib().lbm_strip_compute_rows2_to(self._h, int(first), int(last), int(stream),
                                               up._h, down._h))

    def set_two_step_tile(self, block_x=0, rows=0, tile_order=-1):
        check(lib().lbm_strip_set_two_step_tile(self._h, int(block_x), int(rows),
                                                int(tile_order)))

    def pack_halo(self):
        up = np.empty(self.halo_values, dtype=self.dtype)
        down = np.empty(self.halo_values, dtype=self.dtype)
        check(lib().lbm_strip_pack_halo(self._h, up.ctypes.data, down.ctypes.data))
        return up, down

    def unpack_halo(self, recv_from_down, recv_from_up):
        a = np.ascontiguousarray(recv_from_down, dtype=self.dtype).reshape(-1)
        b = np.ascontiguousarray(recv_from_up, dtype=self.dtype).reshape(-1)
        if a.size != self.halo_values or b.size != self.halo_values:
            raise ValueError(f"halo messages must hold {self.halo_values} values each")
        check(lib().lbm_strip_unpack_halo(self._h, a.ctypes.data, b.ctypes.data))

    def finish_step(self):
        check(lib().lbm_strip_finish_step(self._h))

    def set_block(self, block_x, rows_per_thread):
        check(lib().lbm_strip_set_block(self._h, int(block_x), int(rows_per_thread)))

    def timed_steps(self, steps) -> float:
        sec = C.c_double()
        check(lib().lbm_strip_timed_steps(self._h, int(steps), C.byref(sec)))
        return sec.value


# ---- batched persistent step -----------------------------------------------------------------

def random_domains(seed: int = 0, n: int = 256, nx_range=(32, 1024), ny_range=(32, 512),
                   tau_range=(0.55, 1.5), force_range=(1e-6, 1e-4), max_obstacles=4,
                   radius_range=(4, 24)):
    """Independent walled channels drawn from `seed`: nx, ny uniform integers in the given
    ranges (inclusive), tau uniform, gx and gy log-uniform, 0..max_obstacles obstacles (squares
    or discs of radius radius_range) at uniform positions. Returns dicts with nx, ny, tau, fx,
    fy, solid (flat uint8)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(force_range[0]), np.log(force_range[1])
    out = []
    for _ in range(n):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        tau = float(rng.uniform(*tau_range))
        fx, fy = (float(np.exp(rng.uniform(lo, hi))) for _ in range(2))
        m = make_channel(nx, ny).reshape(ny, nx)
        for _ in range(int(rng.integers(0, max_obstacles + 1))):
            r = int(rng.integers(radius_range[0], radius_range[1] + 1))
            xc, yc = float(rng.uniform(0, nx)), float(rng.uniform(0, ny))
            if rng.integers(0, 2):
                x0, x1 = max(0, int(xc) - r), min(nx, int(xc) + r)
                y0, y1 = max(0, int(yc) - r), min(ny, int(yc) + r)
                m[y0:y1, x0:x1] = 1
            else:
                flat = m.reshape(-1)
                add_disc(flat, nx, ny, xc, yc, r)
        out.append(dict(nx=nx, ny=ny, tau=tau, fx=fx, fy=fy, solid=m.reshape(-1).copy()))
    return out


class Batch:
    """Many independent domains stepped by one persistent kernel per step, in captured HIP
    graphs of `graph_steps` steps. Every domain is bit-identical to a single-domain Solver."""

    def __init__(self, domains, precision="float", init=(1.0, 0.0, 0.0)):
        self.precision = precision_code(precision)
        self.dtype = real_dtype(self.precision)
        self.domains = list(domains)
        self._masks = [np.ascontiguousarray(d["solid"], dtype=np.uint8).reshape(-1)
                       for d in self.domains]
        descs = (DomainDesc * len(self.domains))()
        for k, (d, m) in enumerate(zip(self.domains, self._masks)):
            if m.size != d["nx"] * d["ny"]:
                raise ValueError(f"domain {k}: solid mask size does not match the grid")
            descs[k] = DomainDesc(d["nx"], d["ny"], d["tau"], d["fx"], d["fy"],
                                  m.ctypes.data_as(C.POINTER(C.c_uint8)), *init)
        h = C.c_void_p()
        check(lib().lbm_batch_create(len(self.domains), descs, self.precision, C.byref(h)))
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.lbm_batch_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def set_graph_steps(self, k):
        check(lib().lbm_batch_set_graph_steps(self._h, int(k)))

    def set_shape(self, tile_x, tile_rows, workgroups):
        check(lib().lbm_batch_set_shape(self._h, int(tile_x), int(tile_rows), int(workgroups)))

    def set_order(self, order: int, streaming_stores: bool = True):
        """0 tiles w, w + G, ...; 1 contiguous chunks banded per XCD; 2 wave work-list."""
        check(lib().lbm_batch_set_order(self._h, int(order), int(bool(streaming_stores))))

    def order(self) -> tuple[int, bool]:
        o, nt = C.c_int(), C.c_int()
        check(lib().lbm_batch_get_order(self._h, C.byref(o), C.byref(nt)))
        return o.value, bool(nt.value)

    def run(self, steps):
        check(lib().lbm_batch_run(self._h, int(steps)))

    def timed_run(self, steps) -> float:
        s = C.c_double()
        check(lib().lbm_batch_timed_run(self._h, int(steps), C.byref(s)))
        return s.value

    def populations(self, k) -> np.ndarray:
        d = self.domains[k]
        out = np.empty(9 * d["nx"] * d["ny"], dtype=self.dtype)
        check(lib().lbm_batch_populations(self._h, int(k), out.ctypes.data))
        return out

    def total_cells(self) -> int:
        n = C.c_longlong()
        check(lib().lbm_batch_total_cells(self._h, C.byref(n)))
        return n.value

    def set_steps_per_launch(self, steps_per_launch: int, tile_x: int = 0, tile_rows: int = 0,
                             workgroups: int = -1):
        """1: one step per launch; 2 (default): pairs of steps in one two-step launch (same
        bits); tile_x: sub-tile width (0 keeps); tile_rows: band height (0 keeps, -1 the
        modelled height); workgroups its persistent grid (0: one workgroup per work item, -1
        keeps)."""
        check(lib().lbm_batch_set_steps_per_launch(self._h, int(steps_per_launch), int(tile_x),
                                                   int(tile_rows), int(workgroups)))

    def steps_per_launch(self) -> dict:
        a, b, c, w = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().lbm_batch_get_steps_per_launch(self._h, C.byref(a), C.byref(b), C.byref(c),
                                                   C.byref(w)))
        return {"steps_per_launch": a.value, "tile_x": b.value, "tile_rows": c.value,
                "workgroups": w.value}

    def set_two_step_threads(self, threads: int):
        """Threads per workgroup of the two-step kernel (multiple of 64, at most 512; double
        256); caps the sub-tile width to it."""
        check(lib().lbm_batch_set_two_step_threads(self._h, int(threads)))

    def two_step_layout(self) -> dict:
        """The two-step work list: threads per workgroup, work items, band rows in use, whether
        they are the modelled choice, and the fill (cell rows over lane rows of the items)."""
        th, it, rows, mod = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        fill = C.c_double()
        check(lib().lbm_batch_two_step_layout(self._h, C.byref(th), C.byref(it), C.byref(rows),
                                              C.byref(mod), C.byref(fill)))
        return {"threads": th.value, "items": it.value, "band_rows": rows.value,
                "modelled": bool(mod.value), "fill": fill.value}

    def autotune(self, steps=10):
        tx, tr, wg, m = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(lib().lbm_batch_autotune(self._h, int(steps), C.byref(tx), C.byref(tr), C.byref(wg),
                                       C.byref(m)))
        return tx.value, tr.value, wg.value, m.value


# ---- discrete adjoint ------------------------------------------------------------------------

def checkpoint_forward_steps(states: int, slots: int) -> int:
    """Forward steps of the binomial checkpointing schedule, first pass included (host
    computation only)."""
    n = C.c_longlong()
    check(lib().lbm_checkpoint_forward_steps(int(states), int(slots), C.byref(n)))
    return n.value


SCHEDULES = {"auto": 0, "binomial": 1, "pipelined": 2}
_SCHEDULE_NAMES = {v: k for k, v in SCHEDULES.items()}


def checkpoint_schedule(states: int, slots: int, kind: str = "auto") -> dict:
    """The checkpointing schedule for `states` = T + 1 states within `slots` stored states
    (built and replay-checked on the host): kind, stored states used, levels, forward steps and
    launches of the gradient sweep."""
    used, levels = C.c_int(), C.c_int()
    fsteps, launches = C.c_longlong(), C.c_longlong()
    check(lib().lbm_checkpoint_schedule(int(states), int(slots), SCHEDULES[kind], C.byref(used),
                                        C.byref(levels), C.byref(fsteps), C.byref(launches)))
    return {"kind": "pipelined" if levels.value else "binomial", "slots": used.value,
            "levels": levels.value, "forward_steps": fsteps.value, "launches": launches.value}


class Adjoint:
    """Gradient of J = mean u_x over a probe window, averaged over the last `avg_steps` states of
    a `steps`-step run from rest, w.r.t. a per-cell force field and tau (discrete adjoint with
    checkpointing in at most `snapshots` stored states; `schedule`: "auto", "binomial" or
    "pipelined")."""

    def __init__(self, nx, ny, tau, solid, steps, probe, avg_steps, snapshots,
                 precision="float", schedule="auto"):
        self.precision = precision_code(precision)
        self.dtype = real_dtype(self.precision)
        self.nx, self.ny = int(nx), int(ny)
        self._solid = np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
        if self._solid.size != self.nx * self.ny:
            raise ValueError("solid mask size does not match the grid")
        x0, y0, w, h = (int(v) for v in probe)
        hd = C.c_void_p()
        check(lib().lbm_adjoint_create(self.nx, self.ny, float(tau), self._solid, int(steps), x0,
                                       y0, w, h, int(avg_steps), int(snapshots), self.precision,
                                       C.byref(hd)))
        self._h = hd
        if schedule != "auto":
            self.set_schedule(schedule)

    def set_schedule(self, kind: str):
        check(lib().lbm_adjoint_set_schedule(self._h, SCHEDULES[kind]))

    def schedule_info(self) -> dict:
        kind, slots, levels = C.c_int(), C.c_int(), C.c_int()
        fsteps, launches = C.c_longlong(), C.c_longlong()
        check(lib().lbm_adjoint_schedule_info(self._h, C.byref(kind), C.byref(slots),
                                              C.byref(levels), C.byref(fsteps),
                                              C.byref(launches)))
        return {"kind": _SCHEDULE_NAMES[kind.value], "slots": slots.value, "levels": levels.value,
                "forward_steps": fsteps.value, "launches": launches.value}

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.lbm_adjoint_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def set_force_field(self, fxy):
        a = np.ascontiguousarray(fxy, dtype=self.dtype).reshape(-1)
        if a.size != 2 * self.nx * self.ny:
            raise ValueError("force field must hold 2 * nx * ny values")
        check(lib().lbm_adjoint_set_force_field(self._h, a.ctypes.data))

    def set_tau(self, tau):
        check(lib().lbm_adjoint_set_tau(self._h, float(tau)))

    def forward(self):
        J, s = C.c_double(), C.c_double()
        check(lib().lbm_adjoint_forward(self._h, C.byref(J), C.byref(s)))
        return J.value, s.value

    def gradient(self):
        """Returns (J, dJ/dF (2, ny, nx), dJ/dtau, seconds, forward steps recomputed)."""
        J, dt, s = C.c_double(), C.c_double(), C.c_double()
        fs = C.c_longlong()
        dF = np.empty(2 * self.nx * self.ny, dtype=self.dtype)
        check(lib().lbm_adjoint_gradient(self._h, C.byref(J), dF.ctypes.data, C.byref(dt),
                                         C.byref(s), C.byref(fs)))
        return J.value, dF.reshape(2, self.ny, self.nx), dt.value, s.value, fs.value


# ---- analysis (scripts/strouhal.py:57-81 method) ---------------------------------------------

def equilibrium_field(ux: np.ndarray, uy: np.ndarray, rho: np.ndarray | float = 1.0,
                      dtype=np.float64) -> np.ndarray:
    """Shifted equilibrium g_eq = w (drho + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u^2)) of a velocity
    field (src/common/d2q9.hpp:40-42), as a post-streaming state in the v0 AoS layout."""
    cx = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.float64)
    cy = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.float64)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    u = np.asarray(ux, np.float64).reshape(-1, 1)
    v = np.asarray(uy, np.float64).reshape(-1, 1)
    r = np.broadcast_to(np.asarray(rho, np.float64).reshape(-1, 1), u.shape)
    cu = u * cx + v * cy
    g = w * ((r - 1.0) + r * (3 * cu + 4.5 * cu * cu - 1.5 * (u * u + v * v)))
    return g.astype(dtype).reshape(-1)


def _cylinder_setup(nx, ny, D, xc, yc, Re, U):
    """tau, the mask, and the Poiseuille initial velocity of mean U of the cylinder instance."""
    nu = U * D / Re
    tau = 0.5 + 3.0 * nu
    H = ny - 2
    umax = 1.5 * U
    solid = np.maximum(make_channel(nx, ny), make_cylinder(nx, ny, xc, yc, D))
    yy = np.arange(ny, dtype=np.float64) - 0.5  # distance from the lower halfway wall
    prof = np.where((yy > 0) & (yy < H), 4 * umax * yy * (H - yy) / (H * H), 0.0)
    ux = np.repeat(prof[:, None], nx, axis=1).reshape(-1)
    return nu, tau, solid, ux


def _cylinder_solver(precision, nx, ny, tau, force, solid, ux):
    S = equilibrium_field(ux, np.zeros_like(ux), 1.0, real_dtype(precision))
    S.reshape(-1, 9)[solid.astype(bool)] = 0
    s = Solver(nx, ny, tau, force, 0.0, solid, precision)
    s.set_populations(S)
    return s


def cylinder_force(nx=1024, ny=256, D=32.0, xc=256.0, yc=129.0, Re=100.0, U=0.05,
                   steps=40000, sample_every=20, window=0.5, tol=2e-3, max_runs=8):
    """The body force that holds the cylinder instance at Reynolds number Re: the mean fluid
    velocity over the analysis window (the last `window` fraction of `steps`, sampled every
    `sample_every` steps from the FP64 momentum reduction) equals U, so Re_w = U_w D / nu = Re
    with nu = U D / Re. Found on the host by a secant iteration on FP64 runs of the whole
    instance from its Poiseuille start, to a relative velocity error below `tol`. The force
    that balances the walls alone (8 nu u_max / H^2 for an empty channel) leaves the flow
    decelerating under the cylinder's drag (window mean 0.79 U at Re 100); the iteration
    starts from four times it and a measured window-velocity response of 2.15e4 per unit force.
    Returns (force, window velocity, runs as (force, U_w) pairs)."""
    nu, tau, solid, ux = _cylinder_setup(nx, ny, D, xc, yc, Re, U)
    fluid = float(nx * ny - int(solid.sum()))
    k0 = int(steps // sample_every * (1 - window))

    def window_velocity(force):
        s = _cylinder_solver("double", nx, ny, tau, force, solid, ux)
        s.run(k0 * sample_every)
        u = []
        for _ in range(steps // sample_every - k0):
            s.run(sample_every)
            u.append(s.momentum()[0] / fluid)
        s.close()
        return float(np.mean(u))

    H = ny - 2
    f = 4.0 * 8.0 * nu * 1.5 * U / (H * H)
    slope = 2.15e4
    runs = []
    for _ in range(max_runs):
        uw = window_velocity(f)
        runs.append((f, uw))
        if abs(uw / U - 1.0) < tol:
            return f, uw, runs
        if len(runs) >= 2 and runs[-1][0] != runs[-2][0]:
            (f1, u1), (f2, u2) = runs[-2], runs[-1]
            if u2 != u1:
                slope = (u2 - u1) / (f2 - f1)
        f = f + (U - uw) / slope
    raise RuntimeError(f"cylinder force calibration did not converge: {runs}")


def cylinder_instance(precision="float", nx=1024, ny=256, D=32.0, xc=256.0, yc=129.0,
                      Re=100.0, U=0.05, steps=40000, sample_every=20, window=0.5, force=None):
    """Force-driven walled channel (periodic in x) with a staircase cylinder at Reynolds number
    Re: tau from nu = U D / Re, the flow started from the Poiseuille profile of mean velocity U,
    and the body force that keeps the mean fluid velocity over the analysis window at U
    (cylinder_force(), computed once in FP64 unless `force` is given), so that
    Re_w = U_w D / nu = Re. Every `sample_every` steps three signals are recorded: the lift on
    the cylinder (momentum exchange, FP64), the transverse velocity u_y at the wake probe
    (xc + 3 D, yc), the signal the reference's Strouhal script analyses
    (src/common/driver.hpp:175-214, scripts/strouhal.py:57-78), and the mean fluid velocity
    u_x (FP64 momentum / fluid cells). Over the analysis window (the last `window` fraction,
    the reference script's default), the Strouhal number St = f D / U_w comes from the lift
    (FFT peak and upward zero crossings), with the probe's as a cross-check; the velocity scale
    U_w is the window mean of the fluid velocity.
    Returns a dict with st_lift_fft, st_lift_zero, lift_periods, st_fft / st_zero / periods of
    the probe, U_window, Re_window, tau, force, the calibration runs and the signals."""
    nu, tau, solid, ux = _cylinder_setup(nx, ny, D, xc, yc, Re, U)
    runs = None
    if force is None:
        force, _, runs = cylinder_force(nx, ny, D, xc, yc, Re, U, steps, sample_every, window)
    s = _cylinder_solver(precision, nx, ny, tau, force, solid, ux)
    r = int(np.ceil(D / 2)) + 2
    box = (int(xc) - r, int(yc) - r, 2 * r + 1, 2 * r + 1)
    fluid = float(nx * ny - int(solid.sum()))
    px, py = int(xc + 3.0 * D), int(yc)
    cy = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.float64)
    lift, probe, umean = [], [], []
    for _ in range(steps // sample_every):
        s.run(sample_every)
        f = s.populations_at(px, py).astype(np.float64)
        probe.append(float(cy @ f) / (1.0 + float(f.sum())))
        lift.append(s.momentum_exchange(*box)[1])
        umean.append(s.momentum()[0] / fluid)
    s.close()
    lift, probe, umean = np.array(lift), np.array(probe), np.array(umean)
    k0 = int(len(lift) * (1 - window))
    Uw = float(umean[k0:].mean())
    st_fft, st_zero, periods = strouhal(probe[k0:], float(sample_every), D, Uw)
    sl_fft, sl_zero, sl_periods = strouhal(lift[k0:], float(sample_every), D, Uw)
    return dict(st_lift_fft=sl_fft, st_lift_zero=sl_zero, lift_periods=sl_periods,
                st_fft=st_fft, st_zero=st_zero, periods=periods, U_window=Uw,
                Re_window=Uw * D / nu, tau=tau, force=force, calibration=runs, probe=probe,
                lift=lift, umean=umean)


def strouhal(uy: np.ndarray, dt: float, D: float, U: float) -> tuple[float, float, int]:
    """Shedding Strouhal number of a transverse-velocity probe signal sampled every dt steps.

    FFT peak of the mean-removed signal with a Hann window, 16x zero padding and parabolic
    refinement on the log spectrum; and the mean period between upward zero crossings (linear
    interpolation). Returns (St_fft, St_zero_crossing, periods)."""
    s = np.asarray(uy, dtype=np.float64)
    s = s - s.mean()
    n = s.size
    spec = np.abs(np.fft.rfft(s * np.hanning(n), 16 * n))
    freqs = np.fft.rfftfreq(16 * n, d=dt)
    k = int(np.argmax(spec[1:]) + 1)
    k = min(max(k, 1), spec.size - 2)
    a, b, c = np.log(spec[k - 1:k + 2] + 1e-300)
    den = a - 2 * b + c
    shift = 0.5 * (a - c) / den if den != 0 else 0.0
    f_fft = (k + shift) * (freqs[1] - freqs[0])
    t = np.arange(n) * dt
    up = np.where((s[:-1] < 0) & (s[1:] >= 0))[0]
    tc = t[up] - s[up] * (t[up + 1] - t[up]) / (s[up + 1] - s[up])
    periods = np.diff(tc)
    f_zero = 1.0 / periods.mean() if periods.size else float("nan")
    return f_fft * D / U, f_zero * D / U, int(periods.size)
