"""ORACLE — test infrastructure only (never part of the product path).

ctypes bindings for the CPU restatement of the reference solver (``lbm_oracle.cpp``, built into
``oracle/build/liblbm_oracle.so``) and for the reference's own CPU solvers compiled from their
sources into ``oracle/_ref/`` (``ref_shim.cpp``). Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` import this package, and only as a checker or a timed
CPU baseline (``cpu_protocol.py``: the reference's timing protocol in a child process).

Function by function, what pins it (besides the reference's own outputs in tests/golden/):
  * ``equilibrium`` / ``guo`` — closed-form moment identities (tests/test_oracle.py);
  * ``collide`` — rest fixed point, tau = 1 identity, momentum gain = F per cell;
  * ``PushAoS`` step — golden fixtures of reference v0 (float/double, walls + block,
    cylinder with open x), Poiseuille exactness at tau = 1/2 + sqrt(3)/4, mass conservation;
  * ``PullSoA`` step — bitwise equal to ``PushAoS`` for 1, 2, 3, 4, 8 threads;
  * ``pull_to_aos`` / ``aos_to_pull`` — inverse of each other on fluid slots (round trip);
  * ``vorticity`` — bitwise equal to the reference's io.cpp vorticity (oracle/_ref) on random
    fields, periodic and open; exact on linear shear fields (interior and one-sided columns).
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "liblbm_oracle.so"
REF_DIR = HERE / "_ref"

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_lib = None


def build(ref_root: str | None = "/root/reference") -> None:
    """Compile the oracle (and the reference shim when the reference checkout exists)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "-B", "all"], check=True)
    if ref_root and os.path.isdir(ref_root):
        subprocess.run(["make", "-s", "-C", str(HERE), "-B", "ref", f"REF={ref_root}"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(str(LIB_PATH))
        vp, d, i = C.c_void_p, C.c_double, C.c_int
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [i, i, i, i, d, d, d, _u8p]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_set_inflow_outflow.argtypes = [vp, d]
        L.oracle_set_force_field.argtypes = [vp, vp]
        L.oracle_init_equilibrium.argtypes = [vp, d, d, d]
        L.oracle_set_populations.argtypes = [vp, vp]
        L.oracle_run.argtypes = [vp, i]
        L.oracle_timed_run.argtypes = [vp, i]
        L.oracle_timed_run.restype = d
        L.oracle_populations.argtypes = [vp, vp]
        L.oracle_raw_state.argtypes = [vp, vp]
        L.oracle_raw_state.restype = i
        L.oracle_macroscopic.argtypes = [vp, vp, vp, vp]
        L.oracle_total_mass.argtypes = [vp]
        L.oracle_total_mass.restype = d
        L.oracle_momentum.argtypes = [vp, C.POINTER(d), C.POINTER(d)]
        L.oracle_threads.argtypes = [vp]
        L.oracle_threads.restype = i
        L.oracle_set_threads.argtypes = [i]
        L.oracle_equilibrium_f64.argtypes = [d] * 5
        L.oracle_equilibrium_f64.restype = d
        L.oracle_equilibrium_f32.argtypes = [C.c_float] * 5
        L.oracle_equilibrium_f32.restype = C.c_float
        L.oracle_guo_f64.argtypes = [d] * 7
        L.oracle_guo_f64.restype = d
        L.oracle_collide_f64.argtypes = [vp, d, d, d, d]
        L.oracle_collide_f32.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float]
        L.oracle_weight_f64.argtypes = [i]
        L.oracle_weight_f64.restype = d
        L.oracle_weight_f32.argtypes = [i]
        L.oracle_weight_f32.restype = C.c_float
        L.oracle_pull_to_aos.argtypes = [i, i, i, _u8p, i, d, vp, vp]
        L.oracle_aos_to_pull.argtypes = [i, i, i, _u8p, i, vp, vp]
        L.oracle_make_channel.argtypes = [i, i, _u8p]
        L.oracle_make_cylinder.argtypes = [i, i, d, d, d, _u8p]
        _lib = L
    return _lib


def _dtype(precision: int):
    return np.float32 if precision == 0 else np.float64


class Oracle:
    """Mirror of the reference's solver interface on the CPU restatement.

    kind "v0" = sequential push AoS (bitwise reference); kind "pull" = OpenMP pull SoA.
    """

    def __init__(self, nx, ny, tau, fx, fy, solid, precision="float", kind="v0"):
        self.precision = 0 if precision in ("float", "f32", 0) else 1
        self.dtype = _dtype(self.precision)
        self.nx, self.ny = int(nx), int(ny)
        self.solid = np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
        if self.solid.size != self.nx * self.ny:
            raise ValueError("solid mask size does not match the grid")
        if float(self.dtype(tau)) <= 0.5:
            raise ValueError("tau must be > 0.5")
        self.fx, self.fy = float(fx), float(fy)
        self._force = None
        k = 0 if kind == "v0" else 1
        self.h = lib().oracle_create(k, self.precision, self.nx, self.ny, float(tau), float(fx),
                                     float(fy), self.solid)
        if not self.h:
            raise ValueError("oracle_create failed")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().oracle_destroy(h)
            self.h = None

    @property
    def N(self):
        return self.nx * self.ny

    def set_inflow_outflow(self, u_in):
        lib().oracle_set_inflow_outflow(self.h, float(u_in))

    def set_force_field(self, fxy):
        a = np.ascontiguousarray(fxy, dtype=self.dtype).reshape(-1)
        assert a.size == 2 * self.N
        self._force = a
        lib().oracle_set_force_field(self.h, a.ctypes.data)

    def init_equilibrium(self, rho0, ux0, uy0):
        lib().oracle_init_equilibrium(self.h, float(rho0), float(ux0), float(uy0))

    def set_populations(self, S):
        a = np.ascontiguousarray(S, dtype=self.dtype).reshape(-1)
        assert a.size == 9 * self.N
        lib().oracle_set_populations(self.h, a.ctypes.data)

    def run(self, steps):
        lib().oracle_run(self.h, int(steps))

    def timed_run(self, steps):
        return lib().oracle_timed_run(self.h, int(steps))

    def populations(self):
        out = np.empty(9 * self.N, dtype=self.dtype)
        lib().oracle_populations(self.h, out.ctypes.data)
        return out

    def raw_state(self):
        out = np.empty(9 * self.N, dtype=self.dtype)
        if lib().oracle_raw_state(self.h, out.ctypes.data) != 0:
            raise ValueError("raw_state is only defined for the pull kind")
        return out

    def macroscopic(self):
        r = np.empty(self.N, dtype=self.dtype)
        u = np.empty(self.N, dtype=self.dtype)
        v = np.empty(self.N, dtype=self.dtype)
        lib().oracle_macroscopic(self.h, r.ctypes.data, u.ctypes.data, v.ctypes.data)
        return r, u, v

    def total_mass(self):
        return lib().oracle_total_mass(self.h)

    def momentum(self):
        px, py = C.c_double(), C.c_double()
        lib().oracle_momentum(self.h, C.byref(px), C.byref(py))
        return px.value, py.value

    def threads(self):
        return lib().oracle_threads(self.h)


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def make_channel(nx, ny):
    m = np.empty(nx * ny, dtype=np.uint8)
    lib().oracle_make_channel(nx, ny, m)
    return m


def make_cylinder(nx, ny, xc, yc, D):
    m = np.empty(nx * ny, dtype=np.uint8)
    lib().oracle_make_cylinder(nx, ny, float(xc), float(yc), float(D), m)
    return m


def pull_to_aos(R, nx, ny, solid, open_x=False, u_in=0.0):
    R = np.ascontiguousarray(R)
    p = 0 if R.dtype == np.float32 else 1
    S = np.empty_like(R)
    lib().oracle_pull_to_aos(p, nx, ny, np.ascontiguousarray(solid, np.uint8).reshape(-1),
                             int(open_x), float(u_in), R.ctypes.data, S.ctypes.data)
    return S


def aos_to_pull(S, nx, ny, solid, open_x=False):
    S = np.ascontiguousarray(S)
    p = 0 if S.dtype == np.float32 else 1
    R = np.empty_like(S)
    lib().oracle_aos_to_pull(p, nx, ny, np.ascontiguousarray(solid, np.uint8).reshape(-1),
                             int(open_x), S.ctypes.data, R.ctypes.data)
    return R


def collide(g9, fx, fy, omega, fp):
    g = np.ascontiguousarray(g9).copy()
    if g.dtype == np.float32:
        lib().oracle_collide_f32(g.ctypes.data, float(fx), float(fy), float(omega), float(fp))
    else:
        lib().oracle_collide_f64(g.ctypes.data, float(fx), float(fy), float(omega), float(fp))
    return g


def weight(i, precision="double"):
    return lib().oracle_weight_f32(i) if precision == "float" else lib().oracle_weight_f64(i)


def equilibrium(wi, drho, rho, cu, usq):
    return lib().oracle_equilibrium_f64(wi, drho, rho, cu, usq)


def guo(wi, cxi, cyi, ux, uy, fx, fy):
    return lib().oracle_guo_f64(wi, cxi, cyi, ux, uy, fx, fy)


def vorticity(ux, uy, solid, nx, ny, periodic_x=True):
    """dUy/dx - dUx/dy as src/common/io.cpp:54-74 computes it, in the dtype of ux: central
    differences over (x +- 1) / 2 and (y +- 1) / 2 with periodic wrap; with periodic_x False the
    columns x = 0 and x = nx - 1 take one-sided differences over h = 1; 0 in solid cells."""
    t = np.asarray(ux).dtype.type
    u = np.asarray(ux).reshape(ny, nx)
    v = np.asarray(uy).reshape(ny, nx)
    xp = np.roll(np.arange(nx), -1)
    xm = np.roll(np.arange(nx), 1)
    h = np.full(nx, 2, dtype=t)
    if not periodic_x:
        xm[0], h[0] = 0, 1
        xp[nx - 1], h[nx - 1] = nx - 1, 1
    duy_dx = (v[:, xp] - v[:, xm]) / h[None, :]
    dux_dy = (np.roll(u, -1, axis=0) - np.roll(u, 1, axis=0)) / t(2)
    w = (duy_dx - dux_dy).astype(t)
    w[np.asarray(solid).reshape(ny, nx) != 0] = t(0)
    return w.reshape(-1)


# ---- the reference's own CPU solvers (oracle/_ref), when they were built here ----------------

class Reference:
    """ctypes handle on oracle/_ref/libref_{float,double}.so (reference v0/v1 via ref_shim.cpp)."""

    def __init__(self, precision="float"):
        name = "libref_float.so" if precision in ("float", 0) else "libref_double.so"
        path = REF_DIR / name
        if not path.exists():
            raise FileNotFoundError(path)
        L = C.CDLL(str(path))
        d, i, vp = C.c_double, C.c_int, C.c_void_p
        for f in (L.ref_v0_run, L.ref_v1_run):
            f.argtypes = [i, i, d, d, d, _u8p, i, d, d, d, d, i, vp]
        L.ref_v0_total_mass.argtypes = [i, i, d, d, d, _u8p, d, d, d, i]
        L.ref_v0_total_mass.restype = d
        L.ref_make_cylinder.argtypes = [i, i, d, d, d, _u8p]
        L.ref_v1_channel_timed.argtypes = [i, i, d, d, i, i, C.POINTER(i)]
        L.ref_v1_channel_timed.restype = d
        L.ref_v1_protocol.argtypes = [i, i, d, d, d, _u8p, i, i, i, C.POINTER(d), C.POINTER(i)]
        L.ref_real_size.restype = i
        L.ref_vorticity.argtypes = [vp, vp, _u8p, i, i, i, vp]
        L.ref_write_npy.argtypes = [C.c_char_p, vp, i, i, i]
        self.L = L
        self.dtype = np.float32 if L.ref_real_size() == 4 else np.float64

    @staticmethod
    def available(precision="float"):
        name = "libref_float.so" if precision in ("float", 0) else "libref_double.so"
        return (REF_DIR / name).exists()

    def run(self, which, nx, ny, tau, fx, fy, solid, steps, init=(1.0, 0.0, 0.0), open_x=False,
            u_in=0.0):
        out = np.empty(9 * nx * ny, dtype=self.dtype)
        f = self.L.ref_v0_run if which == "v0" else self.L.ref_v1_run
        f(nx, ny, float(tau), float(fx), float(fy), np.ascontiguousarray(solid, np.uint8).reshape(-1),
          int(open_x), float(u_in), float(init[0]), float(init[1]), float(init[2]), int(steps),
          out.ctypes.data)
        return out

    def vorticity(self, ux, uy, solid, nx, ny, periodic_x=True):
        u = np.ascontiguousarray(ux, dtype=self.dtype).reshape(-1)
        v = np.ascontiguousarray(uy, dtype=self.dtype).reshape(-1)
        out = np.empty(nx * ny, dtype=self.dtype)
        self.L.ref_vorticity(u.ctypes.data, v.ctypes.data,
                             np.ascontiguousarray(solid, np.uint8).reshape(-1), nx, ny,
                             int(bool(periodic_x)), out.ctypes.data)
        return out

    def write_npy(self, path, a, ny, nx):
        """The reference's writer: the solver real, or uint8 for a uint8 array."""
        a = np.ascontiguousarray(a)
        kind = 1 if a.dtype == np.uint8 else 0
        if kind == 0:
            a = np.ascontiguousarray(a, dtype=self.dtype)
        self.L.ref_write_npy(str(path).encode(), a.ctypes.data, int(ny), int(nx), kind)

    def v1_protocol(self, nx, ny, tau, fx, fy, solid, warmup, steps, reps):
        """Seconds of `reps` timed runs of `steps` steps of the reference's v1 after `warmup`
        untimed steps (one solver, rest start), and the OpenMP threads it used."""
        secs = (C.c_double * reps)()
        t = C.c_int()
        self.L.ref_v1_protocol(nx, ny, float(tau), float(fx), float(fy),
                               np.ascontiguousarray(solid, np.uint8).reshape(-1), int(warmup),
                               int(steps), int(reps), secs, C.byref(t))
        return list(secs), t.value

    def v1_channel_timed(self, nx, ny, tau, fx, warmup, steps):
        t = C.c_int()
        sec = self.L.ref_v1_channel_timed(nx, ny, float(tau), float(fx), int(warmup), int(steps),
                                          C.byref(t))
        return sec, t.value
