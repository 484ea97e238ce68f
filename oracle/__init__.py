"""ctypes wrapper of the CPU oracle (oracle/spp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline, never by the product (spp_amd).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libspp_oracle.so")

_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)


def build() -> str:
    """Compile the oracle with its Makefile (gcc -O3)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


NATIVE_FLAGS = ["-O3", "-march=native", "-ffp-contract=off", "-std=c99", "-fPIC"]
NATIVE_LIB_PATH = os.path.join(HERE, "build", "libspp_oracle_native.so")


def build_native(cc: str = "gcc") -> str:
    """Compile the oracle for the host it runs on (-O3 -march=native, contraction off so the
    arithmetic stays that of the portable build): the timed CPU baseline, built where it is timed."""
    os.makedirs(os.path.dirname(NATIVE_LIB_PATH), exist_ok=True)
    tmp = NATIVE_LIB_PATH + ".tmp"
    subprocess.run([cc, *NATIVE_FLAGS, "-shared", "-o", tmp, os.path.join(HERE, "spp_oracle.c"), "-lm"], check=True)
    os.replace(tmp, NATIVE_LIB_PATH)
    return NATIVE_LIB_PATH


def compiler_version(cc: str = "gcc") -> str:
    try:
        return subprocess.run([cc, "--version"], capture_output=True, text=True, check=True).stdout.split("\n")[0]
    except (OSError, subprocess.CalledProcessError):
        return "unknown"


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        _lib = load_library(LIB_PATH)
    return _lib


def load_library(path: str) -> C.CDLL:
    """Load an oracle build and declare its C signatures."""
    L = C.CDLL(path)
    L.orc_error.restype = C.c_char_p
    L.orc_periodic.restype = C.c_double
    L.orc_periodic.argtypes = [C.c_double, C.c_double]
    L.orc_cell_quirk.argtypes = [C.c_double, C.c_double, C.c_int]
    L.orc_cell_exact.argtypes = [C.c_double, C.c_double, C.c_int]
    L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.orc_noise_normal.restype = C.c_double
    L.orc_noise_normal.argtypes = [C.c_uint64, C.c_int32, C.c_int64]
    L.orc_force_law.argtypes = [C.c_int] + [C.c_double] * 6 + [_D, _D]
    L.orc_brute_pairs.restype = C.c_int64
    L.orc_brute_pairs.argtypes = [_D, _D, C.c_int32, C.c_double, C.c_double, C.c_double, _I32, _I32, C.c_int64]
    L.orc_random_state.argtypes = [C.c_uint64, C.c_int32, C.c_double, C.c_double, _D, _D, _D]
    L.orc_create.restype = C.c_void_p
    L.orc_create.argtypes = [C.c_void_p]
    L.orc_destroy.argtypes = [C.c_void_p]
    for nm in ("orc_set_state", "orc_get_state"):
        getattr(L, nm).argtypes = [C.c_void_p, _D, _D, _D, _D, _D, C.c_int32]
    L.orc_update_cells.argtypes = [C.c_void_p]
    L.orc_get_cells.argtypes = [C.c_void_p, _I32, _I32, C.c_int32]
    L.orc_half_list.restype = C.c_int64
    L.orc_half_list.argtypes = [C.c_void_p, _I32, _I32, C.c_int64]
    L.orc_interact.argtypes = [C.c_void_p]
    L.orc_ring_energy.restype = C.c_double
    L.orc_ring_energy.argtypes = [C.c_void_p]
    L.orc_get_forces.argtypes = [C.c_void_p, _D, _D, _D, C.c_int32]
    L.orc_set_noise.argtypes = [C.c_void_p, _D, C.c_int32]
    L.orc_move_stage.argtypes = [C.c_void_p, C.c_int32]
    L.orc_one_step.argtypes = [C.c_void_p]
    L.orc_multi_step.argtypes = [C.c_void_p, C.c_int32]
    L.orc_polarization.restype = C.c_double
    L.orc_polarization.argtypes = [C.c_void_p, _D, _D]
    L.orc_step_count.restype = C.c_int64
    L.orc_step_count.argtypes = [C.c_void_p]
    L.orc_rebuild_count.restype = C.c_int64
    L.orc_rebuild_count.argtypes = [C.c_void_p]
    L.orc_chain_mobilities.argtypes = [C.c_int32, C.c_double, C.c_double, _D, _D, _D]
    return L


def _p(a, t=_D):
    return None if a is None else a.ctypes.data_as(t)


def periodic(x: float, L2: float) -> float:
    return lib().orc_periodic(x, L2)


def cell_quirk(x: float, L: float, divisor: int) -> int:
    return lib().orc_cell_quirk(x, L, divisor)


def cell_exact(x: float, L: float, divisor: int) -> int:
    return lib().orc_cell_exact(x, L, divisor)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def noise_normal(seed: int, pid: int, step: int) -> float:
    return lib().orc_noise_normal(seed, pid, step)


def force_law(kind: int, dx: float, dy: float, d: float, rc: float, sigma: float, A: float):
    fx, fy = C.c_double(), C.c_double()
    lib().orc_force_law(kind, dx, dy, d, rc, sigma, A, C.byref(fx), C.byref(fy))
    return fx.value, fy.value


def brute_pairs(x, y, Lx, Ly, r):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = x.size
    m = lib().orc_brute_pairs(_p(x), _p(y), n, Lx, Ly, r, None, None, 0)
    pi = np.empty(max(m, 1), dtype=np.int32)
    pj = np.empty(max(m, 1), dtype=np.int32)
    lib().orc_brute_pairs(_p(x), _p(y), n, Lx, Ly, r, _p(pi, _I32), _p(pj, _I32), m)
    return pi[:m], pj[:m]


def chain_mobilities(nb: int, sep: float, sigma: float):
    """(m_parallel, m_perpendicular, k_perpendicular) of an nb-bead chain (Set_nb)."""
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    lib().orc_chain_mobilities(nb, sep, sigma, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def random_state(seed: int, n: int, Lx: float, Ly: float):
    x, y, th = np.empty(n), np.empty(n), np.empty(n)
    lib().orc_random_state(seed, n, Lx, Ly, _p(x), _p(y), _p(th))
    return x, y, th


class OracleError(RuntimeError):
    pass


class Oracle:
    """The reference's serial Box, restated in C, driven by an spp_params struct."""

    def __init__(self, params, library=None):
        import ctypes as _C

        self._lib = library if library is not None else lib()
        self._params = params
        self._h = self._lib.orc_create(_C.byref(params))
        if not self._h:
            raise OracleError(self._lib.orc_error().decode())
        self.n = params.n

    def _check(self, rc):
        if rc:
            raise OracleError(f"oracle error {rc}: {self._lib.orc_error().decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, xw, yw, theta, xu=None, yu=None):
        def f(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.shape != (self.n,):
                raise ValueError(f"array of shape {a.shape} for {self.n} particles")
            return a

        xw, yw, theta, xu, yu = f(xw), f(yw), f(theta), f(xu), f(yu)
        self._check(self._lib.orc_set_state(self._h, _p(xw), _p(yw), _p(xu), _p(yu), _p(theta), self.n))

    def get_state(self):
        a = [np.empty(self.n) for _ in range(5)]
        self._check(self._lib.orc_get_state(self._h, *[_p(x) for x in a], self.n))
        return dict(zip(("xw", "yw", "xu", "yu", "theta"), a))

    def update_cells(self):
        self._check(self._lib.orc_update_cells(self._h))

    def get_cells(self):
        cx = np.empty(self.n, dtype=np.int32)
        cy = np.empty(self.n, dtype=np.int32)
        self._check(self._lib.orc_get_cells(self._h, _p(cx, _I32), _p(cy, _I32), self.n))
        return cx, cy

    def half_list(self):
        m = self._lib.orc_half_list(self._h, None, None, 0)
        a = np.empty(max(m, 1), dtype=np.int32)
        b = np.empty(max(m, 1), dtype=np.int32)
        self._lib.orc_half_list(self._h, _p(a, _I32), _p(b, _I32), m)
        return a[:m], b[:m]

    def interact(self):
        self._check(self._lib.orc_interact(self._h))

    def get_forces(self):
        a = [np.empty(self.n) for _ in range(3)]
        self._check(self._lib.orc_get_forces(self._h, *[_p(x) for x in a], self.n))
        return a

    def set_noise(self, dtheta):
        d = np.ascontiguousarray(dtheta, dtype=np.float64)
        if d.shape != (self.n,):
            raise ValueError(f"array of shape {d.shape} for {self.n} particles")
        self._check(self._lib.orc_set_noise(self._h, _p(d), self.n))

    def move_stage(self, stage):
        self._check(self._lib.orc_move_stage(self._h, stage))

    def one_step(self):
        self._check(self._lib.orc_one_step(self._h))

    def multi_step(self, steps):
        self._check(self._lib.orc_multi_step(self._h, steps))

    def polarization(self):
        sx, sy = C.c_double(), C.c_double()
        p = self._lib.orc_polarization(self._h, C.byref(sx), C.byref(sy))
        return p, sx.value, sy.value

    def ring_energy(self):
        return self._lib.orc_ring_energy(self._h)

    def step_count(self):
        return self._lib.orc_step_count(self._h)

    def rebuild_count(self):
        return self._lib.orc_rebuild_count(self._h)
