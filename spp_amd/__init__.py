"""spp_amd -- Python binding of the MI355X SPP particle engine's C ABI (include/spp_engine.h).

The engine itself is C++/HIP (spp_amd/csrc); this module only loads the in-tree
``libspp_engine.so`` with ctypes and exposes it with numpy arrays, for tests and the benchmark.
There is no fallback: if the library is missing, importing the binding raises.

Method names follow the reference's Box surface (serial/box.h:40-57, parallel/box.h:40-53):
``interact`` = Box::Interact, ``move_stage`` = Move / Move_Runge_Kutta*_k, ``one_step`` =
One_Step, ``multi_step`` = Multi_Step, ``snapshot``/``restore`` = Save/Load.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspp_engine.so")

ABI_VERSION = 1
MODEL_REPULSIVE, MODEL_CONTINUOUS, MODEL_ABP = 0, 1, 2
EULER, RK2, RK4 = 1, 2, 4
REP_NONE, REP_R12_TRUNCATED, REP_HARMONIC, REP_R12, REP_YUKAWA_TRUNCATED, REP_YUKAWA = range(6)
BIN_EXACT, BIN_REF_QUIRK = 0, 1
REBUILD_DISPLACEMENT, REBUILD_PERIOD = 0, 1
NOISE_STREAM, NOISE_INJECTED = 0, 1
REBUILD_SORT_AUTO, REBUILD_SORT_GENERAL = 0, 1
KERNEL_SLOT, KERNEL_PERSIST, KERNEL_SLOT_DEEP = 0, 3, 4  # 1 and 2 are retired
PERSIST_MAX_N = 4096  # particles per system of the persistent kernel (KERNEL_PERSIST)

OK, ERR_ARG, ERR_HIP, ERR_STATE, ERR_CAPACITY, ERR_UNSUPPORTED, ERR_NO_DEVICE = 0, -1, -2, -3, -4, -5, -6


class Params(C.Structure):
    """Mirror of ``spp_params``."""

    _fields_ = [
        ("abi_version", C.c_int32),
        ("model", C.c_int32),
        ("integrator", C.c_int32),
        ("repulsion", C.c_int32),
        ("n", C.c_int32),
        ("n_membrane", C.c_int32),
        ("binning", C.c_int32),
        ("rebuild", C.c_int32),
        ("rebuild_period", C.c_int32),
        ("ncell_x", C.c_int32),
        ("ncell_y", C.c_int32),
        ("noise_mode", C.c_int32),
        ("seed", C.c_uint64),
        ("Lx", C.c_double),
        ("Ly", C.c_double),
        ("dt", C.c_double),
        ("v0", C.c_double),
        ("g", C.c_double),
        ("R_align", C.c_double),
        ("alpha", C.c_double),
        ("A_rep", C.c_double),
        ("sigma_rep", C.c_double),
        ("rc_rep", C.c_double),
        ("D_rot", C.c_double),
        ("r_list", C.c_double),
        ("mob_par", C.c_double),
        ("mob_perp", C.c_double),
        ("mob_rot", C.c_double),
        ("torque0", C.c_double),
        ("k_bond", C.c_double),
        ("l_bond", C.c_double),
        ("kappa_bend", C.c_double),
        ("ell_capacity", C.c_int32),
        ("kernel_variant", C.c_int32),
        ("graph_intervals", C.c_int32),
        ("rebuild_sort", C.c_int32),
        ("nb_swimmer", C.c_int32),
        ("device_timing", C.c_int32),
        ("reserved_i", C.c_int32 * 2),
        ("bead_sep", C.c_double),
        ("sigma_bead", C.c_double),
        ("rc_bead", C.c_double),
        ("A_bead", C.c_double),
        ("reserved_d", C.c_double * 7),
    ]

    def update(self, **kw) -> "Params":
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(f"spp_params has no field {k!r}")
            setattr(self, k, v)
        return self

    def copy(self) -> "Params":
        p = Params()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(Params))
        return p


class Info(C.Structure):
    """Mirror of ``spp_info``."""

    _fields_ = [
        ("steps", C.c_int64),
        ("rebuilds", C.c_int64),
        ("rollbacks", C.c_int64),
        ("host_syncs", C.c_int64),
        ("t", C.c_double),
        ("ncell_x", C.c_int32),
        ("ncell_y", C.c_int32),
        ("ell_capacity", C.c_int32),
        ("max_neighbors", C.c_int32),
        ("mean_neighbors", C.c_double),
        ("tile_rebuilds", C.c_int64),
        ("tile_fallbacks", C.c_int64),
        ("device_ms", C.c_double),
        ("wide_rebuilds", C.c_int64),
    ]


class Clock(C.Structure):
    """Mirror of ``spp_clock``."""

    _fields_ = [
        ("step", C.c_int64),
        ("t", C.c_double),
        ("rebuild_estimate", C.c_double * 3),
    ]


class System(C.Structure):
    """Mirror of ``spp_system`` (one box of a batched engine)."""

    _fields_ = [
        ("n", C.c_int32),
        ("ncell_x", C.c_int32),
        ("ncell_y", C.c_int32),
        ("reserved", C.c_int32),
        ("Lx", C.c_double),
        ("Ly", C.c_double),
        ("v0", C.c_double),
        ("g", C.c_double),
        ("D_rot", C.c_double),
        ("seed", C.c_uint64),
    ]


# Exported symbols and their signatures (every declaration of include/spp_engine.h).
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
SIGNATURES = {
    "spp_params_default": (C.c_int, [C.POINTER(Params), C.c_int32]),
    "spp_create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "spp_create_batch": (C.c_int, [C.POINTER(Params), C.POINTER(System), C.c_int32, C.POINTER(C.c_void_p)]),
    "spp_destroy": (None, [C.c_void_p]),
    "spp_last_error": (C.c_char_p, [C.c_void_p]),
    "spp_get_info": (C.c_int, [C.c_void_p, C.POINTER(Info)]),
    "spp_set_state": (C.c_int, [C.c_void_p, _D, _D, _D, _D, _D, C.c_int32]),
    "spp_get_state": (C.c_int, [C.c_void_p, _D, _D, _D, _D, _D, C.c_int32]),
    "spp_build_neighbors": (C.c_int, [C.c_void_p]),
    "spp_get_neighbors": (C.c_int, [C.c_void_p, _I32, _I32, C.c_int64, _I64]),
    "spp_get_cells": (C.c_int, [C.c_void_p, _I32, _I32, C.c_int32]),
    "spp_interact": (C.c_int, [C.c_void_p]),
    "spp_get_forces": (C.c_int, [C.c_void_p, _D, _D, _D, C.c_int32]),
    "spp_inject_noise": (C.c_int, [C.c_void_p, _D, C.c_int32]),
    "spp_move_stage": (C.c_int, [C.c_void_p, C.c_int32]),
    "spp_one_step": (C.c_int, [C.c_void_p]),
    "spp_multi_step": (C.c_int, [C.c_void_p, C.c_int32]),
    "spp_polarization": (C.c_int, [C.c_void_p, _D, _D, _D]),
    "spp_batch_polarization": (C.c_int, [C.c_void_p, _D, C.c_int32]),
    "spp_multi_step_sampled": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _D, C.c_int64]),
    "spp_pair_distribution": (C.c_int, [C.c_void_p, C.c_double, C.c_int32, _I64, _D]),
    "spp_pair_distribution_noalloc": (C.c_int, [C.c_void_p, C.c_double, C.c_int32, _I64, _D]),
    "spp_lyapunov_init": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_uint64]),
    "spp_lyapunov_renormalize": (C.c_int, [C.c_void_p, _D]),
    "spp_lyapunov_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "spp_lyapunov_get": (C.c_int, [C.c_void_p, _D, _D]),
    "spp_lyapunov_vectors": (C.c_int, [C.c_void_p, C.c_int32, _D, C.c_int64]),
    "spp_snapshot": (C.c_int, [C.c_void_p, C.c_int32]),
    "spp_restore": (C.c_int, [C.c_void_p, C.c_int32]),
    "spp_synchronize": (C.c_int, [C.c_void_p]),
    "spp_get_clock": (C.c_int, [C.c_void_p, C.c_void_p]),
    "spp_set_clock": (C.c_int, [C.c_void_p, C.c_void_p]),
    "spp_reset_forces": (C.c_int, [C.c_void_p]),
    "spp_pair_interact": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "spp_graph_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "spp_profile_enable": (C.c_int, [C.c_void_p, C.c_int32]),
    "spp_profile_read": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, _D, _I64, C.c_int32, _I32]),
    "spp_time_stage": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _D]),
    "spp_random_state": (C.c_int, [C.c_uint64, C.c_int32, C.c_double, C.c_double, _D, _D, _D]),
    "spp_random_formation": (C.c_int, [C.c_uint64, C.c_int32, C.c_double, C.c_double, C.c_double, _D, _D, _D]),
    "spp_polar_formation": (C.c_int, [C.c_uint64, C.c_int32, C.c_double, C.c_double, _D, _D, _D]),
    "spp_noise_normal": (C.c_double, [C.c_uint64, C.c_int32, C.c_int64]),
    "spp_triangle_lattice": (C.c_int, [C.c_int32, C.c_double, C.c_double, C.c_double, _D, _D]),
    "spp_ring_membrane": (C.c_int, [C.c_int32, C.c_double, C.c_double, _D, _D]),
    "spp_disk_state": (C.c_int, [C.c_uint64, C.c_int32, C.c_double, _D, _D, _D]),
}

_lib = None


def lib() -> C.CDLL:
    """Load the in-tree engine library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            )
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class SppError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"spp error {code}: {msg}")
        self.code = code


def _ptr(a, t=_D):
    return None if a is None else a.ctypes.data_as(t)


def default_params(model: int = MODEL_REPULSIVE, **kw) -> Params:
    p = Params()
    rc = lib().spp_params_default(C.byref(p), model)
    if rc:
        raise SppError(rc, "spp_params_default")
    return p.update(**kw)


def random_state(seed: int, n: int, Lx: float, Ly: float):
    x = np.empty(n)
    y = np.empty(n)
    th = np.empty(n)
    rc = lib().spp_random_state(seed, n, Lx, Ly, _ptr(x), _ptr(y), _ptr(th))
    if rc:
        raise SppError(rc, "spp_random_state")
    return x, y, th


def random_formation(seed: int, n: int, Lx: float, Ly: float, sigma: float = 0.0):
    """Random_Formation with the sigma inset (positions within L - sigma of the centre)."""
    x, y, th = np.empty(n), np.empty(n), np.empty(n)
    rc = lib().spp_random_formation(seed, n, Lx, Ly, sigma, _ptr(x), _ptr(y), _ptr(th))
    if rc:
        raise SppError(rc, "spp_random_formation")
    return x, y, th


def polar_formation(seed: int, n: int, Lx: float, Ly: float):
    """Polar_Formation: random positions, every heading 0."""
    x, y, th = np.empty(n), np.empty(n), np.empty(n)
    rc = lib().spp_polar_formation(seed, n, Lx, Ly, _ptr(x), _ptr(y), _ptr(th))
    if rc:
        raise SppError(rc, "spp_polar_formation")
    return x, y, th


def triangle_lattice(n: int, Lx: float, Ly: float, sigma: float = 0.0):
    x, y = np.empty(n), np.empty(n)
    rc = lib().spp_triangle_lattice(n, Lx, Ly, sigma, _ptr(x), _ptr(y))
    if rc:
        raise SppError(rc, "spp_triangle_lattice")
    return x, y


def ring_membrane(nm: int, bead_diameter: float = 1.0, radius: float = 0.0):
    x, y = np.empty(nm), np.empty(nm)
    rc = lib().spp_ring_membrane(nm, bead_diameter, radius, _ptr(x), _ptr(y))
    if rc:
        raise SppError(rc, "spp_ring_membrane")
    return x, y


def disk_state(seed: int, n: int, radius: float):
    x, y, th = np.empty(n), np.empty(n), np.empty(n)
    rc = lib().spp_disk_state(seed, n, radius, _ptr(x), _ptr(y), _ptr(th))
    if rc:
        raise SppError(rc, "spp_disk_state")
    return x, y, th


def noise_normal(seed: int, pid: int, step: int) -> float:
    return lib().spp_noise_normal(seed, pid, step)


def system(n: int, Lx: float, Ly: float | None = None, v0: float = 1.0, g: float = 1.0,
           D_rot: float = 0.0, seed: int = 1, ncell_x: int = 0, ncell_y: int = 0) -> System:
    s = System()
    s.n, s.Lx, s.Ly = n, Lx, Lx if Ly is None else Ly
    s.v0, s.g, s.D_rot, s.seed = v0, g, D_rot, seed
    s.ncell_x, s.ncell_y = ncell_x, ncell_y
    return s


class Engine:
    """One engine (device state + stream) for one box of particles, or for a batch of independent
    boxes when ``systems`` is given (caller arrays are then the systems' particles concatenated)."""

    def __init__(self, params: Params, systems=None):
        self._lib = lib()
        self.params = params.copy()
        h = C.c_void_p()
        if systems is None:
            rc = self._lib.spp_create(C.byref(self.params), C.byref(h))
            self.sizes = [params.n]
        else:
            arr = (System * len(systems))(*systems)
            rc = self._lib.spp_create_batch(C.byref(self.params), arr, len(systems), C.byref(h))
            self.sizes = [s.n for s in systems]
        if rc:
            raise SppError(rc, self._lib.spp_last_error(None).decode())
        self._h = h
        self.n = int(sum(self.sizes))
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)

    # -- plumbing ------------------------------------------------------------------------------
    def _check(self, rc: int):
        if rc:
            raise SppError(rc, self._lib.spp_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.spp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state ---------------------------------------------------------------------------------
    def _host_array(self, a, name):
        """A contiguous float64 copy of `a` holding exactly one value per particle (the C side reads
        n values through the pointer)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.n,):
            raise ValueError(f"{name} has shape {a.shape}; the engine holds {self.n} particles")
        return a

    def set_state(self, xw, yw, theta, xu=None, yu=None):
        f = lambda a, name: None if a is None else self._host_array(a, name)
        xw, yw, theta, xu, yu = f(xw, "xw"), f(yw, "yw"), f(theta, "theta"), f(xu, "xu"), f(yu, "yu")
        self._check(self._lib.spp_set_state(self._h, _ptr(xw), _ptr(yw), _ptr(xu), _ptr(yu), _ptr(theta), self.n))

    def get_state(self):
        a = [np.empty(self.n) for _ in range(5)]
        self._check(self._lib.spp_get_state(self._h, *[_ptr(x) for x in a], self.n))
        return dict(zip(("xw", "yw", "xu", "yu", "theta"), a))

    def info(self) -> Info:
        i = Info()
        self._check(self._lib.spp_get_info(self._h, C.byref(i)))
        return i

    # -- neighbours ------------------------------------------------------------------------------
    def build_neighbors(self):
        self._check(self._lib.spp_build_neighbors(self._h))

    def get_neighbors(self):
        nnz = C.c_int64(0)
        rc = self._lib.spp_get_neighbors(self._h, None, None, 0, C.byref(nnz))
        if rc not in (OK, ERR_CAPACITY):
            self._check(rc)
        off = np.empty(self.n + 1, dtype=np.int32)
        idx = np.empty(max(nnz.value, 1), dtype=np.int32)
        self._check(self._lib.spp_get_neighbors(self._h, _ptr(off, _I32), _ptr(idx, _I32), idx.size, C.byref(nnz)))
        return off, idx[: nnz.value]

    def get_cells(self):
        cx = np.empty(self.n, dtype=np.int32)
        cy = np.empty(self.n, dtype=np.int32)
        self._check(self._lib.spp_get_cells(self._h, _ptr(cx, _I32), _ptr(cy, _I32), self.n))
        return cx, cy

    # -- Box surface ---------------------------------------------------------------------------
    def interact(self):
        self._check(self._lib.spp_interact(self._h))

    def get_forces(self):
        a = [np.empty(self.n) for _ in range(3)]
        self._check(self._lib.spp_get_forces(self._h, *[_ptr(x) for x in a], self.n))
        return a

    def inject_noise(self, dtheta):
        d = self._host_array(dtheta, "dtheta")
        self._check(self._lib.spp_inject_noise(self._h, _ptr(d), self.n))

    def move_stage(self, stage: int):
        self._check(self._lib.spp_move_stage(self._h, stage))

    def one_step(self):
        self._check(self._lib.spp_one_step(self._h))

    def multi_step(self, steps: int):
        self._check(self._lib.spp_multi_step(self._h, steps))

    def polarization(self):
        p, sx, sy = C.c_double(), C.c_double(), C.c_double()
        self._check(self._lib.spp_polarization(self._h, C.byref(p), C.byref(sx), C.byref(sy)))
        return p.value, sx.value, sy.value

    def multi_step_sampled(self, steps: int, every: int):
        """Run `steps` steps, sampling every system's polarisation every `every` steps on the
        device; returns [steps // every, systems]."""
        ns = len(self.sizes)
        out = np.empty((steps // every) * ns)
        self._check(self._lib.spp_multi_step_sampled(self._h, steps, every, _ptr(out), out.size))
        return out.reshape(steps // every, ns)

    def batch_polarization(self):
        p = np.empty(len(self.sizes))
        self._check(self._lib.spp_batch_polarization(self._h, _ptr(p), len(self.sizes)))
        return p

    def pair_distribution(self, r_max: float, nbins: int):
        """(bin centres, g(r), ordered-pair counts) over [0, r_max)."""
        counts = np.empty(nbins, dtype=np.int64)
        g = np.empty(nbins)
        self._check(self._lib.spp_pair_distribution(self._h, r_max, nbins, _ptr(counts, _I64), _ptr(g)))
        dr = r_max / nbins
        return (np.arange(nbins) + 0.5) * dr, g, counts

    def pair_distribution_noalloc(self, r_max: float, nbins: int, counts=None, g=None):
        """spp_pair_distribution_noalloc into caller-provided arrays (either may be None)."""
        for a in (counts, g):
            if a is not None and (a.shape != (nbins,) or not a.flags.c_contiguous):
                raise ValueError("output arrays need shape (nbins,), contiguous")
        if counts is not None and counts.dtype != np.int64 or g is not None and g.dtype != np.float64:
            raise ValueError("counts must be int64 and g float64")
        self._check(self._lib.spp_pair_distribution_noalloc(self._h, r_max, nbins, _ptr(counts, _I64), _ptr(g)))

    # -- Lyapunov clones (batched engine: replicas x (base + vectors) systems) -------------------
    def lyapunov_init(self, replicas: int, vectors: int, epsilon: float, rescale: bool = True, seed: int = 1):
        self._lyap = (replicas, vectors)
        self._check(self._lib.spp_lyapunov_init(self._h, replicas, vectors, epsilon, 1 if rescale else 0, seed))

    def lyapunov_renormalize(self, want: bool = False):
        if not want:
            self._check(self._lib.spp_lyapunov_renormalize(self._h, None))
            return None
        R, K = self._lyap
        out = np.empty(R * K)
        self._check(self._lib.spp_lyapunov_renormalize(self._h, _ptr(out)))
        return out.reshape(R, K)

    def lyapunov_run(self, steps: int, interval: int):
        self._check(self._lib.spp_lyapunov_run(self._h, steps, interval))

    def lyapunov_get(self):
        """(accumulated log growth [replica, vector], time covered)."""
        R, K = self._lyap
        out = np.empty(R * K)
        t = C.c_double()
        self._check(self._lib.spp_lyapunov_get(self._h, _ptr(out), C.byref(t)))
        return out.reshape(R, K), t.value

    def lyapunov_vectors(self, replica: int):
        """Orthonormal directions of one replica, shape [vector, 3 (x, y, theta), n]."""
        R, K = self._lyap
        n = self.sizes[0]
        out = np.empty(K * 3 * n)
        self._check(self._lib.spp_lyapunov_vectors(self._h, replica, _ptr(out), out.size))
        return out.reshape(K, 3, n)

    def snapshot(self, slot: int = 0):
        self._check(self._lib.spp_snapshot(self._h, slot))

    def restore(self, slot: int = 0):
        self._check(self._lib.spp_restore(self._h, slot))

    def reset_forces(self):
        """Particle::Reset of every particle (the split path's force accumulators)."""
        self._check(self._lib.spp_reset_forces(self._h))

    def pair_interact(self, i: int, j: int):
        """particles[i].Interact(particles[j]): adds the pair's force and torque to both."""
        self._check(self._lib.spp_pair_interact(self._h, i, j))

    def get_clock(self) -> Clock:
        """Noise-stream step, time and rebuild estimates: with get_state, a copy that replays."""
        c = Clock()
        self._check(self._lib.spp_get_clock(self._h, C.byref(c)))
        return c

    def set_clock(self, clock: Clock):
        """The random-generator and schedule part of Box::Load(State)."""
        self._check(self._lib.spp_set_clock(self._h, C.byref(clock)))

    def synchronize(self):
        self._check(self._lib.spp_synchronize(self._h))

    def graph_stats(self):
        """(intervals captured, intervals launched from a graph, graphs held)."""
        c, r, k = C.c_int64(), C.c_int64(), C.c_int32()
        self._check(self._lib.spp_graph_stats(self._h, C.byref(c), C.byref(r), C.byref(k)))
        return c.value, r.value, k.value

    # -- timing --------------------------------------------------------------------------------
    def profile(self, on: bool):
        self._check(self._lib.spp_profile_enable(self._h, 1 if on else 0))

    def time_stage(self, stage: int, reps: int = 50) -> float:
        """Average microseconds of one fused stage kernel over `reps` back-to-back launches."""
        us = C.c_double()
        self._check(self._lib.spp_time_stage(self._h, stage, reps, C.byref(us)))
        return us.value

    def profile_read(self):
        buf = C.create_string_buffer(4096)
        ms = np.zeros(64)
        cnt = np.zeros(64, dtype=np.int64)
        nk = C.c_int32(0)
        self._check(self._lib.spp_profile_read(self._h, buf, 4096, _ptr(ms), _ptr(cnt, _I64), 64, C.byref(nk)))
        names = buf.value.decode().split("\n")[: nk.value]
        return {n: (float(ms[i]), int(cnt[i])) for i, n in enumerate(names)}
