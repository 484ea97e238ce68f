"""Binding of the surrogate network's C ABI (include/spp_surrogate.h) in the engine library.

The network (4 -> 128 -> 128 -> 128 -> 3, SiLU) runs forward, backward and Adam on the GPU with the
library's BF16-MFMA kernels; this module only moves numpy arrays across the boundary.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import SppError, lib as _engine_lib


class MlpConfig(C.Structure):
    """Mirror of ``spp_mlp_config``."""

    _fields_ = [
        ("n_in", C.c_int32),
        ("hidden", C.c_int32),
        ("n_hidden", C.c_int32),
        ("n_out", C.c_int32),
        ("max_batch", C.c_int32),
        ("reserved", C.c_int32),
        ("seed", C.c_uint64),
        ("lr", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("eps", C.c_float),
    ]


_F = C.POINTER(C.c_float)
_Dp = C.POINTER(C.c_double)
SIGNATURES = {
    "spp_mlp_config_default": (None, [C.POINTER(MlpConfig)]),
    "spp_mlp_create": (C.c_int, [C.POINTER(MlpConfig), C.POINTER(C.c_void_p)]),
    "spp_mlp_destroy": (None, [C.c_void_p]),
    "spp_mlp_last_error": (C.c_char_p, [C.c_void_p]),
    "spp_mlp_param_count": (C.c_int64, [C.c_void_p]),
    "spp_mlp_get_params": (C.c_int, [C.c_void_p, _F, C.c_int64]),
    "spp_mlp_set_params": (C.c_int, [C.c_void_p, _F, C.c_int64]),
    "spp_mlp_forward": (C.c_int, [C.c_void_p, _F, C.c_int32, _F]),
    "spp_mlp_loss_grad": (C.c_int, [C.c_void_p, _F, _F, _F, C.c_int32, _F, _F]),
    "spp_mlp_train_step": (C.c_int, [C.c_void_p, _F, _F, _F, C.c_int32, _F]),
    "spp_mlp_train": (C.c_int, [C.c_void_p, _F, _F, _F, C.c_int32, C.c_int32, C.c_int32, C.c_uint64, _F]),
    "spp_mlp_last_train_ms": (C.c_int, [C.c_void_p, _Dp]),
}

_bound = None


def lib():
    global _bound
    if _bound is None:
        L = _engine_lib()
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _bound = L
    return _bound


def _f(a, shape=None):
    """A contiguous float32 copy and its pointer; `shape` (None entries free) is checked, since
    the C side reads as many values as the shape it is told implies."""
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and (a.ndim != len(shape) or any(w is not None and w != d for w, d in zip(shape, a.shape))):
        raise ValueError(f"array of shape {a.shape}, expected {tuple('*' if w is None else w for w in shape)}")
    return a, a.ctypes.data_as(_F)


def default_config(**kw) -> MlpConfig:
    c = MlpConfig()
    lib().spp_mlp_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Mlp:
    """The surrogate network on the GPU."""

    def __init__(self, config: MlpConfig | None = None, **kw):
        self._lib = lib()
        self.config = config if config is not None else default_config(**kw)
        h = C.c_void_p()
        rc = self._lib.spp_mlp_create(C.byref(self.config), C.byref(h))
        if rc:
            raise SppError(rc, self._lib.spp_mlp_last_error(None).decode())
        self._h = h
        c = self.config
        self.dims = [c.n_in] + [c.hidden] * c.n_hidden + [c.n_out]
        self.n_params = int(self._lib.spp_mlp_param_count(h))

    def _check(self, rc):
        if rc:
            raise SppError(rc, self._lib.spp_mlp_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.spp_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.empty(self.n_params, dtype=np.float32)
        self._check(self._lib.spp_mlp_get_params(self._h, p.ctypes.data_as(_F), p.size))
        return p

    def set_params(self, p):
        p, pp = _f(p, (self.n_params,))
        self._check(self._lib.spp_mlp_set_params(self._h, pp, p.size))

    def forward(self, x):
        x, xp = _f(np.atleast_2d(x), (None, self.dims[0]))  # one point may come as a 1-D row
        y = np.empty((x.shape[0], self.dims[-1]), dtype=np.float32)
        self._check(self._lib.spp_mlp_forward(self._h, xp, x.shape[0], y.ctypes.data_as(_F)))
        return y

    def loss_grad(self, x, t, w=None):
        x, xp = _f(x, (None, self.dims[0]))
        t, tp = _f(t, (x.shape[0], self.dims[-1]))
        w, wp = _f(w, (self.dims[-1],))
        loss = C.c_float()
        g = np.empty(self.n_params, dtype=np.float32)
        self._check(self._lib.spp_mlp_loss_grad(self._h, xp, tp, wp, x.shape[0], C.byref(loss), g.ctypes.data_as(_F)))
        return loss.value, g

    def train_step(self, x, t, w=None):
        x, xp = _f(x, (None, self.dims[0]))
        t, tp = _f(t, (x.shape[0], self.dims[-1]))
        w, wp = _f(w, (self.dims[-1],))
        loss = C.c_float()
        self._check(self._lib.spp_mlp_train_step(self._h, xp, tp, wp, x.shape[0], C.byref(loss)))
        return loss.value

    def train(self, x, t, epochs, batch=256, w=None, seed=1):
        x, xp = _f(x, (None, self.dims[0]))
        t, tp = _f(t, (x.shape[0], self.dims[-1]))
        w, wp = _f(w, (self.dims[-1],))
        loss = C.c_float()
        self._check(self._lib.spp_mlp_train(self._h, xp, tp, wp, x.shape[0], epochs, batch, seed, C.byref(loss)))
        return loss.value

    def last_train_ms(self):
        ms = C.c_double()
        self._check(self._lib.spp_mlp_last_train_ms(self._h, C.byref(ms)))
        return ms.value
