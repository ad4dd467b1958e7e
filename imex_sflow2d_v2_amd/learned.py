"""Learned Newton warm start: the device MLP (imexsw_mlp_*), harvesting of
(state entering Newton, converged solution) pairs from the solver itself,
training on the device, and switching a Solver to warm-started Newton.

The MLP's forward, backward and Adam step run in the library's own HIP
kernels (BF16 operands, FP32 accumulation on the matrix cores); nothing here
computes on the host besides the standardisation statistics.
"""
import ctypes as C

import numpy as np

from .solver import ImexswError, load_library

N_FEATURES = 9
N_TARGETS = 2
PAIR_WIDTH = N_FEATURES + N_TARGETS
FEATURES = ("log_h", "u", "v", "T", "alpha_1", "alpha_2", "alpha_3", "log_dt", "slope")
_F = C.POINTER(C.c_float)


def _fp(a):
    return a.ctypes.data_as(_F)


class Mlp:
    """9 -> 64 -> 64 -> 64 -> 2 GELU network on the device."""

    def __init__(self, seed=0):
        self.lib = load_library()
        h = C.c_void_p()
        if self.lib.imexsw_mlp_create(C.byref(h), int(seed)) != 0:
            raise ImexswError(self.lib.imexsw_mlp_last_error(None).decode())
        self.h = h
        self.n_params = self.lib.imexsw_mlp_num_params()

    def _chk(self, rc):
        if rc < 0:
            raise ImexswError(self.lib.imexsw_mlp_last_error(self.h).decode())
        return rc

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        n = np.zeros(2 * N_FEATURES + 2 * N_TARGETS, dtype=np.float32)
        self._chk(self.lib.imexsw_mlp_get_params(self.h, _fp(p), _fp(n)))
        return p, n

    def set_params(self, params=None, norm=None):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        n = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        self._chk(self.lib.imexsw_mlp_set_params(self.h, None if p is None else _fp(p),
                                                 None if n is None else _fp(n)))

    def forward(self, x):
        """Raw features (n, 9) -> raw targets (n, 2)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], N_TARGETS), dtype=np.float32)
        self._chk(self.lib.imexsw_mlp_forward(self.h, _fp(x), x.shape[0], _fp(y)))
        return y

    def grad(self, xs, ts):
        """Loss and parameter gradient on standardised (n, 9), (n, 2)."""
        xs = np.ascontiguousarray(xs, dtype=np.float32)
        ts = np.ascontiguousarray(ts, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_float(0.0)
        self._chk(self.lib.imexsw_mlp_grad(self.h, _fp(xs), _fp(ts), xs.shape[0], _fp(g),
                                           C.byref(loss)))
        return g, loss.value

    def train(self, pairs, epochs=2, batch=8192, lr=1e-3, seed=0):
        pairs = np.ascontiguousarray(pairs, dtype=np.float32)
        hist = np.zeros(epochs, dtype=np.float32)
        self._chk(self.lib.imexsw_mlp_train(self.h, _fp(pairs), pairs.shape[0], int(epochs),
                                            int(batch), float(lr), int(seed), _fp(hist)))
        return hist

    def close(self):
        if getattr(self, "h", None):
            self.lib.imexsw_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def start_harvest(solver, capacity):
    """Record the pairs of every converged stage-2 Newton solve of the
    following steps (up to `capacity`)."""
    solver._chk(solver.lib.imexsw_harvest(solver.h, int(capacity)))


def harvested(solver, cap=None):
    """(pairs (n, 11) float32, produced) -- produced may exceed n when the
    buffer filled up."""
    n = C.c_int64(0)
    solver._chk(solver.lib.imexsw_harvest_get(solver.h, None, 0, C.byref(n)))
    k = n.value if cap is None else min(n.value, cap)
    out = np.zeros((k, PAIR_WIDTH), dtype=np.float32)
    solver._chk(solver.lib.imexsw_harvest_get(solver.h, _fp(out), k, C.byref(n)))
    keep = min(k, n.value)
    return out[:keep], n.value


def stop_harvest(solver):
    solver._chk(solver.lib.imexsw_harvest(solver.h, 0))


def set_warm_start(solver, mlp):
    """Warm-start the solver's Newton solves with `mlp` (None: plain)."""
    solver._chk(solver.lib.imexsw_set_warm_start(solver.h, None if mlp is None else mlp.h))
