"""Learned shift proposals for wavenumber sweeps.

A small network maps (alpha, Re, Ma) to the 8 leading omega of the BiGlobal
problem (the eigenvalues a fixed-shift solve finds nearest the search centre,
ordered by growth rate Im(omega), largest first).  Its training and inference
run on the GPU through the library's own kernels (tailor_mlp_* in
include/tailor.h); this module only normalises inputs and targets and keeps
the normalisation with the weights.

The sweep uses the centroid of the predicted omegas as the shift centre of
each instance: the wanted cluster then sits around the shift and
Krylov-Schur converges in fewer restarts.  A proposal is only a starting
point: an instance whose pairs do not reach the residual tolerance is solved
again with the plain fixed shift.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import TailorError, library

# fixed affine input normalisation over the training ranges
ALPHA_RANGE = (0.5, 3.0)
RE_RANGE = (500.0, 3000.0)
MA_RANGE = (0.3, 0.9)
N_OUT = 16


def normalise_inputs(alpha, re, ma) -> np.ndarray:
    x = np.stack([np.asarray(alpha, float), np.asarray(re, float), np.asarray(ma, float)], axis=-1)
    lo = np.array([ALPHA_RANGE[0], RE_RANGE[0], MA_RANGE[0]])
    hi = np.array([ALPHA_RANGE[1], RE_RANGE[1], MA_RANGE[1]])
    return ((x - 0.5 * (lo + hi)) / (0.5 * (hi - lo))).astype(np.float32).reshape(-1, 3)


def omegas_to_targets(omegas: np.ndarray) -> np.ndarray:
    """(n, 8) complex -> (n, 16) real: real parts then imaginary parts."""
    w = np.asarray(omegas, dtype=np.complex128).reshape(-1, 8)
    return np.concatenate([w.real, w.imag], axis=1)


def targets_to_omegas(t: np.ndarray) -> np.ndarray:
    t = np.asarray(t, dtype=np.float64).reshape(-1, N_OUT)
    return t[:, :8] + 1j * t[:, 8:]


def leading_order(omegas) -> np.ndarray:
    """The 8 omegas ordered by Im(omega) descending (ties by Re)."""
    w = np.asarray(omegas, dtype=np.complex128)
    return w[np.lexsort((w.real, -w.imag))][:8]


def _bind(lib):
    if getattr(lib, "_mlp_bound", False):
        return
    P = C.POINTER
    f = P(C.c_float)
    lib.tailor_mlp_param_count.argtypes = []
    lib.tailor_mlp_create.argtypes = [P(C.c_void_p), C.c_int, C.c_uint64]
    lib.tailor_mlp_destroy.argtypes = [C.c_void_p]
    lib.tailor_mlp_destroy.restype = None
    lib.tailor_mlp_last_error.argtypes = [C.c_void_p]
    lib.tailor_mlp_last_error.restype = C.c_char_p
    lib.tailor_mlp_get_params.argtypes = [C.c_void_p, f]
    lib.tailor_mlp_set_params.argtypes = [C.c_void_p, f]
    lib.tailor_mlp_predict.argtypes = [C.c_void_p, f, C.c_int32, f]
    lib.tailor_mlp_gradients.argtypes = [C.c_void_p, f, f, C.c_int32, f, f]
    lib.tailor_mlp_train.argtypes = [C.c_void_p, f, f, C.c_int32, C.c_int32, C.c_float, f]
    lib._mlp_bound = True


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Network:
    """The raw network on a device: normalised tensors in and out."""

    def __init__(self, device: int = 0, seed: int = 2):
        self.lib = library()
        _bind(self.lib)
        self.n_params = self.lib.tailor_mlp_param_count()
        ptr = C.c_void_p()
        st = self.lib.tailor_mlp_create(C.byref(ptr), device, seed)
        if st:
            raise TailorError(st, "tailor_mlp_create failed")
        self.ptr = ptr

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.tailor_mlp_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st:
            raise TailorError(st, self.lib.tailor_mlp_last_error(self.ptr).decode())

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        self._check(self.lib.tailor_mlp_get_params(self.ptr, _fptr(out)))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self._check(self.lib.tailor_mlp_set_params(self.ptr, _fptr(p)))

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 3)
        Y = np.zeros((X.shape[0], N_OUT), dtype=np.float32)
        self._check(self.lib.tailor_mlp_predict(self.ptr, _fptr(X), X.shape[0], _fptr(Y)))
        return Y

    def gradients(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 3)
        T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1, N_OUT)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_float()
        self._check(self.lib.tailor_mlp_gradients(self.ptr, _fptr(X), _fptr(T), X.shape[0],
                                                  _fptr(g), C.byref(loss)))
        return loss.value, g

    def train(self, X, T, steps: int, lr: float = 1e-3) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 3)
        T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1, N_OUT)
        hist = np.zeros(steps, dtype=np.float32)
        self._check(self.lib.tailor_mlp_train(self.ptr, _fptr(X), _fptr(T), X.shape[0], steps,
                                              lr, _fptr(hist)))
        return hist


class ShiftModel:
    """(alpha, Re, Ma) -> predicted 8 leading omega, with its normalisation."""

    def __init__(self, device: int = 0, seed: int = 2):
        self.net = Network(device, seed)
        self.t_mean = np.zeros(N_OUT)
        self.t_std = np.ones(N_OUT)

    def fit(self, alpha, re, ma, omegas, steps: int = 3000, lr: float = 3e-3) -> np.ndarray:
        X = normalise_inputs(alpha, re, ma)
        T = omegas_to_targets(np.array([leading_order(w) for w in omegas]))
        self.t_mean = T.mean(axis=0)
        self.t_std = np.maximum(T.std(axis=0), 1e-12)
        Tn = ((T - self.t_mean) / self.t_std).astype(np.float32)
        return self.net.train(X, Tn, steps, lr)

    def predict(self, alpha, re, ma) -> np.ndarray:
        Y = self.net.predict(normalise_inputs(alpha, re, ma)).astype(np.float64)
        return targets_to_omegas(Y * self.t_std + self.t_mean)

    def shift_centres(self, alpha, re, ma) -> np.ndarray:
        """Centroid of the predicted omegas: one shift centre per input."""
        return self.predict(alpha, re, ma).mean(axis=1)

    def save(self, path) -> None:
        np.savez(path, params=self.net.params, t_mean=self.t_mean, t_std=self.t_std)

    def load(self, path) -> "ShiftModel":
        with np.load(Path(path), allow_pickle=False) as z:
            self.net.params = z["params"]
            self.t_mean = z["t_mean"]
            self.t_std = z["t_std"]
        return self
