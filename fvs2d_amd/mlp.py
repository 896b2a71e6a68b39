"""ctypes binding of the warm-start network (include/fvs2d_mlp.h).

The network has no counterpart in the reference, which starts every steady
solve from the free stream (src/initialize.f90:36-43).  It runs only on the
device: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, ptr

RAW, FEATURES, WIDTH, OUT = 6, 68, 256, 4

MLP_EXPORTS = (
    "fvs_mlp_create", "fvs_mlp_destroy", "fvs_mlp_nparams", "fvs_mlp_get_params", "fvs_mlp_set_params",
    "fvs_mlp_features", "fvs_mlp_set_data", "fvs_mlp_train", "fvs_mlp_grad", "fvs_mlp_step", "fvs_mlp_predict",
    "fvs_mlp_kernel_times", "fvs_mlp_last_elapsed_ms",
)

KERNEL_CLASSES = ("features", "gemm_forward", "output_layer", "gemm_weight_grad", "gemm_backward_data", "adam")


class MLPOptions(C.Structure):
    """fvs_mlp_options (include/fvs2d_mlp.h)."""
    _fields_ = [("depth", C.c_int), ("batch", C.c_int), ("lr", C.c_double), ("beta1", C.c_double),
                ("beta2", C.c_double), ("eps", C.c_double), ("seed", C.c_ulonglong)]


_bound = False


def _load():
    global _bound
    L = _lib.load()
    if _bound:
        return L
    V, I, D = C.c_void_p, C.c_int, C.c_double
    L.fvs_mlp_create.argtypes = [C.POINTER(V), C.POINTER(MLPOptions)]
    L.fvs_mlp_destroy.argtypes = [V]
    L.fvs_mlp_nparams.argtypes = [V]
    L.fvs_mlp_nparams.restype = C.c_longlong
    L.fvs_mlp_get_params.argtypes = [V, V]
    L.fvs_mlp_set_params.argtypes = [V, V]
    L.fvs_mlp_features.argtypes = [V, I, V, V]
    L.fvs_mlp_set_data.argtypes = [V, C.c_longlong, V, V]
    L.fvs_mlp_train.argtypes = [V, I, V]
    L.fvs_mlp_grad.argtypes = [V, I, V, V, V, V]
    L.fvs_mlp_step.argtypes = [V, I, V, V, V]
    L.fvs_mlp_predict.argtypes = [V, C.c_longlong, V, V]
    L.fvs_mlp_kernel_times.argtypes = [V, I, V, V, V]
    L.fvs_mlp_last_elapsed_ms.argtypes = [V]
    L.fvs_mlp_last_elapsed_ms.restype = D
    _bound = True
    return L


def _f32(a, cols):
    a = np.ascontiguousarray(a, np.float32)
    if a.ndim != 2 or a.shape[1] != cols:
        raise ValueError(f"expected an (n, {cols}) array, got {a.shape}")
    return a


class WarmStartNet:
    """68 -> depth x 256 (GELU) -> 4 on the matrix cores (bfloat16 operands,
    float32 accumulation), trained with Adam by the library's own kernels."""

    def __init__(self, depth: int = 5, batch: int = 65536, lr: float = 1e-3, beta1: float = 0.9,
                 beta2: float = 0.999, eps: float = 1e-8, seed: int = 0):
        L = _load()
        self._L = L
        self.opts = MLPOptions(depth, batch, lr, beta1, beta2, eps, seed)
        h = C.c_void_p()
        check(L.fvs_mlp_create(C.byref(h), C.byref(self.opts)))
        self._h = h
        self.depth, self.batch = depth, batch
        self.nparams = int(L.fvs_mlp_nparams(h))

    def close(self):
        if getattr(self, "_h", None):
            self._L.fvs_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.empty(self.nparams, np.float32)
        check(self._L.fvs_mlp_get_params(self._h, ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters")
        check(self._L.fvs_mlp_set_params(self._h, ptr(p)))

    def features(self, raw) -> np.ndarray:
        raw = _f32(raw, RAW)
        out = np.empty((raw.shape[0], FEATURES), np.float32)
        check(self._L.fvs_mlp_features(self._h, raw.shape[0], ptr(raw), ptr(out)))
        return out

    def set_data(self, raw, target):
        raw, target = _f32(raw, RAW), _f32(target, OUT)
        if raw.shape[0] != target.shape[0]:
            raise ValueError("raw and target row counts differ")
        check(self._L.fvs_mlp_set_data(self._h, raw.shape[0], ptr(raw), ptr(target)))

    def train(self, steps: int) -> np.ndarray:
        loss = np.empty(steps)
        check(self._L.fvs_mlp_train(self._h, steps, ptr(loss)))
        return loss

    def grad(self, raw, target):
        raw, target = _f32(raw, RAW), _f32(target, OUT)
        g = np.empty(self.nparams, np.float32)
        loss = np.zeros(1)
        check(self._L.fvs_mlp_grad(self._h, raw.shape[0], ptr(raw), ptr(target), ptr(loss), ptr(g)))
        return float(loss[0]), g

    def step(self, raw, target) -> float:
        raw, target = _f32(raw, RAW), _f32(target, OUT)
        loss = np.zeros(1)
        check(self._L.fvs_mlp_step(self._h, raw.shape[0], ptr(raw), ptr(target), ptr(loss)))
        return float(loss[0])

    def predict(self, raw) -> np.ndarray:
        raw = _f32(raw, RAW)
        out = np.empty((raw.shape[0], OUT), np.float32)
        check(self._L.fvs_mlp_predict(self._h, raw.shape[0], ptr(raw), ptr(out)))
        return out

    def kernel_times(self, steps: int):
        """Per kernel class: (mean ms per launch, matrix-core flops per
        launch, launches per step) over ``steps`` training steps (the
        weights advance)."""
        ms, fl = np.zeros(6), np.zeros(6)
        n = np.zeros(6, np.int32)
        check(self._L.fvs_mlp_kernel_times(self._h, steps, ptr(ms), ptr(fl), ptr(n)))
        return {k: (float(ms[i]), float(fl[i]), int(n[i])) for i, k in enumerate(KERNEL_CLASSES)}

    @property
    def last_elapsed_ms(self) -> float:
        return float(self._L.fvs_mlp_last_elapsed_ms(self._h))
