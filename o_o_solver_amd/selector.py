"""The learned precision selector (no counterpart in the reference): an MLP of 8 → 32 → 32 → 1 on the device that
gives, for each cut of a plan's levels, p(the cut pays): factorising the subtrees below the cut in FP32 (the rest in
FP64, refined by conjugate gradients to the target) reaches the target backward error in less time than FP64 fronts
throughout. Its features per cut (oosg_selector_cut_features) are the cut's depth, the share of the flops below it,
the front orders below and above it, log n, the scaled-diagonal dominance of the equations below it and their pivot
growth in the plan's last factorisation. The selector takes the cut of the highest probability when it is at least
the threshold, else FP64 fronts throughout.

Training data are (features, outcome) pairs of real runs: o_o_solver_amd/train_selector.py times a family of
synthetic models with FP64 fronts and with every cut on the GPU.
"""

import ctypes
import os

import numpy as np

from . import _native

FEATURES = 8
HIDDEN = 32
PARAMETERS = HIDDEN * FEATURES + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_weights.npz")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(status):
    if status != 0:
        from . import _check as check
        check(status)


class PrecisionSelector:
    def __init__(self, device=0, seed=1, parameters=None):
        self._lib = _native.lib()
        handle = ctypes.c_void_p()
        _check(self._lib.oosg_selector_create(int(device), int(seed), ctypes.byref(handle)))
        self.handle = handle
        if parameters is not None:
            self.set_parameters(parameters)

    @classmethod
    def trained(cls, device=0, path=WEIGHTS):
        """The selector with the committed weights (None when they are absent)."""
        if not os.path.exists(path):
            return None
        with np.load(path, allow_pickle=False) as data:
            return cls(device, parameters=data["parameters"])

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.oosg_selector_destroy(self.handle)
            self.handle = None

    def parameters(self):
        out = np.zeros(PARAMETERS, np.float32)
        _check(self._lib.oosg_selector_parameters(self.handle, _ptr(out)))
        return out

    def set_parameters(self, parameters):
        p = np.ascontiguousarray(parameters, dtype=np.float32)
        if p.shape != (PARAMETERS,):
            raise ValueError(f"selector: {PARAMETERS} parameters expected")
        _check(self._lib.oosg_selector_set_parameters(self.handle, _ptr(p)))

    def save(self, path=WEIGHTS):
        np.savez(path, parameters=self.parameters())

    @staticmethod
    def cut_features(plan, device_matrix):
        """The features of each cut l = 1 .. L of the plan's levels (row l - 1): the levels below l in FP32."""
        lib = _native.lib()
        count = ctypes.c_int32()
        _check(lib.oosg_selector_cut_features(plan.handle, device_matrix.handle, None, 0, ctypes.byref(count)))
        out = np.zeros((count.value, FEATURES), np.float32)
        if count.value > 0:
            _check(lib.oosg_selector_cut_features(plan.handle, device_matrix.handle, _ptr(out), count.value,
                                                  ctypes.byref(count)))
        return out

    def choose(self, plan, device_matrix, threshold=0.5):
        """The cut (the levels factorised in FP32: 0 for FP64 fronts throughout, L for FP32 throughout) and the
        probability of each cut 1 .. L."""
        p = self.predict(self.cut_features(plan, device_matrix)) if plan.levels > 0 else np.zeros(0, np.float32)
        if p.size == 0:
            return 0, p
        best = int(np.argmax(p))
        return (best + 1 if p[best] >= threshold else 0), p

    def predict(self, features):
        x = np.ascontiguousarray(np.atleast_2d(features), dtype=np.float32)
        out = np.zeros(x.shape[0], np.float32)
        _check(self._lib.oosg_selector_predict(self.handle, _ptr(x), x.shape[0], _ptr(out)))
        return out

    def gradient(self, features, labels):
        x = np.ascontiguousarray(np.atleast_2d(features), dtype=np.float32)
        y = np.ascontiguousarray(labels, dtype=np.float32)
        loss = np.zeros(1, np.float32)
        grad = np.zeros(PARAMETERS, np.float32)
        _check(self._lib.oosg_selector_gradient(self.handle, _ptr(x), _ptr(y), x.shape[0], _ptr(loss), _ptr(grad)))
        return float(loss[0]), grad

    def train(self, features, labels, steps=500, learning_rate=1e-2):
        x = np.ascontiguousarray(np.atleast_2d(features), dtype=np.float32)
        y = np.ascontiguousarray(labels, dtype=np.float32)
        loss = np.zeros(1, np.float32)
        _check(self._lib.oosg_selector_train(self.handle, _ptr(x), _ptr(y), x.shape[0], int(steps),
                                             float(learning_rate), _ptr(loss)))
        return float(loss[0])
