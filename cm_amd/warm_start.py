"""Learned warm start: a per-node MLP, trained and evaluated by liboc_asm's
own HIP kernels, that predicts each load step's nodal displacement increment
as Newton's starting guess.  The plain Newton solve remains the judge: every
accepted answer meets the same FP64 residual tolerance.

Features per quadratic node (9): reference coordinates X (3), current
displacement z - X (3), load increment of the step, c10, c01.  Target (3):
displacement increment of the step divided by the load increment.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import call

N_FEATURES = 9
N_OUTPUTS = 3
# fixed feature normalisation: x' = (x - shift) * scale
FEATURE_SHIFT = np.array([0.5, 0.5, 0.5, 0.0, 0.0, 0.0, 0.0, 3.0, 5.0], np.float32)
FEATURE_SCALE = np.array([2.0, 2.0, 2.0, 5.0, 5.0, 5.0, 25.0, 0.5, 0.2], np.float32)


class WarmStartMLP:
    """9 -> hidden x layers (GELU) -> 3 MLP on the GPU."""

    def __init__(self, hidden=128, layers=4, seed=0):
        h = ctypes.c_void_p()
        call("oc_mlp_create", N_FEATURES, int(hidden), int(layers), N_OUTPUTS, int(seed), ctypes.byref(h))
        self._h = h
        n = ctypes.c_int()
        call("oc_mlp_parameter_count", self._h, ctypes.byref(n))
        self.n_parameters = n.value
        self.hidden, self.layers = hidden, layers
        self.set_normalisation(FEATURE_SHIFT, FEATURE_SCALE)

    def destroy(self):
        if self._h:
            call("oc_mlp_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if self._h and _lib._LIB is not None:
                _lib._LIB.oc_mlp_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_normalisation(self, shift, scale):
        call("oc_mlp_normalisation_set", self._h, np.ascontiguousarray(shift, np.float32),
             np.ascontiguousarray(scale, np.float32))

    def parameters(self):
        p = np.zeros(self.n_parameters, np.float32)
        call("oc_mlp_parameters_get", self._h, p)
        return p

    def set_parameters(self, p):
        call("oc_mlp_parameters_set", self._h, np.ascontiguousarray(p, np.float32))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], N_OUTPUTS), np.float32)
        call("oc_mlp_forward", self._h, int(x.shape[0]), x, y)
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.n_parameters, np.float32)
        loss = ctypes.c_float()
        call("oc_mlp_gradient", self._h, int(x.shape[0]), x, t, ctypes.byref(loss), g)
        return float(loss.value), g

    def set_dataset(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        call("oc_mlp_dataset_set", self._h, int(x.shape[0]), x, t)

    def train(self, steps, batch=4096, learning_rate=1e-3, seed=1):
        loss = ctypes.c_float()
        call("oc_mlp_train", self._h, int(steps), int(batch), float(learning_rate), int(seed), ctypes.byref(loss))
        return float(loss.value)

    def warm_start(self, equations_set, load_increment):
        """z += load_increment * MLP(features) on the free displacement DOFs."""
        call("oc_mlp_warm_start", self._h, equations_set.handle, float(load_increment))

    def save(self, path):
        np.savez(path, parameters=self.parameters(), hidden=self.hidden, layers=self.layers)

    @classmethod
    def load(cls, path):
        d = np.load(path)
        net = cls(hidden=int(d["hidden"]), layers=int(d["layers"]))
        net.set_parameters(d["parameters"])
        return net


def step_features(mesh, dofs, load_increment, c10, c01):
    """Host-side features of every quadratic node (training data)."""
    nq = mesh.n_nodes
    X = mesh.coords
    u = dofs[:3 * nq].reshape(3, nq) - X
    f = np.empty((nq, N_FEATURES), np.float32)
    f[:, 0:3] = X.T
    f[:, 3:6] = u.T
    f[:, 6] = load_increment
    f[:, 7] = c10
    f[:, 8] = c01
    return f
