"""Learned warm start: a per-node MLP, trained and evaluated by liboc_asm's
own HIP kernels, that predicts each load step's nodal displacement increment
as Newton's starting guess.  The Newton solve's own residual test still
decides: every accepted answer meets the same FP64 residual tolerance.

Features per quadratic node (9): reference coordinates X (3), current
displacement z - X (3), load increment of the step, c10, c01.  Targets (4):
displacement increment of the step and the increment of the (trilinear)
hydrostatic pressure interpolated to the node (in units of c10 + c01), both
divided by the load increment.  The pressure output matters: the exact displacement increment with
the previous step's pressure still leaves about 13% of the plain start's
residual (the stress balance needs the new pressure).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import call

N_FEATURES = 9
N_OUTPUTS = 4
# fixed feature normalisation: x' = (x - shift) * scale
FEATURE_SHIFT = np.array([0.5, 0.5, 0.5, 0.0, 0.0, 0.0, 0.0, 3.0, 5.0], np.float32)
FEATURE_SCALE = np.array([2.0, 2.0, 2.0, 5.0, 5.0, 5.0, 25.0, 0.5, 0.2], np.float32)


class WarmStartMLP:
    """9 -> hidden x layers (GELU) -> 3 MLP on the GPU."""

    def __init__(self, hidden=128, layers=4, seed=0, outputs=N_OUTPUTS):
        h = ctypes.c_void_p()
        call("oc_mlp_create", N_FEATURES, int(hidden), int(layers), int(outputs), int(seed), ctypes.byref(h))
        self._h = h
        n = ctypes.c_int()
        call("oc_mlp_parameter_count", self._h, ctypes.byref(n))
        self.n_parameters = n.value
        self.hidden, self.layers, self.outputs = hidden, layers, outputs
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
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_parameters:
            raise ValueError(f"expected {self.n_parameters} parameters, got {p.size}")
        call("oc_mlp_parameters_set", self._h, p)

    def set_inference_precision(self, precision):
        """OC_FP32 (default), OC_BF16 or OC_BF16_HIDDEN (bf16 hidden layers
        only) for forward() and the warm starts."""
        call("oc_mlp_inference_precision_set", self._h, int(precision))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim != 2 or x.shape[1] != N_FEATURES:
            raise ValueError(f"features must be (n, {N_FEATURES})")
        y = np.zeros((x.shape[0], self.outputs), np.float32)
        call("oc_mlp_forward", self._h, int(x.shape[0]), x, y)
        return y

    def gradient(self, x, t):
        x, t = self._checked(x, t)
        g = np.zeros(self.n_parameters, np.float32)
        loss = ctypes.c_float()
        call("oc_mlp_gradient", self._h, int(x.shape[0]), x, t, ctypes.byref(loss), g)
        return float(loss.value), g

    def _checked(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        if x.ndim != 2 or x.shape[1] != N_FEATURES or t.shape != (x.shape[0], self.outputs):
            raise ValueError(f"features must be (n, {N_FEATURES}) and targets (n, {self.outputs})")
        return x, t

    def set_dataset(self, x, t):
        x, t = self._checked(x, t)
        call("oc_mlp_dataset_set", self._h, int(x.shape[0]), x, t)

    def train(self, steps, batch=4096, learning_rate=1e-3, seed=1):
        loss = ctypes.c_float()
        call("oc_mlp_train", self._h, int(steps), int(batch), float(learning_rate), int(seed), ctypes.byref(loss))
        return float(loss.value)

    def warm_start(self, equations_set, load_increment):
        """z += load_increment * MLP(features) on the free DOFs; a batch takes
        one load increment per mesh (scalar or array)."""
        inc = np.asarray(load_increment, np.float64)
        if inc.ndim == 0:
            call("oc_mlp_warm_start", self._h, equations_set.handle, float(inc))
        else:
            call("oc_mlp_warm_start_batch", self._h, equations_set.handle, np.ascontiguousarray(inc))

    def save(self, path):
        np.savez(path, parameters=self.parameters(), hidden=self.hidden, layers=self.layers, outputs=self.outputs)

    @classmethod
    def load(cls, path):
        d = np.load(path)
        net = cls(hidden=int(d["hidden"]), layers=int(d["layers"]), outputs=int(d["outputs"]) if "outputs" in d else 3)
        net.set_parameters(d["parameters"])
        return net


# trilinear interpolation weights of the 8 element corners at the 27
# triquadratic nodes (xi1 fastest in both numberings)
_T = np.array([0.0, 0.5, 1.0])
_W = np.array([[(_T[a % 3] if k & 1 else 1 - _T[a % 3]) * (_T[(a // 3) % 3] if k & 2 else 1 - _T[(a // 3) % 3]) *
                (_T[a // 9] if k & 4 else 1 - _T[a // 9]) for k in range(8)] for a in range(27)])


def pressure_at_quadratic_nodes(mesh, p_lin):
    """The trilinear pressure field evaluated at every quadratic node."""
    pq = np.empty(mesh.n_nodes)
    pq[mesh.elem_nodes] = p_lin[mesh.elem_nodes_l] @ _W.T
    return pq


def step_targets(mesh, d0, d1, load_increment, c10, c01):
    """(z1 - z0, (p1 - p0 at the quadratic nodes) / (c10 + c01)) / load increment."""
    nq = mesh.n_nodes
    t = np.empty((nq, N_OUTPUTS), np.float32)
    t[:, 0:3] = ((d1[:3 * nq] - d0[:3 * nq]).reshape(3, nq).T / load_increment)
    t[:, 3] = pressure_at_quadratic_nodes(mesh, (d1[3 * nq:] - d0[3 * nq:]) / (load_increment * (c10 + c01)))
    return t


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
