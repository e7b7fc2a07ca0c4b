"""Regular generated meshes of the unit cube and their DOF numbering.

Restates, for Lagrange tensor-product hexahedra in rectangular Cartesian
coordinates:

* GENERATED_MESH_REGULAR_CREATE_FINISH (generated_mesh_routines.f90:1386-1612):
  elements numbered xi1 fastest, element nodes xi1 fastest,
  node ``np = 1 + i + j*T1 + k*T1*T2`` on the grid of T = order*n+1 nodes;
* GeneratedMesh_RegularComponentNodesToUserNumbers (:5043-5458) for a
  trilinear component on a triquadratic mesh: corner nodes take the
  quadratic user numbers;
* MeshTopologyNodesCalculate (mesh_routines.f90:9006-9060): a component's
  mesh nodes are its global nodes sorted ascending, so the trilinear
  component's corner (i, j, k) has rank ``i + j*(n1+1) + k*(n1+1)*(n2+1)``;
* FIELD_MAPPINGS_CALCULATE separated-component DOF order
  (field_routines.f90:10345-10628): all DOFs of component 1, then 2, ...;
* GeneratedMesh_RegularGeometricParametersCalculate (:3360-3445):
  ``x = (i-1) * extent/(T-1)``.

The seeded interior-node jitter is the synthetic-input rule of the
benchmark configurations (uniform in [-jitter*h, jitter*h] per component,
h = 1/n, splitmix64 stream), not a reference feature.
"""
from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def uniform(seed, index):
    """splitmix64 uniform doubles in [0, 1) for integer array ``index``."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(index, dtype=np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def regular_elements(order, n1, n2, n3):
    """0-based element-node connectivity, shape (E, (order+1)**3)."""
    nx = order + 1
    T1, T2 = order * n1 + 1, order * n2 + 1
    e1, e2, e3 = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    base = (e1 * order + e2 * order * T1 + e3 * order * T1 * T2).transpose(2, 1, 0).reshape(-1)
    a, b, c = np.meshgrid(np.arange(nx), np.arange(nx), np.arange(nx), indexing="ij")
    local = (a + b * T1 + c * T1 * T2).transpose(2, 1, 0).reshape(-1)
    return (base[:, None] + local[None, :]).astype(np.int32)


def regular_coordinates(order, n1, n2, n3, jitter=0.0, seed=0):
    """Separated coordinates, shape (3, N)."""
    T = np.array([order * n1 + 1, order * n2 + 1, order * n3 + 1])
    N = int(T.prod())
    p = np.arange(N, dtype=np.int64)
    idx = np.stack([p % T[0], (p // T[0]) % T[1], p // (T[0] * T[1])])
    delta = 1.0 / (T - 1).astype(np.float64)
    xyz = 0.0 + idx.astype(np.float64) * delta[:, None]
    if jitter:
        h = 1.0 / np.array([n1, n2, n3], dtype=np.float64)
        interior = np.all((idx > 0) & (idx < (T - 1)[:, None]), axis=0)
        for c in range(3):
            u = uniform(seed, 3 * p + c)
            xyz[c] = np.where(interior, xyz[c] + (2.0 * u - 1.0) * jitter * h[c], xyz[c])
    return xyz


@dataclass
class RegularMesh:
    """A generated regular mesh of the unit cube.

    order 1: trilinear (Laplace); order 2: triquadratic geometry and
    displacement with a trilinear pressure component (``elem_nodes_l``).
    """
    order: int
    n: tuple
    elem_nodes: np.ndarray          # (E, 8 or 27), 0-based
    coords: np.ndarray              # (3, N)
    elem_nodes_l: np.ndarray = None  # (E, 8) pressure component, 0-based

    @classmethod
    def generate(cls, order, n, jitter=0.0, seed=0):
        n1, n2, n3 = (n, n, n) if np.isscalar(n) else tuple(n)
        conn = regular_elements(order, n1, n2, n3)
        xyz = regular_coordinates(order, n1, n2, n3, jitter, seed)
        conn_l = regular_elements(1, n1, n2, n3) if order == 2 else None
        return cls(order, (n1, n2, n3), conn, xyz, conn_l)

    @property
    def n_elements(self):
        return self.elem_nodes.shape[0]

    @property
    def n_nodes(self):
        return self.coords.shape[1]

    @property
    def n_nodes_l(self):
        n1, n2, n3 = self.n
        return (n1 + 1) * (n2 + 1) * (n3 + 1)

    def grid_index(self):
        """Integer grid position (i, j, k) of every node of the main component."""
        T1 = self.order * self.n[0] + 1
        T2 = self.order * self.n[1] + 1
        p = np.arange(self.n_nodes)
        return np.stack([p % T1, (p // T1) % T2, p // (T1 * T2)])

    def face_nodes(self, axis, side):
        """Nodes on the face x_axis = 0 (side 0) or = 1 (side 1)."""
        idx = self.grid_index()[axis]
        last = self.order * self.n[axis]
        return np.nonzero(idx == (0 if side == 0 else last))[0].astype(np.int32)
