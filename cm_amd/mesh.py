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


@dataclass
class UserMesh:
    """A mesh given by user-numbered nodes and elements, numbered as the
    reference numbers a region's meshes.

    * Region nodes (cmfe_Nodes) have global numbers in the order they are
      given; their user numbers may be any distinct integers.
    * Element nodes are given by user number in the basis' local node order
      (cmfe_MeshElements_NodesSet) and translated to global numbers.
    * A mesh component's nodes are the global nodes its elements use, sorted
      ascending, and its local node number is the rank in that list
      (MeshTopologyNodesCalculate, mesh_routines.f90:9006-9119).
    * DOFs are separated by component over those local nodes
      (FIELD_MAPPINGS_CALCULATE, field_routines.f90:10345-10628), which is the
      layout the equations sets take.

    ``order`` 1: trilinear (Laplace); 2: triquadratic geometry and displacement,
    with the trilinear pressure component's elements in ``elem_nodes_l``.
    """
    order: int
    elem_nodes: np.ndarray          # (E, 8 or 27), local node numbers of the main component
    coords: np.ndarray              # (3, N) at the main component's nodes, local order
    node_user_numbers: np.ndarray   # (N,) user number of each local node
    elem_nodes_l: np.ndarray = None  # (E, 8) local numbers of the pressure component
    node_user_numbers_l: np.ndarray = None
    element_user_numbers: np.ndarray = None

    @classmethod
    def from_user_numbers(cls, order, node_user_numbers, coords, element_nodes, element_nodes_l=None,
                          element_user_numbers=None):
        """node_user_numbers (N_region,), coords (3, N_region) in the same
        order, element_nodes (E, (order+1)^3) and, for order 2,
        element_nodes_l (E, 8): user node numbers."""
        user = np.asarray(node_user_numbers, np.int64)
        xyz = np.asarray(coords, np.float64)
        if user.ndim != 1 or xyz.shape != (3, user.size):
            raise ValueError("coords must have shape (3, len(node_user_numbers))")
        if np.unique(user).size != user.size:
            raise ValueError("node user numbers must be distinct")
        order_of = np.argsort(user, kind="stable")
        sorted_user = user[order_of]

        def to_global(e_user, what):
            e_user = np.asarray(e_user, np.int64)
            pos = np.searchsorted(sorted_user, e_user)
            pos = np.minimum(pos, sorted_user.size - 1)
            if not np.array_equal(sorted_user[pos], e_user):
                bad = e_user[sorted_user[pos] != e_user].ravel()[0]
                raise ValueError(f"{what}: user node {int(bad)} is not a region node")
            return order_of[pos]  # 0-based global numbers

        def component(e_user, nen, what):
            e_user = np.asarray(e_user)
            if e_user.ndim != 2 or e_user.shape[1] != nen:
                raise ValueError(f"{what} must have shape (E, {nen})")
            g = to_global(e_user, what)
            used = np.unique(g)            # sorted global numbers of the component
            local = np.searchsorted(used, g).astype(np.int32)
            return used, local

        nen = (order + 1) ** 3
        used, local = component(element_nodes, nen, "element_nodes")
        mesh = cls(order, local, np.ascontiguousarray(xyz[:, used]), user[used])
        if order == 2:
            if element_nodes_l is None:
                raise ValueError("a triquadratic mesh needs the trilinear pressure component's element nodes")
            used_l, local_l = component(element_nodes_l, 8, "element_nodes_l")
            if local_l.shape[0] != local.shape[0]:
                raise ValueError("both components must have the same elements")
            mesh.elem_nodes_l = local_l
            mesh.node_user_numbers_l = user[used_l]
        n_el = mesh.elem_nodes.shape[0]
        mesh.element_user_numbers = (np.arange(1, n_el + 1) if element_user_numbers is None
                                     else np.asarray(element_user_numbers, np.int64))
        return mesh

    @property
    def n_elements(self):
        return self.elem_nodes.shape[0]

    @property
    def n_nodes(self):
        return self.coords.shape[1]

    @property
    def n_nodes_l(self):
        return 0 if self.node_user_numbers_l is None else self.node_user_numbers_l.size

    def face_nodes(self, axis, side, tol=1e-12):
        """Nodes of the main component on the plane x_axis = min (side 0) or
        max (side 1) of the mesh."""
        x = self.coords[axis]
        target = x.min() if side == 0 else x.max()
        return np.nonzero(np.abs(x - target) <= tol * max(1.0, abs(target)))[0].astype(np.int32)


def as_user_mesh(mesh, region_order=None, user_numbers=None):
    """A RegularMesh restated as user-numbered input: the region's nodes in
    ``region_order`` (a permutation of the mesh nodes; default identity) with
    ``user_numbers`` (default 1-based global), elements by user number."""
    n = mesh.n_nodes
    region_order = np.arange(n) if region_order is None else np.asarray(region_order)
    user = np.arange(1, n + 1) if user_numbers is None else np.asarray(user_numbers)
    user_of_node = np.empty(n, np.int64)
    user_of_node[region_order] = user
    e_user = user_of_node[mesh.elem_nodes]
    e_user_l = None
    if mesh.order == 2:
        # corner (i, j, k) of the trilinear component is quadratic grid node (2i, 2j, 2k)
        n1, n2, n3 = mesh.n
        T1, T2 = 2 * n1 + 1, 2 * n2 + 1
        i = np.arange(mesh.n_nodes_l)
        ci, cj, ck = i % (n1 + 1), (i // (n1 + 1)) % (n2 + 1), i // ((n1 + 1) * (n2 + 1))
        quad_of_corner = 2 * ci + 2 * cj * T1 + 2 * ck * T1 * T2
        e_user_l = user_of_node[quad_of_corner[mesh.elem_nodes_l]]
    return UserMesh.from_user_numbers(mesh.order, user, mesh.coords[:, region_order], e_user, e_user_l)
