"""Host-side equations sets: the Python face of the assembly C ABI.

Mirrors the reference's equations-set operator interface for this path:
``assemble`` (EQUATIONS_SET_ASSEMBLE, equations_set_routines.f90:981),
``residual_evaluate`` (EQUATIONS_SET_RESIDUAL_EVALUATE, :5184) and
``jacobian_evaluate`` (EQUATIONS_SET_JACOBIAN_EVALUATE, :4694), with the
matrix accessors of cmfe_Equations_LinearMatrixGet / JacobianMatrixGet
(opencmiss_iron.f90:22861, 23233) and the CSR storage-location / data getters
(:44843, :44920).  Every call goes to the HIP library; a failing status
raises CMISSError.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import call

_initialised_device = None


def initialise(device=0):
    """Select the HIP device (gfx950) and create the library stream."""
    global _initialised_device
    call("oc_initialise", int(device))
    _initialised_device = int(device)


def finalise():
    global _initialised_device
    call("oc_finalise")
    _initialised_device = None


def synchronise():
    call("oc_synchronise")


def event_record(slot):
    call("oc_event_record", int(slot))


def event_elapsed(start, end):
    ms = ctypes.c_float(0.0)
    call("oc_event_elapsed", int(start), int(end), ctypes.byref(ms))
    return float(ms.value)


def kernel_timing_enable(enable=True, only=None):
    """Bracket library kernel launches with events (clears totals): every
    kernel, or only the launches of kernel name `only`."""
    call("oc_kernel_timing_select", only.encode() if only else None)
    call("oc_kernel_timing_enable", 1 if enable else 0)


def kernel_timing(name):
    """(total milliseconds, launches) of a library kernel since the last enable."""
    ms, n = ctypes.c_double(0.0), ctypes.c_longlong(0)
    call("oc_kernel_timing_get", name.encode(), ctypes.byref(ms), ctypes.byref(n))
    return float(ms.value), int(n.value)


class EquationsSet:
    """Common handle management for the two equations-set classes."""

    def __init__(self, handle):
        self._h = handle
        cls, nrows, nnz, ne, ncol = (ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int(),
                                     ctypes.c_int())
        call("oc_equations_set_info", self._h, ctypes.byref(cls), ctypes.byref(nrows), ctypes.byref(nnz),
             ctypes.byref(ne), ctypes.byref(ncol))
        self.set_class = cls.value
        self.n_rows = nrows.value
        self.nnz = nnz.value
        self.n_elements = ne.value
        self.n_colours = ncol.value

    def destroy(self):
        if self._h:
            call("oc_equations_set_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            if self._h and _lib._LIB is not None:
                _lib._LIB.oc_equations_set_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- structure --------------------------------------------------------
    def matrix_structure(self, index_base=0):
        row_ptr = np.zeros(self.n_rows + 1, np.int32)
        col_idx = np.zeros(self.nnz, np.int32)
        call("oc_matrix_structure_get", self._h, int(index_base), row_ptr, col_idx)
        return row_ptr, col_idx

    # -- dependent field --------------------------------------------------
    def set_dofs(self, dofs):
        dofs = np.ascontiguousarray(dofs, dtype=np.float64)
        if dofs.shape != (self.n_rows,):
            raise ValueError(f"dofs must have shape ({self.n_rows},)")
        call("oc_dofs_set", self._h, dofs)

    def get_dofs(self):
        out = np.zeros(self.n_rows)
        call("oc_dofs_get", self._h, out)
        return out

    # -- results ----------------------------------------------------------
    def matrix_values(self):
        out = np.zeros(self.nnz)
        call("oc_matrix_values_get", self._h, out)
        return out

    def matrix_values_sp(self):
        out = np.zeros(self.nnz, np.float32)
        call("oc_matrix_values_get_sp", self._h, out)
        return out

    def residual(self):
        out = np.zeros(self.n_rows)
        call("oc_residual_get", self._h, out)
        return out

    def rhs(self):
        out = np.zeros(self.n_rows)
        call("oc_rhs_get", self._h, out)
        return out

    def device_pointer(self, buffer):
        p = ctypes.c_void_p()
        call("oc_device_pointer_get", self._h, int(buffer), ctypes.byref(p))
        return p.value

    # -- solver matrices ----------------------------------------------------
    def solver_matrices_assemble(self, precision=_lib.OC_FP64):
        """Solver matrix and vector from the last assembled matrix: the rows and
        columns of fixed DOFs removed, free DOFs numbered in ascending order
        (SOLVER_MATRICES_STATIC_ASSEMBLE, solver_routines.f90:13745-14771)."""
        call("oc_solver_matrices_assemble", self._h, int(precision))

    def solver_matrix_info(self):
        n, nnz = ctypes.c_int(), ctypes.c_longlong()
        call("oc_solver_matrix_info", self._h, ctypes.byref(n), ctypes.byref(nnz))
        return n.value, nnz.value

    def solver_matrix_structure(self, index_base=0):
        n, nnz = self.solver_matrix_info()
        row_ptr = np.zeros(n + 1, np.int32)
        col_idx = np.zeros(nnz, np.int32)
        call("oc_solver_matrix_structure_get", self._h, int(index_base), row_ptr, col_idx)
        return row_ptr, col_idx

    def solver_matrix_values(self):
        out = np.zeros(self.solver_matrix_info()[1])
        call("oc_solver_matrix_values_get", self._h, out)
        return out

    def solver_matrix_values_sp(self):
        out = np.zeros(self.solver_matrix_info()[1], np.float32)
        call("oc_solver_matrix_values_get_sp", self._h, out)
        return out

    def solver_vector(self):
        out = np.zeros(self.solver_matrix_info()[0])
        call("oc_solver_vector_get", self._h, out)
        return out

    # -- boundary conditions and solution ------------------------------------
    def set_boundary_conditions(self, dof_indices, full_values):
        """Fixed(-incremented) Dirichlet DOFs (0-based) and their full values."""
        idx = np.ascontiguousarray(dof_indices, dtype=np.int32)
        val = np.ascontiguousarray(full_values, dtype=np.float64)
        if idx.shape != val.shape:
            raise ValueError("dof_indices and full_values must have the same shape")
        call("oc_boundary_conditions_set", self._h, int(idx.size), idx, val, 0)

    def load_increment_apply(self, iteration, max_iterations):
        call("oc_load_increment_apply", self._h, int(iteration), int(max_iterations))


def newton_options(**overrides):
    """oc_newton_options with the reference defaults, then ``overrides``."""
    o = _lib.NewtonOptions()
    call("oc_newton_options_default", ctypes.byref(o))
    for k, v in overrides.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class LaplaceEquationsSet(EquationsSet):
    """Classical field / Laplace / standard subtype on trilinear hexahedra."""

    def __init__(self, mesh):
        if mesh.order != 1:
            raise ValueError("the Laplace equations set uses a trilinear mesh")
        h = ctypes.c_void_p()
        coords = np.ascontiguousarray(mesh.coords, dtype=np.float64).ravel()
        conn = np.ascontiguousarray(mesh.elem_nodes, dtype=np.int32).ravel()
        call("oc_laplace_create", mesh.n_nodes, coords, mesh.n_elements, conn, 0, ctypes.byref(h))
        super().__init__(h)
        self.mesh = mesh

    def assemble(self):
        """Zero and assemble the linear matrix and RHS vector on the GPU."""
        call("oc_equations_set_assemble", self._h)

    def set_fused(self, fused=1):
        """Assembly form: 1 (the default) fused one-pass assembly of a
        structured grid large enough to fill the GPU, 2 fused for any
        structured grid, 0 the element-matrix buffer + row gather; returns
        whether the fused pass now assembles this set."""
        active = ctypes.c_int(0)
        call("oc_laplace_fused_set", self._h, int(fused), ctypes.byref(active))
        return bool(active.value)

    def solve(self, options=None, raise_on_failure=True):
        """Static linear solve with the current Dirichlet conditions."""
        opts = options if options is not None else newton_options()
        st = _lib.NewtonStats()
        rc = _lib.load().oc_linear_solve(self._h, ctypes.byref(opts), ctypes.byref(st))
        if rc != 0 and (raise_on_failure or rc != _lib.OC_ERR_NOT_CONVERGED):
            _lib.check(rc)
        return st.as_dict()


class MooneyRivlinEquationsSet(EquationsSet):
    """Elasticity / finite elasticity / Mooney-Rivlin subtype:
    triquadratic deformed coordinates plus trilinear hydrostatic pressure."""

    def __init__(self, mesh, c10, c01):
        if mesh.order != 2:
            raise ValueError("the Mooney-Rivlin equations set uses a triquadratic mesh")
        h = ctypes.c_void_p()
        coords = np.ascontiguousarray(mesh.coords, dtype=np.float64).ravel()
        cq = np.ascontiguousarray(mesh.elem_nodes, dtype=np.int32).ravel()
        cl = np.ascontiguousarray(mesh.elem_nodes_l, dtype=np.int32).ravel()
        call("oc_mooney_rivlin_create", mesh.n_nodes, coords, mesh.n_nodes_l, mesh.n_elements, cq, cl, 0,
             float(c10), float(c01), ctypes.byref(h))
        super().__init__(h)
        self.mesh = mesh
        self.c10, self.c01 = float(c10), float(c01)

    def set_materials(self, c10, c01):
        call("oc_materials_set", self._h, float(c10), float(c01))
        self.c10, self.c01 = float(c10), float(c01)

    def set_subtype(self, subtype):
        """OC_MOONEY_RIVLIN_SUBTYPE (analytic Jacobian) or
        OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE (2nd Piola form, forward-
        difference FP64 Jacobian)."""
        call("oc_equations_set_subtype_set", self._h, int(subtype))
        self.subtype = int(subtype)

    def undeformed_dofs(self):
        """z = X and p = 0 (the initial dependent field)."""
        nq = self.mesh.n_nodes
        d = np.zeros(self.n_rows)
        d[:3 * nq] = self.mesh.coords.ravel()
        return d

    def residual_evaluate(self):
        call("oc_equations_set_residual_evaluate", self._h)

    def jacobian_evaluate(self, precision=_lib.OC_FP64):
        call("oc_equations_set_jacobian_evaluate", self._h, int(precision))

    def set_multigrid(self, pre_sweeps=1, post_sweeps=2, omega=0.8):
        """Declare the mesh(es) regular n^3 generated meshes and build the
        geometric multigrid hierarchy (OC_PRECONDITIONER_MULTIGRID); the sweep
        counts are the triquadratic level's."""
        meshes = getattr(self, "meshes", None) or [self.mesh]
        ns = []
        for m in meshes:
            if len(set(m.n)) != 1:
                raise ValueError("multigrid needs n x n x n meshes")
            ns.append(m.n[0])
        call("oc_multigrid_structure_set", self._h, np.array(ns, np.int32))
        call("oc_multigrid_options_set", self._h, int(pre_sweeps), int(post_sweeps), float(omega))

    def jacobian_evaluate_fd(self):
        """Forward-difference Jacobian (FP64 values), the reference's rule."""
        call("oc_equations_set_jacobian_evaluate_fd", self._h)

    def jacobian_evaluate_graph(self, precision=_lib.OC_FP64):
        """Jacobian assembly replayed from the set's recorded HIP graph."""
        call("oc_jacobian_evaluate_graph", self._h, int(precision))

    def autotune(self, precision=_lib.OC_FP64):
        """Times the launch-shape candidates on this set, keeps the fastest;
        returns the timings (dict)."""
        import json
        buf = ctypes.create_string_buffer(4096)
        call("oc_autotune", self._h, int(precision), buf, 4096)
        return json.loads(buf.value.decode())

    def residual_norm(self):
        """||R|| over the free DOFs at the current dependent field."""
        v = ctypes.c_double(0.0)
        call("oc_residual_norm", self._h, ctypes.byref(v))
        return v.value

    def newton_solve(self, options=None, raise_on_failure=True):
        """Newton solve at the current Dirichlet values; returns the stats dict."""
        opts = options if options is not None else newton_options()
        st = _lib.NewtonStats()
        rc = _lib.load().oc_newton_solve(self._h, ctypes.byref(opts), ctypes.byref(st))
        if rc != 0 and (raise_on_failure or rc != _lib.OC_ERR_NOT_CONVERGED):
            _lib.check(rc)
        return st.as_dict()

    def uniaxial_boundary_conditions(self, stretch):
        """Symmetry planes z_x = 0 on x = 0, z_y = 0 on y = 0, z_z = 0 on z = 0
        and z_x -> stretch on x = 1 (fixed-incremented)."""
        m = self.mesh
        nq = m.n_nodes
        idx = [m.face_nodes(0, 0), nq + m.face_nodes(1, 0), 2 * nq + m.face_nodes(2, 0), m.face_nodes(0, 1)]
        vals = [np.zeros(len(idx[0])), np.zeros(len(idx[1])), np.zeros(len(idx[2])),
                np.full(len(idx[3]), float(stretch))]
        self.set_boundary_conditions(np.concatenate(idx), np.concatenate(vals))
        return np.concatenate(idx)


class MooneyRivlinBatch(MooneyRivlinEquationsSet):
    """Independent Mooney-Rivlin meshes of uneven size as one equations set
    (oc_mooney_rivlin_batch_create): one launch assembles every element of
    every mesh; DOFs are component-major over the union of the meshes."""

    def __init__(self, meshes, c10, c01):
        for m in meshes:
            if m.order != 2:
                raise ValueError("the Mooney-Rivlin equations set uses triquadratic meshes")
        k = len(meshes)
        c10 = np.ascontiguousarray(np.broadcast_to(np.asarray(c10, np.float64), (k,)))
        c01 = np.ascontiguousarray(np.broadcast_to(np.asarray(c01, np.float64), (k,)))
        nq = np.array([m.n_nodes for m in meshes], np.int32)
        nl = np.array([m.n_nodes_l for m in meshes], np.int32)
        ne = np.array([m.n_elements for m in meshes], np.int32)
        coords = np.concatenate([np.ascontiguousarray(m.coords, np.float64).ravel() for m in meshes])
        cq = np.concatenate([np.ascontiguousarray(m.elem_nodes, np.int32).ravel() for m in meshes])
        cl = np.concatenate([np.ascontiguousarray(m.elem_nodes_l, np.int32).ravel() for m in meshes])
        h = ctypes.c_void_p()
        call("oc_mooney_rivlin_batch_create", k, nq, nl, ne, coords, cq, cl, c10, c01, 0, ctypes.byref(h))
        EquationsSet.__init__(self, h)
        self.meshes = list(meshes)
        self.mesh = None
        self.c10s, self.c01s = c10.copy(), c01.copy()
        self.c10, self.c01 = float(c10[0]), float(c01[0])
        self.q_off = np.concatenate([[0], np.cumsum(nq)]).astype(np.int64)
        self.l_off = np.concatenate([[0], np.cumsum(nl)]).astype(np.int64)
        self.e_off = np.concatenate([[0], np.cumsum(ne)]).astype(np.int64)
        self.n_nodes_total = int(self.q_off[-1])

    @property
    def n_meshes(self):
        return len(self.meshes)

    def mesh_dofs(self, m):
        """Union DOF of each local DOF (z1 nodes, z2, z3, p) of mesh m."""
        NQ = self.n_nodes_total
        q0, q1, l0, l1 = self.q_off[m], self.q_off[m + 1], self.l_off[m], self.l_off[m + 1]
        return np.concatenate([c * NQ + np.arange(q0, q1) for c in range(3)] + [3 * NQ + np.arange(l0, l1)])

    def mesh_matrix(self, m, precision=_lib.OC_FP64, index_base=0):
        """(row_ptr, col_idx, values) of mesh m's block of the last Jacobian of
        `precision`, in the DOF order of a single-mesh set."""
        n, nnz = ctypes.c_int(), ctypes.c_longlong()
        call("oc_mesh_matrix_get", self._h, int(m), int(precision), int(index_base), ctypes.byref(n),
             ctypes.byref(nnz), None, None, None)
        rp = np.zeros(n.value + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        vals = np.zeros(nnz.value, np.float64 if precision == _lib.OC_FP64 else np.float32)
        call("oc_mesh_matrix_get", self._h, int(m), int(precision), int(index_base), ctypes.byref(n),
             ctypes.byref(nnz), rp.ctypes.data, ci.ctypes.data, vals.ctypes.data)
        return rp, ci, vals

    def set_materials(self, c10, c01):
        c10 = np.ascontiguousarray(np.broadcast_to(np.asarray(c10, np.float64), (self.n_meshes,)))
        c01 = np.ascontiguousarray(np.broadcast_to(np.asarray(c01, np.float64), (self.n_meshes,)))
        call("oc_mesh_materials_set", self._h, c10, c01)
        self.c10s, self.c01s = c10.copy(), c01.copy()

    def undeformed_dofs(self):
        d = np.zeros(self.n_rows)
        for m, mesh in enumerate(self.meshes):
            idx = self.mesh_dofs(m)[:3 * mesh.n_nodes]
            d[idx] = mesh.coords.ravel()
        return d

    def uniaxial_boundary_conditions(self, stretches):
        """The uniaxial conditions of every mesh (stretch per mesh); returns
        the union's fixed DOFs."""
        stretches = np.broadcast_to(np.asarray(stretches, np.float64), (self.n_meshes,))
        idx, vals = [], []
        for m, mesh in enumerate(self.meshes):
            nq = mesh.n_nodes
            loc = [mesh.face_nodes(0, 0), nq + mesh.face_nodes(1, 0), 2 * nq + mesh.face_nodes(2, 0),
                   mesh.face_nodes(0, 1)]
            v = [np.zeros(len(loc[0])), np.zeros(len(loc[1])), np.zeros(len(loc[2])),
                 np.full(len(loc[3]), float(stretches[m]))]
            idx.append(self.mesh_dofs(m)[np.concatenate(loc)])
            vals.append(np.concatenate(v))
        idx, vals = np.concatenate(idx), np.concatenate(vals)
        self.set_boundary_conditions(idx, vals)
        return idx

    def mesh_residual_norms(self):
        out = np.zeros(self.n_meshes)
        call("oc_mesh_residual_norms", self._h, out)
        return out

    def newton_solve(self, options=None, reference_norms=None, raise_on_failure=True):
        """Newton solve of every mesh; returns (stats dict, per-mesh final ||F||)."""
        opts = options if options is not None else newton_options()
        st = _lib.NewtonStats()
        finals = np.zeros(self.n_meshes)
        refs = None
        if reference_norms is not None:
            refs = np.ascontiguousarray(reference_norms, np.float64)
            if refs.shape != (self.n_meshes,):
                raise ValueError(f"reference_norms must have shape ({self.n_meshes},)")
        rc = _lib.load().oc_newton_solve_batch(self._h, ctypes.byref(opts),
                                               None if refs is None else refs.ctypes.data, finals.ctypes.data,
                                               ctypes.byref(st))
        if rc != 0 and (raise_on_failure or rc != _lib.OC_ERR_NOT_CONVERGED):
            _lib.check(rc)
        return st.as_dict(), finals
