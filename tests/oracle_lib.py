"""ctypes binding of the C oracle (oracle/liboc_oracle.so).

Test infrastructure only: the tests and the CPU-baseline leg of bench.py use
it as the checker.  The product package never imports this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB_PATH = os.path.join(ROOT, "oracle", "liboc_oracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def load():
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
    lib = ctypes.CDLL(_LIB_PATH)
    c_int, c_double, c_u64, c_i64 = ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_int64
    sig = {
        "oco_gauss_legendre": (c_int, [c_int, _dp, _dp]),
        "oco_basis_tables": (c_int, [c_int, c_int, _dp, _dp, _dp]),
        "oco_invert3": (c_int, [_dp, _dp, _dp]),
        "oco_regular_num_nodes": (c_i64, [c_int, c_int, c_int, c_int]),
        "oco_regular_elements": (c_int, [c_int, c_int, c_int, c_int, _i32]),
        "oco_regular_coordinates": (c_int, [c_int, c_int, c_int, c_int, c_double, c_u64, _dp]),
        "oco_regular_pressure_elements": (c_int, [c_int, c_int, c_int, _i32]),
        "oco_csr_add": (c_int, [_i64, _i32, _dp, c_int, _i32, c_int, _i32, _dp]),
        "oco_laplace_element": (c_int, [_dp, _dp]),
        "oco_laplace_assemble": (c_int, [c_int, _i32, c_int, _dp, _i64, _i32, _dp, _dp]),
        "oco_mr_stress": (c_int, [_dp, c_double, c_double, c_double, c_double, _dp]),
        "oco_mr_elasticity": (c_int, [_dp, c_double, c_double, c_double, c_double, _dp, _dp]),
        "oco_mr_element_residual": (c_int, [_dp, _dp, _dp, c_double, c_double, _dp]),
        "oco_mr_element_jacobian": (c_int, [_dp, _dp, _dp, c_double, c_double, _dp]),
        "oco_mr_element_jacobian_fd": (c_int, [_dp, _dp, _dp, c_double, c_double, c_double, _dp]),
        "oco_mr_residual": (c_int, [c_int, _i32, _i32, c_int, c_int, _dp, _dp, c_double, c_double, _dp]),
        "oco_imr_stress": (c_int, [_dp, c_double, c_double, c_double, _dp, ctypes.POINTER(c_double)]),
        "oco_imr_element_residual": (c_int, [_dp, _dp, _dp, c_double, c_double, _dp]),
        "oco_imr_element_jacobian_fd": (c_int, [_dp, _dp, _dp, c_double, c_double, c_double, _dp]),
        "oco_imr_residual": (c_int, [c_int, _i32, _i32, c_int, c_int, _dp, _dp, c_double, c_double, _dp]),
        "oco_mr_jacobian": (c_int, [c_int, _i32, _i32, c_int, c_int, _dp, _dp, c_double, c_double, c_int,
                                    _i64, _i32, _dp]),
        "oco_mr_assemble_parallel": (c_int, [c_int, _i32, _i32, c_int, c_int, _dp, _dp, c_double, c_double,
                                             _i64, _i32, _dp, _dp, c_int]),
        "oco_bc_increment": (c_int, [c_int, _i32, _dp, _dp, c_int, c_int]),
        "oco_uniform": (c_double, [c_u64, c_u64]),
        "oco_solver_system": (c_int, [c_int, _i64, _i32, _dp, _u8, _dp, _dp, c_int, ctypes.POINTER(c_i64),
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    lib.oco_pattern.restype = c_int
    lib.oco_pattern.argtypes = [c_int, c_int, _i32, _i32, ctypes.POINTER(ctypes.c_void_p),
                                ctypes.POINTER(c_i64), ctypes.c_void_p, ctypes.c_void_p]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def gauss_legendre(n):
    x, w = np.zeros(n), np.zeros(n)
    _check(lib().oco_gauss_legendre(n, x, w), "gauss_legendre")
    return x, w


def basis_tables(order, ng1):
    G, ns = ng1 ** 3, (order + 1) ** 3
    w, pos, fns = np.zeros(G), np.zeros(3 * G), np.zeros(G * ns * 4)
    _check(lib().oco_basis_tables(order, ng1, w, pos, fns), "basis_tables")
    return w, pos.reshape(G, 3), fns.reshape(G, ns, 4)


def regular_mesh(order, n1, n2, n3, jitter=0.0, seed=0):
    nen = (order + 1) ** 3
    E = n1 * n2 * n3
    N = int(lib().oco_regular_num_nodes(order, n1, n2, n3))
    conn = np.zeros(E * nen, np.int32)
    _check(lib().oco_regular_elements(order, n1, n2, n3, conn), "regular_elements")
    xyz = np.zeros(3 * N)
    _check(lib().oco_regular_coordinates(order, n1, n2, n3, float(jitter), int(seed), xyz), "coords")
    return conn.reshape(E, nen), xyz.reshape(3, N)


def pressure_elements(n1, n2, n3):
    E = n1 * n2 * n3
    conn = np.zeros(E * 8, np.int32)
    _check(lib().oco_regular_pressure_elements(n1, n2, n3, conn), "pressure_elements")
    return conn.reshape(E, 8)


def pattern(components):
    """components: list of (n_nodes, conn[E, nen]).  Returns 0-based (row_ptr, col_idx)."""
    ncomp = len(components)
    nelem = components[0][1].shape[0]
    nn = np.array([c[0] for c in components], np.int32)
    nen = np.array([c[1].shape[1] for c in components], np.int32)
    conns = [np.ascontiguousarray(c[1], dtype=np.int32).ravel() for c in components]
    ptrs = (ctypes.c_void_p * ncomp)(*[c.ctypes.data for c in conns])
    nnz = ctypes.c_int64(0)
    _check(lib().oco_pattern(ncomp, nelem, nn, nen, ptrs, ctypes.byref(nnz), None, None), "pattern")
    ndof = int(nn.sum())
    row_ptr = np.zeros(ndof + 1, np.int64)
    col_idx = np.zeros(nnz.value, np.int32)
    _check(lib().oco_pattern(ncomp, nelem, nn, nen, ptrs, ctypes.byref(nnz),
                             ctypes.c_void_p(row_ptr.ctypes.data), ctypes.c_void_p(col_idx.ctypes.data)),
           "pattern")
    return row_ptr, col_idx


def laplace_element(xe):
    K = np.zeros(64)
    _check(lib().oco_laplace_element(np.ascontiguousarray(xe, dtype=np.float64).ravel(), K), "laplace_element")
    return K.reshape(8, 8)


def laplace_assemble(conn, xyz, row_ptr, col_idx):
    E = conn.shape[0]
    N = xyz.shape[1]
    vals = np.zeros(col_idx.shape[0])
    rhs = np.zeros(N)
    _check(lib().oco_laplace_assemble(E, np.ascontiguousarray(conn, np.int32).ravel(), N,
                                      np.ascontiguousarray(xyz).ravel(), row_ptr, col_idx, vals, rhs),
           "laplace_assemble")
    return vals, rhs


def mr_stress(F, p, c10, c01):
    F = np.asfortranarray(F, dtype=np.float64)
    Fc = np.ascontiguousarray(F.T).ravel()  # column-major flat
    J = abs(np.linalg.det(F))
    s = np.zeros(6)
    _check(lib().oco_mr_stress(Fc, J, p, c10, c01, s), "mr_stress")
    return s


def imr_stress(F, p, c10, c01):
    """Cauchy stress (Voigt) and Jznu of the incompressible MR subtype."""
    Fc = np.ascontiguousarray(np.asarray(F, dtype=np.float64).T).ravel()
    s, J = np.zeros(6), ctypes.c_double(0.0)
    _check(lib().oco_imr_stress(Fc, p, c10, c01, s, ctypes.byref(J)), "imr_stress")
    return s, J.value


def imr_element_residual(Xe, ze, pe, c10, c01):
    R = np.zeros(89)
    _check(lib().oco_imr_element_residual(np.ascontiguousarray(Xe).ravel(), np.ascontiguousarray(ze).ravel(),
                                          np.ascontiguousarray(pe, dtype=np.float64), c10, c01, R), "imr_res")
    return R


def imr_residual(elem_q, elem_l, Xq, dofs, c10, c01):
    E = elem_q.shape[0]
    nq = Xq.shape[1]
    nl = dofs.shape[0] - 3 * nq
    res = np.zeros(dofs.shape[0])
    _check(lib().oco_imr_residual(E, np.ascontiguousarray(elem_q, np.int32).ravel(),
                                  np.ascontiguousarray(elem_l, np.int32).ravel(), nq, nl,
                                  np.ascontiguousarray(Xq).ravel(), np.ascontiguousarray(dofs), c10, c01, res),
           "imr_residual")
    return res


def mr_elasticity(F, p, c10, c01):
    Fc = np.ascontiguousarray(np.asarray(F, dtype=np.float64).T).ravel()
    J = abs(np.linalg.det(F))
    c, s = np.zeros(36), np.zeros(6)
    _check(lib().oco_mr_elasticity(Fc, J, p, c10, c01, c, s), "mr_elasticity")
    return c.reshape(6, 6).T, s  # back to (i, j) indexing


def mr_element_residual(Xe, ze, pe, c10, c01):
    R = np.zeros(89)
    _check(lib().oco_mr_element_residual(np.ascontiguousarray(Xe).ravel(), np.ascontiguousarray(ze).ravel(),
                                         np.ascontiguousarray(pe, dtype=np.float64), c10, c01, R), "mr_res")
    return R


def mr_element_jacobian(Xe, ze, pe, c10, c01):
    J = np.zeros(89 * 89)
    _check(lib().oco_mr_element_jacobian(np.ascontiguousarray(Xe).ravel(), np.ascontiguousarray(ze).ravel(),
                                         np.ascontiguousarray(pe, dtype=np.float64), c10, c01, J), "mr_jac")
    return J.reshape(89, 89)


def mr_element_jacobian_fd(Xe, ze, pe, c10, c01, delta):
    J = np.zeros(89 * 89)
    _check(lib().oco_mr_element_jacobian_fd(np.ascontiguousarray(Xe).ravel(), np.ascontiguousarray(ze).ravel(),
                                            np.ascontiguousarray(pe, dtype=np.float64), c10, c01, delta, J),
           "mr_jac_fd")
    return J.reshape(89, 89)


def mr_residual(elem_q, elem_l, Xq, dofs, c10, c01):
    E = elem_q.shape[0]
    nq = Xq.shape[1]
    nl = dofs.shape[0] - 3 * nq
    res = np.zeros(dofs.shape[0])
    _check(lib().oco_mr_residual(E, np.ascontiguousarray(elem_q, np.int32).ravel(),
                                 np.ascontiguousarray(elem_l, np.int32).ravel(), nq, nl,
                                 np.ascontiguousarray(Xq).ravel(), np.ascontiguousarray(dofs), c10, c01, res),
           "mr_residual")
    return res


def mr_jacobian(elem_q, elem_l, Xq, dofs, c10, c01, row_ptr, col_idx, mode=0):
    """mode 0: analytic MR Jacobian; 1: its forward differences; 2: forward
    differences of the incompressible MR subtype's residual."""
    E = elem_q.shape[0]
    nq = Xq.shape[1]
    nl = dofs.shape[0] - 3 * nq
    vals = np.zeros(col_idx.shape[0])
    _check(lib().oco_mr_jacobian(E, np.ascontiguousarray(elem_q, np.int32).ravel(),
                                 np.ascontiguousarray(elem_l, np.int32).ravel(), nq, nl,
                                 np.ascontiguousarray(Xq).ravel(), np.ascontiguousarray(dofs), c10, c01, mode,
                                 row_ptr, col_idx, vals), "mr_jacobian")
    return vals


def mr_assemble_parallel(elem_q, elem_l, Xq, dofs, c10, c01, row_ptr, col_idx, threads):
    """Residual + analytic Jacobian over element colours on `threads` threads."""
    E = elem_q.shape[0]
    nq = Xq.shape[1]
    nl = dofs.shape[0] - 3 * nq
    vals = np.zeros(col_idx.shape[0])
    res = np.zeros(dofs.shape[0])
    _check(lib().oco_mr_assemble_parallel(E, np.ascontiguousarray(elem_q, np.int32).ravel(),
                                          np.ascontiguousarray(elem_l, np.int32).ravel(), nq, nl,
                                          np.ascontiguousarray(Xq).ravel(), np.ascontiguousarray(dofs), c10, c01,
                                          row_ptr, col_idx, vals, res, int(threads)), "mr_assemble_parallel")
    return res, vals


def bc_increment(dof_idx, full, current, iteration, max_iterations):
    cur = np.array(current, dtype=np.float64, copy=True)
    _check(lib().oco_bc_increment(len(dof_idx), np.ascontiguousarray(dof_idx, np.int32),
                                  np.ascontiguousarray(full, np.float64), cur, iteration, max_iterations),
           "bc_increment")
    return cur


def solver_system(row_ptr, col_idx, values, free_mask, vec, dofs, linear):
    """Solver matrix (CSR, int64 offsets) and vector of one equations set:
    SOLVER_MATRICES_STATIC_ASSEMBLE restated (oracle/oc_oracle.c)."""
    row_ptr = np.ascontiguousarray(row_ptr, np.int64)
    col_idx = np.ascontiguousarray(col_idx, np.int32)
    values = np.ascontiguousarray(values, np.float64)
    mask = np.ascontiguousarray(free_mask, np.uint8)
    vec = np.ascontiguousarray(vec, np.float64)
    dofs = np.ascontiguousarray(dofs, np.float64)
    n_rows = len(row_ptr) - 1
    nnz = ctypes.c_int64()
    n = lib().oco_solver_system(n_rows, row_ptr, col_idx, values, mask, vec, dofs, int(linear), ctypes.byref(nnz),
                                None, None, None, None)
    _check(0 if n >= 0 else n, "solver_system")
    rp = np.zeros(n + 1, np.int64)
    ci = np.zeros(max(nnz.value, 1), np.int32)
    sv = np.zeros(max(nnz.value, 1))
    sb = np.zeros(max(n, 1))
    rc = lib().oco_solver_system(n_rows, row_ptr, col_idx, values, mask, vec, dofs, int(linear), ctypes.byref(nnz),
                                 rp.ctypes.data, ci.ctypes.data, sv.ctypes.data, sb.ctypes.data)
    _check(0 if rc >= 0 else rc, "solver_system")
    return rp, ci[:nnz.value], sv[:nnz.value], sb[:n]
