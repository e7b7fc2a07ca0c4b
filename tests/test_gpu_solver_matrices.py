"""GPU solver matrices (the equations matrix or Jacobian restricted to the
free DOFs, SOLVER_MATRICES_STATIC_ASSEMBLE) against the C oracle.

Bars: structure bit-exact; values bit-exact against the oracle applied to the
same equations values (each solver entry is one equations entry added to
zero); the linear solver vector bit-exact (the same products and adds in the
same order); against the oracle's own assembly the FP64 tolerance of the
assembly tests (1e-12 relative Frobenius).
"""
import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _mask(n, fixed):
    m = np.ones(n, np.uint8)
    m[np.asarray(fixed, np.int64)] = 0
    return m


def _check_against(eq, rp, ci, vals, mask, vec, dofs, linear):
    orp, oci, ov, ob = orc.solver_system(rp.astype(np.int64), ci, vals, mask, vec, dofs, linear)
    n, nnz = eq.solver_matrix_info()
    assert n == len(orp) - 1 and nnz == len(oci)
    srp, sci = eq.solver_matrix_structure(0)
    assert np.array_equal(srp.astype(np.int64), orp) and np.array_equal(sci, oci)
    srp1, sci1 = eq.solver_matrix_structure(1)
    assert np.array_equal(srp1, srp + 1) and np.array_equal(sci1, sci + 1)
    return ov, ob


@pytest.mark.parametrize("n,jitter", [((8, 8, 8), 0.0), ((5, 3, 4), 0.2)])
def test_laplace_solver_system(gpu, n, jitter):
    """Config-1 Dirichlet problem (u = 0 on x = 0, u = 1 on x = 1): reduced
    matrix and lifted right-hand side, and the patch solution u = x."""
    cm = gpu
    mesh = cm.RegularMesh.generate(1, n, jitter=jitter, seed=9)
    eq = cm.LaplaceEquationsSet(mesh)
    f0, f1 = mesh.face_nodes(0, 0), mesh.face_nodes(0, 1)
    fixed = np.concatenate([f0, f1])
    eq.set_boundary_conditions(fixed, np.concatenate([np.zeros(len(f0)), np.ones(len(f1))]))
    eq.load_increment_apply(1, 1)  # the fixed DOFs take their values
    eq.assemble()
    eq.solver_matrices_assemble(cm.OC_FP64)
    rp, ci = eq.matrix_structure(0)
    vals, rhs, dofs = eq.matrix_values(), eq.rhs(), eq.get_dofs()
    mask = _mask(mesh.n_nodes, fixed)
    ov, ob = _check_against(eq, rp, ci, vals, mask, rhs, dofs, True)
    sv, sb = eq.solver_matrix_values(), eq.solver_vector()
    assert np.array_equal(sv, ov)
    assert np.array_equal(sb, ob)
    # against the oracle's own assembly
    orp, oci = orc.pattern([(mesh.n_nodes, mesh.elem_nodes)])
    ovals, orhs = orc.laplace_assemble(mesh.elem_nodes, mesh.coords, orp, oci)
    _, _, ov2, ob2 = orc.solver_system(orp, oci, ovals, mask, orhs, dofs, True)
    assert rel(sv, ov2) < FP64_TOL and rel(sb, ob2) < FP64_TOL
    # the reduced system is solved by the x coordinates of the free nodes
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    srp, sci = eq.solver_matrix_structure(0)
    S = sp.csr_matrix((sv, sci, srp), shape=(len(srp) - 1,) * 2)
    u = sla.spsolve(S.tocsc(), sb)
    assert np.abs(u - mesh.coords[0][mask.astype(bool)]).max() < 1e-10
    # the device buffers hold the same arrays
    for buf in (cm.OC_BUFFER_SOLVER_VALUES, cm.OC_BUFFER_SOLVER_ROW_PTR, cm.OC_BUFFER_SOLVER_COL_IDX,
                cm.OC_BUFFER_SOLVER_VECTOR):
        assert eq.device_pointer(buf)
    eq.destroy()


def test_mooney_rivlin_solver_system(gpu):
    """Uniaxial conditions at a deformed state: reduced Jacobian (FP64 and FP32)
    and the residual of the free rows."""
    cm = gpu
    mesh = cm.RegularMesh.generate(2, (3, 2, 2), jitter=0.1, seed=4)
    c10, c01 = 2.0, 6.0
    eq = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    fixed = eq.uniaxial_boundary_conditions(1.05)
    rng = np.random.default_rng(8)
    nq, nl = mesh.n_nodes, mesh.n_nodes_l
    d = np.concatenate([mesh.coords.ravel(), np.zeros(nl)])
    d[:3 * nq] += 0.02 * rng.uniform(-1, 1, 3 * nq)
    d[3 * nq:] = rng.standard_normal(nl)
    eq.set_dofs(d)
    eq.residual_evaluate()
    eq.jacobian_evaluate(cm.OC_FP64)
    eq.solver_matrices_assemble(cm.OC_FP64)
    rp, ci = eq.matrix_structure(0)
    mask = _mask(eq.n_rows, fixed)
    vals, res = eq.matrix_values(), eq.residual()
    ov, ob = _check_against(eq, rp, ci, vals, mask, res, d, False)
    assert np.array_equal(eq.solver_matrix_values(), ov)
    assert np.array_equal(eq.solver_vector(), ob)
    oJ = orc.mr_jacobian(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, c10, c01, rp.astype(np.int64), ci)
    oR = orc.mr_residual(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, c10, c01)
    _, _, ov2, ob2 = orc.solver_system(rp.astype(np.int64), ci, oJ, mask, oR, d, False)
    assert rel(eq.solver_matrix_values(), ov2) < FP64_TOL and rel(eq.solver_vector(), ob2) < FP64_TOL
    # FP32 Jacobian: the same restriction of the float values
    eq.jacobian_evaluate(cm.OC_FP32)
    eq.solver_matrices_assemble(cm.OC_FP32)
    sp32 = eq.solver_matrix_values_sp()
    keep = np.concatenate([np.flatnonzero(mask[ci[rp[r]:rp[r + 1]]]) + rp[r] for r in np.flatnonzero(mask)])
    assert np.array_equal(sp32, eq.matrix_values_sp()[keep])
    eq.destroy()


def test_solver_system_follows_boundary_conditions(gpu):
    """A new set of fixed DOFs rebuilds the structure; no fixed DOFs gives the
    equations system itself; getters before an assembly fail loudly."""
    cm = gpu
    mesh = cm.RegularMesh.generate(1, (3, 3, 2), jitter=0.1, seed=2)
    eq = cm.LaplaceEquationsSet(mesh)
    eq.assemble()
    with pytest.raises(cm.CMISSError):
        eq.solver_matrix_values()
    eq.solver_matrices_assemble(cm.OC_FP64)
    rp, ci = eq.matrix_structure(0)
    srp, sci = eq.solver_matrix_structure(0)
    assert np.array_equal(srp, rp) and np.array_equal(sci, ci)
    assert np.array_equal(eq.solver_matrix_values(), eq.matrix_values())
    assert np.array_equal(eq.solver_vector(), eq.rhs())
    rng = np.random.default_rng(3)
    for rep in range(2):
        fixed = np.sort(rng.choice(mesh.n_nodes, size=7 + 5 * rep, replace=False))
        eq.set_boundary_conditions(fixed, rng.standard_normal(fixed.size))
        with pytest.raises(cm.CMISSError):
            eq.solver_matrix_info()  # structure is stale until the next assembly
        eq.load_increment_apply(1, 1)
        eq.solver_matrices_assemble(cm.OC_FP64)
        mask = _mask(mesh.n_nodes, fixed)
        ov, ob = _check_against(eq, rp, ci, eq.matrix_values(), mask, eq.rhs(), eq.get_dofs(), True)
        assert np.array_equal(eq.solver_matrix_values(), ov)
        assert np.array_equal(eq.solver_vector(), ob)
    with pytest.raises(cm.CMISSError):
        eq.solver_matrices_assemble(cm.OC_FP32)  # a linear matrix is FP64 only
    eq.destroy()


def test_batch_solver_system(gpu):
    """Two meshes of a batch: the union's reduced Jacobian and residual."""
    cm = gpu
    meshes = [cm.RegularMesh.generate(2, (2, 2, 1), jitter=0.1, seed=11),
              cm.RegularMesh.generate(2, (1, 2, 2), jitter=0.1, seed=12)]
    batch = cm.MooneyRivlinBatch(meshes, [2.0, 1.0], [6.0, 3.0])
    fixed = batch.uniaxial_boundary_conditions([1.04, 1.08])
    batch.set_dofs(batch.undeformed_dofs())
    batch.load_increment_apply(1, 2)
    batch.residual_evaluate()
    batch.jacobian_evaluate(cm.OC_FP64)
    batch.solver_matrices_assemble(cm.OC_FP64)
    rp, ci = batch.matrix_structure(0)
    mask = _mask(batch.n_rows, fixed)
    ov, ob = _check_against(batch, rp, ci, batch.matrix_values(), mask, batch.residual(), batch.get_dofs(), False)
    assert np.all(np.isfinite(ov)) and np.all(np.isfinite(ob))
    assert np.array_equal(batch.solver_matrix_values(), ov)
    assert np.array_equal(batch.solver_vector(), ob)
    batch.destroy()
