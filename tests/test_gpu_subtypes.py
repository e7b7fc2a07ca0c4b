"""The incompressible Mooney-Rivlin subtype (EQUATIONS_SET_INCOMPRESSIBLE_
MOONEY_RIVLIN_SUBTYPE: Cauchy stress from the 2nd Piola tensor with the
pressure inside it, forward-difference Jacobian) on the GPU against the
oracle, and as a cross-check of the analytic Mooney-Rivlin subtype: where the
two laws coincide (c01 = 0, and every law on isochoric states) their residuals
agree and the forward-difference Jacobian matches the analytic one."""
import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _state(mesh, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    z = mesh.coords + amp / max(mesh.n) * rng.uniform(-1, 1, mesh.coords.shape)
    return np.concatenate([z.ravel(), 0.5 * rng.standard_normal(mesh.n_nodes_l)])


def _pattern(mesh):
    return orc.pattern([(mesh.n_nodes, mesh.elem_nodes)] * 3 + [(mesh.n_nodes_l, mesh.elem_nodes_l)])


def test_incompressible_subtype_parity(gpu):
    cm = gpu
    mesh = cm.RegularMesh.generate(2, (3, 2, 4), jitter=0.1, seed=41)
    c10, c01 = 2.0, 6.0
    eq = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    eq.set_subtype(cm.OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE)
    d = _state(mesh, 42)
    eq.set_dofs(d)
    eq.residual_evaluate()
    oR = orc.imr_residual(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, c10, c01)
    assert rel(eq.residual(), oR) < 1e-12
    # the subtype's Jacobian is the forward-difference one (FP64 only)
    orp, oci = _pattern(mesh)
    eq.jacobian_evaluate(cm.OC_FP64)
    oJ = orc.mr_jacobian(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, c10, c01, orp, oci, mode=2)
    assert rel(eq.matrix_values(), oJ) < 1e-7
    with pytest.raises(cm.CMISSError):
        eq.jacobian_evaluate(cm.OC_FP32)
    # the analytic subtype differs where the laws differ (J != 1, c01 != 0)
    mr = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    mr.set_dofs(d)
    mr.residual_evaluate()
    assert rel(mr.residual(), oR) > 1e-4
    eq.destroy()
    mr.destroy()


def test_incompressible_subtype_cross_checks_the_analytic_subtype(gpu):
    """c01 = 0: one neo-Hookean law, two formulations.  The residuals agree to
    rounding; the forward-difference Jacobian of the PK2 form matches the
    analytic Jacobian of the deviatoric form to the one-sided truncation error."""
    cm = gpu
    mesh = cm.RegularMesh.generate(2, 3, jitter=0.1, seed=43)
    c10 = 3.0
    a = cm.MooneyRivlinEquationsSet(mesh, c10, 0.0)
    b = cm.MooneyRivlinEquationsSet(mesh, c10, 0.0)
    b.set_subtype(cm.OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE)
    d = _state(mesh, 44)
    for eq in (a, b):
        eq.set_dofs(d)
        eq.residual_evaluate()
        eq.jacobian_evaluate(cm.OC_FP64)
    assert rel(b.residual(), a.residual()) < 1e-12
    e = rel(b.matrix_values(), a.matrix_values())
    print("FD (PK2 form) vs analytic (deviatoric form):", e)
    assert e < 1e-4
    a.destroy()
    b.destroy()


def test_incompressible_subtype_newton_solve(gpu):
    """Newton with forward-difference Jacobians: the homogeneous uniaxial
    stretch on an affine mesh (exact discrete solution of both subtypes), and at
    c01 = 0 on a jittered mesh the same displacements as the analytic subtype,
    each judged by the oracle residual of its own law."""
    cm = gpu
    opts = cm.newton_options(solution_tolerance=0.0, relative_tolerance=1e-10, absolute_tolerance=1e-13,
                             krylov_restart=200, krylov_relative_tolerance=1e-8,
                             preconditioner=cm.OC_PRECONDITIONER_BLOCK_DIAGONAL, maximum_iterations=30)
    c10, c01, lam = 2.0, 6.0, 1.1
    mesh = cm.RegularMesh.generate(2, 2)
    eq = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    eq.set_subtype(cm.OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE)
    eq.set_dofs(eq.undeformed_dofs())
    fixed = eq.uniaxial_boundary_conditions(lam)
    for step in (1, 2):
        eq.load_increment_apply(step, 2)
        st = eq.newton_solve(opts)
        assert st["converged"], st
    z = eq.get_dofs()[:3 * mesh.n_nodes].reshape(3, mesh.n_nodes)
    assert np.abs(z - mesh.coords * np.array([[lam], [lam ** -0.5], [lam ** -0.5]])).max() < 1e-9
    r = orc.imr_residual(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, eq.get_dofs(), c10, c01)
    r[fixed] = 0.0
    assert np.linalg.norm(r) < 1e-10
    eq.destroy()

    mesh = cm.RegularMesh.generate(2, 3, jitter=0.1, seed=45)
    sols = []
    for subtype in (cm.OC_MOONEY_RIVLIN_SUBTYPE, cm.OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE):
        eq = cm.MooneyRivlinEquationsSet(mesh, 2.5, 0.0)
        eq.set_subtype(subtype)
        eq.set_dofs(eq.undeformed_dofs())
        fixed = eq.uniaxial_boundary_conditions(1.08)
        eq.load_increment_apply(1, 1)
        ref = eq.residual_norm()
        st = eq.newton_solve(opts)
        assert st["converged"], (subtype, st)
        d = eq.get_dofs()
        law = orc.imr_residual if subtype == cm.OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE else orc.mr_residual
        r = law(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, 2.5, 0.0)
        r[fixed] = 0.0
        assert np.linalg.norm(r) <= 1.05e-10 * ref
        sols.append(d)
        eq.destroy()
    assert np.abs(sols[0] - sols[1]).max() < 1e-8
