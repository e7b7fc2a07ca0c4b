"""GPU solver (the end-to-end check of the assembly path): every accepted answer is
checked against the FP64 oracle residual, plus closed-form solutions."""
import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu


def test_laplace_config1_patch_solution(gpu):
    """8^3 trilinear, u = 0 on x = 0 and u = 1 on x = 1: the discrete solution
    is u = x exactly."""
    cm = gpu
    mesh = cm.RegularMesh.generate(1, 8)
    eq = cm.LaplaceEquationsSet(mesh)
    left, right = mesh.face_nodes(0, 0), mesh.face_nodes(0, 1)
    eq.set_boundary_conditions(np.concatenate([left, right]),
                               np.concatenate([np.zeros(len(left)), np.ones(len(right))]))
    st = eq.solve(cm.newton_options(krylov_relative_tolerance=1e-13, krylov_restart=60, absolute_tolerance=0.0))
    assert st["converged"]
    u = eq.get_dofs()
    assert np.abs(u - mesh.coords[0]).max() < 1e-10


@pytest.mark.parametrize("restart", [2, 3, 5])
def test_laplace_restarted_gmres(gpu, restart):
    """Restarted GMRES(m): many restart cycles (each ends on its m-th Arnoldi
    step, read at once) still reach the patch solution u = x (trilinear
    elements reproduce linear fields on jittered meshes too)."""
    cm = gpu
    mesh = cm.RegularMesh.generate(1, 8, jitter=0.2, seed=5)
    eq = cm.LaplaceEquationsSet(mesh)
    left, right = mesh.face_nodes(0, 0), mesh.face_nodes(0, 1)
    eq.set_boundary_conditions(np.concatenate([left, right]),
                               np.concatenate([np.zeros(len(left)), np.ones(len(right))]))
    st = eq.solve(cm.newton_options(krylov_relative_tolerance=1e-12, krylov_restart=restart, absolute_tolerance=0.0,
                                    krylov_maximum_iterations=20000))
    assert st["converged"]
    assert st["krylov_iterations"] > restart  # at least one restart
    assert np.abs(eq.get_dofs() - mesh.coords[0]).max() < 1e-9


def _oracle_free_residual(eq, mesh, fixed, c10, c01):
    d = eq.get_dofs()
    r = orc.mr_residual(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, c10, c01)
    r[fixed] = 0.0
    return np.linalg.norm(r)


def test_mr_homogeneous_uniaxial_solution(gpu):
    """Unjittered mesh, symmetry planes and z_x -> lambda on x = 1: the exact
    discrete solution is the homogeneous incompressible stretch."""
    cm = gpu
    c10, c01, lam = 2.0, 6.0, 1.1
    mesh = cm.RegularMesh.generate(2, 2)
    eq = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    eq.set_dofs(eq.undeformed_dofs())
    fixed = eq.uniaxial_boundary_conditions(lam)
    opts = cm.newton_options(solution_tolerance=0.0, relative_tolerance=1e-12, absolute_tolerance=1e-13, krylov_restart=200,
                             krylov_relative_tolerance=-1.0, preconditioner=cm.OC_PRECONDITIONER_BLOCK_DIAGONAL)
    for step in (1, 2):
        eq.load_increment_apply(step, 2)
        st = eq.newton_solve(opts)
        assert st["converged"] and st["newton_iterations"] <= 8
    d = eq.get_dofs()
    nq = mesh.n_nodes
    z = d[:3 * nq].reshape(3, nq)
    expected = mesh.coords * np.array([[lam], [lam ** -0.5], [lam ** -0.5]])
    assert np.abs(z - expected).max() < 1e-9
    s0 = orc.mr_stress(np.diag([lam, lam ** -0.5, lam ** -0.5]), 0.0, c10, c01)
    assert np.abs(d[3 * nq:] + s0[1]).max() < 1e-8
    assert _oracle_free_residual(eq, mesh, fixed, c10, c01) < 1e-11


@pytest.mark.parametrize("precision,precond", [(0, 1), (1, 1), (0, 2), (0, 0)])
def test_mr_newton_jittered_judged_by_oracle(gpu, precision, precond):
    cm = gpu
    c10, c01 = 2.0, 6.0
    mesh = cm.RegularMesh.generate(2, 3, jitter=0.1, seed=4)
    eq = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    eq.set_dofs(eq.undeformed_dofs())
    fixed = eq.uniaxial_boundary_conditions(1.1)
    opts = cm.newton_options(solution_tolerance=0.0, relative_tolerance=1e-10, absolute_tolerance=0.0, krylov_restart=300,
                             krylov_relative_tolerance=-1.0, jacobian_precision=precision, preconditioner=precond,
                             maximum_iterations=30)
    for step in (1, 2, 3):
        eq.load_increment_apply(step, 3)
        st = eq.newton_solve(opts)
        assert st["converged"], st
        ref = st["initial_residual_norm"]
        assert _oracle_free_residual(eq, mesh, fixed, c10, c01) <= 1.01e-10 * ref + 1e-14


def test_newton_failure_modes_forced_from_host(gpu):
    cm = gpu
    mesh = cm.RegularMesh.generate(2, 2, jitter=0.1, seed=1)
    eq = cm.MooneyRivlinEquationsSet(mesh, 2.0, 6.0)
    eq.set_dofs(eq.undeformed_dofs())
    eq.uniaxial_boundary_conditions(1.2)
    eq.load_increment_apply(1, 1)
    # iteration limit (SNES_DIVERGED_MAX_IT analogue)
    with pytest.raises(cm.CMISSError) as exc:
        eq.newton_solve(cm.newton_options(maximum_iterations=1, relative_tolerance=1e-12))
    assert exc.value.code == -5 and "did not converge" in str(exc.value)
    st = eq.newton_solve(cm.newton_options(maximum_iterations=1, relative_tolerance=1e-12),
                         raise_on_failure=False)
    assert not st["converged"] and st["reason"] == cm.OC_DIVERGED_MAX_IT
    # function-evaluation limit (SNES_DIVERGED_FUNCTION_COUNT)
    st = eq.newton_solve(cm.newton_options(maximum_function_evaluations=2, relative_tolerance=1e-12,
                                           solution_tolerance=0.0), raise_on_failure=False)
    assert not st["converged"] and st["reason"] == cm.OC_DIVERGED_FUNCTION_COUNT
    assert st["residual_evaluations"] >= 2
    # NaN material constant (DIVERGED_FNORM_NAN analogue): stops, reports, never hangs
    eq.set_materials(float("nan"), 6.0)
    st = eq.newton_solve(cm.newton_options(maximum_iterations=5), raise_on_failure=False)
    assert not st["converged"] and not np.isfinite(st["final_residual_norm"])
    assert st["reason"] == cm.OC_DIVERGED_FNORM_NAN


def test_newton_petsc_default_convergence_tests(gpu):
    """SNESConvergedDefault as the reference selects it: the step test
    (||step|| < stol ||x||) ends a solve whose residual target is out of reach;
    an absolute tolerance above the starting residual ends it before a step."""
    cm = gpu
    mesh = cm.RegularMesh.generate(2, 2, jitter=0.1, seed=8)
    eq = cm.MooneyRivlinEquationsSet(mesh, 2.0, 6.0)
    eq.set_dofs(eq.undeformed_dofs())
    eq.uniaxial_boundary_conditions(1.05)
    eq.load_increment_apply(1, 1)
    start = eq.get_dofs()
    ref = eq.residual_norm()
    st = eq.newton_solve(cm.newton_options(relative_tolerance=1e-30, absolute_tolerance=0.0, solution_tolerance=1e-3,
                                           krylov_restart=100))
    assert st["converged"] and st["reason"] == cm.OC_CONVERGED_SNORM_RELATIVE
    assert 1 <= st["newton_iterations"] < 50 and st["final_residual_norm"] > 1e-30 * ref
    eq.set_dofs(start)
    st = eq.newton_solve(cm.newton_options(absolute_tolerance=10.0 * ref))
    assert st["converged"] and st["reason"] == cm.OC_CONVERGED_FNORM_ABS and st["newton_iterations"] == 0
    eq.set_dofs(start)
    st = eq.newton_solve(cm.newton_options(solution_tolerance=0.0, relative_tolerance=1e-8, absolute_tolerance=0.0,
                                           krylov_restart=100))
    assert st["converged"] and st["reason"] == cm.OC_CONVERGED_FNORM_RELATIVE
    eq.destroy()


def test_multigrid_preconditioner_judged_by_oracle(gpu):
    """Geometric multigrid for the displacement block: the same FP64 residual
    bar as the block preconditioners, in far fewer Krylov iterations."""
    cm = gpu
    c10, c01 = 2.0, 6.0
    counts = {}
    for precond in (cm.OC_PRECONDITIONER_MULTIGRID, cm.OC_PRECONDITIONER_BLOCK_TRIANGULAR):
        mesh = cm.RegularMesh.generate(2, 6, jitter=0.1, seed=77)
        eq = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
        eq.set_dofs(eq.undeformed_dofs())
        fixed = eq.uniaxial_boundary_conditions(1.1)
        if precond == cm.OC_PRECONDITIONER_MULTIGRID:
            eq.set_multigrid()
        opts = cm.newton_options(solution_tolerance=0.0, relative_tolerance=1e-10, absolute_tolerance=0.0, krylov_restart=100,
                                 krylov_relative_tolerance=-1.0, preconditioner=precond, maximum_iterations=30)
        counts[precond] = 0
        for step in (1, 2):
            eq.load_increment_apply(step, 4)
            ref = eq.residual_norm()
            st = eq.newton_solve(opts)
            assert st["converged"], st
            counts[precond] += st["krylov_iterations"]
            assert _oracle_free_residual(eq, mesh, fixed, c10, c01) <= 1.05e-10 * ref
        eq.destroy()
    print("Krylov iterations", counts)
    assert counts[cm.OC_PRECONDITIONER_MULTIGRID] * 3 < counts[cm.OC_PRECONDITIONER_BLOCK_TRIANGULAR]


@pytest.mark.parametrize("precision", [0, 1])
def test_multigrid_restarted_gmres_judged_by_oracle(gpu, precision):
    """GMRES(4) under the multigrid preconditioner: every Newton step's linear
    solve runs through several restart cycles and the answer still meets the
    FP64 oracle residual."""
    cm = gpu
    c10, c01 = 2.0, 6.0
    mesh = cm.RegularMesh.generate(2, 4, jitter=0.1, seed=31)
    eq = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    eq.set_dofs(eq.undeformed_dofs())
    fixed = eq.uniaxial_boundary_conditions(1.1)
    eq.set_multigrid()
    opts = cm.newton_options(solution_tolerance=0.0, relative_tolerance=1e-10, absolute_tolerance=0.0, krylov_restart=4,
                             krylov_relative_tolerance=-1.0, jacobian_precision=precision,
                             preconditioner=cm.OC_PRECONDITIONER_MULTIGRID, maximum_iterations=30,
                             krylov_maximum_iterations=20000)
    for step in (1, 2):
        eq.load_increment_apply(step, 4)
        ref = eq.residual_norm()
        st = eq.newton_solve(opts)
        assert st["converged"], st
        assert st["krylov_iterations"] > 4 * st["newton_iterations"]
        assert _oracle_free_residual(eq, mesh, fixed, c10, c01) <= 1.05e-10 * ref
    eq.destroy()


def test_multigrid_rejects_irregular_numbering(gpu):
    cm = gpu
    mesh = cm.RegularMesh.generate(2, 2)
    eq = cm.MooneyRivlinEquationsSet(mesh, 2.0, 6.0)
    with pytest.raises(cm.CMISSError) as exc:
        cm._lib.call("oc_multigrid_structure_set", eq.handle, np.array([3], np.int32))
    assert exc.value.code == -2
    opts = cm.newton_options(preconditioner=cm.OC_PRECONDITIONER_MULTIGRID)
    with pytest.raises(cm.CMISSError):
        eq.newton_solve(opts)
    eq.destroy()


def test_boundary_conditions_changed_after_a_solve(gpu):
    """Solve, set different Dirichlet conditions (another DOF count), solve
    again: the set's recorded Newton-step graph (Jacobian assembly and
    multigrid setup, which read the free-DOF mask) follows the new conditions,
    and the answer matches a fresh set's, judged by the oracle residual."""
    cm = gpu
    c10, c01 = 2.0, 6.0
    mesh = cm.RegularMesh.generate(2, 3, jitter=0.1, seed=9)
    opts = cm.newton_options(solution_tolerance=0.0, relative_tolerance=1e-10, absolute_tolerance=0.0,
                             krylov_restart=100, krylov_relative_tolerance=-1.0,
                             preconditioner=cm.OC_PRECONDITIONER_MULTIGRID, maximum_iterations=30)
    nq = mesh.n_nodes
    x0 = mesh.face_nodes(0, 0)
    clamp = np.concatenate([x0, nq + x0, 2 * nq + x0, mesh.face_nodes(0, 1)])
    clamp_vals = np.concatenate([mesh.coords[0, x0], mesh.coords[1, x0], mesh.coords[2, x0],
                                 np.full(len(mesh.face_nodes(0, 1)), 1.08)])

    def solve(eq):
        eq.set_dofs(eq.undeformed_dofs())
        eq.load_increment_apply(1, 1)
        ref = eq.residual_norm()
        st = eq.newton_solve(opts)
        assert st["converged"], st
        return ref

    eq = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    eq.uniaxial_boundary_conditions(1.05)
    eq.set_multigrid()
    solve(eq)
    eq.set_boundary_conditions(clamp, clamp_vals)
    ref = solve(eq)
    assert _oracle_free_residual(eq, mesh, clamp, c10, c01) <= 1.05e-10 * ref
    fresh = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    fresh.set_boundary_conditions(clamp, clamp_vals)
    fresh.set_multigrid()
    solve(fresh)
    assert np.abs(eq.get_dofs() - fresh.get_dofs()).max() < 1e-8
    eq.destroy()
    fresh.destroy()


def test_plain_fallback_after_a_failed_warm_start(gpu):
    """A learned start that fails (the shipped network's weights scaled by 100
    from the host, then a learned attempt allowed no Newton iteration) restarts
    from the plain start with FP64 Jacobians when the options ask for the
    fallback; the answer meets the plain method's FP64 residual by the oracle."""
    cm = gpu
    from cm_amd import _lib
    from cm_amd.train_warm_start import DEFAULT_WEIGHTS, make_problem, solver_options
    from cm_amd.warm_start import WarmStartMLP
    net = WarmStartMLP.load(DEFAULT_WEIGHTS)
    good = net.parameters()
    case = dict(n=4, c10=2.0, c01=6.0, stretch=0.15, seed=4242)
    mesh, eq = make_problem(case)
    dload = case["stretch"] / 5
    for s, (weights, max_it) in enumerate([(good * 100.0, 8), (good, 0)], start=1):
        net.set_parameters(weights)
        d0 = eq.get_dofs()
        eq.load_increment_apply(s, 5)
        ref = eq.residual_norm()
        eq.set_dofs(d0)
        net.warm_start(eq, dload)
        eq.load_increment_apply(s, 5)
        start = eq.get_dofs()
        st = eq.newton_solve(solver_options(_lib.OC_FP32, reference_norm=ref, maximum_iterations=max_it),
                             raise_on_failure=False)
        assert not st["converged"] and st["fallback_used"] == 0
        eq.set_dofs(start)  # a DOF upload drops the saved plain start: no fallback possible
        st = eq.newton_solve(solver_options(_lib.OC_FP32, reference_norm=ref, maximum_iterations=max_it, fallback=1),
                             raise_on_failure=False)
        assert not st["converged"] and st["fallback_used"] == 0
        eq.set_dofs(d0)
        net.warm_start(eq, dload)
        eq.load_increment_apply(s, 5)
        st = eq.newton_solve(solver_options(_lib.OC_FP32, reference_norm=ref, maximum_iterations=max_it, fallback=1,
                                            fallback_maximum_iterations=50))
        print(s, st)
        assert st["converged"] and st["fallback_used"] == 1 and st["warm_start_reason"] < 0
        assert _oracle_free_residual(eq, mesh, eq.fixed_dofs, case["c10"], case["c01"]) <= 1.05e-10 * ref
    net.destroy()
    eq.destroy()


def test_solution_export_and_ndiff(gpu, tmp_path):
    """The GPU solution of the 8^3 Laplace case (u = 0 on x = 0, u = 1 on x = 1)
    exported as exnode/exelem agrees with the export of the exact field u = x
    within 1e-10 under the ndiff rule, and the homogeneous uniaxial
    Mooney-Rivlin solution with the export of its closed form; a perturbed
    field does not."""
    import os
    from cm_amd import field_io as fio

    class Exact:
        def __init__(self, mesh, dofs):
            self.mesh, self._d = mesh, dofs

        def get_dofs(self):
            return self._d

    cm = gpu
    mesh = cm.RegularMesh.generate(1, 8)
    eq = cm.LaplaceEquationsSet(mesh)
    left, right = mesh.face_nodes(0, 0), mesh.face_nodes(0, 1)
    eq.set_boundary_conditions(np.concatenate([left, right]),
                               np.concatenate([np.zeros(len(left)), np.ones(len(right))]))
    assert eq.solve(cm.newton_options(krylov_relative_tolerance=1e-13, krylov_restart=60,
                                      absolute_tolerance=0.0))["converged"]
    a, b, c = (os.path.join(tmp_path, k) for k in ("gpu", "exact", "off"))
    fio.export_laplace(a, eq)
    fio.export_laplace(b, Exact(mesh, mesh.coords[0].copy()))
    fio.export_laplace(c, Exact(mesh, mesh.coords[0] + 1e-6))
    assert fio.ndiff(a + ".part0.exnode", b + ".part0.exnode", "1e-10") == []
    assert fio.ndiff(a + ".part0.exelem", b + ".part0.exelem", "0") == []
    assert len(fio.ndiff(a + ".part0.exnode", c + ".part0.exnode", "1e-10")) > 0
    eq.destroy()

    c10, c01, lam = 2.0, 6.0, 1.1
    mesh = cm.RegularMesh.generate(2, 2)
    mr = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    mr.set_dofs(mr.undeformed_dofs())
    mr.uniaxial_boundary_conditions(lam)
    mr.load_increment_apply(1, 1)
    assert mr.newton_solve(cm.newton_options(solution_tolerance=0.0, relative_tolerance=1e-13, absolute_tolerance=1e-14,
                                             krylov_restart=200, krylov_relative_tolerance=-1.0,
                                             preconditioner=cm.OC_PRECONDITIONER_BLOCK_DIAGONAL))["converged"]
    s0 = orc.mr_stress(np.diag([lam, lam ** -0.5, lam ** -0.5]), 0.0, c10, c01)
    z = mesh.coords * np.array([[lam], [lam ** -0.5], [lam ** -0.5]])
    exact = np.concatenate([z.ravel(), np.full(mesh.n_nodes_l, -s0[1])])
    fio.export_mooney_rivlin(a, mr)
    fio.export_mooney_rivlin(b, Exact(mesh, exact))
    assert fio.ndiff(a + ".part0.exnode", b + ".part0.exnode", "1e-9") == []
    assert fio.ndiff(a + ".part0.exelem", b + ".part0.exelem", "0") == []
    mr.destroy()
