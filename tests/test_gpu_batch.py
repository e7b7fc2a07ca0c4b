"""Batches of independent Mooney-Rivlin meshes assembled and solved as one
equations set: every per-mesh result must equal the single-mesh set's
(bit-exact for assembly), graph replays equal eager launches, the autotuner
changes no assembled value, and the batched Newton meets every mesh's own
tolerance, judged by the oracle residual."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as orc

pytestmark = pytest.mark.gpu

CASES = [dict(n=2, seed=11, c10=2.0, c01=6.0), dict(n=3, seed=12, c10=1.5, c01=0.5),
         dict(n=1, seed=13, c10=4.0, c01=9.0)]


def make_meshes(cm, cases):
    return [cm.RegularMesh.generate(2, c["n"], jitter=0.1, seed=c["seed"]) for c in cases]


def perturbed(mesh, seed, amp=0.02):
    rng = np.random.default_rng(seed)
    z = mesh.coords * np.array([1.05, 0.97, 0.98])[:, None] + amp * rng.uniform(-1, 1, mesh.coords.shape)
    return np.concatenate([z.ravel(), rng.standard_normal(mesh.n_nodes_l)])


def union_csr(eq, values):
    rp, ci = eq.matrix_structure()
    return sp.csr_matrix((values, ci, rp), shape=(eq.n_rows, eq.n_rows))


def test_batch_assembly_equals_single_meshes_bit_exact(gpu):
    cm = gpu
    meshes = make_meshes(cm, CASES)
    batch = cm.MooneyRivlinBatch(meshes, [c["c10"] for c in CASES], [c["c01"] for c in CASES])
    d = np.zeros(batch.n_rows)
    singles = []
    for m, (mesh, c) in enumerate(zip(meshes, CASES)):
        dm = perturbed(mesh, 100 + m)
        d[batch.mesh_dofs(m)] = dm
        eq = cm.MooneyRivlinEquationsSet(mesh, c["c10"], c["c01"])
        eq.set_dofs(dm)
        eq.residual_evaluate()
        eq.jacobian_evaluate(cm.OC_FP64)
        eq.jacobian_evaluate(cm.OC_FP32)
        singles.append((eq, eq.residual(), eq.matrix_values(), eq.matrix_values_sp()))
    batch.set_dofs(d)
    batch.residual_evaluate()
    batch.jacobian_evaluate(cm.OC_FP64)
    batch.jacobian_evaluate(cm.OC_FP32)
    r = batch.residual()
    A64, A32 = union_csr(batch, batch.matrix_values()), union_csr(batch, batch.matrix_values_sp())
    assert A64.nnz == sum(s[0].nnz for s in singles)
    for m, (eq, r1, v64, v32) in enumerate(singles):
        idx = batch.mesh_dofs(m)
        np.testing.assert_array_equal(r[idx], r1)
        rp, ci = eq.matrix_structure()
        for A, v in ((A64, v64), (A32, v32)):
            sub = A[idx][:, idx].tocsr()
            sub.sort_indices()
            np.testing.assert_array_equal(sub.indptr, rp)
            np.testing.assert_array_equal(sub.indices, ci)
            np.testing.assert_array_equal(sub.data, v)
        eq.destroy()
    batch.destroy()


def test_graph_replay_equals_eager_launches(gpu):
    cm = gpu
    meshes = make_meshes(cm, CASES[:2])
    batch = cm.MooneyRivlinBatch(meshes, [2.0, 1.0], [6.0, 3.0])
    for rep in range(3):
        d = np.zeros(batch.n_rows)
        for m, mesh in enumerate(meshes):
            d[batch.mesh_dofs(m)] = perturbed(mesh, 7 * rep + m)
        batch.set_dofs(d)
        batch.jacobian_evaluate(cm.OC_FP64)
        eager = batch.matrix_values()
        batch.residual_evaluate()  # eager work between the replays
        batch.jacobian_evaluate_graph(cm.OC_FP64)
        cm.synchronise()
        np.testing.assert_array_equal(batch.matrix_values(), eager)
    batch.destroy()


def test_autotuned_launch_shapes_change_no_value(gpu):
    cm = gpu
    mesh = cm.RegularMesh.generate(2, 4, jitter=0.1, seed=5)
    eq = cm.MooneyRivlinEquationsSet(mesh, 2.0, 6.0)
    eq.set_dofs(perturbed(mesh, 3))
    eq.jacobian_evaluate(cm.OC_FP64)
    before = eq.matrix_values()
    report = eq.autotune(cm.OC_FP64)
    print(report)
    chosen = report["chosen"]
    assert chosen["element_blocks_per_cu"] in (1, 2, 3, 4, 6, 8)
    assert chosen["gather_chunk"] in (4, 8, 16)
    assert chosen["spmv_lanes"] in (8, 16, 32, 64)
    eq.jacobian_evaluate(cm.OC_FP64)
    np.testing.assert_array_equal(eq.matrix_values(), before)
    eq.destroy()


def test_batched_newton_meets_every_mesh_tolerance(gpu):
    cm = gpu
    cases = [dict(n=2, seed=21, c10=2.0, c01=6.0, s=1.10), dict(n=3, seed=22, c10=1.0, c01=0.2, s=1.05),
             dict(n=2, seed=23, c10=5.0, c01=10.0, s=1.20)]
    meshes = make_meshes(cm, cases)
    batch = cm.MooneyRivlinBatch(meshes, [c["c10"] for c in cases], [c["c01"] for c in cases])
    batch.set_dofs(batch.undeformed_dofs())
    fixed = batch.uniaxial_boundary_conditions([c["s"] for c in cases])
    opts = cm.newton_options(solution_tolerance=0.0, relative_tolerance=1e-10, absolute_tolerance=0.0, krylov_restart=100,
                             krylov_relative_tolerance=-1.0, preconditioner=cm.OC_PRECONDITIONER_BLOCK_TRIANGULAR,
                             maximum_iterations=30)
    for step in (1, 2):
        batch.load_increment_apply(step, 2)
        refs = batch.mesh_residual_norms()
        st, finals = batch.newton_solve(opts)
        assert st["converged"], st
        assert np.all(finals <= 1e-10 * refs)
        d = batch.get_dofs()
        for m, (mesh, c) in enumerate(zip(meshes, cases)):
            idx = batch.mesh_dofs(m)
            r = orc.mr_residual(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d[idx], c["c10"], c["c01"])
            local_fixed = np.isin(idx, fixed)
            r[local_fixed] = 0.0
            assert np.linalg.norm(r) <= 1.05e-10 * refs[m], (step, m)
    batch.destroy()


def test_warm_start_on_a_batch_equals_per_mesh(gpu):
    cm = gpu
    from cm_amd.warm_start import WarmStartMLP
    net = WarmStartMLP(hidden=32, layers=2, seed=9)  # 4 outputs: displacement and pressure
    meshes = make_meshes(cm, CASES)
    batch = cm.MooneyRivlinBatch(meshes, [c["c10"] for c in CASES], [c["c01"] for c in CASES])
    d = np.zeros(batch.n_rows)
    expect = []
    for m, (mesh, c) in enumerate(zip(meshes, CASES)):
        dm = perturbed(mesh, 50 + m)
        d[batch.mesh_dofs(m)] = dm
        eq = cm.MooneyRivlinEquationsSet(mesh, c["c10"], c["c01"])
        eq.set_dofs(dm)
        net.warm_start(eq, 0.01 * (m + 1))
        expect.append(eq.get_dofs())
        eq.destroy()
    batch.set_dofs(d)
    net.warm_start(batch, 0.01 * np.arange(1, len(CASES) + 1))
    out = batch.get_dofs()
    for m in range(len(CASES)):
        np.testing.assert_array_equal(out[batch.mesh_dofs(m)], expect[m])
    batch.destroy()
    net.destroy()


def test_batched_newton_with_multigrid(gpu):
    cm = gpu
    cases = [dict(n=2, seed=31, c10=2.0, c01=6.0, s=1.10), dict(n=5, seed=32, c10=1.0, c01=0.2, s=1.05),
             dict(n=3, seed=33, c10=5.0, c01=10.0, s=1.20), dict(n=1, seed=34, c10=3.0, c01=1.0, s=1.15)]
    meshes = make_meshes(cm, cases)
    batch = cm.MooneyRivlinBatch(meshes, [c["c10"] for c in cases], [c["c01"] for c in cases])
    batch.set_dofs(batch.undeformed_dofs())
    fixed = batch.uniaxial_boundary_conditions([c["s"] for c in cases])
    batch.set_multigrid()
    opts = cm.newton_options(solution_tolerance=0.0, relative_tolerance=1e-10, absolute_tolerance=0.0, krylov_restart=100,
                             krylov_relative_tolerance=-1.0, preconditioner=cm.OC_PRECONDITIONER_MULTIGRID,
                             maximum_iterations=30)
    for step in (1, 2):
        batch.load_increment_apply(step, 3)
        refs = batch.mesh_residual_norms()
        st, finals = batch.newton_solve(opts)
        assert st["converged"], st
        d = batch.get_dofs()
        for m, (mesh, c) in enumerate(zip(meshes, cases)):
            idx = batch.mesh_dofs(m)
            r = orc.mr_residual(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d[idx], c["c10"], c["c01"])
            r[np.isin(idx, fixed)] = 0.0
            assert np.linalg.norm(r) <= 1.05e-10 * refs[m], (step, m)
    batch.destroy()
