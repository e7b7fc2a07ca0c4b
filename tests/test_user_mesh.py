"""General (user-numbered) mesh ingestion: the reference's topology rule
(mesh component nodes = the global nodes its elements use, sorted; local
number = rank) restated in cm_amd.mesh.UserMesh."""
import numpy as np
import pytest

import oracle_lib as orc
from cm_amd.mesh import RegularMesh, UserMesh, as_user_mesh


@pytest.mark.parametrize("order", [1, 2])
def test_identity_region_order_reproduces_generated_numbering(order):
    m = RegularMesh.generate(order, (2, 3, 2), jitter=0.1, seed=3)
    u = as_user_mesh(m)
    assert np.array_equal(u.elem_nodes, m.elem_nodes) and np.array_equal(u.coords, m.coords)
    if order == 2:
        assert u.n_nodes_l == m.n_nodes_l and np.array_equal(u.elem_nodes_l, m.elem_nodes_l)


@pytest.mark.parametrize("order", [1, 2])
def test_shuffled_region_keeps_geometry_and_ranks_global_numbers(order):
    m = RegularMesh.generate(order, (3, 2, 2), jitter=0.15, seed=7)
    rng = np.random.default_rng(11)
    region = rng.permutation(m.n_nodes)
    user = rng.choice(10 * m.n_nodes, size=m.n_nodes, replace=False) + 1000
    u = as_user_mesh(m, region_order=region, user_numbers=user)
    # every element sees the same nodal coordinates in the same local order
    assert np.array_equal(u.coords[:, u.elem_nodes], m.coords[:, m.elem_nodes])
    # local node j is the j-th smallest global number (region position) in use
    pos = {int(v): p for p, v in enumerate(user)}
    gl = np.array([pos[int(v)] for v in u.node_user_numbers])
    assert np.all(np.diff(gl) > 0)
    if order == 2:
        gl_l = np.array([pos[int(v)] for v in u.node_user_numbers_l])
        assert np.all(np.diff(gl_l) > 0) and u.n_nodes_l == m.n_nodes_l


def test_user_mesh_assembly_is_the_permuted_generated_assembly():
    """The oracle's Laplace matrix on the shuffled mesh is the generated
    mesh's matrix with rows and columns permuted, entry for entry."""
    m = RegularMesh.generate(1, (3, 2, 2), jitter=0.15, seed=7)
    rng = np.random.default_rng(2)
    region = rng.permutation(m.n_nodes)
    u = as_user_mesh(m, region_order=region, user_numbers=np.arange(m.n_nodes) * 3 + 5)
    rp, ci = orc.pattern([(m.n_nodes, m.elem_nodes)])
    v, _ = orc.laplace_assemble(m.elem_nodes, m.coords, rp, ci)
    urp, uci = orc.pattern([(u.n_nodes, u.elem_nodes)])
    uv, _ = orc.laplace_assemble(u.elem_nodes, u.coords, urp, uci)
    # node p of the generated mesh is local node perm[p] of the user mesh
    perm = np.empty(m.n_nodes, np.int64)
    perm[m.elem_nodes.ravel()] = u.elem_nodes.ravel()
    K = {(r, int(c)): v[k] for r in range(m.n_nodes) for k, c in zip(range(rp[r], rp[r + 1]), ci[rp[r]:rp[r + 1]])}
    inv = np.argsort(perm)
    for r in range(m.n_nodes):
        ur = perm[r]
        cols = uci[urp[ur]:urp[ur + 1]]
        vals = uv[urp[ur]:urp[ur + 1]]
        got = {int(inv[c]): x for c, x in zip(cols, vals)}
        want = {c: K[(r, c)] for c in ci[rp[r]:rp[r + 1]]}
        assert got == want


def test_user_mesh_errors():
    with pytest.raises(ValueError):
        UserMesh.from_user_numbers(1, [1, 2, 2], np.zeros((3, 3)), [[1, 2, 3, 1, 2, 3, 1, 2]])
    with pytest.raises(ValueError):
        UserMesh.from_user_numbers(1, [1, 2, 3], np.zeros((3, 3)), [[1, 2, 3, 4, 1, 2, 3, 1]])
    with pytest.raises(ValueError):
        UserMesh.from_user_numbers(2, np.arange(1, 28), np.zeros((3, 27)), [np.arange(1, 28)])


@pytest.mark.gpu
def test_gpu_sets_take_a_user_mesh(gpu):
    """Both equations sets assembled on the GPU from user-numbered meshes
    (shuffled region order, arbitrary user numbers) match the oracle."""
    cm = gpu
    rng = np.random.default_rng(21)
    m1 = RegularMesh.generate(1, (3, 2, 2), jitter=0.1, seed=4)
    u1 = as_user_mesh(m1, region_order=rng.permutation(m1.n_nodes),
                      user_numbers=rng.choice(100000, m1.n_nodes, replace=False) + 1)
    lap = cm.LaplaceEquationsSet(u1)
    lap.assemble()
    rp, ci = lap.matrix_structure(0)
    orp, oci = orc.pattern([(u1.n_nodes, u1.elem_nodes)])
    assert np.array_equal(rp.astype(np.int64), orp) and np.array_equal(ci, oci)
    ov, _ = orc.laplace_assemble(u1.elem_nodes, u1.coords, orp, oci)
    assert np.linalg.norm(lap.matrix_values() - ov) <= 1e-12 * np.linalg.norm(ov)
    lap.destroy()
    m2 = RegularMesh.generate(2, (2, 2, 1), jitter=0.1, seed=6)
    u2 = as_user_mesh(m2, region_order=rng.permutation(m2.n_nodes),
                      user_numbers=rng.choice(100000, m2.n_nodes, replace=False) + 1)
    eq = cm.MooneyRivlinEquationsSet(u2, 2.0, 6.0)
    d = np.concatenate([u2.coords.ravel() * 1.01, 0.3 * rng.standard_normal(u2.n_nodes_l)])
    eq.set_dofs(d)
    eq.residual_evaluate()
    eq.jacobian_evaluate(cm.OC_FP64)
    rp, ci = eq.matrix_structure(0)
    orp, oci = orc.pattern([(u2.n_nodes, u2.elem_nodes)] * 3 + [(u2.n_nodes_l, u2.elem_nodes_l)])
    assert np.array_equal(rp.astype(np.int64), orp) and np.array_equal(ci, oci)
    oR = orc.mr_residual(u2.elem_nodes, u2.elem_nodes_l, u2.coords, d, 2.0, 6.0)
    oJ = orc.mr_jacobian(u2.elem_nodes, u2.elem_nodes_l, u2.coords, d, 2.0, 6.0, orp, oci)
    assert np.linalg.norm(eq.residual() - oR) <= 1e-12 * np.linalg.norm(oR)
    assert np.linalg.norm(eq.matrix_values() - oJ) <= 1e-12 * np.linalg.norm(oJ)
    eq.destroy()
