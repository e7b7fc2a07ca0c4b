"""GPU parity of the assembly path against the C oracle (FP64 restatement of
the reference's element loops).

Bars: CSR row_ptr/col_idx bit-exact (0- and 1-based); FP64 values and
residuals within 1e-12 relative Frobenius error; FP32 Jacobians within
2e-6 relative Frobenius error of the FP64 oracle (float rounding of the
contraction, FP64 kinematics).
"""
import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-6


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _pattern_oracle_laplace(mesh):
    return orc.pattern([(mesh.n_nodes, mesh.elem_nodes)])


def _pattern_oracle_mr(mesh):
    return orc.pattern([(mesh.n_nodes, mesh.elem_nodes)] * 3 + [(mesh.n_nodes_l, mesh.elem_nodes_l)])


@pytest.mark.parametrize("n", [(1, 1, 1), (3, 5, 4), (8, 8, 8)])
def test_laplace_parity(gpu, n):
    cm = gpu
    mesh = cm.RegularMesh.generate(1, n)
    eq = cm.LaplaceEquationsSet(mesh)
    eq.assemble()
    rp, ci = eq.matrix_structure(0)
    orp, oci = _pattern_oracle_laplace(mesh)
    assert np.array_equal(rp.astype(np.int64), orp) and np.array_equal(ci, oci)
    rp1, ci1 = eq.matrix_structure(1)
    assert np.array_equal(rp1, rp + 1) and np.array_equal(ci1, ci + 1)
    n1, n2, n3 = n
    assert eq.nnz == (3 * n1 + 1) * (3 * n2 + 1) * (3 * n3 + 1) or min(n) == 1
    vals = eq.matrix_values()
    ovals, orhs = orc.laplace_assemble(mesh.elem_nodes, mesh.coords, orp, oci)
    assert rel(vals, ovals) < FP64_TOL
    assert np.array_equal(eq.rhs(), orhs)


def test_laplace_jittered_parity(gpu):
    cm = gpu
    mesh = cm.RegularMesh.generate(1, (6, 4, 5), jitter=0.2, seed=11)
    eq = cm.LaplaceEquationsSet(mesh)
    eq.assemble()
    rp, ci = eq.matrix_structure(0)
    ovals, _ = orc.laplace_assemble(mesh.elem_nodes, mesh.coords, rp.astype(np.int64), ci)
    assert rel(eq.matrix_values(), ovals) < FP64_TOL


@pytest.mark.parametrize("n,jitter", [((1, 1, 1), 0.0), ((3, 5, 4), 0.2), ((17, 9, 40), 0.1), ((33, 2, 65), 0.15)])
def test_laplace_fused_equals_two_pass(gpu, n, jitter):
    """The fused structured-grid assembly (node columns of 15 x 7 over
    8-layer segments: these sizes give partial tiles in x and y and several
    z segments) writes the same CSR values bit for bit as the element buffer +
    row gather, and both match the oracle."""
    cm = gpu
    mesh = cm.RegularMesh.generate(1, n, jitter=jitter, seed=31)
    eq = cm.LaplaceEquationsSet(mesh)
    assert not eq.set_fused(0)
    eq.assemble()
    two_pass = eq.matrix_values()
    assert eq.set_fused(2)  # a structured grid: fused at any size
    eq.assemble()
    fused = eq.matrix_values()
    assert np.array_equal(fused, two_pass)
    assert np.array_equal(eq.rhs(), np.zeros(eq.n_rows))
    rp, ci = eq.matrix_structure(0)
    ovals, _ = orc.laplace_assemble(mesh.elem_nodes, mesh.coords, rp.astype(np.int64), ci)
    assert rel(fused, ovals) < FP64_TOL
    eq.destroy()


def _deformed_state(mesh, scale, seed):
    rng = np.random.default_rng(seed)
    nq, nl = mesh.n_nodes, mesh.n_nodes_l
    d = np.concatenate([mesh.coords.ravel(), np.zeros(nl)])
    d[:3 * nq] += scale * rng.uniform(-1, 1, 3 * nq) / max(mesh.n)
    d[3 * nq:] = rng.standard_normal(nl)
    return d


@pytest.mark.parametrize("n,jitter", [((1, 1, 1), 0.0), ((2, 3, 2), 0.1), ((4, 4, 4), 0.1)])
def test_mooney_rivlin_parity(gpu, n, jitter):
    cm = gpu
    mesh = cm.RegularMesh.generate(2, n, jitter=jitter, seed=5)
    c10, c01 = 2.0, 6.0
    eq = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    rp, ci = eq.matrix_structure(0)
    orp, oci = _pattern_oracle_mr(mesh)
    assert np.array_equal(rp.astype(np.int64), orp) and np.array_equal(ci, oci)
    d = _deformed_state(mesh, 0.1, 3)
    eq.set_dofs(d)
    assert np.array_equal(eq.get_dofs(), d)
    eq.residual_evaluate()
    oref = orc.mr_residual(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, c10, c01)
    assert rel(eq.residual(), oref) < FP64_TOL
    eq.jacobian_evaluate(cm.OC_FP64)
    oJ = orc.mr_jacobian(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, c10, c01, orp, oci)
    assert rel(eq.matrix_values(), oJ) < FP64_TOL
    eq.jacobian_evaluate(cm.OC_FP32)
    assert rel(eq.matrix_values_sp().astype(np.float64), oJ) < FP32_TOL


def _permuted(cm, mesh, seed):
    """The same mesh with its node numbering (both components) shuffled."""
    rng = np.random.default_rng(seed)
    pq = rng.permutation(mesh.n_nodes)
    coords = np.empty_like(mesh.coords)
    coords[:, pq] = mesh.coords
    conn_l = None
    if mesh.elem_nodes_l is not None:
        pl = rng.permutation(mesh.n_nodes_l)
        conn_l = pl[mesh.elem_nodes_l].astype(np.int32)
    return cm.RegularMesh(mesh.order, mesh.n, pq[mesh.elem_nodes].astype(np.int32), coords, conn_l)


def test_laplace_parity_shuffled_numbering(gpu):
    """Scatter maps and gathers must not rely on the generated node order."""
    cm = gpu
    mesh = _permuted(cm, cm.RegularMesh.generate(1, (5, 4, 3), jitter=0.1, seed=4), 17)
    eq = cm.LaplaceEquationsSet(mesh)
    assert not eq.set_fused(2)  # not numbered as a structured grid: the row gather serves it
    eq.assemble()
    rp, ci = eq.matrix_structure(0)
    orp, oci = _pattern_oracle_laplace(mesh)
    assert np.array_equal(rp.astype(np.int64), orp) and np.array_equal(ci, oci)
    ovals, _ = orc.laplace_assemble(mesh.elem_nodes, mesh.coords, orp, oci)
    assert rel(eq.matrix_values(), ovals) < FP64_TOL


def test_mooney_rivlin_parity_shuffled_numbering(gpu):
    """The neighbour-rank scatter map on an arbitrary node numbering."""
    cm = gpu
    mesh = _permuted(cm, cm.RegularMesh.generate(2, (3, 2, 4), jitter=0.1, seed=6), 23)
    c10, c01 = 3.0, 1.5
    eq = cm.MooneyRivlinEquationsSet(mesh, c10, c01)
    rp, ci = eq.matrix_structure(0)
    orp, oci = _pattern_oracle_mr(mesh)
    assert np.array_equal(rp.astype(np.int64), orp) and np.array_equal(ci, oci)
    d = _deformed_state(mesh, 0.1, 8)
    eq.set_dofs(d)
    eq.residual_evaluate()
    assert rel(eq.residual(), orc.mr_residual(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, c10, c01)) < FP64_TOL
    oJ = orc.mr_jacobian(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, c10, c01, orp, oci)
    eq.jacobian_evaluate(cm.OC_FP64)
    assert rel(eq.matrix_values(), oJ) < FP64_TOL
    eq.jacobian_evaluate(cm.OC_FP32)
    assert rel(eq.matrix_values_sp().astype(np.float64), oJ) < FP32_TOL


def test_mooney_rivlin_16_config(gpu):
    """Config 2 size: 16^3 triquadratic/trilinear, jitter 0.1h, c10=2, c01=6."""
    cm = gpu
    n = 16
    mesh = cm.RegularMesh.generate(2, n, jitter=0.1, seed=2024)
    eq = cm.MooneyRivlinEquationsSet(mesh, 2.0, 6.0)
    assert eq.n_rows == 112724
    assert eq.nnz == 9 * (8 * n + 1) ** 3 + 6 * (5 * n + 1) ** 3 + (3 * n + 1) ** 3 == 22626496
    d = _deformed_state(mesh, 0.05, 9)
    eq.set_dofs(d)
    eq.residual_evaluate()
    eq.jacobian_evaluate(cm.OC_FP64)
    rp, ci = eq.matrix_structure(0)
    orp, oci = _pattern_oracle_mr(mesh)
    assert np.array_equal(rp.astype(np.int64), orp) and np.array_equal(ci, oci)
    oref = orc.mr_residual(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, 2.0, 6.0)
    assert rel(eq.residual(), oref) < FP64_TOL
    oJ = orc.mr_jacobian(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, 2.0, 6.0, orp, oci)
    J = eq.matrix_values()
    assert rel(J, oJ) < FP64_TOL
    # repeated assembly is deterministic (ordered row gather)
    eq.jacobian_evaluate(cm.OC_FP64)
    assert np.array_equal(eq.matrix_values(), J)


def test_mooney_rivlin_32_config(gpu):
    """Config 3 size: 32^3 (32,768 elements, 859,812 DOFs, 178,723,696 nonzeros)."""
    cm = gpu
    n = 32
    mesh = cm.RegularMesh.generate(2, n, jitter=0.1, seed=2025)
    eq = cm.MooneyRivlinEquationsSet(mesh, 2.0, 6.0)
    assert eq.n_rows == 859812
    assert eq.nnz == 9 * (8 * n + 1) ** 3 + 6 * (5 * n + 1) ** 3 + (3 * n + 1) ** 3 == 178723696
    d = _deformed_state(mesh, 0.05, 10)
    eq.set_dofs(d)
    eq.residual_evaluate()
    assert rel(eq.residual(), orc.mr_residual(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, 2.0, 6.0)) < FP64_TOL
    rp, ci = eq.matrix_structure(0)
    orp, oci = _pattern_oracle_mr(mesh)
    assert np.array_equal(rp.astype(np.int64), orp) and np.array_equal(ci, oci)
    del rp, ci
    oJ = orc.mr_jacobian(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, 2.0, 6.0, orp, oci)
    eq.jacobian_evaluate(cm.OC_FP64)
    assert rel(eq.matrix_values(), oJ) < FP64_TOL
    eq.jacobian_evaluate(cm.OC_FP32)
    assert rel(eq.matrix_values_sp().astype(np.float64), oJ) < FP32_TOL
    eq.destroy()


def test_undeformed_residual_vanishes(gpu):
    cm = gpu
    mesh = cm.RegularMesh.generate(2, 5, jitter=0.1, seed=1)
    eq = cm.MooneyRivlinEquationsSet(mesh, 2.0, 6.0)
    eq.set_dofs(eq.undeformed_dofs())
    eq.residual_evaluate()
    assert np.abs(eq.residual()).max() < 1e-13


def test_fortran_host_assembles_laplace(gpu):
    """A Fortran program assembles the 2x2x2 Laplace matrix through the
    ISO_C_BINDING module with 1-based indices: 27 rows, 343 = 7^3 nonzeros,
    first column 1, K_11 = h/3 with h = 1/2, zero row sums."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "drive_laplace")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (no amdflang)")
    out = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    line = [l for l in out.stdout.splitlines() if l.startswith("rows")][0].split()
    assert line[1] == "27" and line[3] == "343" and line[5] == "1"
    assert abs(float(line[7]) - 0.5 / 3) < 1e-15
    assert float(line[9]) < 1e-14


def test_finite_difference_jacobian_against_oracle(gpu):
    """The forward-difference Jacobian (the reference's rule for subtypes
    without an analytic one) against the oracle's forward differences, and
    against the analytic Jacobian to the truncation error of the scheme."""
    cm = gpu
    mesh = cm.RegularMesh.generate(2, 2, jitter=0.1, seed=6)
    eq = cm.MooneyRivlinEquationsSet(mesh, 2.0, 6.0)
    rp, ci = eq.matrix_structure(0)
    d = _deformed_state(mesh, 0.1, 4)
    eq.set_dofs(d)
    eq.jacobian_evaluate_fd()
    Jfd = eq.matrix_values()
    oJfd = orc.mr_jacobian(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, 2.0, 6.0, rp.astype(np.int64), ci,
                           mode=1)
    # both are (R(x + delta e) - R(x)) / delta in FP64: rounding of the two
    # residuals (~1e-16 relative) is amplified by 1/delta ~ 1e7
    assert rel(Jfd, oJfd) < 1e-7
    # one-sided differences with delta = (1 + ||x||) 1e-7 ~ 1e-6 here: O(delta) truncation
    eq.jacobian_evaluate(cm.OC_FP64)
    assert rel(Jfd, eq.matrix_values()) < 1e-4


def test_fortran_host_mooney_rivlin_residual_and_jacobians(gpu, tmp_path):
    """A Fortran host (tests/fortran/drive_mooney_rivlin.f90, the shape of the
    Iron-side residual and Jacobian evaluation drivers) passes 1-based
    connectivity and its dependent field through ISO_C_BINDING and gets the
    residual and its own 1-based CSR (ROW_INDICES, COLUMN_INDICES, DATA_DP) of
    the FP64 and FP32 Jacobians: structure bit-exact, FP64 values and residual
    within 1e-12, FP32 within 2e-6 of the oracle."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "drive_mooney_rivlin")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (no amdflang)")
    cm = gpu
    mesh = cm.RegularMesh.generate(2, (3, 2, 4), jitter=0.1, seed=21)
    c10, c01 = 2.5, 4.0
    d = _deformed_state(mesh, 0.1, 22)
    fin, fout = os.path.join(tmp_path, "in.bin"), os.path.join(tmp_path, "out.bin")
    with open(fin, "wb") as f:
        np.array([mesh.n_nodes, mesh.n_nodes_l, mesh.n_elements], np.int32).tofile(f)
        np.array([c10, c01], np.float64).tofile(f)
        np.ascontiguousarray(mesh.coords, np.float64).tofile(f)
        (np.ascontiguousarray(mesh.elem_nodes, np.int32) + 1).tofile(f)
        (np.ascontiguousarray(mesh.elem_nodes_l, np.int32) + 1).tofile(f)
        d.astype(np.float64).tofile(f)
    out = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    with open(fout, "rb") as f:
        n, nnz = np.fromfile(f, np.int32, 2)
        res = np.fromfile(f, np.float64, n)
        rp = np.fromfile(f, np.int32, n + 1)
        ci = np.fromfile(f, np.int32, nnz)
        v64 = np.fromfile(f, np.float64, nnz)
        v32 = np.fromfile(f, np.float64, nnz)
    orp, oci = _pattern_oracle_mr(mesh)
    assert np.array_equal(rp.astype(np.int64), orp + 1) and np.array_equal(ci, oci + 1)
    oR = orc.mr_residual(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, c10, c01)
    oJ = orc.mr_jacobian(mesh.elem_nodes, mesh.elem_nodes_l, mesh.coords, d, c10, c01, orp, oci)
    assert rel(res, oR) < FP64_TOL
    assert rel(v64, oJ) < FP64_TOL
    assert rel(v32, oJ) < FP32_TOL
