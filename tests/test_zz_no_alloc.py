"""oc_residual_evaluate_noalloc (the residual entry point that allocates
nothing) against oc_equations_set_residual_evaluate, through the C ABI, on two
calls in a row, the second on a smaller set; the caller's device arrays are
allocated here with the HIP runtime."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


H2D, D2H = 1, 2


def _deformed(mesh, seed):
    rng = np.random.default_rng(seed)
    z = mesh.coords * np.array([1.04, 0.98, 0.98])[:, None] + 0.02 / max(mesh.n) * rng.uniform(-1, 1, mesh.coords.shape)
    return np.concatenate([z.ravel(), rng.standard_normal(mesh.n_nodes_l)])


def test_noalloc_residual_matches_the_allocating_entry_point(gpu):
    cm = gpu
    from cm_amd import _lib
    lib = _lib.load()
    hip = _hip()
    for n, seed in ((12, 3), (5, 4)):  # the second call on a smaller set
        mesh = cm.RegularMesh.generate(2, n, jitter=0.1, seed=seed)
        eq = cm.MooneyRivlinEquationsSet(mesh, 2.0 + seed, 6.0 - seed)
        d = _deformed(mesh, seed)
        eq.set_dofs(d)
        eq.residual_evaluate()
        expect = eq.residual()
        nbytes = 8 * eq.n_rows
        xd, rd = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(xd), nbytes) == 0 and hip.hipMalloc(ctypes.byref(rd), nbytes) == 0
        try:
            assert hip.hipMemcpy(xd, d.ctypes.data, nbytes, H2D) == 0
            fill = np.full(eq.n_rows, np.nan)
            assert hip.hipMemcpy(rd, fill.ctypes.data, nbytes, H2D) == 0
            assert lib.oc_residual_evaluate_noalloc(eq.handle, xd, rd) == 0, _lib.last_error()
            cm.synchronise()
            got = np.empty(eq.n_rows)
            assert hip.hipMemcpy(got.ctypes.data, rd, nbytes, D2H) == 0
            np.testing.assert_array_equal(got, expect)
        finally:
            hip.hipFree(xd)
            hip.hipFree(rd)
        eq.destroy()


def test_noalloc_residual_rejects_sets_beyond_its_scratch(gpu):
    cm = gpu
    from cm_amd import _lib
    lib = _lib.load()
    mesh = cm.RegularMesh.generate(2, (40, 40, 30))  # 48,000 elements > OC_NOALLOC_MAX_ELEMENTS
    eq = cm.MooneyRivlinEquationsSet(mesh, 2.0, 6.0)
    p = eq.device_pointer(_lib.OC_BUFFER_DOFS)
    r = eq.device_pointer(_lib.OC_BUFFER_RESIDUAL)
    assert lib.oc_residual_evaluate_noalloc(eq.handle, ctypes.c_void_p(p), ctypes.c_void_p(r)) == _lib.OC_ERR_UNSUPPORTED
    assert "47000" in _lib.last_error()
    assert lib.oc_residual_evaluate_noalloc(eq.handle, None, ctypes.c_void_p(r)) == _lib.OC_ERR_INVALID_ARGUMENT
    eq.destroy()
