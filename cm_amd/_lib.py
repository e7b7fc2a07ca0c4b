"""ctypes binding of liboc_asm.so (the C ABI declared in include/oc_asm.h).

Status codes become :class:`CMISSError`, as the reference's Python package
turns a nonzero ``cmfe`` status into an exception
(bindings/python/opencmiss/_iron_utils.py:23-78).  The library must be the
in-tree build; there is no fallback implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboc_asm.so")

OC_NO_ERROR = 0
OC_ERR_INVALID_HANDLE = -1
OC_ERR_INVALID_ARGUMENT = -2
OC_ERR_HIP = -3
OC_ERR_ALLOCATION = -4
OC_ERR_NOT_CONVERGED = -5
OC_ERR_PATTERN = -6
OC_ERR_NOT_INITIALISED = -7
OC_ERR_UNSUPPORTED = -8

OC_EQUATIONS_SET_LAPLACE = 1
OC_EQUATIONS_SET_MOONEY_RIVLIN = 2
OC_MOONEY_RIVLIN_SUBTYPE = 1
OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE = 21
OC_FP64 = 0
OC_FP32 = 1
OC_BF16 = 2  # MLP inference only
OC_BF16_HIDDEN = 3  # MLP inference only: bf16 hidden layers, FP32 first and last layers

OC_BUFFER_MATRIX_VALUES = 0
OC_BUFFER_MATRIX_VALUES_SP = 1
OC_BUFFER_ROW_PTR = 2
OC_BUFFER_COL_IDX = 3
OC_BUFFER_DOFS = 4
OC_BUFFER_RESIDUAL = 5
OC_BUFFER_RHS = 6
OC_BUFFER_SOLVER_VALUES = 7
OC_BUFFER_SOLVER_VALUES_SP = 8
OC_BUFFER_SOLVER_ROW_PTR = 9
OC_BUFFER_SOLVER_COL_IDX = 10
OC_BUFFER_SOLVER_VECTOR = 11


class CMISSError(RuntimeError):
    """Raised when a liboc_asm entry point returns a nonzero status."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


OC_PRECONDITIONER_JACOBI = 0
OC_PRECONDITIONER_BLOCK_DIAGONAL = 1
OC_PRECONDITIONER_BLOCK_TRIANGULAR = 2
OC_PRECONDITIONER_MULTIGRID = 3

OC_CONVERGED_FNORM_ABS = 2
OC_CONVERGED_FNORM_RELATIVE = 3
OC_CONVERGED_SNORM_RELATIVE = 4
OC_DIVERGED_FUNCTION_COUNT = -2
OC_DIVERGED_LINEAR_SOLVE = -3
OC_DIVERGED_FNORM_NAN = -4
OC_DIVERGED_MAX_IT = -5
OC_DIVERGED_LINE_SEARCH = -6
OC_SOLVER_NO_OUTPUT = 0
OC_SOLVER_PROGRESS_OUTPUT = 1


class NewtonOptions(ctypes.Structure):
    """oc_newton_options (include/oc_asm.h)."""
    _fields_ = [
        ("absolute_tolerance", ctypes.c_double),
        ("relative_tolerance", ctypes.c_double),
        ("reference_norm", ctypes.c_double),
        ("maximum_iterations", ctypes.c_int),
        ("jacobian_precision", ctypes.c_int),
        ("krylov_restart", ctypes.c_int),
        ("krylov_maximum_iterations", ctypes.c_int),
        ("krylov_relative_tolerance", ctypes.c_double),
        ("line_search", ctypes.c_int),
        ("line_search_alpha", ctypes.c_double),
        ("preconditioner", ctypes.c_int),
        ("solution_tolerance", ctypes.c_double),
        ("maximum_function_evaluations", ctypes.c_int),
        ("maximum_linear_solve_failures", ctypes.c_int),
        ("maximum_step_failures", ctypes.c_int),
        ("output_type", ctypes.c_int),
        ("fallback", ctypes.c_int),
        ("fallback_maximum_iterations", ctypes.c_int),
    ]


class NewtonStats(ctypes.Structure):
    """oc_newton_stats (include/oc_asm.h)."""
    _fields_ = [
        ("converged", ctypes.c_int),
        ("newton_iterations", ctypes.c_int),
        ("krylov_iterations", ctypes.c_int),
        ("residual_evaluations", ctypes.c_int),
        ("jacobian_evaluations", ctypes.c_int),
        ("initial_residual_norm", ctypes.c_double),
        ("final_residual_norm", ctypes.c_double),
        ("seconds", ctypes.c_double),
        ("reason", ctypes.c_int),
        ("linear_solve_failures", ctypes.c_int),
        ("step_failures", ctypes.c_int),
        ("fallback_used", ctypes.c_int),
        ("warm_start_reason", ctypes.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_H = ctypes.c_void_p
_c_int, _c_double, _c_float = ctypes.c_int, ctypes.c_double, ctypes.c_float

# name -> argtypes (all return int status)
SIGNATURES = {
    "oc_initialise": [_c_int],
    "oc_finalise": [],
    "oc_last_error": [ctypes.c_char_p, _c_int],
    "oc_device_get": [ctypes.POINTER(_c_int)],
    "oc_laplace_create": [_c_int, _dp, _c_int, _i32, _c_int, ctypes.POINTER(_H)],
    "oc_mooney_rivlin_create": [_c_int, _dp, _c_int, _c_int, _i32, _i32, _c_int, _c_double, _c_double,
                                ctypes.POINTER(_H)],
    "oc_equations_set_destroy": [_H],
    "oc_equations_set_info": [_H, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), ctypes.POINTER(ctypes.c_longlong),
                              ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)],
    "oc_materials_set": [_H, _c_double, _c_double],
    "oc_matrix_structure_get": [_H, _c_int, _i32, _i32],
    "oc_dofs_set": [_H, _dp],
    "oc_dofs_get": [_H, _dp],
    "oc_dofs_set_device": [_H, ctypes.c_void_p],
    "oc_equations_set_assemble": [_H],
    "oc_laplace_fused_set": [_H, _c_int, ctypes.POINTER(_c_int)],
    "oc_equations_set_residual_evaluate": [_H],
    "oc_equations_set_jacobian_evaluate": [_H, _c_int],
    "oc_equations_set_jacobian_evaluate_fd": [_H],
    "oc_matrix_values_get": [_H, _dp],
    "oc_matrix_values_get_sp": [_H, _fp],
    "oc_residual_get": [_H, _dp],
    "oc_rhs_get": [_H, _dp],
    "oc_device_pointer_get": [_H, _c_int, ctypes.POINTER(ctypes.c_void_p)],
    "oc_solver_matrices_assemble": [_H, _c_int],
    "oc_solver_matrix_info": [_H, ctypes.POINTER(_c_int), ctypes.POINTER(ctypes.c_longlong)],
    "oc_solver_matrix_structure_get": [_H, _c_int, _i32, _i32],
    "oc_solver_matrix_values_get": [_H, _dp],
    "oc_solver_matrix_values_get_sp": [_H, _fp],
    "oc_solver_vector_get": [_H, _dp],
    "oc_event_record": [_c_int],
    "oc_event_elapsed": [_c_int, _c_int, ctypes.POINTER(_c_float)],
    "oc_synchronise": [],
    "oc_boundary_conditions_set": [_H, _c_int, _i32, _dp, _c_int],
    "oc_load_increment_apply": [_H, _c_int, _c_int],
    "oc_newton_options_default": [ctypes.POINTER(NewtonOptions)],
    "oc_newton_solve": [_H, ctypes.POINTER(NewtonOptions), ctypes.POINTER(NewtonStats)],
    "oc_linear_solve": [_H, ctypes.POINTER(NewtonOptions), ctypes.POINTER(NewtonStats)],
    "oc_residual_norm": [_H, ctypes.POINTER(ctypes.c_double)],
    "oc_mlp_create": [_c_int, _c_int, _c_int, _c_int, ctypes.c_ulonglong, ctypes.POINTER(_H)],
    "oc_mlp_destroy": [_H],
    "oc_mlp_parameter_count": [_H, ctypes.POINTER(_c_int)],
    "oc_mlp_parameters_get": [_H, _fp],
    "oc_mlp_parameters_set": [_H, _fp],
    "oc_mlp_normalisation_set": [_H, _fp, _fp],
    "oc_mlp_forward": [_H, _c_int, _fp, _fp],
    "oc_mlp_inference_precision_set": [_H, _c_int],
    "oc_mlp_gradient": [_H, _c_int, _fp, _fp, ctypes.POINTER(_c_float), _fp],
    "oc_mlp_dataset_set": [_H, _c_int, _fp, _fp],
    "oc_mlp_train": [_H, _c_int, _c_int, _c_float, ctypes.c_ulonglong, ctypes.POINTER(_c_float)],
    "oc_mlp_warm_start": [_H, _H, _c_double],
    "oc_mlp_warm_start_batch": [_H, _H, _dp],
    "oc_mooney_rivlin_batch_create": [_c_int, _i32, _i32, _i32, _dp, _i32, _i32, _dp, _dp, _c_int,
                                      ctypes.POINTER(_H)],
    "oc_mesh_count": [_H, ctypes.POINTER(_c_int)],
    "oc_equations_set_subtype_set": [_H, _c_int],
    "oc_residual_evaluate_noalloc": [_H, ctypes.c_void_p, ctypes.c_void_p],
    "oc_mesh_matrix_get": [_H, _c_int, _c_int, _c_int, ctypes.POINTER(_c_int), ctypes.POINTER(ctypes.c_longlong),
                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "oc_equations_set_mesh_get": [_H, _c_int] + [ctypes.POINTER(_c_int)] * 6,
    "oc_mesh_materials_set": [_H, _dp, _dp],
    "oc_newton_solve_batch": [_H, ctypes.POINTER(NewtonOptions), ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.POINTER(NewtonStats)],
    "oc_mesh_residual_norms": [_H, _dp],
    "oc_autotune": [_H, _c_int, ctypes.c_char_p, _c_int],
    "oc_multigrid_structure_set": [_H, _i32],
    "oc_multigrid_options_set": [_H, _c_int, _c_int, _c_double],
    "oc_jacobian_evaluate_graph": [_H, _c_int],
    "oc_kernel_timing_enable": [_c_int],
    "oc_kernel_timing_select": [ctypes.c_char_p],
    "oc_kernel_timing_get": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)],
}

_LIB = None


def load():
    """Load the in-tree library; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not found: build it with __graft_entry__.build() "
                          "(make -C cm_amd/csrc); there is no fallback path")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = _c_int
        f.argtypes = args
    _LIB = lib
    return lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().oc_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(code):
    if code != OC_NO_ERROR:
        raise CMISSError(code, last_error())
    return code


def call(name, *args):
    return check(getattr(load(), name)(*args))
