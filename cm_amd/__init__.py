"""cm_amd -- MI355X (gfx950) finite element assembly for OpenCMISS-Iron's
Laplace and Mooney-Rivlin equations sets.

The compute path is liboc_asm.so (hand-written HIP kernels behind the C ABI
in include/oc_asm.h); this package is its Python host side.
"""
from ._lib import (OC_PRECONDITIONER_BLOCK_DIAGONAL, OC_PRECONDITIONER_BLOCK_TRIANGULAR,  # noqa: F401
                   OC_PRECONDITIONER_JACOBI, OC_PRECONDITIONER_MULTIGRID)
from ._lib import (OC_CONVERGED_FNORM_ABS, OC_CONVERGED_FNORM_RELATIVE, OC_CONVERGED_SNORM_RELATIVE,  # noqa: F401
                   OC_DIVERGED_FNORM_NAN, OC_DIVERGED_FUNCTION_COUNT, OC_DIVERGED_LINE_SEARCH,
                   OC_DIVERGED_LINEAR_SOLVE, OC_DIVERGED_MAX_IT, OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE,
                   OC_MOONEY_RIVLIN_SUBTYPE, OC_SOLVER_NO_OUTPUT, OC_SOLVER_PROGRESS_OUTPUT)
from ._lib import (CMISSError, OC_BF16, OC_BF16_HIDDEN, OC_FP32, OC_FP64, OC_BUFFER_COL_IDX, OC_BUFFER_DOFS,  # noqa: F401
                   OC_BUFFER_MATRIX_VALUES, OC_BUFFER_MATRIX_VALUES_SP, OC_BUFFER_RESIDUAL, OC_BUFFER_RHS,
                   OC_BUFFER_ROW_PTR, OC_BUFFER_SOLVER_COL_IDX, OC_BUFFER_SOLVER_ROW_PTR, OC_BUFFER_SOLVER_VALUES,
                   OC_BUFFER_SOLVER_VALUES_SP, OC_BUFFER_SOLVER_VECTOR, LIB_PATH)
from .equations_set import (EquationsSet, LaplaceEquationsSet, MooneyRivlinBatch, MooneyRivlinEquationsSet,  # noqa: F401
                            event_elapsed, event_record, finalise, initialise, kernel_timing,
                            kernel_timing_enable, newton_options, synchronise)
from .mesh import RegularMesh, UserMesh, as_user_mesh  # noqa: F401

__all__ = ["CMISSError", "EquationsSet", "LaplaceEquationsSet", "MooneyRivlinBatch", "MooneyRivlinEquationsSet",
           "RegularMesh", "UserMesh", "as_user_mesh",
           "initialise", "finalise", "synchronise", "event_record", "event_elapsed", "OC_FP32", "OC_FP64",
           "newton_options"]
