!> ISO_C_BINDING interface of liboc_asm (the C ABI in oc_asm.h) for a Fortran
!> host such as OpenCMISS-Iron.  A thin replacement of EQUATIONS_SET_ASSEMBLE,
!> EQUATIONS_SET_RESIDUAL_EVALUATE and EQUATIONS_SET_JACOBIAN_EVALUATE
!> (equations_set_routines.f90:981, 5184, 4694) calls these and maps a nonzero
!> status to FlagError (base_routines.f90:480-520).
MODULE OC_ASM
  USE ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(C_INT), PARAMETER :: OC_NO_ERROR = 0
  INTEGER(C_INT), PARAMETER :: OC_ERR_INVALID_HANDLE = -1
  INTEGER(C_INT), PARAMETER :: OC_ERR_INVALID_ARGUMENT = -2
  INTEGER(C_INT), PARAMETER :: OC_ERR_HIP = -3
  INTEGER(C_INT), PARAMETER :: OC_ERR_ALLOCATION = -4
  INTEGER(C_INT), PARAMETER :: OC_ERR_NOT_CONVERGED = -5
  INTEGER(C_INT), PARAMETER :: OC_ERR_PATTERN = -6
  INTEGER(C_INT), PARAMETER :: OC_ERR_NOT_INITIALISED = -7
  INTEGER(C_INT), PARAMETER :: OC_ERR_UNSUPPORTED = -8
  INTEGER(C_INT), PARAMETER :: OC_FP64 = 0, OC_FP32 = 1, OC_BF16 = 2, OC_BF16_HIDDEN = 3
  INTEGER(C_INT), PARAMETER :: OC_MOONEY_RIVLIN_SUBTYPE = 1, OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE = 21
  INTEGER(C_INT), PARAMETER :: OC_EQUATIONS_SET_LAPLACE = 1, OC_EQUATIONS_SET_MOONEY_RIVLIN = 2
  INTEGER(C_INT), PARAMETER :: OC_BUFFER_MATRIX_VALUES = 0, OC_BUFFER_MATRIX_VALUES_SP = 1, &
      & OC_BUFFER_ROW_PTR = 2, OC_BUFFER_COL_IDX = 3, OC_BUFFER_DOFS = 4, OC_BUFFER_RESIDUAL = 5, OC_BUFFER_RHS = 6, &
      & OC_BUFFER_SOLVER_VALUES = 7, OC_BUFFER_SOLVER_VALUES_SP = 8, OC_BUFFER_SOLVER_ROW_PTR = 9, &
      & OC_BUFFER_SOLVER_COL_IDX = 10, OC_BUFFER_SOLVER_VECTOR = 11
  INTEGER(C_INT), PARAMETER :: OC_PRECONDITIONER_JACOBI = 0, OC_PRECONDITIONER_BLOCK_DIAGONAL = 1, &
      & OC_PRECONDITIONER_BLOCK_TRIANGULAR = 2, OC_PRECONDITIONER_MULTIGRID = 3
  INTEGER(C_INT), PARAMETER :: OC_CONVERGED_FNORM_ABS = 2, OC_CONVERGED_FNORM_RELATIVE = 3, &
      & OC_CONVERGED_SNORM_RELATIVE = 4, OC_DIVERGED_FUNCTION_COUNT = -2, OC_DIVERGED_LINEAR_SOLVE = -3, &
      & OC_DIVERGED_FNORM_NAN = -4, OC_DIVERGED_MAX_IT = -5, OC_DIVERGED_LINE_SEARCH = -6
  INTEGER(C_INT), PARAMETER :: OC_SOLVER_NO_OUTPUT = 0, OC_SOLVER_PROGRESS_OUTPUT = 1
  INTEGER(C_INT), PARAMETER :: OC_NOALLOC_MAX_ELEMENTS = 47000

  !> oc_newton_options (field order and types as in oc_asm.h)
  TYPE, BIND(C) :: OC_NEWTON_OPTIONS
    REAL(C_DOUBLE) :: absolute_tolerance, relative_tolerance, reference_norm
    INTEGER(C_INT) :: maximum_iterations, jacobian_precision, krylov_restart, krylov_maximum_iterations
    REAL(C_DOUBLE) :: krylov_relative_tolerance
    INTEGER(C_INT) :: line_search
    REAL(C_DOUBLE) :: line_search_alpha
    INTEGER(C_INT) :: preconditioner
    REAL(C_DOUBLE) :: solution_tolerance
    INTEGER(C_INT) :: maximum_function_evaluations
    INTEGER(C_INT) :: maximum_linear_solve_failures, maximum_step_failures, output_type
    INTEGER(C_INT) :: fallback, fallback_maximum_iterations
  END TYPE OC_NEWTON_OPTIONS

  !> oc_newton_stats
  TYPE, BIND(C) :: OC_NEWTON_STATS
    INTEGER(C_INT) :: converged, newton_iterations, krylov_iterations, residual_evaluations, jacobian_evaluations
    REAL(C_DOUBLE) :: initial_residual_norm, final_residual_norm, seconds
    INTEGER(C_INT) :: reason
    INTEGER(C_INT) :: linear_solve_failures, step_failures, fallback_used, warm_start_reason
  END TYPE OC_NEWTON_STATS

  INTERFACE
    INTEGER(C_INT) FUNCTION oc_initialise(device) BIND(C, NAME="oc_initialise")
      IMPORT :: C_INT
      INTEGER(C_INT), VALUE :: device
    END FUNCTION oc_initialise

    INTEGER(C_INT) FUNCTION oc_finalise() BIND(C, NAME="oc_finalise")
      IMPORT :: C_INT
    END FUNCTION oc_finalise

    INTEGER(C_INT) FUNCTION oc_last_error(buffer, buffer_size) BIND(C, NAME="oc_last_error")
      IMPORT :: C_INT, C_CHAR
      CHARACTER(KIND=C_CHAR), DIMENSION(*) :: buffer
      INTEGER(C_INT), VALUE :: buffer_size
    END FUNCTION oc_last_error

    INTEGER(C_INT) FUNCTION oc_mesh_matrix_get(handle, mesh, precision, index_base, n_rows, nnz, row_ptr, col_idx, &
        & values) BIND(C, NAME="oc_mesh_matrix_get")
      IMPORT :: C_INT, C_LONG_LONG, C_PTR
      TYPE(C_PTR), VALUE :: handle, row_ptr, col_idx, values
      INTEGER(C_INT), VALUE :: mesh, precision, index_base
      INTEGER(C_INT), INTENT(OUT) :: n_rows
      INTEGER(C_LONG_LONG), INTENT(OUT) :: nnz
    END FUNCTION oc_mesh_matrix_get

    INTEGER(C_INT) FUNCTION oc_residual_evaluate_noalloc(handle, dofs_device, residual_device) &
        & BIND(C, NAME="oc_residual_evaluate_noalloc")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle, dofs_device, residual_device
    END FUNCTION oc_residual_evaluate_noalloc

    INTEGER(C_INT) FUNCTION oc_equations_set_subtype_set(handle, subtype) BIND(C, NAME="oc_equations_set_subtype_set")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: subtype
    END FUNCTION oc_equations_set_subtype_set

    INTEGER(C_INT) FUNCTION oc_laplace_create(n_nodes, coords, n_elements, elem_nodes, index_base, handle) &
        & BIND(C, NAME="oc_laplace_create")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      INTEGER(C_INT), VALUE :: n_nodes, n_elements, index_base
      REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: coords
      INTEGER(C_INT), DIMENSION(*), INTENT(IN) :: elem_nodes
      TYPE(C_PTR), INTENT(OUT) :: handle
    END FUNCTION oc_laplace_create

    INTEGER(C_INT) FUNCTION oc_mooney_rivlin_create(n_nodes_q, coords_q, n_nodes_l, n_elements, elem_nodes_q, &
        & elem_nodes_l, index_base, c10, c01, handle) BIND(C, NAME="oc_mooney_rivlin_create")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      INTEGER(C_INT), VALUE :: n_nodes_q, n_nodes_l, n_elements, index_base
      REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: coords_q
      INTEGER(C_INT), DIMENSION(*), INTENT(IN) :: elem_nodes_q, elem_nodes_l
      REAL(C_DOUBLE), VALUE :: c10, c01
      TYPE(C_PTR), INTENT(OUT) :: handle
    END FUNCTION oc_mooney_rivlin_create

    INTEGER(C_INT) FUNCTION oc_equations_set_destroy(handle) BIND(C, NAME="oc_equations_set_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
    END FUNCTION oc_equations_set_destroy

    INTEGER(C_INT) FUNCTION oc_equations_set_info(handle, set_class, n_rows, nnz, n_elements, n_colours) &
        & BIND(C, NAME="oc_equations_set_info")
      IMPORT :: C_INT, C_LONG_LONG, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), INTENT(OUT) :: set_class, n_rows, n_elements, n_colours
      INTEGER(C_LONG_LONG), INTENT(OUT) :: nnz
    END FUNCTION oc_equations_set_info

    INTEGER(C_INT) FUNCTION oc_matrix_structure_get(handle, index_base, row_ptr, col_idx) &
        & BIND(C, NAME="oc_matrix_structure_get")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: index_base
      INTEGER(C_INT), DIMENSION(*), INTENT(OUT) :: row_ptr, col_idx
    END FUNCTION oc_matrix_structure_get

    INTEGER(C_INT) FUNCTION oc_dofs_set(handle, dofs) BIND(C, NAME="oc_dofs_set")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: dofs
    END FUNCTION oc_dofs_set

    INTEGER(C_INT) FUNCTION oc_dofs_get(handle, dofs) BIND(C, NAME="oc_dofs_get")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: dofs
    END FUNCTION oc_dofs_get

    INTEGER(C_INT) FUNCTION oc_equations_set_assemble(handle) BIND(C, NAME="oc_equations_set_assemble")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
    END FUNCTION oc_equations_set_assemble

    INTEGER(C_INT) FUNCTION oc_laplace_fused_set(handle, fused, active) BIND(C, NAME="oc_laplace_fused_set")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: fused
      INTEGER(C_INT), INTENT(OUT) :: active
    END FUNCTION oc_laplace_fused_set

    INTEGER(C_INT) FUNCTION oc_equations_set_residual_evaluate(handle) &
        & BIND(C, NAME="oc_equations_set_residual_evaluate")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
    END FUNCTION oc_equations_set_residual_evaluate

    INTEGER(C_INT) FUNCTION oc_equations_set_jacobian_evaluate(handle, precision) &
        & BIND(C, NAME="oc_equations_set_jacobian_evaluate")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: precision
    END FUNCTION oc_equations_set_jacobian_evaluate

    INTEGER(C_INT) FUNCTION oc_matrix_values_get(handle, values) BIND(C, NAME="oc_matrix_values_get")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: values
    END FUNCTION oc_matrix_values_get

    INTEGER(C_INT) FUNCTION oc_residual_get(handle, residual) BIND(C, NAME="oc_residual_get")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: residual
    END FUNCTION oc_residual_get

    INTEGER(C_INT) FUNCTION oc_rhs_get(handle, rhs) BIND(C, NAME="oc_rhs_get")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: rhs
    END FUNCTION oc_rhs_get

    ! ---- boundary conditions, load steps, materials --------------------------
    INTEGER(C_INT) FUNCTION oc_boundary_conditions_set(handle, n, dof_indices, full_values, index_base) &
        & BIND(C, NAME="oc_boundary_conditions_set")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: n, index_base
      INTEGER(C_INT), DIMENSION(*), INTENT(IN) :: dof_indices
      REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: full_values
    END FUNCTION oc_boundary_conditions_set

    INTEGER(C_INT) FUNCTION oc_load_increment_apply(handle, iteration, max_iterations) &
        & BIND(C, NAME="oc_load_increment_apply")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: iteration, max_iterations
    END FUNCTION oc_load_increment_apply

    INTEGER(C_INT) FUNCTION oc_materials_set(handle, c10, c01) BIND(C, NAME="oc_materials_set")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), VALUE :: c10, c01
    END FUNCTION oc_materials_set

    INTEGER(C_INT) FUNCTION oc_equations_set_jacobian_evaluate_fd(handle) &
        & BIND(C, NAME="oc_equations_set_jacobian_evaluate_fd")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
    END FUNCTION oc_equations_set_jacobian_evaluate_fd

    INTEGER(C_INT) FUNCTION oc_matrix_values_get_sp(handle, values) BIND(C, NAME="oc_matrix_values_get_sp")
      IMPORT :: C_INT, C_FLOAT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_FLOAT), DIMENSION(*), INTENT(OUT) :: values
    END FUNCTION oc_matrix_values_get_sp

    INTEGER(C_INT) FUNCTION oc_device_pointer_get(handle, buffer, device_pointer) &
        & BIND(C, NAME="oc_device_pointer_get")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: buffer
      TYPE(C_PTR), INTENT(OUT) :: device_pointer
    END FUNCTION oc_device_pointer_get

    ! solver matrices (SOLVER_MATRICES_STATIC_ASSEMBLE, solver_routines.f90:13745-14771)
    INTEGER(C_INT) FUNCTION oc_solver_matrices_assemble(handle, precision) &
        & BIND(C, NAME="oc_solver_matrices_assemble")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: precision
    END FUNCTION oc_solver_matrices_assemble

    INTEGER(C_INT) FUNCTION oc_solver_matrix_info(handle, n_rows, nnz) BIND(C, NAME="oc_solver_matrix_info")
      IMPORT :: C_INT, C_LONG_LONG, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), INTENT(OUT) :: n_rows
      INTEGER(C_LONG_LONG), INTENT(OUT) :: nnz
    END FUNCTION oc_solver_matrix_info

    INTEGER(C_INT) FUNCTION oc_solver_matrix_structure_get(handle, index_base, row_ptr, col_idx) &
        & BIND(C, NAME="oc_solver_matrix_structure_get")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: index_base
      INTEGER(C_INT), DIMENSION(*), INTENT(OUT) :: row_ptr, col_idx
    END FUNCTION oc_solver_matrix_structure_get

    INTEGER(C_INT) FUNCTION oc_solver_matrix_values_get(handle, values) BIND(C, NAME="oc_solver_matrix_values_get")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: values
    END FUNCTION oc_solver_matrix_values_get

    INTEGER(C_INT) FUNCTION oc_solver_matrix_values_get_sp(handle, values) &
        & BIND(C, NAME="oc_solver_matrix_values_get_sp")
      IMPORT :: C_INT, C_FLOAT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_FLOAT), DIMENSION(*), INTENT(OUT) :: values
    END FUNCTION oc_solver_matrix_values_get_sp

    INTEGER(C_INT) FUNCTION oc_solver_vector_get(handle, vector) BIND(C, NAME="oc_solver_vector_get")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: vector
    END FUNCTION oc_solver_vector_get

    INTEGER(C_INT) FUNCTION oc_dofs_set_device(handle, dofs_device) BIND(C, NAME="oc_dofs_set_device")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle, dofs_device
    END FUNCTION oc_dofs_set_device

    INTEGER(C_INT) FUNCTION oc_device_get(device) BIND(C, NAME="oc_device_get")
      IMPORT :: C_INT
      INTEGER(C_INT), INTENT(OUT) :: device
    END FUNCTION oc_device_get

    INTEGER(C_INT) FUNCTION oc_synchronise() BIND(C, NAME="oc_synchronise")
      IMPORT :: C_INT
    END FUNCTION oc_synchronise

    ! ---- solver (SOLVER_NEWTON_LINESEARCH_SOLVE analogue) ---------------------
    INTEGER(C_INT) FUNCTION oc_newton_options_default(options) BIND(C, NAME="oc_newton_options_default")
      IMPORT :: C_INT, OC_NEWTON_OPTIONS
      TYPE(OC_NEWTON_OPTIONS), INTENT(OUT) :: options
    END FUNCTION oc_newton_options_default

    INTEGER(C_INT) FUNCTION oc_newton_solve(handle, options, stats) BIND(C, NAME="oc_newton_solve")
      IMPORT :: C_INT, C_PTR, OC_NEWTON_OPTIONS, OC_NEWTON_STATS
      TYPE(C_PTR), VALUE :: handle
      TYPE(OC_NEWTON_OPTIONS), INTENT(IN) :: options
      TYPE(OC_NEWTON_STATS), INTENT(OUT) :: stats
    END FUNCTION oc_newton_solve

    INTEGER(C_INT) FUNCTION oc_linear_solve(handle, options, stats) BIND(C, NAME="oc_linear_solve")
      IMPORT :: C_INT, C_PTR, OC_NEWTON_OPTIONS, OC_NEWTON_STATS
      TYPE(C_PTR), VALUE :: handle
      TYPE(OC_NEWTON_OPTIONS), INTENT(IN) :: options
      TYPE(OC_NEWTON_STATS), INTENT(OUT) :: stats
    END FUNCTION oc_linear_solve

    INTEGER(C_INT) FUNCTION oc_residual_norm(handle, norm) BIND(C, NAME="oc_residual_norm")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), INTENT(OUT) :: norm
    END FUNCTION oc_residual_norm

    ! ---- batches of meshes (no counterpart in the reference) ------------------
    INTEGER(C_INT) FUNCTION oc_mooney_rivlin_batch_create(n_meshes, n_nodes_q, n_nodes_l, n_elements, coords_q, &
        & elem_nodes_q, elem_nodes_l, c10, c01, index_base, handle) BIND(C, NAME="oc_mooney_rivlin_batch_create")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      INTEGER(C_INT), VALUE :: n_meshes, index_base
      INTEGER(C_INT), DIMENSION(*), INTENT(IN) :: n_nodes_q, n_nodes_l, n_elements, elem_nodes_q, elem_nodes_l
      REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: coords_q, c10, c01
      TYPE(C_PTR), INTENT(OUT) :: handle
    END FUNCTION oc_mooney_rivlin_batch_create

    INTEGER(C_INT) FUNCTION oc_mesh_count(handle, n_meshes) BIND(C, NAME="oc_mesh_count")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), INTENT(OUT) :: n_meshes
    END FUNCTION oc_mesh_count

    INTEGER(C_INT) FUNCTION oc_equations_set_mesh_get(handle, mesh, node_offset_q, n_nodes_q, node_offset_l, &
        & n_nodes_l, element_offset, n_elements) BIND(C, NAME="oc_equations_set_mesh_get")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: mesh
      INTEGER(C_INT), INTENT(OUT) :: node_offset_q, n_nodes_q, node_offset_l, n_nodes_l, element_offset, n_elements
    END FUNCTION oc_equations_set_mesh_get

    INTEGER(C_INT) FUNCTION oc_mesh_materials_set(handle, c10, c01) BIND(C, NAME="oc_mesh_materials_set")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: c10, c01
    END FUNCTION oc_mesh_materials_set

    INTEGER(C_INT) FUNCTION oc_mesh_residual_norms(handle, norms) BIND(C, NAME="oc_mesh_residual_norms")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: norms
    END FUNCTION oc_mesh_residual_norms

    INTEGER(C_INT) FUNCTION oc_newton_solve_batch(handle, options, reference_norms, final_norms, stats) &
        & BIND(C, NAME="oc_newton_solve_batch")
      IMPORT :: C_INT, C_DOUBLE, C_PTR, OC_NEWTON_OPTIONS, OC_NEWTON_STATS
      TYPE(C_PTR), VALUE :: handle
      TYPE(OC_NEWTON_OPTIONS), INTENT(IN) :: options
      REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: reference_norms
      REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: final_norms
      TYPE(OC_NEWTON_STATS), INTENT(OUT) :: stats
    END FUNCTION oc_newton_solve_batch

    INTEGER(C_INT) FUNCTION oc_jacobian_evaluate_graph(handle, precision) BIND(C, NAME="oc_jacobian_evaluate_graph")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: precision
    END FUNCTION oc_jacobian_evaluate_graph

    INTEGER(C_INT) FUNCTION oc_autotune(handle, precision, report, report_size) BIND(C, NAME="oc_autotune")
      IMPORT :: C_INT, C_CHAR, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: precision, report_size
      CHARACTER(KIND=C_CHAR), DIMENSION(*) :: report
    END FUNCTION oc_autotune

    INTEGER(C_INT) FUNCTION oc_multigrid_structure_set(handle, n_per_mesh) BIND(C, NAME="oc_multigrid_structure_set")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), DIMENSION(*), INTENT(IN) :: n_per_mesh
    END FUNCTION oc_multigrid_structure_set

    INTEGER(C_INT) FUNCTION oc_multigrid_options_set(handle, pre_sweeps, post_sweeps, omega) &
        & BIND(C, NAME="oc_multigrid_options_set")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT), VALUE :: pre_sweeps, post_sweeps
      REAL(C_DOUBLE), VALUE :: omega
    END FUNCTION oc_multigrid_options_set

    ! ---- measurement ----------------------------------------------------------
    INTEGER(C_INT) FUNCTION oc_event_record(slot) BIND(C, NAME="oc_event_record")
      IMPORT :: C_INT
      INTEGER(C_INT), VALUE :: slot
    END FUNCTION oc_event_record

    INTEGER(C_INT) FUNCTION oc_event_elapsed(start, end, milliseconds) BIND(C, NAME="oc_event_elapsed")
      IMPORT :: C_INT, C_FLOAT
      INTEGER(C_INT), VALUE :: start, end
      REAL(C_FLOAT), INTENT(OUT) :: milliseconds
    END FUNCTION oc_event_elapsed

    INTEGER(C_INT) FUNCTION oc_kernel_timing_enable(enable) BIND(C, NAME="oc_kernel_timing_enable")
      IMPORT :: C_INT
      INTEGER(C_INT), VALUE :: enable
    END FUNCTION oc_kernel_timing_enable

    INTEGER(C_INT) FUNCTION oc_kernel_timing_select(kernel_name) BIND(C, NAME="oc_kernel_timing_select")
      IMPORT :: C_INT, C_CHAR
      CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: kernel_name
    END FUNCTION oc_kernel_timing_select

    INTEGER(C_INT) FUNCTION oc_kernel_timing_get(kernel_name, total_ms, launches) BIND(C, NAME="oc_kernel_timing_get")
      IMPORT :: C_INT, C_CHAR, C_DOUBLE, C_LONG_LONG
      CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: kernel_name
      REAL(C_DOUBLE), INTENT(OUT) :: total_ms
      INTEGER(C_LONG_LONG), INTENT(OUT) :: launches
    END FUNCTION oc_kernel_timing_get

    ! ---- learned warm start (no counterpart in the reference) ------------------
    INTEGER(C_INT) FUNCTION oc_mlp_create(n_inputs, n_hidden, n_hidden_layers, n_outputs, seed, mlp) &
        & BIND(C, NAME="oc_mlp_create")
      IMPORT :: C_INT, C_LONG_LONG, C_PTR
      INTEGER(C_INT), VALUE :: n_inputs, n_hidden, n_hidden_layers, n_outputs
      INTEGER(C_LONG_LONG), VALUE :: seed
      TYPE(C_PTR), INTENT(OUT) :: mlp
    END FUNCTION oc_mlp_create

    INTEGER(C_INT) FUNCTION oc_mlp_destroy(mlp) BIND(C, NAME="oc_mlp_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: mlp
    END FUNCTION oc_mlp_destroy

    INTEGER(C_INT) FUNCTION oc_mlp_parameter_count(mlp, count) BIND(C, NAME="oc_mlp_parameter_count")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: mlp
      INTEGER(C_INT), INTENT(OUT) :: count
    END FUNCTION oc_mlp_parameter_count

    INTEGER(C_INT) FUNCTION oc_mlp_parameters_set(mlp, parameters) BIND(C, NAME="oc_mlp_parameters_set")
      IMPORT :: C_INT, C_FLOAT, C_PTR
      TYPE(C_PTR), VALUE :: mlp
      REAL(C_FLOAT), DIMENSION(*), INTENT(IN) :: parameters
    END FUNCTION oc_mlp_parameters_set

    INTEGER(C_INT) FUNCTION oc_mlp_parameters_get(mlp, parameters) BIND(C, NAME="oc_mlp_parameters_get")
      IMPORT :: C_INT, C_FLOAT, C_PTR
      TYPE(C_PTR), VALUE :: mlp
      REAL(C_FLOAT), DIMENSION(*), INTENT(OUT) :: parameters
    END FUNCTION oc_mlp_parameters_get

    INTEGER(C_INT) FUNCTION oc_mlp_normalisation_set(mlp, shift, scale) BIND(C, NAME="oc_mlp_normalisation_set")
      IMPORT :: C_INT, C_FLOAT, C_PTR
      TYPE(C_PTR), VALUE :: mlp
      REAL(C_FLOAT), DIMENSION(*), INTENT(IN) :: shift, scale
    END FUNCTION oc_mlp_normalisation_set

    INTEGER(C_INT) FUNCTION oc_mlp_forward(mlp, n, x, y) BIND(C, NAME="oc_mlp_forward")
      IMPORT :: C_INT, C_FLOAT, C_PTR
      TYPE(C_PTR), VALUE :: mlp
      INTEGER(C_INT), VALUE :: n
      REAL(C_FLOAT), DIMENSION(*), INTENT(IN) :: x
      REAL(C_FLOAT), DIMENSION(*), INTENT(OUT) :: y
    END FUNCTION oc_mlp_forward

    INTEGER(C_INT) FUNCTION oc_mlp_inference_precision_set(mlp, precision) &
        BIND(C, NAME="oc_mlp_inference_precision_set")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: mlp
      INTEGER(C_INT), VALUE :: precision
    END FUNCTION oc_mlp_inference_precision_set

    INTEGER(C_INT) FUNCTION oc_mlp_gradient(mlp, n, x, t, loss, gradient) BIND(C, NAME="oc_mlp_gradient")
      IMPORT :: C_INT, C_FLOAT, C_PTR
      TYPE(C_PTR), VALUE :: mlp
      INTEGER(C_INT), VALUE :: n
      REAL(C_FLOAT), DIMENSION(*), INTENT(IN) :: x, t
      REAL(C_FLOAT), INTENT(OUT) :: loss
      REAL(C_FLOAT), DIMENSION(*), INTENT(OUT) :: gradient
    END FUNCTION oc_mlp_gradient

    INTEGER(C_INT) FUNCTION oc_mlp_dataset_set(mlp, n, x, t) BIND(C, NAME="oc_mlp_dataset_set")
      IMPORT :: C_INT, C_FLOAT, C_PTR
      TYPE(C_PTR), VALUE :: mlp
      INTEGER(C_INT), VALUE :: n
      REAL(C_FLOAT), DIMENSION(*), INTENT(IN) :: x, t
    END FUNCTION oc_mlp_dataset_set

    INTEGER(C_INT) FUNCTION oc_mlp_train(mlp, steps, batch, learning_rate, seed, mean_loss) &
        & BIND(C, NAME="oc_mlp_train")
      IMPORT :: C_INT, C_FLOAT, C_LONG_LONG, C_PTR
      TYPE(C_PTR), VALUE :: mlp
      INTEGER(C_INT), VALUE :: steps, batch
      REAL(C_FLOAT), VALUE :: learning_rate
      INTEGER(C_LONG_LONG), VALUE :: seed
      REAL(C_FLOAT), INTENT(OUT) :: mean_loss
    END FUNCTION oc_mlp_train

    INTEGER(C_INT) FUNCTION oc_mlp_warm_start(mlp, handle, load_increment) BIND(C, NAME="oc_mlp_warm_start")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: mlp, handle
      REAL(C_DOUBLE), VALUE :: load_increment
    END FUNCTION oc_mlp_warm_start

    INTEGER(C_INT) FUNCTION oc_mlp_warm_start_batch(mlp, handle, load_increments) &
        & BIND(C, NAME="oc_mlp_warm_start_batch")
      IMPORT :: C_INT, C_DOUBLE, C_PTR
      TYPE(C_PTR), VALUE :: mlp, handle
      REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: load_increments
    END FUNCTION oc_mlp_warm_start_batch
  END INTERFACE

CONTAINS

  !> Copies the library's last error message into a Fortran string.
  SUBROUTINE OC_ERROR_MESSAGE(message)
    CHARACTER(LEN=*), INTENT(OUT) :: message
    CHARACTER(KIND=C_CHAR), DIMENSION(1024) :: buffer
    INTEGER(C_INT) :: rc
    INTEGER :: i
    rc = oc_last_error(buffer, 1024_C_INT)
    message = ""
    DO i = 1, MIN(LEN(message), 1024)
      IF (buffer(i) == C_NULL_CHAR) EXIT
      message(i:i) = buffer(i)
    END DO
  END SUBROUTINE OC_ERROR_MESSAGE

END MODULE OC_ASM
