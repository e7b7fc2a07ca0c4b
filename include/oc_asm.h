/*
 * oc_asm.h -- C ABI of the MI355X (gfx950) finite element assembly library
 * (liboc_asm.so).  It replaces, at the assembly level, the equations-set
 * element loops of OpenCMISS-Iron for the Laplace (standard subtype) and
 * Mooney-Rivlin finite elasticity equations sets, plus the scatter-add of
 * element matrices into the CSR equations matrices.
 *
 * Conventions (mirroring the reference's generated C bindings,
 * bindings/generate_bindings/c.py:20-40, 234-590):
 *   - every entry point returns int: OC_NO_ERROR (0) or a negative code;
 *     the message of the last failure on the calling thread is available
 *     through oc_last_error (cmfe_ExtractErrorMessage analogue,
 *     bindings/python/opencmiss/_iron_utils.py:23-78);
 *   - objects are opaque handles; scalars IN by value, OUT by pointer;
 *     arrays are (pointer, size) pairs; indices are 32-bit ints, values
 *     double (float for FP32 Jacobians);
 *   - host arrays are caller-owned and copied in; device buffers are owned by
 *     the library and may be borrowed with oc_device_pointer_get
 *     (cmfe_DistributedMatrix_DataGetDPObj analogue, opencmiss_iron.f90:
 *     44920-44945);
 *   - field vectors use separated-component DOF order
 *     (field_routines.f90:10345-10347); coordinates are separated
 *     (x[0..N) y[N..2N) z[2N..3N));
 *   - index_base selects 0- or 1-based node/DOF indices on input and output
 *     (the reference stores CSR 1-based, equations_matrices_routines.f90:
 *     5133,5194-5198).
 * A handle is not thread-safe; distinct handles may be used concurrently.
 * Every call is ordered on the library's stream; copy-out calls synchronise.
 */
#ifndef OC_ASM_H
#define OC_ASM_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (CMFE_NO_ERROR = 0 analogue) ------------------------- */
#define OC_NO_ERROR 0
#define OC_ERR_INVALID_HANDLE (-1)
#define OC_ERR_INVALID_ARGUMENT (-2)
#define OC_ERR_HIP (-3)
#define OC_ERR_ALLOCATION (-4)
#define OC_ERR_NOT_CONVERGED (-5)
#define OC_ERR_PATTERN (-6)
#define OC_ERR_NOT_INITIALISED (-7)
#define OC_ERR_UNSUPPORTED (-8)

/* ---- equations-set classes ---------------------------------------------- */
#define OC_EQUATIONS_SET_LAPLACE 1         /* classical field, Laplace, standard subtype */
#define OC_EQUATIONS_SET_MOONEY_RIVLIN 2   /* elasticity, finite elasticity, MR subtype */

/* ---- Mooney-Rivlin subtypes (equations_set_constants.f90:121, 141) -------- */
#define OC_MOONEY_RIVLIN_SUBTYPE 1                  /* analytic Jacobian (the default) */
#define OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE 21  /* PK2 form, forward-difference Jacobian */

/* ---- precision flags ----------------------------------------------------- */
#define OC_FP64 0
#define OC_FP32 1
#define OC_BF16 2   /* MLP inference only */
#define OC_BF16_HIDDEN 3  /* MLP inference only: bf16 hidden layers, FP32 first and last layers */

/* ---- device buffers that can be borrowed --------------------------------- */
#define OC_BUFFER_MATRIX_VALUES 0     /* double[nnz]: linear matrix or FP64 Jacobian */
#define OC_BUFFER_MATRIX_VALUES_SP 1  /* float[nnz]: FP32 Jacobian */
#define OC_BUFFER_ROW_PTR 2           /* int64_t[n_rows+1], 0-based (device offsets are 64-bit) */
#define OC_BUFFER_COL_IDX 3           /* int[nnz], 0-based */
#define OC_BUFFER_DOFS 4              /* double[n_rows]: dependent field values */
#define OC_BUFFER_RESIDUAL 5          /* double[n_rows] */
#define OC_BUFFER_RHS 6               /* double[n_rows] */
#define OC_BUFFER_SOLVER_VALUES 7     /* double[solver nnz]: solver matrix (oc_solver_matrices_assemble) */
#define OC_BUFFER_SOLVER_VALUES_SP 8  /* float[solver nnz]: solver matrix from the FP32 Jacobian */
#define OC_BUFFER_SOLVER_ROW_PTR 9    /* int64_t[solver rows+1], 0-based */
#define OC_BUFFER_SOLVER_COL_IDX 10   /* int[solver nnz], 0-based solver columns */
#define OC_BUFFER_SOLVER_VECTOR 11    /* double[solver rows] */

typedef struct oc_equations_set_s *oc_equations_set;

/* ---- library ------------------------------------------------------------- */
/* Select the HIP device and create the library stream.  Replaces the
 * computational-environment start-up (computational_environment.f90:539-633)
 * for the single-GPU path. */
int oc_initialise(int device);
int oc_finalise(void);
/* Copies the last error message of the calling thread (NUL-terminated). */
int oc_last_error(char *buffer, int buffer_size);
/* Number of the HIP device in use, or -1. */
int oc_device_get(int *device);

/* ---- equations sets ------------------------------------------------------ */
/* Laplace, standard subtype, trilinear Lagrange hexahedra, 2x2x2 Gauss.
 * Replaces the setup of Laplace_equations_routines.f90:1872-1900 plus the
 * linear-matrix structure calculation (equations_matrices_routines.f90:
 * 4572-4780).  coords: 3*n_nodes separated; elem_nodes: 8*n_elements,
 * element-major, local node order xi1 fastest. */
int oc_laplace_create(int n_nodes, const double *coords, int n_elements, const int *elem_nodes,
                      int index_base, oc_equations_set *equations_set);

/* Mooney-Rivlin finite elasticity (EQUATIONS_SET_MOONEY_RIVLIN_SUBTYPE):
 * triquadratic displacement (deformed coordinates z, 3 components) and
 * trilinear hydrostatic pressure (component 4), 3x3x3 Gauss; geometric field
 * triquadratic.  Replaces finite_elasticity_routines.f90:4696-4711,
 * 6962-7025 and the Jacobian structure calculation
 * (equations_matrices_routines.f90:5048-5230).
 * coords_q: 3*n_nodes_q; elem_nodes_q: 27*n_elements; elem_nodes_l:
 * 8*n_elements indexing the pressure component's own nodes (0..n_nodes_l-1
 * for base 0).  c10, c01: materials field components 1 and 2
 * (finite_elasticity_routines.f90:6609-6613). */
int oc_mooney_rivlin_create(int n_nodes_q, const double *coords_q, int n_nodes_l, int n_elements,
                            const int *elem_nodes_q, const int *elem_nodes_l, int index_base, double c10,
                            double c01, oc_equations_set *equations_set);
/* A batch of independent Mooney-Rivlin meshes of uneven size as one equations
 * set (no single counterpart in the reference, where each mesh is its own
 * equations set; the element routines are those of oc_mooney_rivlin_create).
 * Per-mesh arrays: n_nodes_q, n_nodes_l, n_elements, c10, c01 [n_meshes];
 * coords_q concatenates each mesh's 3*n_nodes_q (separated) array;
 * elem_nodes_q / elem_nodes_l concatenate each mesh's connectivity with
 * mesh-local node numbers.  The set is the disjoint union: mesh m's nodes are
 * shifted by its node offset (oc_equations_set_mesh_get) and the DOFs are
 * component-major over the union (z1 of all nodes, z2, z3, then p), so the
 * equations matrix is block diagonal and each assembly is one launch over
 * every element of every mesh. */
int oc_mooney_rivlin_batch_create(int n_meshes, const int *n_nodes_q, const int *n_nodes_l, const int *n_elements,
                                  const double *coords_q, const int *elem_nodes_q, const int *elem_nodes_l,
                                  const double *c10, const double *c01, int index_base,
                                  oc_equations_set *equations_set);
/* Number of meshes of a set (1 unless created as a batch). */
int oc_mesh_count(oc_equations_set equations_set, int *n_meshes);
/* Mesh `mesh` of a batch: first union quadratic node, quadratic node count,
 * first union pressure node, pressure node count, first element, elements.
 * DOF of (component c < 3, mesh node i): c * n_nodes_total + node_offset_q + i;
 * pressure node i: 3 * n_nodes_total + node_offset_l + i. */
int oc_equations_set_mesh_get(oc_equations_set equations_set, int mesh, int *node_offset_q, int *n_nodes_q,
                              int *node_offset_l, int *n_nodes_l, int *element_offset, int *n_elements);
/* The equations-matrix block of mesh `mesh` of a batch as a CSR matrix of its
 * own, in the DOF order of a single-mesh set (z1 of the mesh's nodes, z2, z3,
 * then p): n_rows and nnz receive its sizes; row_ptr[n_rows+1] and
 * col_idx[nnz] (index_base 0 or 1) and values[nnz] (double for OC_FP64, float
 * for OC_FP32: the last Jacobian of that precision) are filled when not NULL.
 * No counterpart in the reference, where each mesh is its own equations set
 * and this is cmfe_DistributedMatrix_StorageLocationsGetObj / DataGetDPObj
 * (opencmiss_iron.f90:44843, 44920) of that set. */
int oc_mesh_matrix_get(oc_equations_set equations_set, int mesh, int precision, int index_base, int *n_rows,
                       long long *nnz, int *row_ptr, int *col_idx, void *values);
int oc_equations_set_destroy(oc_equations_set equations_set);
/* class (OC_EQUATIONS_SET_*), number of rows (= DOFs), number of nonzeros,
 * number of elements, number of element colours (0: the assembly needs none). */
int oc_equations_set_info(oc_equations_set equations_set, int *set_class, int *n_rows, long long *nnz,
                          int *n_elements, int *n_colours);
/* Materials (c10, c01) of a Mooney-Rivlin set (every mesh of a batch). */
int oc_materials_set(oc_equations_set equations_set, double c10, double c01);
/* Constitutive subtype of a Mooney-Rivlin set (every mesh of a batch):
 * OC_MOONEY_RIVLIN_SUBTYPE (EQUATIONS_SET_MOONEY_RIVLIN_SUBTYPE: deviatoric
 * Cauchy stress plus p, analytic Jacobian) or
 * OC_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE (EQUATIONS_SET_INCOMPRESSIBLE_
 * MOONEY_RIVLIN_SUBTYPE: Cauchy stress pushed forward from the 2nd Piola tensor
 * with the pressure inside it, finite_elasticity_routines.f90:3522-3549, in the
 * generic residual loop :1654-1761; its Jacobian is the forward-difference one,
 * FP64 only, as the reference evaluates it). */
int oc_equations_set_subtype_set(oc_equations_set equations_set, int subtype);
/* Per-mesh materials of a batch: c10[n_meshes], c01[n_meshes]. */
int oc_mesh_materials_set(oc_equations_set equations_set, const double *c10, const double *c01);

/* ---- CSR structure -------------------------------------------------------- */
/* Copies the equations-matrix structure (row_ptr[n_rows+1], col_idx[nnz]),
 * columns sorted ascending and unique per row (cmfe_DistributedMatrix_
 * StorageLocationsGetObj analogue, opencmiss_iron.f90:44843).  Fails with
 * OC_ERR_UNSUPPORTED when nnz does not fit the 4-byte integers. */
int oc_matrix_structure_get(oc_equations_set equations_set, int index_base, int *row_ptr, int *col_idx);

/* ---- dependent field ------------------------------------------------------- */
int oc_dofs_set(oc_equations_set equations_set, const double *dofs);
int oc_dofs_get(oc_equations_set equations_set, double *dofs);
/* Device-to-device set from a borrowed/foreign device pointer. */
int oc_dofs_set_device(oc_equations_set equations_set, const double *dofs_device);

/* ---- assembly -------------------------------------------------------------- */
/* Linear static assembly (EQUATIONS_SET_ASSEMBLE -> ..._STATIC_LINEAR_FEM,
 * equations_set_routines.f90:981, 1295-1492): zero and assemble the linear
 * matrix values and the RHS vector. */
int oc_equations_set_assemble(oc_equations_set equations_set);
/* Laplace assembly form (no counterpart in the reference).  A structured grid
 * (nodes and elements numbered x fastest, as the regular generated mesh) can
 * assemble in one fused pass with no element-matrix buffer: fused = 1 (the
 * default) does so when the grid is large enough to fill the GPU, 2 whenever
 * the mesh is such a grid, 0 never (element-matrix buffer and CSR row gather).
 * Other meshes always use the latter; the values are the same bit for bit.
 * active (may be NULL) returns whether the fused pass now assembles the set. */
int oc_laplace_fused_set(oc_equations_set equations_set, int fused, int *active);
/* Nonlinear residual (EQUATIONS_SET_RESIDUAL_EVALUATE ->
 * ..._STATIC_FEM, equations_set_routines.f90:5184, 5483-5629). */
int oc_equations_set_residual_evaluate(oc_equations_set equations_set);
/* Analytic Jacobian (EQUATIONS_SET_JACOBIAN_EVALUATE -> ..._STATIC_FEM,
 * equations_set_routines.f90:4694, 4878-5024).  precision OC_FP64 fills
 * OC_BUFFER_MATRIX_VALUES, OC_FP32 fills OC_BUFFER_MATRIX_VALUES_SP. */
int oc_equations_set_jacobian_evaluate(oc_equations_set equations_set, int precision);
/* Forward-difference Jacobian into the FP64 values: column k of each element
 * matrix is (R_e(x + delta e_k) - R_e(x)) / delta with delta = (1 + ||x||_2)
 * 1e-7 over the whole dependent vector, assembled like the analytic one
 * (EquationsSet_FiniteElementJacobianEvaluateFD, equations_set_routines.f90:
 * 3032-3150; the default for subtypes without an analytic Jacobian). */
int oc_equations_set_jacobian_evaluate_fd(oc_equations_set equations_set);
/* The same assembly replayed from a HIP graph recorded on the first call
 * (kernel nodes only; re-recorded when the precision or launch shapes
 * change).  Results are identical to oc_equations_set_jacobian_evaluate. */
int oc_jacobian_evaluate_graph(oc_equations_set equations_set, int precision);

/* Residual evaluation that allocates nothing, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  R(x)
 * of a Mooney-Rivlin set, as oc_equations_set_residual_evaluate computes it
 * (same kernels, same values), with x = dofs_device and the result in
 * residual_device: caller-owned device arrays of n_rows doubles in the set's
 * DOF order.  The element residuals pass through a static device array of the
 * library (OC_NOALLOC_MAX_ELEMENTS * 89 doubles, 31.9 MiB, one per GPU, held
 * until the process ends), so sets of more than OC_NOALLOC_MAX_ELEMENTS
 * elements fail with OC_ERR_UNSUPPORTED.  Calls that overlap on one GPU share
 * that array and race; calls on the library stream are ordered. */
#define OC_NOALLOC_MAX_ELEMENTS 47000
int oc_residual_evaluate_noalloc(oc_equations_set equations_set, const double *dofs_device, double *residual_device);

/* ---- copy-out / borrow ----------------------------------------------------- */
int oc_matrix_values_get(oc_equations_set equations_set, double *values);
int oc_matrix_values_get_sp(oc_equations_set equations_set, float *values);
int oc_residual_get(oc_equations_set equations_set, double *residual);
int oc_rhs_get(oc_equations_set equations_set, double *rhs);
int oc_device_pointer_get(oc_equations_set equations_set, int buffer, void **device_pointer);

/* ---- solver matrices ------------------------------------------------------- */
/* The solver system built from the last assembled equations matrix
 * (SOLVER_MATRICES_STATIC_ASSEMBLE, solver_routines.f90:13745-14771, with
 * SOLVER_MATRIX_JACOBIAN_MATRIX_ADD and its equations-matrix counterpart,
 * solver_matrices_routines.f90:1258-1483).  Rows and columns of the DOFs fixed
 * by oc_boundary_conditions_set are removed and the free DOFs are numbered in
 * ascending order (INCLUDE_ROW / INCLUDE_COLUMN, solver_mapping_routines.f90:
 * 434-520); the coupling coefficients are 1.  Laplace: matrix = the linear
 * matrix (FP64), vector = RHS minus the fixed columns times the fixed DOFs'
 * current values (solver_routines.f90:14380-14620).  Mooney-Rivlin: matrix =
 * the Jacobian of `precision`, vector = the residual of the free rows.  The
 * structure is built on the first call after a boundary-condition change. */
int oc_solver_matrices_assemble(oc_equations_set equations_set, int precision);
int oc_solver_matrix_info(oc_equations_set equations_set, int *n_rows, long long *nnz);
/* 4-byte structure (OC_ERR_UNSUPPORTED when the solver nnz does not fit). */
int oc_solver_matrix_structure_get(oc_equations_set equations_set, int index_base, int *row_ptr, int *col_idx);
int oc_solver_matrix_values_get(oc_equations_set equations_set, double *values);
int oc_solver_matrix_values_get_sp(oc_equations_set equations_set, float *values);
int oc_solver_vector_get(oc_equations_set equations_set, double *vector);

/* ---- boundary conditions and solution --------------------------------------- */
/* Fixed (or fixed-incremented) Dirichlet conditions on `n` DOFs with their
 * full values (BOUNDARY_CONDITION_FIXED[_INCREMENTED],
 * boundary_condition_routines.f90).  Replaces any previous set.  Does not
 * change the DOFs; see oc_load_increment_apply. */
int oc_boundary_conditions_set(oc_equations_set equations_set, int n, const int *dof_indices,
                               const double *full_values, int index_base);
/* EQUATIONS_SET_BOUNDARY_CONDITIONS_INCREMENT (equations_set_routines.f90:
 * 6491-6610): for load step `iteration` of `max_iterations`, each Dirichlet
 * DOF becomes cur + (full - cur)/(max - iteration + 1); the full value at the
 * last step. */
int oc_load_increment_apply(oc_equations_set equations_set, int iteration, int max_iterations);

#define OC_PRECONDITIONER_JACOBI 0          /* point Jacobi, zero diagonal -> 1 (PETSc PCJACOBI) */
#define OC_PRECONDITIONER_BLOCK_DIAGONAL 1  /* nodal 3x3 block Jacobi + diagonal Schur estimate for p */
#define OC_PRECONDITIONER_BLOCK_TRIANGULAR 2 /* block upper-triangular variant of the above */
#define OC_PRECONDITIONER_MULTIGRID 3       /* block upper-triangular, displacement block by one geometric
                                               multigrid V-cycle (needs oc_multigrid_structure_set) */

typedef struct {
    double absolute_tolerance;   /* Newton: converged when ||F|| <= max(abs, rel * ||F_ref||) */
    double relative_tolerance;   /* relative to ||F_ref|| (the step's initial residual, or reference_norm) */
    double reference_norm;       /* > 0: use instead of the initial residual norm */
    int maximum_iterations;      /* Newton iterations */
    int jacobian_precision;      /* OC_FP64 or OC_FP32 */
    int krylov_restart;          /* GMRES restart length */
    int krylov_maximum_iterations;
    /* > 0: fixed forcing term; <= 0: Eisenstat-Walker capped by ||F||/||F_ref||.  Either way a linear solve also
     * stops once its residual is below half the Newton target, which the step then meets */
    double krylov_relative_tolerance;
    int line_search;             /* 0: full steps, 1: backtracking (cubic/quadratic, alpha below) */
    double line_search_alpha;    /* sufficient-decrease constant (reference default 1e-4) */
    int preconditioner;          /* OC_PRECONDITIONER_* */
    double solution_tolerance;   /* stol: converged when ||step|| < stol * ||x|| over the free DOFs; 0 disables */
    int maximum_function_evaluations; /* residual evaluations allowed (< 0: no limit) */
    /* GMRES solves allowed to end above their tolerance (at krylov_maximum_iterations) before the solve stops
     * with OC_DIVERGED_LINEAR_SOLVE; a failed solve below the limit still gives the step (PETSc's
     * SNESSetMaxLinearSolveFailures, default 1: the first failure stops) */
    int maximum_linear_solve_failures;
    /* line searches allowed to end without sufficient decrease before the solve stops with
     * OC_DIVERGED_LINE_SEARCH (PETSc's SNESSetMaxNonlinearStepFailures, default 1) */
    int maximum_step_failures;
    /* OC_SOLVER_NO_OUTPUT or OC_SOLVER_PROGRESS_OUTPUT: one monitor line per Newton iteration on stderr
     * (SOLVER_PROGRESS_OUTPUT and the nonlinear monitor, solver_routines.f90:350, 18628-18631, 20520-20635) */
    int output_type;
    /* 1: when a solve started from a learned warm start (oc_mlp_warm_start) fails, the meshes that did not
     * converge restart from the plain start (the DOFs before the warm start, with the same load increments
     * applied) with an FP64 Jacobian and fallback_maximum_iterations Newton iterations (0: maximum_iterations);
     * oc_newton_stats.fallback_used reports it.  No counterpart in the reference: its divergence handling
     * stops with an error (solver_routines.f90:20161-20268). */
    int fallback;
    int fallback_maximum_iterations;
} oc_newton_options;

#define OC_SOLVER_NO_OUTPUT 0
#define OC_SOLVER_PROGRESS_OUTPUT 1

/* Why a Newton solve stopped (oc_newton_stats.reason): the codes of PETSc's
 * SNESConvergedReason for the default convergence test the reference selects
 * (SOLVER_NEWTON_CONVERGENCE_PETSC_DEFAULT, solver_routines.f90:17780-17851). */
#define OC_CONVERGED_FNORM_ABS 2       /* ||F|| <= absolute_tolerance */
#define OC_CONVERGED_FNORM_RELATIVE 3  /* ||F|| <= relative_tolerance * ||F_ref|| */
#define OC_CONVERGED_SNORM_RELATIVE 4  /* ||step|| < solution_tolerance * ||x|| */
#define OC_DIVERGED_FUNCTION_COUNT -2  /* maximum_function_evaluations reached */
#define OC_DIVERGED_LINEAR_SOLVE -3    /* maximum_linear_solve_failures GMRES solves ended above their tolerance */
#define OC_DIVERGED_FNORM_NAN -4       /* non-finite residual */
#define OC_DIVERGED_MAX_IT -5          /* maximum_iterations reached */
#define OC_DIVERGED_LINE_SEARCH -6     /* maximum_step_failures line searches ended without sufficient decrease */

typedef struct {
    int converged;
    int newton_iterations;
    int krylov_iterations;       /* total over the Newton iterations */
    int residual_evaluations;
    int jacobian_evaluations;
    double initial_residual_norm;
    double final_residual_norm;
    double seconds;
    int reason;                  /* OC_CONVERGED_* (> 0) or OC_DIVERGED_* (< 0) */
    int linear_solve_failures;   /* GMRES solves that ended above their tolerance */
    int step_failures;           /* line searches that ended without sufficient decrease */
    int fallback_used;           /* 1: the answer comes from the plain-start fallback (options.fallback) */
    int warm_start_reason;       /* reason of the warm-started attempt when the fallback ran, else 0 */
} oc_newton_stats;

/* Reference defaults (solver_routines.f90:17780-17851, 10957-11006): abs tol
 * 1e-10, rel tol 1e-5, solution tol 1e-5, 50 iterations, 1000 function
 * evaluations, line search alpha 1e-4, GMRES(30), rtol 1e-5, Jacobi; PETSc's
 * one linear-solve failure and one step failure; no output; no fallback. */
int oc_newton_options_default(oc_newton_options *options);
/* Newton solve of R(x) = 0 at the current Dirichlet values for a Mooney-
 * Rivlin set (SOLVER_NEWTON_LINESEARCH_SOLVE, solver_routines.f90:18871-
 * 19028, with the residual and Jacobian callbacks of problem_routines.f90:
 * 4122-4483).  Dirichlet DOFs are eliminated (their rows and columns are
 * dropped from the linear systems, solver_matrices_routines.f90:1258-1483). */
int oc_newton_solve(oc_equations_set equations_set, const oc_newton_options *options, oc_newton_stats *stats);
/* Newton solve of every mesh of a batch at once (no counterpart in the
 * reference).  Mesh m has converged when ||F_m|| <= max(absolute_tolerance,
 * relative_tolerance * ref_m), ref_m = reference_norms[m] or, when
 * reference_norms is NULL, mesh m's own starting residual norm; GMRES and
 * the line search weigh each mesh's residual by 1/ref_m.  final_norms
 * (optional, n_meshes) receives ||F_m||.  The Jacobian assembly and
 * preconditioner setup of each Newton step replay one recorded HIP graph. */
int oc_newton_solve_batch(oc_equations_set equations_set, const oc_newton_options *options,
                          const double *reference_norms, double *final_norms, oc_newton_stats *stats);
/* Static linear solve of a Laplace set (Problem_SolverEquationsStaticLinear
 * Solve, problem_routines.f90:2740-2830): assemble, set the Dirichlet values
 * and solve the free rows with the Krylov settings of `options`. */
int oc_linear_solve(oc_equations_set equations_set, const oc_newton_options *options, oc_newton_stats *stats);
/* Geometric multigrid for OC_PRECONDITIONER_MULTIGRID (no counterpart in the
 * reference, whose KSP preconditioner is chosen by the user): declares every
 * mesh of the set a regular generated mesh of n_per_mesh[m]^3 elements
 * (node numbering as GENERATED_MESH_REGULAR_CREATE_FINISH; checked, else
 * OC_ERR_INVALID_ARGUMENT) and builds the level hierarchy: triquadratic ->
 * trilinear on the same elements -> coarsening by 2 per direction down to
 * <= 3 nodes per direction.  Coarse operators are Galerkin products of the
 * current Jacobian (formed with the preconditioner setup of each Newton step). */
int oc_multigrid_structure_set(oc_equations_set equations_set, const int *n_per_mesh);
/* V-cycle smoothing on the triquadratic level: damped nodal block-Jacobi
 * sweeps before and after the coarse correction (defaults 1, 2, 0.8).  One
 * post-sweep needs omega <= 0.6: with larger omega it amplifies the
 * highest modes of the element layer that takes a load step's boundary
 * increment, and GMRES stalls on the first Jacobian of the step; two pre-sweeps
 * at 0.8 stall it too.  The trilinear and coarser levels take one sweep each
 * way. */
int oc_multigrid_options_set(oc_equations_set equations_set, int pre_sweeps, int post_sweeps, double omega);
/* Launch-shape autotuner (no counterpart in the reference): times the
 * element-matrix kernel's grid, the CSR row gather's prefetch depth, the
 * SpMV's lanes per row and the level-0 multigrid kernel's shape on this set at
 * its current DOFs, keeps the fastest of each for later launches, and writes
 * the timings as JSON into report. */
int oc_autotune(oc_equations_set equations_set, int precision, char *report, int report_size);
/* Euclidean norm of the residual over the free DOFs at the current DOFs. */
int oc_residual_norm(oc_equations_set equations_set, double *norm);
/* The same per mesh of a batch: norms[n_meshes]. */
int oc_mesh_residual_norms(oc_equations_set equations_set, double *norms);

/* ---- learned warm start (no counterpart in the reference) ------------------
 * A per-node MLP predicts each load step's nodal displacement increment as
 * Newton's starting guess; Newton's own residual test still decides.  Inputs
 * per quadratic node (9): X (3), current displacement z - X (3), load
 * increment, c10, c01; output (3): displacement increment per unit load
 * increment.  FP32 parameters, trained with Adam by the library's kernels. */
typedef struct oc_mlp_s *oc_mlp;
int oc_mlp_create(int n_inputs, int n_hidden, int n_hidden_layers, int n_outputs, unsigned long long seed,
                  oc_mlp *mlp);
int oc_mlp_destroy(oc_mlp mlp);
int oc_mlp_parameter_count(oc_mlp mlp, int *count);
int oc_mlp_parameters_get(oc_mlp mlp, float *parameters);
int oc_mlp_parameters_set(oc_mlp mlp, const float *parameters);
/* Per-feature affine normalisation x' = (x - shift) * scale (n_inputs each). */
int oc_mlp_normalisation_set(oc_mlp mlp, const float *shift, const float *scale);
/* y = MLP(x) for n samples (host arrays, x: n*n_inputs, y: n*n_outputs). */
int oc_mlp_forward(oc_mlp mlp, int n, const float *x, float *y);
/* Inference precision of oc_mlp_forward and the warm starts: OC_FP32 (the
 * default), OC_BF16 (bf16 weights and activations on MFMA with FP32
 * accumulation, bias and GELU; layer widths up to 128) or OC_BF16_HIDDEN (the
 * same for the hidden layers only: the first and last layers, and so the input
 * features and the output increment, stay FP32).  Training stays FP32.
 * No counterpart in the reference: the warm start is this library's own. */
int oc_mlp_inference_precision_set(oc_mlp mlp, int precision);
/* Loss 1/(n*n_out) sum (y - t)^2 and its gradient w.r.t. the parameters
 * (host arrays; gradient in parameter order). */
int oc_mlp_gradient(oc_mlp mlp, int n, const float *x, const float *t, float *loss, float *gradient);
/* Uploads a training set (n samples) to the device. */
int oc_mlp_dataset_set(oc_mlp mlp, int n, const float *x, const float *t);
/* `steps` Adam steps on minibatches drawn (seeded) from the device training
 * set; returns the mean loss of the last `steps` minibatches. */
int oc_mlp_train(oc_mlp mlp, int steps, int batch, float learning_rate, unsigned long long seed, float *mean_loss);
/* Warm start of a Mooney-Rivlin set: z += load_increment * MLP(features)[0:3]
 * on every free displacement DOF, features built on the device from the set's
 * coordinates and current DOFs.  A 4-output MLP also predicts the pressure
 * increment per unit load at the quadratic nodes; each free pressure DOF takes
 * the value of its coincident (element corner) quadratic node.  Call before
 * oc_load_increment_apply.  The pressure output is in units of c10 + c01. */
int oc_mlp_warm_start(oc_mlp mlp, oc_equations_set equations_set, double load_increment);
/* The same for a batch of meshes: load_increments[n_meshes], one per mesh. */
int oc_mlp_warm_start_batch(oc_mlp mlp, oc_equations_set equations_set, const double *load_increments);

/* ---- timing on the library stream ----------------------------------------- */
/* Records event `slot` (0..63) on the library stream. */
int oc_event_record(int slot);
/* Milliseconds between two recorded events (synchronises on `end`). */
int oc_event_elapsed(int start, int end, float *milliseconds);
int oc_synchronise(void);

/* ---- per-kernel timing (TAU phases analogue, equations_set_routines.f90:
 * 1332-1336,1366,1414) ------------------------------------------------------ */
/* When enabled, every kernel launch is bracketed by events on the library
 * stream and its duration is accumulated under its kernel name
 * ("laplace_element", "laplace_row_gather", "mr_residual", "dof_gather",
 * "mr_gauss_jacobian", "mr_element_matrix_fp64", "mr_element_matrix_fp32",
 * "csr_row_gather_fp64", "csr_row_gather_fp32", "solver_matrix_fp64",
 * "solver_matrix_fp32", "fill", ...).  Enabling or
 * resetting clears the accumulators. */
int oc_kernel_timing_enable(int enable);
/* Restricts the timing to the launches of one kernel name (NULL: all), so
 * that a timed region carries one event pair per launch of that kernel. */
int oc_kernel_timing_select(const char *kernel_name);
/* Synchronises, then returns the total milliseconds and launch count of
 * `kernel_name` since the last enable/reset. */
int oc_kernel_timing_get(const char *kernel_name, double *total_ms, long long *launches);

#ifdef __cplusplus
}
#endif
#endif
