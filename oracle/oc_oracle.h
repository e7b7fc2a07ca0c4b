/*
 * oc_oracle.h -- FP64 CPU restatement of OpenCMISS-Iron's finite element
 * assembly path (Laplace standard subtype and Mooney-Rivlin finite
 * elasticity), used ONLY as test infrastructure: the parity checker for the
 * HIP product library and the timed CPU baseline.  Nothing in the product
 * path links or calls this library.
 *
 * Conventions (all restated from the reference, see oc_oracle.c for the
 * file:line each function follows):
 *   - indices are 0-based here; the reference stores CSR 1-based;
 *   - field vectors use separated-component order: all DOFs of component 1,
 *     then component 2, ... (field_routines.f90:10345-10347);
 *   - coordinates are separated too: x[0..N), y[N..2N), z[2N..3N);
 *   - element-node arrays are element-major: nodes[e*nen + k];
 *   - element matrices are row-major: Ke[m*ne + n] is MATRIX(m+1,n+1).
 *
 * Return value of every int function: 0 = OK, negative = error.
 */
#ifndef OC_ORACLE_H
#define OC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- basis and quadrature ------------------------------------------------ */
int oco_gauss_legendre(int n, double *x, double *w);
/* order 1 = Lagrange linear (2 nodes per xi), 2 = quadratic (3 per xi).
 * ng1 = Gauss points per xi direction (1..4).
 * weights[G], positions[3*G], fns[(g*ns + s)*4 + nu] with nu = 0 value,
 * 1..3 = d/dxi1..3.  G = ng1^3, ns = (order+1)^3. */
int oco_basis_tables(int order, int ng1, double *weights, double *positions, double *fns);

/* --- 3x3 helpers (maths.f90) ---------------------------------------------- */
int oco_invert3(const double *A, double *B, double *det); /* column-major */

/* --- regular generated mesh ---------------------------------------------- */
/* Node counts of a regular n1 x n2 x n3 mesh with the given order. */
int64_t oco_regular_num_nodes(int order, int n1, int n2, int n3);
/* Element-node connectivity (0-based), element numbering xi1 fastest. */
int oco_regular_elements(int order, int n1, int n2, int n3, int32_t *elem_nodes);
/* Coordinates origin 0, extent 1 in each direction (separated layout),
 * with interior nodes jittered componentwise by U[-jitter*h, jitter*h]
 * (h = 1/n per direction) from the splitmix64 stream of `seed`. */
int oco_regular_coordinates(int order, int n1, int n2, int n3, double jitter, uint64_t seed,
                            double *coords);
/* Corner (trilinear) component of a quadratic mesh: element connectivity into
 * the trilinear component's own sorted node numbering. */
int oco_regular_pressure_elements(int n1, int n2, int n3, int32_t *elem_nodes_l);

/* --- CSR pattern (equations_matrices_routines.f90:4572-5230) -------------- */
/* ncomp components; comp c has nn[c] nodes, nen[c] nodes per element and
 * connectivity conn[c] (E*nen[c]); DOF offset of comp c = sum nn[<c].
 * Pass row_ptr NULL to get only *nnz_out.  col_idx must hold *nnz. */
int oco_pattern(int ncomp, int nelem, const int32_t *nn, const int32_t *nen,
                const int32_t *const *conn, int64_t *nnz_out, int64_t *row_ptr, int32_t *col_idx);

/* --- CSR add (matrix_vector.f90:2839-2878) -------------------------------- */
int oco_csr_add(const int64_t *row_ptr, const int32_t *col_idx, double *values, int nrow,
                const int32_t *rows, int ncol, const int32_t *cols, const double *Ke);

/* --- Laplace (Laplace_equations_routines.f90:428-571) --------------------- */
int oco_laplace_element(const double *xe /*3x8 [c*8+n]*/, double *Ke /*8x8*/);
int oco_laplace_assemble(int nelem, const int32_t *elem_nodes, int nnode, const double *coords,
                         const int64_t *row_ptr, const int32_t *col_idx, double *values, double *rhs);

/* --- Mooney-Rivlin (finite_elasticity_routines.f90) ----------------------- */
/* Cauchy stress (Voigt 11,22,33,12,13,23) from F (column-major 3x3), p. */
int oco_mr_stress(const double *F, double Jznu, double p, double c10, double c01, double *sigma);
/* Spatial elasticity tensor (6x6 column-major) and stress. */
int oco_mr_elasticity(const double *F, double Jznu, double p, double c10, double c01,
                      double *ctan, double *sigma);
/* Element residual R_e(89).  Xe, ze: [c*27+n]; pe[8]. */
int oco_mr_element_residual(const double *Xe, const double *ze, const double *pe,
                            double c10, double c01, double *Re);
/* Analytic element Jacobian (89x89 row-major). */
int oco_mr_element_jacobian(const double *Xe, const double *ze, const double *pe,
                            double c10, double c01, double *Je);
/* Forward-difference element Jacobian with a given delta. */
int oco_mr_element_jacobian_fd(const double *Xe, const double *ze, const double *pe,
                               double c10, double c01, double delta, double *Je);
/* Global assembly.  dofs = [z1(nq) z2(nq) z3(nq) p(nl)]. */
int oco_mr_residual(int nelem, const int32_t *elem_q, const int32_t *elem_l, int nq, int nl,
                    const double *Xq, const double *dofs, double c10, double c01, double *res);
/* Incompressible MR subtype: Cauchy stress (returns Jznu), element residual,
 * forward-difference element Jacobian, residual assembly. */
int oco_imr_stress(const double *F, double p, double c10, double c01, double *sigma, double *jznu);
int oco_imr_element_residual(const double *Xe, const double *ze, const double *pe,
                             double c10, double c01, double *Re);
int oco_imr_element_jacobian_fd(const double *Xe, const double *ze, const double *pe,
                                double c10, double c01, double delta, double *Je);
int oco_imr_residual(int nelem, const int32_t *elem_q, const int32_t *elem_l, int nq, int nl,
                     const double *Xq, const double *dofs, double c10, double c01, double *res);
/* mode 0 analytic, 1 forward difference (delta=(1+||dofs||_2)*1e-7), 2 forward
 * difference of the incompressible MR subtype's residual (same delta). */
int oco_mr_jacobian(int nelem, const int32_t *elem_q, const int32_t *elem_l, int nq, int nl,
                    const double *Xq, const double *dofs, double c10, double c01, int mode,
                    const int64_t *row_ptr, const int32_t *col_idx, double *values);

/* Multi-threaded baseline: residual + analytic Jacobian, greedy element
 * colours in sequence, each colour's elements over `threads` OpenMP threads. */
int oco_mr_assemble_parallel(int nelem, const int32_t *elem_q, const int32_t *elem_l, int nq, int nl,
                             const double *Xq, const double *dofs, double c10, double c01,
                             const int64_t *row_ptr, const int32_t *col_idx, double *values, double *res,
                             int threads);

/* --- boundary-condition increment (equations_set_routines.f90:6491-6610) -- */
int oco_bc_increment(int ndof, const int32_t *dof_idx, const double *full, double *current,
                     int iteration, int max_iterations);

/* --- solver matrices (SOLVER_MATRICES_STATIC_ASSEMBLE, solver_routines.f90:
 * 13745-14771; SOLVER_MATRIX_JACOBIAN_MATRIX_ADD, solver_matrices_routines.f90:
 * 1258-1483) ----------------------------------------------------------------
 * Solver rows/columns are the free DOFs (free_mask[d] != 0) in ascending
 * order; coupling coefficients 1.  Returns the solver row count; the solver
 * nnz goes to *nnz.  With out_row_ptr == NULL only the counts are returned;
 * otherwise out_row_ptr[n+1], out_col_idx[nnz], out_values[nnz] and
 * out_vector[n] are filled.  linear != 0: vector = RHS vec with the fixed
 * columns times dofs moved over; linear == 0: vector = vec (the residual) on
 * the free rows. */
int oco_solver_system(int n_rows, const int64_t *row_ptr, const int32_t *col_idx, const double *values,
                      const uint8_t *free_mask, const double *vec, const double *dofs, int linear, int64_t *nnz,
                      int64_t *out_row_ptr, int32_t *out_col_idx, double *out_values, double *out_vector);

/* --- splitmix64 uniform stream shared with the host generator ------------- */
double oco_uniform(uint64_t seed, uint64_t index);

#ifdef __cplusplus
}
#endif
#endif
