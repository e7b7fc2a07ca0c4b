!> Drives the nonlinear (Mooney-Rivlin) path of liboc_asm from Fortran through
!> the OC_ASM ISO_C_BINDING module, shaped like the Iron-side replacement of
!> EQUATIONS_SET_RESIDUAL_EVALUATE_STATIC_FEM (equations_set_routines.f90:
!> 5483-5629) and EQUATIONS_SET_JACOBIAN_EVALUATE_STATIC_FEM (:4878-5024): the
!> host keeps the dependent field and a 1-based CSR matrix of its own
!> (ROW_INDICES, COLUMN_INDICES, DATA_DP, as the CMISS distributed matrix,
!> matrix_vector.f90), and one call per assembly fills them.
!>
!> Usage: drive_mooney_rivlin <input> <output>.  The input (stream, native
!> byte order) holds nq, nl, ne (int32), c10, c01 (real64), the separated
!> coordinates (3*nq), the 1-based connectivity of the displacement (27*ne)
!> and pressure (8*ne) components and the dependent DOFs (3*nq + nl).  The
!> output holds n_rows, nnz (int32), the residual (n_rows real64), ROW_INDICES
!> (n_rows+1), COLUMN_INDICES (nnz), the FP64 Jacobian's DATA_DP (nnz) and the
!> FP32 Jacobian widened to DATA_DP (nnz).
MODULE MR_HOST
  USE ISO_C_BINDING
  USE OC_ASM
  IMPLICIT NONE
CONTAINS
  SUBROUTINE CHECK(rc, what)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: what
    CHARACTER(LEN=512) :: msg
    IF (rc /= OC_NO_ERROR) THEN
      CALL OC_ERROR_MESSAGE(msg)
      WRITE(*,'(A,A,I0,A,A)') what, " failed: ", rc, " ", TRIM(msg)
      STOP 1
    END IF
  END SUBROUTINE CHECK

  !> Residual evaluation of the static nonlinear equations set: the dependent
  !> field goes in, the nonlinear (residual) vector comes out.
  SUBROUTINE RESIDUAL_EVALUATE_STATIC_FEM(h, dofs, residual)
    TYPE(C_PTR), INTENT(IN) :: h
    REAL(C_DOUBLE), INTENT(IN) :: dofs(:)
    REAL(C_DOUBLE), INTENT(OUT) :: residual(:)
    CALL CHECK(oc_dofs_set(h, dofs), "oc_dofs_set")
    CALL CHECK(oc_equations_set_residual_evaluate(h), "oc_equations_set_residual_evaluate")
    CALL CHECK(oc_residual_get(h, residual), "oc_residual_get")
  END SUBROUTINE RESIDUAL_EVALUATE_STATIC_FEM

  !> Jacobian evaluation into the host's CSR storage (1-based ROW_INDICES /
  !> COLUMN_INDICES, values in DATA_DP), FP64 or FP32 element Jacobians.
  SUBROUTINE JACOBIAN_EVALUATE_STATIC_FEM(h, precision, row_indices, column_indices, data_dp)
    TYPE(C_PTR), INTENT(IN) :: h
    INTEGER(C_INT), INTENT(IN) :: precision
    INTEGER(C_INT), INTENT(OUT) :: row_indices(:), column_indices(:)
    REAL(C_DOUBLE), INTENT(OUT) :: data_dp(:)
    REAL(C_FLOAT), ALLOCATABLE :: data_sp(:)
    CALL CHECK(oc_equations_set_jacobian_evaluate(h, precision), "oc_equations_set_jacobian_evaluate")
    CALL CHECK(oc_matrix_structure_get(h, 1_C_INT, row_indices, column_indices), "oc_matrix_structure_get")
    IF (precision == OC_FP64) THEN
      CALL CHECK(oc_matrix_values_get(h, data_dp), "oc_matrix_values_get")
    ELSE
      ALLOCATE(data_sp(SIZE(data_dp)))
      CALL CHECK(oc_matrix_values_get_sp(h, data_sp), "oc_matrix_values_get_sp")
      data_dp = REAL(data_sp, C_DOUBLE)
    END IF
  END SUBROUTINE JACOBIAN_EVALUATE_STATIC_FEM
END MODULE MR_HOST

PROGRAM DRIVE_MOONEY_RIVLIN
  USE ISO_C_BINDING
  USE OC_ASM
  USE MR_HOST
  IMPLICIT NONE
  TYPE(C_PTR) :: h
  INTEGER(C_INT) :: nq, nl, ne, cls, nrows, nel, ncol
  INTEGER(C_LONG_LONG) :: nnz
  REAL(C_DOUBLE) :: c10, c01
  INTEGER(C_INT), ALLOCATABLE :: conn_q(:), conn_l(:), row_indices(:), column_indices(:)
  REAL(C_DOUBLE), ALLOCATABLE :: coords(:), dofs(:), residual(:), data64(:), data32(:)
  CHARACTER(LEN=1024) :: fin, fout
  INTEGER :: u

  IF (COMMAND_ARGUMENT_COUNT() < 2) STOP 2
  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  OPEN(NEWUNIT=u, FILE=TRIM(fin), ACCESS="STREAM", FORM="UNFORMATTED", STATUS="OLD")
  READ(u) nq, nl, ne, c10, c01
  ALLOCATE(coords(3*nq), conn_q(27*ne), conn_l(8*ne), dofs(3*nq + nl))
  READ(u) coords, conn_q, conn_l, dofs
  CLOSE(u)

  CALL CHECK(oc_initialise(0_C_INT), "oc_initialise")
  CALL CHECK(oc_mooney_rivlin_create(nq, coords, nl, ne, conn_q, conn_l, 1_C_INT, c10, c01, h), &
    & "oc_mooney_rivlin_create")
  CALL CHECK(oc_equations_set_info(h, cls, nrows, nnz, nel, ncol), "oc_equations_set_info")
  ALLOCATE(residual(nrows), row_indices(nrows + 1), column_indices(nnz), data64(nnz), data32(nnz))
  CALL RESIDUAL_EVALUATE_STATIC_FEM(h, dofs, residual)
  CALL JACOBIAN_EVALUATE_STATIC_FEM(h, OC_FP64, row_indices, column_indices, data64)
  CALL JACOBIAN_EVALUATE_STATIC_FEM(h, OC_FP32, row_indices, column_indices, data32)

  OPEN(NEWUNIT=u, FILE=TRIM(fout), ACCESS="STREAM", FORM="UNFORMATTED", STATUS="REPLACE")
  WRITE(u) nrows, INT(nnz, C_INT), residual, row_indices, column_indices, data64, data32
  CLOSE(u)
  WRITE(*,'(A,I0,A,I0,A,I0)') "rows ", nrows, " nnz ", nnz, " first_col ", column_indices(1)
  CALL CHECK(oc_equations_set_destroy(h), "oc_equations_set_destroy")
  CALL CHECK(oc_finalise(), "oc_finalise")
END PROGRAM DRIVE_MOONEY_RIVLIN
