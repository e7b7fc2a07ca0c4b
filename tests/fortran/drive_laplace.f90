!> Drives liboc_asm from Fortran through the OC_ASM ISO_C_BINDING module, the
!> way an Iron host would: 1-based connectivity in, 1-based CSR out.
!> Without arguments it only exercises the error path (no device needed);
!> with the argument "gpu" it assembles the Laplace matrix of a 2x2x2 unit
!> cube and prints the CSR sizes, the diagonal entry of node 1 and the
!> element-free row sums.
PROGRAM DRIVE_LAPLACE
  USE ISO_C_BINDING
  USE OC_ASM
  IMPLICIT NONE
  TYPE(C_PTR) :: h
  INTEGER(C_INT) :: rc, cls, nrows, ne, ncol, i, j, k, e, a, b, c, n, np
  INTEGER(C_LONG_LONG) :: nnz
  INTEGER(C_INT), ALLOCATABLE :: conn(:), row_ptr(:), col_idx(:)
  REAL(C_DOUBLE), ALLOCATABLE :: coords(:), vals(:)
  CHARACTER(LEN=256) :: msg, arg
  REAL(C_DOUBLE) :: rowsum, maxrow

  rc = oc_equations_set_info(C_NULL_PTR, cls, nrows, nnz, ne, ncol)
  CALL OC_ERROR_MESSAGE(msg)
  WRITE(*,'(A,I0,A,A)') "null handle: ", rc, " ", TRIM(msg)
  IF (COMMAND_ARGUMENT_COUNT() < 1) STOP
  CALL GET_COMMAND_ARGUMENT(1, arg)
  IF (TRIM(arg) /= "gpu") STOP

  rc = oc_initialise(0_C_INT)
  IF (rc /= OC_NO_ERROR) THEN
    CALL OC_ERROR_MESSAGE(msg)
    WRITE(*,*) "initialise failed: ", TRIM(msg)
    STOP 1
  END IF
  n = 2
  np = n + 1
  ALLOCATE(coords(3*np**3), conn(8*n**3))
  DO k = 0, n
    DO j = 0, n
      DO i = 0, n
        coords(1 + i + j*np + k*np*np) = REAL(i, C_DOUBLE) / n
        coords(1 + np**3 + i + j*np + k*np*np) = REAL(j, C_DOUBLE) / n
        coords(1 + 2*np**3 + i + j*np + k*np*np) = REAL(k, C_DOUBLE) / n
      END DO
    END DO
  END DO
  e = 0
  DO k = 0, n - 1
    DO j = 0, n - 1
      DO i = 0, n - 1
        DO c = 0, 1
          DO b = 0, 1
            DO a = 0, 1
              e = e + 1
              conn(e) = 1 + (i + a) + (j + b)*np + (k + c)*np*np
            END DO
          END DO
        END DO
      END DO
    END DO
  END DO
  rc = oc_laplace_create(np**3, coords, n**3, conn, 1_C_INT, h)
  IF (rc /= OC_NO_ERROR) STOP 2
  rc = oc_equations_set_info(h, cls, nrows, nnz, ne, ncol)
  ALLOCATE(row_ptr(nrows + 1), col_idx(nnz), vals(nnz))
  rc = oc_equations_set_assemble(h)
  IF (rc /= OC_NO_ERROR) STOP 3
  rc = oc_matrix_structure_get(h, 1_C_INT, row_ptr, col_idx)
  rc = oc_matrix_values_get(h, vals)
  maxrow = 0.0_C_DOUBLE
  DO i = 1, nrows
    rowsum = SUM(vals(row_ptr(i):row_ptr(i+1)-1))
    maxrow = MAX(maxrow, ABS(rowsum))
  END DO
  WRITE(*,'(A,I0,A,I0,A,I0,A,ES23.16,A,ES10.3)') "rows ", nrows, " nnz ", nnz, " first_col ", col_idx(1), &
    & " K11 ", vals(1), " max_rowsum ", maxrow
  rc = oc_equations_set_destroy(h)
  rc = oc_finalise()
END PROGRAM DRIVE_LAPLACE
