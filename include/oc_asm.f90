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
  INTEGER(C_INT), PARAMETER :: OC_FP64 = 0, OC_FP32 = 1

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
