!------------------------------------------------------------------------------
! Kind parameters and the C-ABI interface blocks of libsgw_hip.so (include/sgw.h).
!
! In a Quantum ESPRESSO build `dp` is QE's `kinds` dp; both are the C double.
!------------------------------------------------------------------------------
MODULE sgw_kinds

  USE, INTRINSIC :: iso_c_binding, ONLY: c_double

  IMPLICIT NONE

  INTEGER, PARAMETER :: dp = c_double

END MODULE sgw_kinds

MODULE sgw_capi

  USE, INTRINSIC :: iso_c_binding

  IMPLICIT NONE

  !> sgw_select_solver_config (include/sgw.h)
  TYPE, BIND(C) :: sgw_select_solver_config
    TYPE(c_ptr)        :: priority
    INTEGER(c_int32_t) :: npriority
    INTEGER(c_int32_t) :: max_iter
    REAL(c_double)     :: threshold
    INTEGER(c_int32_t) :: bicg_lmax
  END TYPE sgw_select_solver_config

  !> sgw_bicgstab_config
  TYPE, BIND(C) :: sgw_bicgstab_config
    INTEGER(c_int32_t) :: lmax
    REAL(c_double)     :: threshold
    INTEGER(c_int32_t) :: max_iter
  END TYPE sgw_bicgstab_config

  !> sgw_linear_solver_config
  TYPE, BIND(C) :: sgw_linear_solver_config
    INTEGER(c_int32_t) :: recover_subspace
    INTEGER(c_int32_t) :: max_iter
    REAL(c_double)     :: threshold
    REAL(c_double)     :: abs_threshold
  END TYPE sgw_linear_solver_config

  INTERFACE

    FUNCTION sgw_select_solver(config, AA, ctx, bb, n, sigma, nshift, xx) &
        BIND(C, name="sgw_select_solver") RESULT(status)
      IMPORT :: c_int32_t, c_int64_t, c_funptr, c_ptr, c_double_complex, &
                sgw_select_solver_config
      TYPE(sgw_select_solver_config), INTENT(IN) :: config
      TYPE(c_funptr), VALUE :: AA
      TYPE(c_ptr), VALUE :: ctx
      INTEGER(c_int64_t), VALUE :: n
      COMPLEX(c_double_complex), INTENT(IN) :: bb(n)
      INTEGER(c_int32_t), VALUE :: nshift
      COMPLEX(c_double_complex), INTENT(IN) :: sigma(nshift)
      COMPLEX(c_double_complex), INTENT(OUT) :: xx(n, nshift)
      INTEGER(c_int32_t) :: status
    END FUNCTION sgw_select_solver

    FUNCTION sgw_bicgstab(config, AA, ctx, bb, n, sigma, nshift, xx, iters) &
        BIND(C, name="sgw_bicgstab") RESULT(status)
      IMPORT :: c_int32_t, c_int64_t, c_funptr, c_ptr, c_double_complex, sgw_bicgstab_config
      TYPE(sgw_bicgstab_config), INTENT(IN) :: config
      TYPE(c_funptr), VALUE :: AA
      TYPE(c_ptr), VALUE :: ctx
      INTEGER(c_int64_t), VALUE :: n
      COMPLEX(c_double_complex), INTENT(IN) :: bb(n)
      INTEGER(c_int32_t), VALUE :: nshift
      COMPLEX(c_double_complex), INTENT(IN) :: sigma(nshift)
      COMPLEX(c_double_complex), INTENT(OUT) :: xx(n, nshift)
      TYPE(c_ptr), VALUE :: iters
      INTEGER(c_int32_t) :: status
    END FUNCTION sgw_bicgstab

    FUNCTION sgw_linear_solver(config, AA, ctx, bb, n, sigma, nshift, xx) &
        BIND(C, name="sgw_linear_solver") RESULT(status)
      IMPORT :: c_int32_t, c_int64_t, c_funptr, c_ptr, c_double_complex, &
                sgw_linear_solver_config
      TYPE(sgw_linear_solver_config), INTENT(INOUT) :: config
      TYPE(c_funptr), VALUE :: AA
      TYPE(c_ptr), VALUE :: ctx
      INTEGER(c_int64_t), VALUE :: n
      COMPLEX(c_double_complex), INTENT(IN) :: bb(n)
      INTEGER(c_int32_t), VALUE :: nshift
      COMPLEX(c_double_complex), INTENT(IN) :: sigma(nshift)
      COMPLEX(c_double_complex), INTENT(OUT) :: xx(n, nshift)
      INTEGER(c_int32_t) :: status
    END FUNCTION sgw_linear_solver

    FUNCTION sgw_last_error() BIND(C, name="sgw_last_error") RESULT(msg)
      IMPORT :: c_ptr
      TYPE(c_ptr) :: msg
    END FUNCTION sgw_last_error

  END INTERFACE

  !> Fortran-side interface of the operator procedure argument AA
  !! (algo/linear_solver/src/select_solver.f90:77-87).
  ABSTRACT INTERFACE
    SUBROUTINE op_interface(sigma, xx, Ax)
      IMPORT :: c_double_complex
      COMPLEX(c_double_complex), INTENT(IN)  :: sigma
      COMPLEX(c_double_complex), INTENT(IN)  :: xx(:)
      COMPLEX(c_double_complex), INTENT(OUT) :: Ax(:)
    END SUBROUTINE op_interface
  END INTERFACE

  !> the operator of the solve in progress (the reference's solvers are
  !! serial per rank; nested solves save and restore it)
  PROCEDURE(op_interface), POINTER :: active_op => NULL()

CONTAINS

  !> C-callable trampoline: calls the active Fortran operator with
  !! assumed-shape views of the C arrays.
  SUBROUTINE op_trampoline(sigma, x, ax, n, ctx) BIND(C)
    COMPLEX(c_double_complex), INTENT(IN) :: sigma
    INTEGER(c_int64_t), VALUE :: n
    COMPLEX(c_double_complex), INTENT(IN) :: x(n)
    COMPLEX(c_double_complex), INTENT(OUT) :: ax(n)
    TYPE(c_ptr), VALUE :: ctx
    CALL active_op(sigma, x, ax)
  END SUBROUTINE op_trampoline

  !> Report an API error and abort, as the reference's errore does.
  SUBROUTINE sgw_abort(routine, status)
    CHARACTER(*), INTENT(IN) :: routine
    INTEGER, INTENT(IN) :: status
    CHARACTER(len=512), POINTER :: msg
    CHARACTER(kind=c_char), POINTER :: cmsg(:)
    INTEGER :: i
    TYPE(c_ptr) :: p
    p = sgw_last_error()
    CALL c_f_pointer(p, cmsg, [512])
    WRITE(*, '(a)', ADVANCE='NO') ' Error in routine ' // routine // ': '
    DO i = 1, 512
      IF (cmsg(i) == c_null_char) EXIT
      WRITE(*, '(a)', ADVANCE='NO') cmsg(i)
    END DO
    WRITE(*, '(a,i0,a)') ' (status ', status, ')'
    NULLIFY(msg)
    ERROR STOP 1
  END SUBROUTINE sgw_abort

END MODULE sgw_capi
