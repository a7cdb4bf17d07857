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

  !> sgw_pw_desc: the plane-wave operator's data (include/sgw.h)
  TYPE, BIND(C) :: sgw_pw_desc
    INTEGER(c_int32_t) :: n1, n2, n3
    TYPE(c_ptr)        :: vloc      ! REAL(dp) vloc(n1*n2*n3)
    INTEGER(c_int32_t) :: ngroup
    TYPE(c_ptr)        :: npw       ! INTEGER(c_int32_t) npw(ngroup)
    INTEGER(c_int64_t) :: ld
    TYPE(c_ptr)        :: mill      ! INTEGER(c_int32_t) mill(3, ld, ngroup)
    TYPE(c_ptr)        :: kin       ! REAL(dp) kin(ld, ngroup), Ry
    INTEGER(c_int32_t) :: nproj
    TYPE(c_ptr)        :: proj      ! COMPLEX(dp) proj(ld, nproj, ngroup)
    TYPE(c_ptr)        :: dcoef     ! REAL(dp) dcoef(nproj, ngroup), Ry
  END TYPE sgw_pw_desc

  !> sgw_batch_config
  TYPE, BIND(C) :: sgw_batch_config
    INTEGER(c_int32_t) :: nsys
    INTEGER(c_int32_t) :: nshift
    INTEGER(c_int64_t) :: ld
    INTEGER(c_int32_t) :: lmax
    INTEGER(c_int32_t) :: max_iter
    REAL(c_double)     :: threshold
    INTEGER(c_int32_t) :: shift_test   ! 0: seed residual only (reference); 1: every shift
    REAL(c_double)     :: relax_threshold  ! 0: FP64; > 0: FP32 FFTs below this residual
  END TYPE sgw_batch_config

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

    !> Dense device operator (the unit test's linear_problem_apply)
    FUNCTION sgw_operator_dense(a, n) BIND(C, name="sgw_operator_dense") RESULT(op)
      IMPORT :: c_ptr, c_int64_t, c_double_complex
      INTEGER(c_int64_t), VALUE :: n
      COMPLEX(c_double_complex), INTENT(IN) :: a(n, n)
      TYPE(c_ptr) :: op
    END FUNCTION sgw_operator_dense

    !> Plane-wave device operator (h_psi + alpha_pv P_v of linear_op)
    FUNCTION sgw_operator_pw(desc) BIND(C, name="sgw_operator_pw") RESULT(op)
      IMPORT :: c_ptr, sgw_pw_desc
      TYPE(sgw_pw_desc), INTENT(IN) :: desc
      TYPE(c_ptr) :: op
    END FUNCTION sgw_operator_pw

    SUBROUTINE sgw_operator_destroy(op) BIND(C, name="sgw_operator_destroy")
      IMPORT :: c_ptr
      TYPE(c_ptr), VALUE :: op
    END SUBROUTINE sgw_operator_destroy

    !> All (system, shift) pairs of a band/q or G' loop in one call;
    !! rhs(ld, nsys), sigma(nshift, nsys), x(ld, nshift, nsys) in Fortran order
    FUNCTION sgw_sternheimer_solve_batched(op, cfg, group, rhs, sigma, x, ierr, iters) &
        BIND(C, name="sgw_sternheimer_solve_batched") RESULT(status)
      IMPORT :: c_ptr, c_int32_t, c_double_complex, sgw_batch_config
      TYPE(c_ptr), VALUE :: op
      TYPE(sgw_batch_config), INTENT(IN) :: cfg
      INTEGER(c_int32_t), INTENT(IN) :: group(*)
      COMPLEX(c_double_complex), INTENT(IN) :: rhs(*)
      COMPLEX(c_double_complex), INTENT(IN) :: sigma(*)
      COMPLEX(c_double_complex), INTENT(OUT) :: x(*)
      INTEGER(c_int32_t), INTENT(OUT) :: ierr(*)
      INTEGER(c_int32_t), INTENT(OUT) :: iters(*)
      INTEGER(c_int32_t) :: status
    END FUNCTION sgw_sternheimer_solve_batched

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
