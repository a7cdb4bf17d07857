!------------------------------------------------------------------------------
! Drop-in Fortran modules for SternheimerGW's linear-solver path.
!
! The modules keep the public names, derived types, defaults, argument lists
! and ierr codes of
!   bicgstab_module       algo/linear_solver/src/bicgstab.f90:33-146,154
!   linear_solver_module  algo/linear_solver/src/linear_solver.f90:29-56,82
!   select_solver_module  algo/linear_solver/src/select_solver.f90:32-67
! and forward every solve through ISO_C_BINDING to libsgw_hip.so, whose
! solvers keep their state on the GPU.  The operator procedure argument AA is
! called back through a BIND(C) trampoline with assumed-shape views.
!------------------------------------------------------------------------------
MODULE bicgstab_module

  USE sgw_kinds, ONLY: dp

  IMPLICIT NONE

  PRIVATE

  PUBLIC bicgstab, bicgstab_type

  !> Configuration of the BiCGstab solver (bicgstab.f90:44-55)
  TYPE bicgstab_type
    INTEGER  :: lmax = 4
    REAL(dp) :: threshold = 1e-4_dp
    INTEGER  :: max_iter = 10000
  END TYPE bicgstab_type

CONTAINS

  SUBROUTINE bicgstab(config, AA, bb, sigma, xx, ierr)

    USE, INTRINSIC :: iso_c_binding
    USE sgw_capi

    TYPE(bicgstab_type), INTENT(IN) :: config
    INTERFACE
      SUBROUTINE AA(sigma, xx, Ax)
        USE sgw_kinds, ONLY: dp
        COMPLEX(dp), INTENT(IN)  :: sigma
        COMPLEX(dp), INTENT(IN)  :: xx(:)
        COMPLEX(dp), INTENT(OUT) :: Ax(:)
      END SUBROUTINE AA
    END INTERFACE
    COMPLEX(dp), INTENT(IN)  :: bb(:)
    COMPLEX(dp), INTENT(IN)  :: sigma(:)
    COMPLEX(dp), INTENT(OUT) :: xx(:,:)
    INTEGER,     INTENT(OUT) :: ierr

    TYPE(sgw_bicgstab_config) :: cfg
    PROCEDURE(op_interface), POINTER :: saved
    INTEGER(c_int32_t) :: status
    COMPLEX(dp), ALLOCATABLE :: xbuf(:,:)

    IF (SIZE(xx, 1) /= SIZE(bb)) &
      CALL errore_shim("bicgstab", "right-hand side and solution must have same size")
    IF (SIZE(xx, 2) /= SIZE(sigma)) &
      CALL errore_shim("bicgstab", "we need one solution vector per shift for the result")

    cfg%lmax = config%lmax
    cfg%threshold = config%threshold
    cfg%max_iter = config%max_iter
    saved => active_op
    active_op => AA
    ALLOCATE(xbuf(SIZE(bb), SIZE(sigma)))
    status = sgw_bicgstab(cfg, C_FUNLOC(op_trampoline), C_NULL_PTR, bb, &
                          INT(SIZE(bb), c_int64_t), sigma, INT(SIZE(sigma), c_int32_t), &
                          xbuf, C_NULL_PTR)
    active_op => saved
    IF (status < 0) CALL sgw_abort("bicgstab", INT(status))
    xx = xbuf
    ierr = INT(status)

  END SUBROUTINE bicgstab

  SUBROUTINE errore_shim(routine, msg)
    CHARACTER(*), INTENT(IN) :: routine, msg
    WRITE(*, '(a)') ' Error in routine ' // routine // ': ' // msg
    ERROR STOP 1
  END SUBROUTINE errore_shim

END MODULE bicgstab_module

MODULE linear_solver_module

  USE sgw_kinds, ONLY: dp

  IMPLICIT NONE

  PRIVATE
  PUBLIC linear_solver, linear_solver_config

  INTEGER, PARAMETER :: LINEAR_SOLVER_SCRATCH = 1

  !> linear_solver.f90:42-56
  TYPE linear_solver_config
    INTEGER  :: recover_subspace = LINEAR_SOLVER_SCRATCH
    INTEGER  :: max_iter = 10000
    REAL(dp) :: threshold = 1e-4_dp
    REAL(dp) :: abs_threshold
  END TYPE linear_solver_config

CONTAINS

  SUBROUTINE linear_solver(config, AA, bb, sigma, xx, ierr)

    USE, INTRINSIC :: iso_c_binding
    USE sgw_capi

    TYPE(linear_solver_config), INTENT(INOUT) :: config
    INTERFACE
      SUBROUTINE AA(sigma, xx, Ax)
        USE sgw_kinds, ONLY: dp
        COMPLEX(dp), INTENT(IN)  :: sigma
        COMPLEX(dp), INTENT(IN)  :: xx(:)
        COMPLEX(dp), INTENT(OUT) :: Ax(:)
      END SUBROUTINE AA
    END INTERFACE
    COMPLEX(dp), INTENT(IN)  :: bb(:)
    COMPLEX(dp), INTENT(IN)  :: sigma(:)
    COMPLEX(dp), INTENT(OUT) :: xx(:,:)
    INTEGER,     INTENT(OUT) :: ierr

    TYPE(sgw_linear_solver_config) :: cfg
    PROCEDURE(op_interface), POINTER :: saved
    INTEGER(c_int32_t) :: status
    COMPLEX(dp), ALLOCATABLE :: xbuf(:,:)

    cfg%recover_subspace = config%recover_subspace
    cfg%max_iter = config%max_iter
    cfg%threshold = config%threshold
    cfg%abs_threshold = 0.0_dp
    saved => active_op
    active_op => AA
    ALLOCATE(xbuf(SIZE(bb), SIZE(sigma)))
    status = sgw_linear_solver(cfg, C_FUNLOC(op_trampoline), C_NULL_PTR, bb, &
                               INT(SIZE(bb), c_int64_t), sigma, INT(SIZE(sigma), c_int32_t), xbuf)
    active_op => saved
    IF (status < 0) CALL sgw_abort("linear_solver", INT(status))
    config%abs_threshold = cfg%abs_threshold
    xx = xbuf
    ierr = INT(status)

  END SUBROUTINE linear_solver

END MODULE linear_solver_module

MODULE select_solver_module

  USE sgw_kinds, ONLY: dp

  IMPLICIT NONE

  !> select_solver.f90:39-45
  INTEGER,  PARAMETER :: bicgstab_multi = 1
  INTEGER,  PARAMETER :: bicgstab_no_multi = 2
  INTEGER,  PARAMETER :: sgw_linear_solver = 3

  !> select_solver.f90:48-62
  TYPE select_solver_type
    INTEGER, ALLOCATABLE :: priority(:)
    INTEGER  :: max_iter = 10000
    REAL(dp) :: threshold = 1e-4
    INTEGER  :: bicg_lmax = 4
  END TYPE select_solver_type

CONTAINS

  SUBROUTINE select_solver(config, AA, bb, sigma, xx, ierr)

    USE, INTRINSIC :: iso_c_binding
    USE sgw_capi, ONLY: sgw_select_solver_config, op_interface, active_op, op_trampoline, &
                        sgw_abort, sgw_select_solver => sgw_select_solver

    TYPE(select_solver_type), INTENT(IN) :: config
    INTERFACE
      SUBROUTINE AA(sigma, xx, Ax)
        USE sgw_kinds, ONLY: dp
        COMPLEX(dp), INTENT(IN)  :: sigma
        COMPLEX(dp), INTENT(IN)  :: xx(:)
        COMPLEX(dp), INTENT(OUT) :: Ax(:)
      END SUBROUTINE AA
    END INTERFACE
    COMPLEX(dp), INTENT(IN)  :: bb(:)
    COMPLEX(dp), INTENT(IN)  :: sigma(:)
    COMPLEX(dp), INTENT(OUT) :: xx(:,:)
    INTEGER,     INTENT(OUT) :: ierr

    TYPE(sgw_select_solver_config) :: cfg
    INTEGER(c_int32_t), ALLOCATABLE, TARGET :: prio(:)
    PROCEDURE(op_interface), POINTER :: saved
    INTEGER(c_int32_t) :: status
    COMPLEX(dp), ALLOCATABLE :: xbuf(:,:)

    IF (.NOT.ALLOCATED(config%priority)) THEN
      WRITE(*, '(a)') ' Error in routine select_solver: priority of the solvers not specified'
      ERROR STOP 1
    END IF
    ALLOCATE(prio(SIZE(config%priority)))
    prio = INT(config%priority, c_int32_t)
    cfg%priority = C_LOC(prio)
    cfg%npriority = INT(SIZE(prio), c_int32_t)
    cfg%max_iter = config%max_iter
    cfg%threshold = config%threshold
    cfg%bicg_lmax = config%bicg_lmax
    saved => active_op
    active_op => AA
    ALLOCATE(xbuf(SIZE(bb), SIZE(sigma)))
    status = sgw_select_solver(cfg, C_FUNLOC(op_trampoline), C_NULL_PTR, bb, &
                               INT(SIZE(bb), c_int64_t), sigma, INT(SIZE(sigma), c_int32_t), xbuf)
    active_op => saved
    IF (status < 0) CALL sgw_abort("select_solver", INT(status))
    xx = xbuf
    ierr = INT(status)

  END SUBROUTINE select_solver

END MODULE select_solver_module
