!------------------------------------------------------------------------------
! Fortran driver of the reference's solver unit test through the drop-in
! modules: algo/linear_solver/test/linear_solver.pf (test_bicgstab :125-259,
! test_linear_solver :264-345) with the same operator, bounds and loops,
! calling select_solver_module / bicgstab_module / linear_solver_module of
! this package (GPU solvers behind ISO_C_BINDING).
!
! Usage: sgw_fixture_test <problem.bin> [lmin lmax]
! problem.bin: int32 n, int32 nshift, complex128 sigma(nshift),
!              complex128 A(n,n) (column-major), complex128 b(n).
!------------------------------------------------------------------------------
MODULE fixture_problem

  USE sgw_kinds, ONLY: dp

  IMPLICIT NONE

  COMPLEX(dp), ALLOCATABLE :: AA(:,:)

CONTAINS

  !> linear_solver.pf:100-113
  SUBROUTINE linear_problem_apply(sigma, xx, Ax)
    COMPLEX(dp), INTENT(IN)  :: sigma
    COMPLEX(dp), INTENT(IN)  :: xx(:)
    COMPLEX(dp), INTENT(OUT) :: Ax(:)
    Ax = MATMUL(AA, xx) + sigma * xx
  END SUBROUTINE linear_problem_apply

END MODULE fixture_problem

PROGRAM sgw_fixture_test

  USE, INTRINSIC :: iso_c_binding, ONLY: c_int32_t, c_int64_t, c_ptr, c_associated
  USE sgw_kinds,            ONLY: dp
  USE sgw_capi,             ONLY: sgw_batch_config, sgw_operator_dense, sgw_operator_destroy, &
                                  sgw_sternheimer_solve_batched
  USE fixture_problem,      ONLY: AA, linear_problem_apply
  USE bicgstab_module,      ONLY: bicgstab, bicgstab_type
  USE linear_solver_module, ONLY: linear_solver, linear_solver_config
  USE select_solver_module, ONLY: select_solver, select_solver_type, bicgstab_multi, &
                                  sgw_linear_solver

  IMPLICIT NONE

  CHARACTER(len=512) :: path, arg
  INTEGER(c_int32_t) :: n32, ns32
  INTEGER :: vec_size, num_shift, ll, ishift, ierr, lmin, lmax, nfail, u
  COMPLEX(dp), ALLOCATABLE :: sigma(:), bb(:), solution(:,:), work(:)
  REAL(dp), PARAMETER :: threshold = 1e-6_dp
  REAL(dp) :: residual, worst
  TYPE(bicgstab_type) :: config
  TYPE(linear_solver_config) :: lconfig
  TYPE(select_solver_type) :: sconfig

  CALL GET_COMMAND_ARGUMENT(1, path)
  lmin = 1
  lmax = 15
  IF (COMMAND_ARGUMENT_COUNT() >= 3) THEN
    CALL GET_COMMAND_ARGUMENT(2, arg); READ(arg, *) lmin
    CALL GET_COMMAND_ARGUMENT(3, arg); READ(arg, *) lmax
  END IF

  OPEN(NEWUNIT=u, FILE=TRIM(path), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(u) n32, ns32
  vec_size = n32
  num_shift = ns32
  ALLOCATE(sigma(num_shift), AA(vec_size, vec_size), bb(vec_size), work(vec_size))
  READ(u) sigma
  READ(u) AA
  READ(u) bb
  CLOSE(u)
  nfail = 0

  ! test the solver without multishift (linear_solver.pf:194-222)
  ALLOCATE(solution(vec_size, 1))
  worst = 0
  DO ll = lmin, lmax
    DO ishift = 1, num_shift
      config%threshold = threshold
      config%lmax = ll
      CALL bicgstab(config, linear_problem_apply, bb, sigma(ishift:ishift), solution, ierr)
      IF (ierr /= 0) nfail = nfail + 1
      IF (ANY(solution /= solution)) nfail = nfail + 1
      CALL linear_problem_apply(sigma(ishift), solution(:, 1), work)
      work = work - bb
      residual = SQRT(SUM(ABS(work)**2))
      worst = MAX(worst, residual)
      IF (residual > threshold) nfail = nfail + 1
    END DO
  END DO
  WRITE(*, '(a,es10.3,a,i0)') 'single-shift bicgstab: max residual ', worst, ' failures ', nfail
  DEALLOCATE(solution)

  ! test the solver with multishift (linear_solver.pf:224-254)
  ALLOCATE(solution(vec_size, num_shift))
  worst = 0
  DO ll = lmin, lmax
    config%threshold = threshold
    config%lmax = ll
    CALL bicgstab(config, linear_problem_apply, bb, sigma, solution, ierr)
    IF (ierr /= 0) nfail = nfail + 1
    DO ishift = 1, num_shift
      CALL linear_problem_apply(sigma(ishift), solution(:, ishift), work)
      work = work - bb
      residual = SQRT(SUM(ABS(work)**2))
      worst = MAX(worst, residual)
      IF (residual > 10.0_dp * threshold) nfail = nfail + 1
    END DO
  END DO
  WRITE(*, '(a,es10.3,a,i0)') 'multishift bicgstab:   max residual ', worst, ' failures ', nfail

  ! test the linear solver with default configuration (linear_solver.pf:264-345)
  CALL linear_solver(lconfig, linear_problem_apply, bb, sigma, solution, ierr)
  IF (ierr /= 0) nfail = nfail + 1
  worst = 0
  DO ishift = 1, num_shift
    CALL linear_problem_apply(sigma(ishift), solution(:, ishift), work)
    work = work - bb
    residual = SQRT(SUM(ABS(work)**2)) / SQRT(SUM(ABS(bb)**2))
    worst = MAX(worst, residual)
    IF (residual > lconfig%threshold) nfail = nfail + 1
  END DO
  WRITE(*, '(a,es10.3,a,i0)') 'linear_solver:         max rel. residual ', worst, ' failures ', nfail

  ! select_solver with the default priority list [1, 3] (main/src/gw_input.yml:151-158)
  ALLOCATE(sconfig%priority(2))
  sconfig%priority = [bicgstab_multi, sgw_linear_solver]
  sconfig%threshold = threshold
  CALL select_solver(sconfig, linear_problem_apply, bb, sigma, solution, ierr)
  IF (ierr /= 0) nfail = nfail + 1
  worst = 0
  DO ishift = 1, num_shift
    CALL linear_problem_apply(sigma(ishift), solution(:, ishift), work)
    work = work - bb
    worst = MAX(worst, SQRT(SUM(ABS(work)**2)))
  END DO
  IF (worst > 10.0_dp * threshold) nfail = nfail + 1
  WRITE(*, '(a,es10.3,a,i0)') 'select_solver [1,3]:   max residual ', worst, ' failures ', nfail

  ! the batched entry on a dense device operator: two systems (bb and i*bb),
  ! all shifts each, in the layout a green_function / solve_linter shim passes
  BLOCK
    TYPE(c_ptr) :: op
    TYPE(sgw_batch_config) :: bcfg
    INTEGER(c_int32_t) :: grp(2), berr(2), bit(2), st
    COMPLEX(dp), ALLOCATABLE :: rhs2(:,:), sig2(:,:), x2(:,:,:)
    INTEGER :: isys
    op = sgw_operator_dense(AA, INT(vec_size, c_int64_t))
    IF (.NOT. c_associated(op)) THEN
      nfail = nfail + 1
    ELSE
      ALLOCATE(rhs2(vec_size, 2), sig2(num_shift, 2), x2(vec_size, num_shift, 2))
      rhs2(:, 1) = bb
      rhs2(:, 2) = (0.0_dp, 1.0_dp) * bb
      sig2(:, 1) = sigma
      sig2(:, 2) = sigma
      grp = 0
      bcfg = sgw_batch_config(2, num_shift, INT(vec_size, c_int64_t), 4, 10000, threshold, 0, 0.0_dp)
      st = sgw_sternheimer_solve_batched(op, bcfg, grp, rhs2, sig2, x2, berr, bit)
      IF (st /= 0 .OR. ANY(berr /= 0)) nfail = nfail + 1
      worst = 0
      DO isys = 1, 2
        DO ishift = 1, num_shift
          CALL linear_problem_apply(sig2(ishift, isys), x2(:, ishift, isys), work)
          work = work - rhs2(:, isys)
          worst = MAX(worst, SQRT(SUM(ABS(work)**2)))
        END DO
      END DO
      IF (worst > 10.0_dp * threshold) nfail = nfail + 1
      IF (ANY(ABS(x2(:, :, 2) - (0.0_dp, 1.0_dp) * x2(:, :, 1)) > 1e-12_dp)) nfail = nfail + 1
      CALL sgw_operator_destroy(op)
    END IF
    WRITE(*, '(a,es10.3,a,i0)') 'batched (dense op):    max residual ', worst, ' failures ', nfail
  END BLOCK

  IF (nfail /= 0) THEN
    WRITE(*, '(a)') 'FAIL'
    ERROR STOP 1
  END IF
  WRITE(*, '(a)') 'PASS'

END PROGRAM sgw_fixture_test
