!------------------------------------------------------------------------------
! The solver's two callers as batched plain-array routines on a device
! operator (sgw_operator_pw / sgw_operator_dense of include/sgw.h):
!
!   green_columns  the G' column loop of green_function,
!                  phys/green/src/green.f90:196-218: for every listed column
!                  (H - omega) G(:, G') = -delta_{G'}, shifts sigma = -omega;
!   linter_bands   the band loop of the first solve_linter iteration,
!                  phys/coul/src/solve_linter.f90:237-252, 325-374, 462-480:
!                  shifts -(et(v) + omega) over [freq, -freq(2:)], the wg/wk
!                  factor and the +/- omega average.
!
! Every column (band) of one call is one system of a single batched solve
! (sgw_sternheimer_solve_batched), so the whole loop stays on the GPU.  The
! priority list of select_solver_type keeps its meaning: when it starts with
! the multishift BiCGStab (1, or 2, which gives the same result), every system
! is solved batched first; systems left with ierr /= 0 go down the rest of the
! list one by one (select_solver.f90:121-159) through sgw_select_solver_op.
! ierr is returned per system; the reference's callers abort on a non-zero
! one (green.f90:208, solve_linter.f90:370), which the caller does here.
!------------------------------------------------------------------------------
MODULE sgw_callers_module

  USE, INTRINSIC :: iso_c_binding
  USE sgw_kinds,            ONLY: dp
  USE select_solver_module, ONLY: select_solver_type, bicgstab_multi, bicgstab_no_multi

  IMPLICIT NONE

  PRIVATE

  PUBLIC green_columns, linter_bands

  INTERFACE
    FUNCTION c_select_solver_op(config, op, group, bb, n, sigma, nshift, xx) &
        BIND(C, name="sgw_select_solver_op") RESULT(status)
      IMPORT :: c_ptr, c_int32_t, c_int64_t, c_double_complex
      TYPE(c_ptr), VALUE :: config
      TYPE(c_ptr), VALUE :: op
      INTEGER(c_int32_t), VALUE :: group
      COMPLEX(c_double_complex), INTENT(IN) :: bb(*)
      INTEGER(c_int64_t), VALUE :: n
      COMPLEX(c_double_complex), INTENT(IN) :: sigma(*)
      INTEGER(c_int32_t), VALUE :: nshift
      COMPLEX(c_double_complex), INTENT(OUT) :: xx(*)
      INTEGER(c_int32_t) :: status
    END FUNCTION c_select_solver_op
  END INTERFACE

CONTAINS

  !> select_solver over a batch of systems on one device operator.
  !! rhs(ld, nsys), sigma(nshift, nsys), x(ld, nshift, nsys); npw(g) plane
  !! waves of k+q group g (0-based group(s), as in the C interface).
  SUBROUTINE batched_select(op, config, ld, npw, group, rhs, sigma, x, ierr)

    USE sgw_capi, ONLY: sgw_batch_config, sgw_select_solver_config, &
                        sgw_sternheimer_solve_batched, sgw_abort

    TYPE(c_ptr),              INTENT(IN)  :: op
    TYPE(select_solver_type), INTENT(IN)  :: config
    INTEGER,                  INTENT(IN)  :: ld
    INTEGER,                  INTENT(IN)  :: npw(:)
    INTEGER(c_int32_t),       INTENT(IN)  :: group(:)
    COMPLEX(dp),              INTENT(IN)  :: rhs(:,:), sigma(:,:)
    COMPLEX(dp),              INTENT(OUT) :: x(:,:,:)
    INTEGER,                  INTENT(OUT) :: ierr(:)

    TYPE(sgw_batch_config) :: bcfg
    TYPE(sgw_select_solver_config), TARGET :: scfg
    INTEGER(c_int32_t), ALLOCATABLE, TARGET :: rest(:)
    INTEGER(c_int32_t), ALLOCATABLE :: berr(:), bit(:)
    INTEGER(c_int32_t) :: status
    INTEGER :: nsys, nshift, first, isys, n
    COMPLEX(dp), ALLOCATABLE :: xs(:,:)

    IF (.NOT. ALLOCATED(config%priority)) THEN
      WRITE(*, '(a)') ' Error in routine batched_select: priority of the solvers not specified'
      ERROR STOP 1
    END IF
    nsys = SIZE(rhs, 2)
    nshift = SIZE(sigma, 1)
    x = (0.0_dp, 0.0_dp)
    ierr = 1
    first = 1
    IF (config%priority(1) == bicgstab_multi .OR. config%priority(1) == bicgstab_no_multi) THEN
      ALLOCATE(berr(nsys), bit(nsys))
      bcfg = sgw_batch_config(INT(nsys, c_int32_t), INT(nshift, c_int32_t), &
                              INT(ld, c_int64_t), INT(config%bicg_lmax, c_int32_t), &
                              INT(config%max_iter, c_int32_t), config%threshold, 0_c_int32_t, &
                              0.0_dp)
      status = sgw_sternheimer_solve_batched(op, bcfg, group, rhs, sigma, x, berr, bit)
      IF (status < 0) CALL sgw_abort("batched_select", INT(status))
      ierr = INT(berr)
      first = 2
    END IF
    IF (first > SIZE(config%priority) .OR. ALL(ierr == 0)) RETURN
    ! the rest of the priority list, system by system
    ALLOCATE(rest(SIZE(config%priority) - first + 1))
    rest = INT(config%priority(first:), c_int32_t)
    scfg%priority = c_loc(rest)
    scfg%npriority = INT(SIZE(rest), c_int32_t)
    scfg%max_iter = INT(config%max_iter, c_int32_t)
    scfg%threshold = config%threshold
    scfg%bicg_lmax = INT(config%bicg_lmax, c_int32_t)
    DO isys = 1, nsys
      IF (ierr(isys) == 0) CYCLE
      WRITE(*, '(a)') 'First choice of solver did not converge, try a different one'
      n = npw(group(isys) + 1)
      ALLOCATE(xs(n, nshift))
      status = c_select_solver_op(c_loc(scfg), op, group(isys), rhs(1:n, isys), &
                                  INT(n, c_int64_t), sigma(:, isys), INT(nshift, c_int32_t), xs)
      IF (status < 0) CALL sgw_abort("batched_select", INT(status))
      x(:, :, isys) = (0.0_dp, 0.0_dp)
      x(1:n, :, isys) = xs
      ierr(isys) = INT(status)
      DEALLOCATE(xs)
    END DO

  END SUBROUTINE batched_select

  !> green.f90:196-218 for the columns `columns` (1-based plane-wave indices
  !! of the k+q sphere of group `group`, 0-based) at the frequencies omega:
  !! green(1:npw(group+1), iomega, icol) = G(:, columns(icol); omega(iomega)).
  !! `op` is the Green operator (no valence projector, green.f90:282-283).
  SUBROUTINE green_columns(op, config, ld, npw, group, columns, omega, green, ierr)

    TYPE(c_ptr),              INTENT(IN)  :: op
    TYPE(select_solver_type), INTENT(IN)  :: config
    INTEGER,                  INTENT(IN)  :: ld
    INTEGER,                  INTENT(IN)  :: npw(:)
    INTEGER,                  INTENT(IN)  :: group
    INTEGER,                  INTENT(IN)  :: columns(:)
    COMPLEX(dp),              INTENT(IN)  :: omega(:)
    COMPLEX(dp),              INTENT(OUT) :: green(:,:,:)
    INTEGER,                  INTENT(OUT) :: ierr(:)

    COMPLEX(dp), ALLOCATABLE :: rhs(:,:), sigma(:,:), x(:,:,:)
    INTEGER(c_int32_t), ALLOCATABLE :: grp(:)
    INTEGER :: ncol, icol, n

    ncol = SIZE(columns)
    n = npw(group + 1)
    IF (SIZE(green, 1) < n .OR. SIZE(green, 2) /= SIZE(omega) .OR. SIZE(green, 3) /= ncol) THEN
      WRITE(*, '(a)') ' Error in routine green_columns: green has the wrong shape'
      ERROR STOP 1
    END IF
    ALLOCATE(rhs(ld, ncol), sigma(SIZE(omega), ncol), x(ld, SIZE(omega), ncol), grp(ncol))
    rhs = (0.0_dp, 0.0_dp)
    DO icol = 1, ncol
      rhs(columns(icol), icol) = (-1.0_dp, 0.0_dp)             ! green.f90:203-204
      sigma(:, icol) = -omega                                 ! green.f90:207
    END DO
    grp = INT(group, c_int32_t)
    CALL batched_select(op, config, ld, npw, grp, rhs, sigma, x, ierr)
    green = (0.0_dp, 0.0_dp)
    green(1:n, :, :) = x(1:n, :, :)

  END SUBROUTINE green_columns

  !> First solve_linter iteration for the systems s (band et(s) of k+q group
  !! group(s), 0-based): dvpsi(:, s) is -P_c^+ dV psi_v, dpsi(:, ifreq, s)
  !! receives the solution averaged over +/- freq(ifreq) and scaled by
  !! wg(s) / wk(s) (solve_linter.f90:237-252, 325-374, 462-480).
  SUBROUTINE linter_bands(op, config, ld, npw, group, et, freq, wg, wk, dvpsi, dpsi, ierr)

    TYPE(c_ptr),              INTENT(IN)  :: op
    TYPE(select_solver_type), INTENT(IN)  :: config
    INTEGER,                  INTENT(IN)  :: ld
    INTEGER,                  INTENT(IN)  :: npw(:)
    INTEGER(c_int32_t),       INTENT(IN)  :: group(:)
    REAL(dp),                 INTENT(IN)  :: et(:), wg(:), wk(:)
    COMPLEX(dp),              INTENT(IN)  :: freq(:)
    COMPLEX(dp),              INTENT(IN)  :: dvpsi(:,:)
    COMPLEX(dp),              INTENT(OUT) :: dpsi(:,:,:)
    INTEGER,                  INTENT(OUT) :: ierr(:)

    COMPLEX(dp), ALLOCATABLE :: omega(:), sigma(:,:), x(:,:,:)
    LOGICAL :: zero_freq
    INTEGER :: nf, nshift, nsys, isys, ifreq, ineg

    nf = SIZE(freq)
    nsys = SIZE(dvpsi, 2)
    zero_freq = ABS(freq(1)) < 1e-14_dp                       ! solve_linter.f90:237-252
    IF (zero_freq) THEN
      nshift = 2 * nf - 1
    ELSE
      nshift = 2 * nf
    END IF
    ALLOCATE(omega(nshift), sigma(nshift, nsys), x(ld, nshift, nsys))
    omega(1:nf) = freq
    IF (zero_freq) THEN
      omega(nf + 1:) = -freq(2:)
    ELSE
      omega(nf + 1:) = -freq
    END IF
    DO isys = 1, nsys
      sigma(:, isys) = -(et(isys) + omega)                     ! solve_linter.f90:369
    END DO
    CALL batched_select(op, config, ld, npw, group, dvpsi, sigma, x, ierr)
    DO isys = 1, nsys
      x(:, :, isys) = x(:, :, isys) * (wg(isys) / wk(isys))   ! solve_linter.f90:373
      dpsi(:, 1, isys) = x(:, 1, isys)
      DO ifreq = 1, nf                                         ! solve_linter.f90:462-480
        IF (zero_freq .AND. ifreq == 1) CYCLE
        IF (zero_freq) THEN
          ineg = nf + ifreq - 1
        ELSE
          ineg = nf + ifreq
        END IF
        dpsi(:, ifreq, isys) = 0.5_dp * x(:, ifreq, isys) + 0.5_dp * x(:, ineg, isys)
      END DO
    END DO

  END SUBROUTINE linter_bands

END MODULE sgw_callers_module
