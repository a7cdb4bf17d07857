!------------------------------------------------------------------------------
! Driver of sgw_callers_module on a plane-wave device operator: reads an
! operator description and the callers' inputs, runs green_columns (the G'
! loop of green.f90:196-218) on the Green operator and linter_bands (the band
! loop of solve_linter.f90:325-374 with the +/- omega average) on the Coulomb
! operator, and writes the results.
!
! Usage: sgw_callers_test <in.bin> <out.bin>   (little-endian stream files)
!   in:  int32 n1 n2 n3 ngroup nproj nproj_beta; int64 ld; int32 npw(ngroup);
!        real64 vloc(n1 n2 n3); int32 mill(3, ld, ngroup); real64 kin(ld, ngroup);
!        complex128 proj(ld, nproj, ngroup); real64 dcoef(nproj, ngroup);
!        int32 group ncol nomega; int32 columns(ncol); complex128 omega(nomega);
!        int32 nsys nf; int32 grp(nsys); real64 et(nsys) wg(nsys) wk(nsys);
!        complex128 freq(nf); complex128 dvpsi(ld, nsys)
!   out: complex128 green(npw(group+1), nomega, ncol); int32 ierr(ncol);
!        complex128 dpsi(ld, nf, nsys); int32 ierr(nsys)
!------------------------------------------------------------------------------
PROGRAM sgw_callers_test

  USE, INTRINSIC :: iso_c_binding
  USE sgw_kinds,            ONLY: dp
  USE sgw_capi,             ONLY: sgw_pw_desc, sgw_operator_pw, sgw_operator_destroy, sgw_abort
  USE select_solver_module, ONLY: select_solver_type, bicgstab_multi, sgw_linear_solver
  USE sgw_callers_module,   ONLY: green_columns, linter_bands

  IMPLICIT NONE

  CHARACTER(len=512) :: pin, pout
  INTEGER :: u, ngroup, nproj, nb, n, ng, ncol, nomega, nsys, nf, ld
  INTEGER(c_int32_t) :: hdr(6), gh(3), lh(2)
  INTEGER(c_int64_t) :: ld64
  INTEGER(c_int32_t), ALLOCATABLE, TARGET :: npw32(:), mill(:,:,:), cols32(:), grp(:)
  REAL(dp), ALLOCATABLE, TARGET :: vloc(:), kin(:,:), dcoef(:,:), dcoef_b(:,:)
  REAL(dp), ALLOCATABLE :: et(:), wg(:), wk(:)
  COMPLEX(dp), ALLOCATABLE, TARGET :: proj(:,:,:), proj_b(:,:,:)
  COMPLEX(dp), ALLOCATABLE :: omega(:), freq(:), dvpsi(:,:), green(:,:,:), dpsi(:,:,:)
  INTEGER, ALLOCATABLE :: npw(:), columns(:), ierr_g(:), ierr_l(:)
  TYPE(sgw_pw_desc) :: desc
  TYPE(c_ptr) :: op_c, op_g
  TYPE(select_solver_type) :: config

  CALL GET_COMMAND_ARGUMENT(1, pin)
  CALL GET_COMMAND_ARGUMENT(2, pout)
  OPEN(NEWUNIT=u, FILE=TRIM(pin), ACCESS='stream', FORM='unformatted', STATUS='old')
  READ(u) hdr
  READ(u) ld64
  ngroup = hdr(4); nproj = hdr(5); nb = hdr(6); ld = INT(ld64)
  n = hdr(1) * hdr(2) * hdr(3)
  ALLOCATE(npw32(ngroup), vloc(n), mill(3, ld, ngroup), kin(ld, ngroup), &
           proj(ld, nproj, ngroup), dcoef(nproj, ngroup))
  READ(u) npw32, vloc, mill, kin, proj, dcoef
  READ(u) gh
  ng = gh(1); ncol = gh(2); nomega = gh(3)
  ALLOCATE(cols32(ncol), omega(nomega))
  READ(u) cols32, omega
  READ(u) lh
  nsys = lh(1); nf = lh(2)
  ALLOCATE(grp(nsys), et(nsys), wg(nsys), wk(nsys), freq(nf), dvpsi(ld, nsys))
  READ(u) grp, et, wg, wk, freq, dvpsi
  CLOSE(u)
  ALLOCATE(npw(ngroup), columns(ncol))
  npw = INT(npw32)
  columns = INT(cols32)

  ! the Coulomb operator (beta projectors, then the valence states with alpha_pv)
  desc%n1 = hdr(1); desc%n2 = hdr(2); desc%n3 = hdr(3)
  desc%vloc = c_loc(vloc)
  desc%ngroup = INT(ngroup, c_int32_t)
  desc%npw = c_loc(npw32)
  desc%ld = ld64
  desc%mill = c_loc(mill)
  desc%kin = c_loc(kin)
  desc%nproj = INT(nproj, c_int32_t)
  desc%proj = c_loc(proj)
  desc%dcoef = c_loc(dcoef)
  op_c = sgw_operator_pw(desc)
  IF (.NOT. c_associated(op_c)) CALL sgw_abort("sgw_operator_pw", -1)
  ! the Green operator: the beta projectors only (green.f90:282-283)
  ALLOCATE(proj_b(ld, nb, ngroup), dcoef_b(nb, ngroup))
  proj_b = proj(:, 1:nb, :)
  dcoef_b = dcoef(1:nb, :)
  desc%nproj = INT(nb, c_int32_t)
  desc%proj = c_loc(proj_b)
  desc%dcoef = c_loc(dcoef_b)
  op_g = sgw_operator_pw(desc)
  IF (.NOT. c_associated(op_g)) CALL sgw_abort("sgw_operator_pw", -1)

  ! the default priority list [1, 3] (main/src/gw_input.yml:151-158)
  ALLOCATE(config%priority(2))
  config%priority = [bicgstab_multi, sgw_linear_solver]
  config%threshold = 1e-12_dp

  ALLOCATE(green(npw(ng + 1), nomega, ncol), ierr_g(ncol))
  CALL green_columns(op_g, config, ld, npw, ng, columns, omega, green, ierr_g)
  ALLOCATE(dpsi(ld, nf, nsys), ierr_l(nsys))
  CALL linter_bands(op_c, config, ld, npw, grp, et, freq, wg, wk, dvpsi, dpsi, ierr_l)

  OPEN(NEWUNIT=u, FILE=TRIM(pout), ACCESS='stream', FORM='unformatted', STATUS='replace')
  WRITE(u) green, INT(ierr_g, c_int32_t), dpsi, INT(ierr_l, c_int32_t)
  CLOSE(u)
  CALL sgw_operator_destroy(op_c)
  CALL sgw_operator_destroy(op_g)
  WRITE(*, '(a,i0,a,i0,a)') 'green_columns: ', ncol, ' columns; linter_bands: ', nsys, ' bands'
  WRITE(*, '(a)') 'DONE'

END PROGRAM sgw_callers_test
