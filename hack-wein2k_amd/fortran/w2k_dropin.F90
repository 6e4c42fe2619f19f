! HAMILT, HNS and SECLR4 of lapw1 on one MI355X through the C-ABI of
! libw2klapw1.so (include/w2k_lapw1.h), reached with ISO_C_BINDING.
!
! Drop-in for SRC_lapw1/hamilt.F:1-2, SRC_lapw1/hns.F:1-3 and
! SRC_lapw1/seclr4.F:3: the same subroutine names and argument lists, the
! same module inputs (matrices, albl, loabc, lolog, rotmat, struk, out,
! atspdt, loint, comi, comr) and the same module outputs (HS, HSDIAG, AL, BL,
! EIGVAL, Z, NE, NBELW), so a sequential lapw1 build links this file in place
! of those three sources and CALKPT (calkpt.F:143-197) runs unchanged.
!
! Each routine hands the host arrays over and back, exactly as the reference
! leaves them between the calls (CALKPT swaps the diagonal on the host between
! HAMILT and HNS).  The per-run tables are uploaded once, at the first HAMILT
! call of the process, into one device context kept for the whole run.
!
! Errors: a positive status from the eigensolver is a Cholesky INFO and is
! reported like seclr4.F:200-205 (OUTERR, then STOP); negative statuses are
! argument or device errors and print the library's error text.
!
! Lines starting with the variant markers are selected by the same sed step as
! the reference sources (SRC_lapw1/Makefile:227-232).

module w2k_capi
  use iso_c_binding
  implicit none

  ! struct w2k_run_desc of include/w2k_lapw1.h, member for member
  type, bind(c) :: w2k_run_desc
    integer(c_int) :: is_complex, nat, ndf, ltop, lmaxd, nnsl, lmmxd, ngaud, ins
    type(c_ptr) :: mult, pos, rmt, v
    real(c_double) :: vi
    type(c_ptr) :: u, up, du, dup, chn, e
    type(c_ptr) :: lapw
    integer(c_int) :: nlo
    type(c_ptr) :: ilo, loor
    type(c_ptr) :: alo, blo, clo, elo
    type(c_ptr) :: pi12lo, pe12lo, dplo
    type(c_ptr) :: rotloc, rotij
    integer(c_int) :: warp1, warp2, warp3
    type(c_ptr) :: warp
    type(c_ptr) :: lqind, lqns, gfac
    type(c_ptr) :: vns1, vns2, vns3
    type(c_ptr) :: vns1lo, vns2lo, vns3lo
  end type w2k_run_desc

  interface
    integer(c_int) function w2k_create(ctx, device, run) bind(c, name='w2k_create')
      import :: c_int, c_ptr, w2k_run_desc
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
      type(w2k_run_desc), intent(in) :: run
    end function
    integer(c_int) function w2k_set_basis(ctx, nv, nlo, kzz, xk, yk, zk, rk) bind(c, name='w2k_set_basis')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nv, nlo
      integer(c_int), intent(in) :: kzz(*)
      real(c_double), intent(in) :: xk(*), yk(*), zk(*), rk(*)
    end function
    integer(c_int) function w2k_hamilt(ctx) bind(c, name='w2k_hamilt')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function w2k_hns(ctx) bind(c, name='w2k_hns')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function w2k_seclr4(ctx, el, eu, nume, ne, nbelw) bind(c, name='w2k_seclr4')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: el, eu
      integer(c_int), value :: nume
      integer(c_int), intent(out) :: ne, nbelw
    end function
    integer(c_int) function w2k_get_hs(ctx, hs, ld, hsdiag) bind(c, name='w2k_get_hs')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(*), intent(inout) :: hs(*), hsdiag(*)
      integer(c_int), value :: ld
    end function
    integer(c_int) function w2k_put_hs(ctx, hs, ld, hsdiag) bind(c, name='w2k_put_hs')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(*), intent(in) :: hs(*), hsdiag(*)
      integer(c_int), value :: ld
    end function
    integer(c_int) function w2k_get_albl(ctx, al, bl) bind(c, name='w2k_get_albl')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: al(*), bl(*)
    end function
    integer(c_int) function w2k_get_eig(ctx, eigval, z, ldz) bind(c, name='w2k_get_eig')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: eigval(*)
      type(*), intent(inout) :: z(*)
      integer(c_int), value :: ldz
    end function
    type(c_ptr) function w2k_last_error() bind(c, name='w2k_last_error')
      import :: c_ptr
    end function
  end interface

contains

  ! text of a NUL-terminated C string
  function c_text(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=256) :: s
    character(kind=c_char), pointer :: a(:)
    integer :: i
    s = ' '
    if (.not. c_associated(p)) return
    call c_f_pointer(p, a, [256])
    do i = 1, 256
      if (a(i) == c_null_char) exit
      s(i:i) = a(i)
    end do
  end function c_text

end module w2k_capi


module w2k_dropin_state
  use iso_c_binding
  use w2k_capi
  implicit none
  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: device = 0
  ! per-run tables in the C-ABI's element types (kept alive for the upload)
  integer(c_int), allocatable, target, save :: t_mult(:), t_lapw(:,:), t_ilo(:,:), t_loor(:,:)
  integer(c_int), allocatable, target, save :: t_lqind(:), t_lqns(:,:,:)
  real(c_double), allocatable, target, save :: t_pos(:,:), t_rmt(:), t_v(:)
  real(c_double), allocatable, target, save :: t_u(:,:), t_up(:,:), t_du(:,:), t_dup(:,:), t_chn(:,:), t_e(:,:)
  real(c_double), allocatable, target, save :: t_alo(:,:,:), t_blo(:,:,:), t_clo(:,:,:), t_elo(:,:,:)
  real(c_double), allocatable, target, save :: t_pi12lo(:,:), t_pe12lo(:,:), t_dplo(:,:)
  real(c_double), allocatable, target, save :: t_rotloc(:,:,:), t_rotij(:,:,:)
  complex(c_double_complex), allocatable, target, save :: t_warp(:,:,:), t_gfac(:,:)
  real(c_double), allocatable, target, save :: t_vns1(:,:,:,:), t_vns2(:,:,:,:), t_vns3(:,:,:,:)
  real(c_double), allocatable, target, save :: t_vns1lo(:,:,:,:), t_vns2lo(:,:,:,:), t_vns3lo(:,:,:,:)

contains

  subroutine w2k_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc == 0) return
    write(6,*) where, ': MI355X library status ', rc, ' ', trim(c_text(w2k_last_error()))
    write(99,*) where, ': MI355X library status ', rc, ' ', trim(c_text(w2k_last_error()))
    call OUTERR(where, 'device library call failed.')
    stop 'W2K DROP-IN - Error'
  end subroutine w2k_check

  ! Upload of the k-independent state that HAMILT/HNS read (first call only).
  subroutine w2k_ensure_context(LTOP, NAT, U, UP, DU, DUP, MULT, R, POS, V, VOLUME, CHN, E, INS)
    use lolog, only : nlo, loor, lapw, ilo
    use loabc, only : ALO, BLO, CLO, ELO, PI12LO, PE12LO, DPLO
    use rotmat, only : ROTIJ, ROTLOC
    use struk, only : NDF
    use out, only : WARP, WARP1, WARP2, WARP3
    use atspdt, only : LQIND, LQNS, GFAC, VNS1, VNS2, VNS3
    use loint, only : VNS1LO, VNS2LO, VNS3LO
    use comi, only : LNSMAX
    implicit none
    include 'param.inc'
    integer :: LTOP, NAT, INS
    integer :: MULT(NAT)
    double precision :: U(0:LMAX-1,NAT), UP(0:LMAX-1,NAT), DU(0:LMAX-1,NAT), DUP(0:LMAX-1,NAT)
    double precision :: CHN(0:LMAX-1,NAT), E(0:LMAX-1,NAT)
    double precision :: R(NAT), POS(3,NDF), V(NAT), VOLUME
    type(w2k_run_desc) :: d
    integer(c_int) :: rc
    character(len=16) :: env
    integer :: st

    if (c_associated(ctx)) return
    call get_environment_variable('W2K_DEVICE', env, status=st)
    if (st == 0) read(env, *) device

    t_mult = MULT(1:NAT)
    t_pos = POS(:, 1:NDF)
    t_rmt = R(1:NAT)
    t_v = V(1:NAT)
    t_u = U(:, 1:NAT); t_up = UP(:, 1:NAT); t_du = DU(:, 1:NAT); t_dup = DUP(:, 1:NAT)
    t_chn = CHN(:, 1:NAT); t_e = E(:, 1:NAT)
    t_lapw = merge(1, 0, lapw(:, 1:NAT))
    t_loor = merge(1, 0, loor(:, 1:NAT))
    t_ilo = ilo(:, 1:NAT)
    t_alo = ALO(:, :, 1:NAT); t_blo = BLO(:, :, 1:NAT); t_clo = CLO(:, :, 1:NAT); t_elo = ELO(:, :, 1:NAT)
    t_pi12lo = PI12LO(:, 1:NAT); t_pe12lo = PE12LO(:, 1:NAT); t_dplo = DPLO(:, 1:NAT)
    t_rotloc = ROTLOC(:, :, 1:NAT)
    t_rotij = ROTIJ(:, :, 1:NDF)
    if (INS == 0) then
      t_warp = WARP(-WARP1:WARP1, -WARP2:WARP2, -WARP3:WARP3)
    else
      allocate(t_warp(1,1,1)); t_warp = (0.0d0, 0.0d0)
    end if
    t_lqind = LQIND(1:NAT)
    t_lqns = LQNS(:, :, 1:NAT)
    t_gfac = GFAC(:, 1:NAT)
    t_vns1 = VNS1(:, :, :, 1:NAT); t_vns2 = VNS2(:, :, :, 1:NAT); t_vns3 = VNS3(:, :, :, 1:NAT)
    t_vns1lo = VNS1LO(:, :, :, 1:NAT); t_vns2lo = VNS2LO(:, :, :, 1:NAT); t_vns3lo = VNS3LO(:, :, :, 1:NAT)

!_REAL    d%is_complex = 0
!_COMPLEX    d%is_complex = 1
    d%nat = NAT
    d%ndf = NDF
    d%ltop = LTOP
    d%lmaxd = LMAX
    d%nnsl = LNSMAX + 1
    d%lmmxd = size(t_vns1, 2)
    d%ngaud = size(t_lqns, 2)
    d%ins = INS
    d%mult = c_loc(t_mult); d%pos = c_loc(t_pos); d%rmt = c_loc(t_rmt); d%v = c_loc(t_v)
    d%vi = VOLUME
    d%u = c_loc(t_u); d%up = c_loc(t_up); d%du = c_loc(t_du); d%dup = c_loc(t_dup)
    d%chn = c_loc(t_chn); d%e = c_loc(t_e)
    d%lapw = c_loc(t_lapw)
    d%nlo = nlo
    d%ilo = c_loc(t_ilo); d%loor = c_loc(t_loor)
    d%alo = c_loc(t_alo); d%blo = c_loc(t_blo); d%clo = c_loc(t_clo); d%elo = c_loc(t_elo)
    d%pi12lo = c_loc(t_pi12lo); d%pe12lo = c_loc(t_pe12lo); d%dplo = c_loc(t_dplo)
    d%rotloc = c_loc(t_rotloc); d%rotij = c_loc(t_rotij)
    if (INS == 0) then
      d%warp1 = WARP1; d%warp2 = WARP2; d%warp3 = WARP3
    else
      d%warp1 = 0; d%warp2 = 0; d%warp3 = 0
    end if
    d%warp = c_loc(t_warp)
    d%lqind = c_loc(t_lqind); d%lqns = c_loc(t_lqns); d%gfac = c_loc(t_gfac)
    d%vns1 = c_loc(t_vns1); d%vns2 = c_loc(t_vns2); d%vns3 = c_loc(t_vns3)
    d%vns1lo = c_loc(t_vns1lo); d%vns2lo = c_loc(t_vns2lo); d%vns3lo = c_loc(t_vns3lo)
    rc = w2k_create(ctx, int(device, c_int), d)
    call w2k_check(rc, 'HAMILT')
  end subroutine w2k_ensure_context

end module w2k_dropin_state


SUBROUTINE HAMILT(LTOP,NAT,NV,U,UP,DU,DUP,MULT,R,POS,V,VOLUME,CHN,E,INS)
  use iso_c_binding
  use w2k_capi
  use w2k_dropin_state
  use matrices, only : HS, HSDIAG, HSROWS, KZZ, XK, YK, ZK, RK
  use albl, only : AL, BL, init_albl
  use lolog, only : nlo
  use struk, only : NDF
  implicit none
  include 'param.inc'
  INTEGER            INS, LTOP, NAT, NV
  INTEGER            MULT(NAT)
  DOUBLE PRECISION   VOLUME
  DOUBLE PRECISION   R(NAT)
  DOUBLE PRECISION   CHN(0:LMAX-1,NAT), DU(0:LMAX-1,NAT)
  DOUBLE PRECISION   DUP(0:LMAX-1,NAT), E(0:LMAX-1,NAT)
  DOUBLE PRECISION   POS(3,NDF)
  DOUBLE PRECISION   U(0:LMAX-1,NAT), UP(0:LMAX-1,NAT), V(NAT)
  integer(c_int) :: rc

  ! AL/BL live in module albl as in the reference (hamilt.F:158); CALKPT
  ! releases them with END_ALBL after HNS
  call init_albl(HSROWS, LTOP, NAT)
  call w2k_ensure_context(LTOP, NAT, U, UP, DU, DUP, MULT, R, POS, V, VOLUME, CHN, E, INS)
  rc = w2k_set_basis(ctx, int(NV, c_int), int(nlo, c_int), KZZ, XK, YK, ZK, RK)
  call w2k_check(rc, 'HAMILT')
  rc = w2k_hamilt(ctx)
  call w2k_check(rc, 'HAMILT')
  rc = w2k_get_hs(ctx, HS, int(HSROWS, c_int), HSDIAG)
  call w2k_check(rc, 'HAMILT')
  rc = w2k_get_albl(ctx, AL, BL)
  call w2k_check(rc, 'HAMILT')
END SUBROUTINE HAMILT


SUBROUTINE HNS(NV,NAT,NNSL,POS,VI,R,MULT,LQIND,LQNS,GFAC,ROTLOC,ROTIJ,VNS1,VNS2,VNS3)
  use iso_c_binding
  use w2k_capi
  use w2k_dropin_state
  use matrices, only : HS, HSDIAG, HSROWS
  use struk, only : NDF
  implicit none
  include 'param.inc'
  INTEGER            NNSL, NV, NAT
  INTEGER            MULT(NAT), LQIND(NAT)
  INTEGER            LQNS(6,NGAU,NAT)
  DOUBLE PRECISION   VI
  DOUBLE PRECISION   R(NAT)
  DOUBLE PRECISION   ROTLOC(3,3,NAT), ROTIJ(3,3,NDF)
  DOUBLE PRECISION   POS(3,NDF)
  DOUBLE PRECISION   VNS1(0:NSLMAX-1,LMMX,0:NSLMAX-1,NAT)
  DOUBLE PRECISION   VNS2(0:NSLMAX-1,LMMX,0:NSLMAX-1,NAT)
  DOUBLE PRECISION   VNS3(0:NSLMAX-1,LMMX,0:NSLMAX-1,NAT)
  DOUBLE COMPLEX     GFAC(NGAU,NAT)
  integer(c_int) :: rc

  ! the run-level tables were uploaded by HAMILT; the host HS carries the
  ! diagonal swap of calkpt.F:160-166
  if (.not. c_associated(ctx)) then
    write(6,*) 'HNS: HAMILT has not run in this process'
    call OUTERR('HNS', 'device context missing.')
    stop 'W2K DROP-IN - Error'
  end if
  rc = w2k_put_hs(ctx, HS, int(HSROWS, c_int), HSDIAG)
  call w2k_check(rc, 'HNS')
  rc = w2k_hns(ctx)
  call w2k_check(rc, 'HNS')
  rc = w2k_get_hs(ctx, HS, int(HSROWS, c_int), HSDIAG)
  call w2k_check(rc, 'HNS')
END SUBROUTINE HNS


SUBROUTINE SECLR4
  use iso_c_binding
  use w2k_capi
  use w2k_dropin_state
  use matrices, only : HS, HSDIAG, HSROWS, Z, EIGVAL
  use comi, only : NE, NBELW
  implicit none
  include 'param.inc'
  integer(c_int) :: rc, ne_c, nbelw_c
  real(c_double) :: el1, eu1
  INTEGER            STDERR
  PARAMETER          (STDERR = 99)

  if (.not. c_associated(ctx)) then
    write(6,*) 'SECLR4: HAMILT has not run in this process'
    call OUTERR('SECLR4', 'device context missing.')
    stop 'SECLR4 - Error'
  end if
  ! window (EL,EU] from EIGVAL(1:2), seclr4.F:149-154
  IF (EIGVAL(1) .GE. EIGVAL(2)) THEN
     EIGVAL(1) = -1.0D+70
     EIGVAL(2) =  1.0D+70
  ENDIF
  el1 = EIGVAL(1)
  eu1 = EIGVAL(2)
  rc = w2k_put_hs(ctx, HS, int(HSROWS, c_int), HSDIAG)
  call w2k_check(rc, 'SECLR4')
  rc = w2k_seclr4(ctx, el1, eu1, int(min(NUME, size(Z, 2)), c_int), ne_c, nbelw_c)
  if (rc > 0) then
    ! Cholesky of S failed (seclr4.F:200-205)
    write(6,*) 'Cholesky INFO = ', rc
    write(stderr,*) 'Cholesky INFO = ', rc
    CALL OUTERR('SECLR4','POTRF (Scalapack/LAPACK) failed.')
    STOP 'SECLR4 - Error'
  end if
  call w2k_check(rc, 'SECLR4')
  NE = ne_c
  NBELW = nbelw_c
  rc = w2k_get_eig(ctx, EIGVAL, Z, int(HSROWS, c_int))
  call w2k_check(rc, 'SECLR4')
END SUBROUTINE SECLR4
