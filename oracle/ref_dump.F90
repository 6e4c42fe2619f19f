! Dump driver for the reference lapw1 (test infrastructure: builds only into
! oracle/_ref/ and is linked with the reference objects compiled there).
!
! Runs the program sequence of lapw1.F:33-84 with CALKPT (calkpt.F:143-197)
! unrolled, so that the module state can be written out (argument 'nodump':
! timings only; 'eigonly': per k-point NE, NBELW and EIGVAL only; 'basis':
! the run state and per k-point the basis, window, NE, NBELW and EIGVAL, but
! no matrices)
!   - once per run after INILPW/INIKPT            -> run.bin
!   - per k-point at CALKPT entry (basis), after HAMILT, after HNS and after
!     SECLR4                                       -> k<NNNN>.bin
! Lines starting with the variant markers are selected by the same sed step
! as the reference sources (SRC_lapw1/Makefile:227-232).
!
! Record layout (stream, little endian): name(32 chars), kind(int32: 1=int32,
! 2=float64, 3=complex128), ndim(int32), dims(ndim x int32), data (column-major).
module dumpio
  implicit none
contains
  subroutine hdr(u, name, kind, dims)
    integer, intent(in) :: u, kind
    character(len=*), intent(in) :: name
    integer, intent(in) :: dims(:)
    character(len=32) :: nm
    nm = name
    write(u) nm, int(kind,4), int(size(dims),4), int(dims,4)
  end subroutine
  subroutine wi(u, name, a, dims)
    integer, intent(in) :: u
    character(len=*), intent(in) :: name
    integer, intent(in) :: dims(:)
    integer, intent(in) :: a(*)
    call hdr(u, name, 1, dims)
    write(u) int(a(1:product(dims)),4)
  end subroutine
  subroutine wd(u, name, a, dims)
    integer, intent(in) :: u
    character(len=*), intent(in) :: name
    integer, intent(in) :: dims(:)
    double precision, intent(in) :: a(*)
    call hdr(u, name, 2, dims)
    write(u) a(1:product(dims))
  end subroutine
  subroutine wz(u, name, a, dims)
    integer, intent(in) :: u
    character(len=*), intent(in) :: name
    integer, intent(in) :: dims(:)
    double complex, intent(in) :: a(*)
    call hdr(u, name, 3, dims)
    write(u) a(1:product(dims))
  end subroutine
  subroutine ws(u, name, x)
    integer, intent(in) :: u
    character(len=*), intent(in) :: name
    integer, intent(in) :: x
    integer :: b(1)
    b(1) = x
    call wi(u, name, b, [1])
  end subroutine
  subroutine wsd(u, name, x)
    integer, intent(in) :: u
    character(len=*), intent(in) :: name
    double precision, intent(in) :: x
    double precision :: b(1)
    b(1) = x
    call wd(u, name, b, [1])
  end subroutine
end module dumpio

program ref_dump
  use dumpio
  use parallel, only : init_parallel, close_parallel
  use lapw_timer, only : init_all_timer, start_timer, stop_timer, time_total, &
                         time_hamilt, time_hns, time_diag, read_cpu_time, read_wall_time
  use cut1, only : cut
  use matrices, only : HS, HSDIAG, HSROWS, Z, EIGVAL, KZZ, XK, YK, ZK, RK, END_MATRICES, END_matrice1
  use lolog, only : nlo, loor, lapw, ilo
  use atspdt, only : INS, LQIND, LQNS, DP, DPE, E, P, PE, PEI, VNS1, VNS2, VNS3, GFAC
  use comi, only : NVAA, NE, NBELW, NAT, NT, LNSMAX
  use coml, only : ITER, NOHNS
  use comr, only : RKM, EL, EU
  use lstapw, only : NV
  use onekpt, only : SX1, SY1, SZ1
  use rotmat, only : ROTIJ, ROTLOC
  use struk, only : POS, RMT, V, VI, MULT, NDF, multmax
  use albl, only : AL, BL, END_ALBL
  use loabc, only : ALO, BLO, CLO, ELO, PLO, DPLO, PELO, DPELO, PEILO, PI12LO, PE12LO
  use loint, only : VNS1LO, VNS2LO, VNS3LO
  use out, only : WARP, WARP1, WARP2, WARP3, IORD, IMAT, TAU, NKK
  use totpot, only : JA, JB, JC, POTK
  use potnlc, only : JRJ, DH, RO, VR
  use atspdt, only : LMMAX, LM
  use coml, only : REL
  use gener, only : BR2
  use orth, only : ORTHO
  use char, only : LATTIC
  use struk, only : PIA
  implicit none
  include 'param.inc'
  logical :: nmat_only, dodump, eigdump, basisdump, coorsonly
  integer :: numwk, info, kpoint, u, nn, i, j
  integer, allocatable :: ltmp(:,:)
  character(len=80) :: deffn, fname
  double precision :: t0, t1, th, tn, td
  integer(8) :: c0, c1, crate
  character(len=16) :: arg

  cut = .false.
  deffn = 'lapw1.def'
  dodump = .true.
  eigdump = .false.
  basisdump = .false.
  coorsonly = .false.
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    if (arg(1:6) == 'nodump') dodump = .false.
    if (arg(1:7) == 'eigonly') then
      dodump = .false.
      eigdump = .true.
    end if
    if (arg(1:5) == 'basis') then
      dodump = .false.
      basisdump = .true.
    end if
    ! 'coors': run state and per-k basis only (COORS/LOPW), no HAMILT/HNS/SECLR4
    if (arg(1:5) == 'coors') then
      dodump = .false.
      basisdump = .true.
      coorsonly = .true.
    end if
  end if
  call init_parallel
  call init_all_timer
  call start_timer(time_total)
  call inilpw(numwk, deffn, info, nmat_only)
  if (info /= 0) stop 'ref_dump: INILPW failed'
  call inikpt(info)
  if (info /= 0) stop 'ref_dump: INIKPT failed'

  u = 77
  if (dodump .or. basisdump) then
  open(u, file='run.bin', access='stream', form='unformatted', status='replace')
  call ws(u, 'NAT', NAT)
  call ws(u, 'NDF', NDF)
  call ws(u, 'LTOP', NT)
  call ws(u, 'LNSMAX', LNSMAX)
  call ws(u, 'INS', INS)
  call ws(u, 'NLO', NLO)
  call ws(u, 'MULTMAX', multmax)
  call ws(u, 'LMAXP', LMAX)
  call ws(u, 'LOMAXP', LOMAX)
  call ws(u, 'NLOATP', nloat)
  call ws(u, 'NSLMAXP', NSLMAX)
  call ws(u, 'LMMXP', LMMX)
  call ws(u, 'NGAUP', NGAU)
  call wi(u, 'MULT', MULT, [NAT])
  call wd(u, 'POS', POS, [3, NDF])
  call wd(u, 'RMT', RMT, [NAT])
  call wd(u, 'V', V, [NAT])
  call wsd(u, 'VI', VI)
  call wd(u, 'P', P, [LMAX, NAT])
  call wd(u, 'PE', PE, [LMAX, NAT])
  call wd(u, 'DP', DP, [LMAX, NAT])
  call wd(u, 'DPE', DPE, [LMAX, NAT])
  call wd(u, 'PEI', PEI, [LMAX, NAT])
  call wd(u, 'E', E, [LMAX, NAT])
  allocate(ltmp(0:LMAX-1, NAT))
  ltmp = merge(1, 0, lapw)
  call wi(u, 'LAPW', ltmp, [LMAX, NAT])
  deallocate(ltmp)
  allocate(ltmp(0:LOMAX, NAT))
  ltmp = merge(1, 0, loor)
  call wi(u, 'LOOR', ltmp, [LOMAX+1, NAT])
  deallocate(ltmp)
  call wi(u, 'ILO', ILO, [LOMAX+1, NAT])
  call wd(u, 'ALO', ALO, [LOMAX+1, nloat, NAT])
  call wd(u, 'BLO', BLO, [LOMAX+1, nloat, NAT])
  call wd(u, 'CLO', CLO, [LOMAX+1, nloat, NAT])
  call wd(u, 'ELO', ELO, [LOMAX+1, nloat, NAT])
  call wd(u, 'PLO', PLO, [LOMAX+1, NAT])
  call wd(u, 'DPLO', DPLO, [LOMAX+1, NAT])
  call wd(u, 'PELO', PELO, [LOMAX+1, NAT])
  call wd(u, 'DPELO', DPELO, [LOMAX+1, NAT])
  call wd(u, 'PEILO', PEILO, [LOMAX+1, NAT])
  call wd(u, 'PI12LO', PI12LO, [LOMAX+1, NAT])
  call wd(u, 'PE12LO', PE12LO, [LOMAX+1, NAT])
  call wd(u, 'ROTLOC', ROTLOC, [3, 3, NAT])
  call wd(u, 'ROTIJ', ROTIJ, [3, 3, NDF])
  call wi(u, 'LQIND', LQIND, [NAT])
  call wi(u, 'LQNS', LQNS, [6, NGAU, NAT])
  call wz(u, 'GFAC', GFAC, [NGAU, NAT])
  call wd(u, 'VNS1', VNS1, [NSLMAX, LMMX, NSLMAX, NAT])
  call wd(u, 'VNS2', VNS2, [NSLMAX, LMMX, NSLMAX, NAT])
  call wd(u, 'VNS3', VNS3, [NSLMAX, LMMX, NSLMAX, NAT])
  call wd(u, 'VNS1LO', VNS1LO, shape(VNS1LO))
  call wd(u, 'VNS2LO', VNS2LO, shape(VNS2LO))
  call wd(u, 'VNS3LO', VNS3LO, shape(VNS3LO))
  ! radial mesh and spherical potential (r V in Hartree after ATPAR,
  ! atpar.f:98-100), the LM list of each atom's V_LM, REL
  call wi(u, 'JRJ', JRJ, [NAT])
  call wd(u, 'DH', DH, [NAT])
  call wd(u, 'RO', RO, [NAT])
  call wd(u, 'VR', VR, [size(VR,1), NAT])
  call wi(u, 'LMMAX', LMMAX, [NAT])
  call wi(u, 'LM', LM, [2, LMMX, NAT])
  call ws(u, 'REL', merge(1, 0, REL))
  ! reciprocal lattice and cutoff of COORS (coors.f:61-148)
  call wd(u, 'BR2', BR2, [3, 3])
  call wd(u, 'PIA', PIA, [3])
  call ws(u, 'ORTHO', merge(1, 0, ORTHO))
  call ws(u, 'CXZ', merge(1, 0, LATTIC(1:3) == 'CXZ'))
  call wsd(u, 'RKM', RKM)
  if (INS == 0) then
    ! interstitial stars (rdswar.f) and the space group (module out)
    call ws(u, 'IORD', IORD)
    call wi(u, 'IMAT', IMAT, [3, 3, IORD])
    call wd(u, 'TAU', TAU, [3, IORD])
    call ws(u, 'NKK', NKK)
    call wi(u, 'JA', JA, [NKK])
    call wi(u, 'JB', JB, [NKK])
    call wi(u, 'JC', JC, [NKK])
    call wz(u, 'POTK', POTK, [NKK])
    call ws(u, 'WARP1', WARP1)
    call ws(u, 'WARP2', WARP2)
    call ws(u, 'WARP3', WARP3)
    call wz(u, 'WARP', WARP, [2*WARP1+1, 2*WARP2+1, 2*WARP3+1])
  end if
  close(u)
  end if

  th = 0; tn = 0; td = 0
  do kpoint = 1, numwk
    call setkpt(kpoint)
    ! --- CALKPT (calkpt.F:143-197), unrolled ---
    call coors(RKM, NT, SX1, SY1, SZ1, NAT, nmat_only)
    NVAA = NV
    EIGVAL(1) = EL
    EIGVAL(2) = EU
    if (dodump .or. basisdump) then
    write(fname, '(a,i4.4,a)') 'k', kpoint, '.bin'
    open(u, file=fname, access='stream', form='unformatted', status='replace')
    call ws(u, 'NV', NV)
    call ws(u, 'NLO', NLO)
    call ws(u, 'HSROWS', HSROWS)
    call wi(u, 'KZZ', KZZ, [3, HSROWS])
    call wd(u, 'XK', XK, [HSROWS])
    call wd(u, 'YK', YK, [HSROWS])
    call wd(u, 'ZK', ZK, [HSROWS])
    call wd(u, 'RK', RK, [HSROWS])
    call wsd(u, 'SX', SX1)
    call wsd(u, 'SY', SY1)
    call wsd(u, 'SZ', SZ1)
    call wsd(u, 'EL', EL)
    call wsd(u, 'EU', EU)
    call ws(u, 'NUME', NUME)
    end if
    if (coorsonly) then
      call ws(u, 'NE', 0)
      call ws(u, 'NBELW', 0)
      call wd(u, 'EIGVAL', EIGVAL, [1])
      close(u)
      call END_matrice1
      call END_MATRICES
      cycle
    end if
    call system_clock(c0, crate)
    call HAMILT(NT, NAT, NV, P, PE, DP, DPE, MULT, RMT, POS, V, VI, PEI, E, INS)
    call system_clock(c1); th = th + dble(c1 - c0)/dble(crate)
    nn = HSROWS
    if (dodump) then
!_REAL    call wd(u, 'HS_HAMILT', HS, [nn, nn])
!_REAL    call wd(u, 'HSDIAG_HAMILT', HSDIAG, [nn])
!_COMPLEX    call wz(u, 'HS_HAMILT', HS, [nn, nn])
!_COMPLEX    call wz(u, 'HSDIAG_HAMILT', HSDIAG, [nn])
    call wd(u, 'AL', AL, [nn, NT, NAT])
    call wd(u, 'BL', BL, [nn, NT, NAT])
    end if
!_REAL    call DSWAP(HSROWS, HS, HSROWS+1, HSDIAG, 1)
!_COMPLEX    call ZSWAP(HSROWS, HS, HSROWS+1, HSDIAG, 1)
    call system_clock(c0, crate)
    if (.not. nohns) &
      call HNS(NV, NAT, LNSMAX+1, POS, VI, RMT, MULT, LQIND, LQNS, GFAC, ROTLOC, ROTIJ, &
               VNS1(1,1,1,1), VNS2(1,1,1,1), VNS3(1,1,1,1))
    call system_clock(c1); tn = tn + dble(c1 - c0)/dble(crate)
    call END_ALBL
    if (dodump) then
!_REAL    call wd(u, 'HS_HNS', HS, [nn, nn])
!_REAL    call wd(u, 'HSDIAG_HNS', HSDIAG, [nn])
!_COMPLEX    call wz(u, 'HS_HNS', HS, [nn, nn])
!_COMPLEX    call wz(u, 'HSDIAG_HNS', HSDIAG, [nn])
    end if
    call system_clock(c0, crate)
    if (iter) stop 'ref_dump: iterative diagonalisation is not part of the dump'
    call SECLR4
    call system_clock(c1); td = td + dble(c1 - c0)/dble(crate)
    if (dodump .or. basisdump) then
    call ws(u, 'NE', NE)
    call ws(u, 'NBELW', NBELW)
    call wd(u, 'EIGVAL', EIGVAL, [max(NE,1)])
    if (dodump) then
!_REAL    call wd(u, 'Z', Z, [nn, max(NE,1)])
!_COMPLEX    call wz(u, 'Z', Z, [nn, max(NE,1)])
    end if
    close(u)
    end if
    if (eigdump) then
      write(fname, '(a,i4.4,a)') 'k', kpoint, '.bin'
      open(u, file=fname, access='stream', form='unformatted', status='replace')
      call ws(u, 'NV', NV)
      call ws(u, 'NLO', NLO)
      call ws(u, 'HSROWS', HSROWS)
      call ws(u, 'NE', NE)
      call ws(u, 'NBELW', NBELW)
      call wd(u, 'EIGVAL', EIGVAL, [max(NE,1)])
      close(u)
    end if
    call prtkpt(kpoint, info)
    if (info /= 0) stop 'ref_dump: PRTKPT failed'
  end do
  call prtres(numwk, info)
  call stop_timer(time_total)
  open(u, file='ref_times.txt', status='replace')
  write(u, '(a,i6)') 'kpoints ', numwk
  write(u, '(a,f14.6)') 'hamilt_wall_s ', th
  write(u, '(a,f14.6)') 'hns_wall_s ', tn
  write(u, '(a,f14.6)') 'seclr4_wall_s ', td
  write(u, '(a,f14.6)') 'total_wall_s ', read_wall_time(time_total)
  close(u)
  call close_parallel
end program ref_dump
