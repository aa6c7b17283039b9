! fmc_gcmc: the GCMC move loop of the reference (gcmc.f:589-1960) in
! Fortran, calling the MI355X trial-move path through the C ABI
! (fmc_abi.f90 / include/fastmc_amd.h).
!
! Reads CONTROL, FIELD and CONFIG from the working directory in the
! reference's formats (readinputs.f, utility_pack.f:176-457) and writes
! OUTPUT, branch01/numguests.out and, on request, branch01/trace.out.
!
! Moves: insert, delete, displace (+ translate-only / rotate-only), with the
! reference's duni draw order (gcmc.f:745-1338), acceptance rules
! (energy_eval, utility_pack.f:3049-3095) and adaptive step sizes
! (gcmc.f:794-834).  Driver extensions (not in the reference): CONTROL
! keywords "seed i j k l" (fixed duni seed instead of the clock) and
! "trace" (one line per move).

module fmc_parse
  implicit none
  integer, parameter :: lenrec = 256
contains
  subroutine lowcase(s)
    character(len=*), intent(inout) :: s
    integer :: i, c
    do i = 1, len(s)
      c = iachar(s(i:i))
      if (c >= 65 .and. c <= 90) s(i:i) = achar(c + 32)
    end do
  end subroutine

  logical function has(s, key)
    character(len=*), intent(in) :: s, key
    has = index(s, key) > 0
  end function

  ! parse_module.f:100-169 (intstr): first integer, record shifted past it
  integer function intstr(s)
    character(len=*), intent(inout) :: s
    integer :: lst, isn, n
    logical :: cnt, fin, flag
    character :: ksn, ch
    n = len(s); isn = 1; lst = 0; ksn = '+'; intstr = 0
    cnt = .false.; fin = .false.
    do while (lst < n .and. .not. fin)
      lst = lst + 1
      ch = s(lst:lst)
      flag = .false.
      if (ch >= '0' .and. ch <= '9') then
        intstr = 10 * intstr + (iachar(ch) - 48)
        cnt = .true.; flag = .true.
      end if
      if (cnt .and. .not. flag) fin = .true.
      if (flag .and. ksn == '-') isn = -1
      ksn = ch
    end do
    intstr = isn * intstr
    s = s(lst:)
  end function

  ! parse_module.f:171-288 (dblstr): decimals built digit by digit
  real(8) function dblstr(s)
    character(len=*), intent(inout) :: s
    integer :: lst, n, iexp
    real(8) :: ten, one, sn
    logical :: start, ldot, fin, flag
    character :: ksn, ch
    character(len=lenrec) :: rest
    n = len(s); lst = 0; sn = 1.d0; ksn = '+'; ten = 10.d0; one = 1.d0
    dblstr = 0.d0; iexp = 0; start = .false.; ldot = .false.; fin = .false.
    do while (lst < n .and. .not. fin)
      lst = lst + 1
      ch = s(lst:lst)
      flag = .false.
      if (ch >= '0' .and. ch <= '9') then
        dblstr = ten * dblstr + one * dble(iachar(ch) - 48)
        flag = .true.; start = .true.
      end if
      if (ch == '.') then
        flag = .true.; ten = 1.d0; ldot = .true.; start = .true.
      end if
      if (flag .and. ksn == '-') sn = -1.d0
      if (ldot) one = one / 10.d0
      ksn = ch
      if (ksn == 'D') ksn = 'd'
      if (ksn == 'E') ksn = 'e'
      if (start) then
        if (ksn == 'd' .or. ksn == 'e') then
          rest = s(lst+1:)
          iexp = intstr(rest)
          s = rest
          lst = 0
          fin = .true.
          exit
        end if
        if (.not. flag) fin = .true.
      end if
    end do
    dblstr = sn * dblstr * (10.d0 ** iexp)
    if (lst > 0) s = s(lst:)
  end function

  function getword(s) result(w)
    character(len=*), intent(inout) :: s
    character(len=8) :: w
    integer :: i, j
    s = adjustl(s)
    j = index(s, ' ')
    if (j == 0) j = len(s) + 1
    w = s(1:min(j - 1, 8))
    if (j <= len(s)) then
      s = s(j:)
    else
      s = ' '
    end if
    i = 0
  end function
end module fmc_parse

module fmc_rng
  implicit none
  real(4), save :: u(97), c, cd, cm
  integer, save :: ir, jr
contains
  ! gcmc.f:2760-2784
  subroutine duni_init(i0, j0, k0, l0)
    integer, intent(in) :: i0, j0, k0, l0
    integer :: i, j, k, l, m, ii, jj
    real(4) :: s, t
    i = i0; j = j0; k = k0; l = l0
    ir = 97; jr = 33
    do ii = 1, 97
      s = 0.0; t = 0.5
      do jj = 1, 24
        m = mod(mod(i * j, 179) * k, 179)
        i = j; j = k; k = m
        l = mod(53 * l + 1, 169)
        if (mod(l * m, 64) >= 32) s = s + t
        t = 0.5 * t
      end do
      u(ii) = s
    end do
    c = 362436.0 / 16777216.0
    cd = 7654321.0 / 16777216.0
    cm = 16777213.0 / 16777216.0
  end subroutine

  ! gcmc.f:2786-2799
  real(8) function duni()
    real(4) :: uni
    uni = u(ir) - u(jr)
    if (uni < 0.0) uni = uni + 1.0
    u(ir) = uni
    ir = ir - 1
    if (ir == 0) ir = 97
    jr = jr - 1
    if (jr == 0) jr = 97
    c = c - cd
    if (c < 0.0) c = c + cm
    uni = uni - c
    if (uni < 0.0) uni = uni + 1.0
    duni = dble(uni)
  end function
end module fmc_rng

program fmc_gcmc
  use iso_c_binding
  use fmc_abi
  use fmc_parse
  use fmc_rng
  implicit none

  real(8), parameter :: pi = 3.141592653589793d0
  real(8), parameter :: kboltz = 1.9872041d-3, boltz = 1.3806503d-23
  real(8), parameter :: rgas = 8.314472d0 / 4184.d0
  real(8), parameter :: delrmin = 0.1d0, minangle = pi / 18.d0
  integer, parameter :: maxs = 8, maxg = 4, maxsamp = 500000

  type(c_ptr) :: ctx
  character(len=lenrec) :: rec, rec2, title
  character(len=8) :: a1, a2, kw
  character(len=8), allocatable :: unqatm(:)
  integer :: ntpatm, ntpvdw, ntpmls, ntpguest, imcon, levcfg
  integer :: i, j, k, m, n, s, ig, ios, rc, nrept, ifrz, ksite, numsit
  integer, allocatable :: vdw_a(:), vdw_b(:), vdw_key(:)
  real(8), allocatable :: vdw_prm(:)
  real(8) :: engunit, cell(9), rcell(9), det, temp, rcut, delr_in, rvdw
  real(8) :: alpha, eps_prec, volm, statvolm, beta
  integer :: kmax(3), nk, mxgrid
  logical :: have_sum, have_prec, lmove_general(9), do_trace, have_seed
  real(8) :: drewd, dlrpot
  ! molecules
  logical, allocatable :: mol_guest(:)
  integer, allocatable :: mol_nummols(:), mol_nsite(:), mol_gi(:)
  real(8), allocatable :: mol_q(:,:), mol_m(:,:), mol_x(:,:), mol_y(:,:), mol_z(:,:)
  integer, allocatable :: mol_t(:,:), mol_frz(:,:)
  character(len=8), allocatable :: mol_name(:,:)
  ! guests
  integer :: gns(maxg), nmol(maxg), gmol(maxg), gcap
  real(8) :: gtx(maxs,maxg), gty(maxs,maxg), gtz(maxs,maxg), gq(maxs,maxg), gm(maxs,maxg)
  integer :: gtt(maxs,maxg)
  real(8), allocatable :: gx(:,:), gy(:,:), gz(:,:)
  real(8) :: press(maxg), freq(9,maxg), cum(9,maxg), delr(maxg), disp_ratio(maxg)
  real(8) :: tran_ratio(maxg), rota_ratio, e3(maxg), energy(maxg)
  real(8) :: gfreq(9), gdelr, gdisp, gtran
  logical :: gfreq_set(maxg), gdelr_set(maxg), gdisp_set(maxg), gtran_set(maxg)
  ! framework
  real(8), allocatable :: fx(:), fy(:), fz(:), fq(:)
  integer, allocatable :: ft(:), ffz(:)
  integer :: nfw
  ! control
  integer :: mcsteps, eqsteps, nnumg, maxguest, seed(4), values(8)
  real(8) :: tim(6), t_rst
  integer(8) :: cw0, cw1
  ! the rigid framework's REVCON lines (label, atom number, coordinates),
  ! formatted once; the atom numbers are rewritten when they shift
  character(len=1), allocatable :: fw_text(:)
  integer :: fw_first = -1
  ! selectivity windows (gcmc.f:1887-1916)
  integer :: nwind, avcount
  integer(8) :: rollcount
  real(8) :: wsum(maxg), seltot(maxg,maxg), selwin(maxg,maxg), weight, ij_select, aN, delta
  ! move loop
  integer(8) :: gcmccount, prodcount
  integer :: randchoice, kind, natms, nmols, ndraw
  logical :: ins, del, dsp, tran, rota, accepted, production, lgchk, tick(5)
  real(8) :: randmov, ra, rb, rcc, r1, r2, r3, r4, rande, dd, da
  real(8) :: nx(maxs), ny(maxs), nz(maxs), out(3), e6(6), estep, delrc, elrc
  real(8) :: delrdisp(maxg), tran_delr(maxg), rotangle, rota_rotangle
  integer :: disp_count(maxg), accept_disp(maxg), tran_count(maxg), accept_tran(maxg)
  integer :: rota_count, accept_rota, ins_count, del_count, accept_ins, accept_del
  real(8) :: spenergy, ecoulg, evdwg, dlpeng
  real(8), allocatable :: gasmol(:,:), gaseng(:,:)
  real(8) :: st_n(maxg), st_e(maxg), st_en(maxg), st_n2(maxg), st_e2(maxg)
  integer :: totalguests
  ! restart files (gcmc.f:349-410,737-740; utility_pack.f:1772-1862)
  logical :: lrestart, have_rev, rev_prod
  character(len=80) :: cfgname
  real(8) :: chainstats(1 + 14 * maxg), dele_last
  real(8) :: t0, t1
  integer(8) :: c0, c1, crate

  ! ------------------------------------------------------------------
  ! CONTROL pre-scan + read (utility_pack.f:310-414, readinputs.f:515-914)
  ! ------------------------------------------------------------------
  temp = 0.d0; rcut = 0.d0; delr_in = 1.d0; rvdw = 0.d0; have_sum = .false.
  have_prec = .false.; mcsteps = 1; eqsteps = 0; nnumg = 1000; maxguest = 3000
  nwind = 100000
  do_trace = .false.; have_seed = .false.; rota_ratio = 0.5d0; lrestart = .false.
  gfreq = -1.d0; gdelr = -1.d0; gdisp = 0.5d0; gtran = 0.5d0
  press = 0.d0; freq = 0.d0; gfreq_set = .false.; gdelr_set = .false.
  gdisp_set = .false.; gtran_set = .false.; ntpguest = 0
  open(10, file='CONTROL', status='old', action='read')
  read(10, '(a)') title
  do
    read(10, '(a)', iostat=ios) rec
    if (ios /= 0) exit
    call lowcase(rec)
    rec = adjustl(rec)
    if (rec(1:1) == '#' .or. len_trim(rec) == 0) cycle
    if (has(rec, 'seed')) then
      have_seed = .true.
      do k = 1, 4
        seed(k) = intstr(rec)
      end do
    else if (has(rec, 'trace')) then
      do_trace = .true.
    else if (has(rec, 'averaging window')) then
      nwind = intstr(rec)
      if (nwind <= 0) nwind = 100000
    else if (has(rec, 'restart')) then
      lrestart = .true.
    else if (has(rec, 'temp')) then
      temp = dblstr(rec)
    else if (has(rec, '&guest')) then
      ig = intstr(rec)
      ntpguest = max(ntpguest, ig)
      do
        read(10, '(a)') rec2
        call lowcase(rec2)
        rec2 = adjustl(rec2)
        if (rec2(1:1) == '#' .or. len_trim(rec2) == 0) cycle
        if (has(rec2, 'fract')) then
          continue
        else if (has(rec2, 'press')) then
          press(ig) = dblstr(rec2) * 1.d5
        else if (has(rec2, 'move')) then
          gfreq_set(ig) = .true.
          call set_move(rec2, freq(:, ig))
        else if (has(rec2, 'delr')) then
          gdelr_set(ig) = .true.
          delr(ig) = dblstr(rec2)
        else if (has(rec2, 'accep')) then
          if (has(rec2, 'dis')) then
            gdisp_set(ig) = .true.; disp_ratio(ig) = dblstr(rec2)
          else if (has(rec2, 'tra')) then
            gtran_set(ig) = .true.; tran_ratio(ig) = dblstr(rec2)
          end if
        else if (has(rec2, '&end')) then
          exit
        end if
      end do
    else if (has(rec, 'steps')) then
      mcsteps = intstr(rec)
    else if (has(rec, 'numg')) then
      nnumg = intstr(rec)
      if (nnumg <= 0) nnumg = 1000
    else if (has(rec, 'equilibr')) then
      eqsteps = intstr(rec)
    else if (has(rec, 'cut')) then
      rcut = dblstr(rec)
    else if (has(rec, 'delr')) then
      delr_in = dblstr(rec)
    else if (has(rec, 'ewald')) then
      if (has(rec, 'precision')) then
        have_prec = .true.
        eps_prec = dblstr(rec)
      else
        have_sum = .true.
        alpha = dblstr(rec)
        do k = 1, 3
          kmax(k) = intstr(rec)
        end do
      end if
    else if (has(rec, 'rvdw')) then
      rvdw = dblstr(rec)
    else if (has(rec, 'max guest atoms')) then
      maxguest = intstr(rec)
    else if (has(rec, 'move')) then
      call set_move(rec, gfreq)
    else if (has(rec, 'accep')) then
      if (has(rec, 'dis')) then
        gdisp = dblstr(rec)
      else if (has(rec, 'tra')) then
        gtran = dblstr(rec)
      else if (has(rec, 'rot')) then
        rota_ratio = dblstr(rec)
      end if
    else if (has(rec, 'finish')) then
      exit
    end if
  end do
  close(10)
  if (temp <= 0.d0) call fail('no temperature in CONTROL (reference error 380)')
  do ig = 1, ntpguest
    if (.not. gfreq_set(ig)) freq(:, ig) = max(gfreq, 0.d0)
    if (.not. gdelr_set(ig)) delr(ig) = delr_in
    if (.not. gdisp_set(ig)) disp_ratio(ig) = gdisp
    if (.not. gtran_set(ig)) tran_ratio(ig) = gtran
  end do

  ! ------------------------------------------------------------------
  ! FIELD (readinputs.f:11-307, vdw_terms.f:1-187)
  ! ------------------------------------------------------------------
  engunit = 1.d0; ntpatm = 0; ntpvdw = 0; ntpmls = 0
  allocate(unqatm(256))
  open(11, file='FIELD', status='old', action='read')
  read(11, '(a)') rec
  do
    read(11, '(a)', iostat=ios) rec
    if (ios /= 0) exit
    rec2 = rec
    call lowcase(rec2)
    rec2 = adjustl(rec2)
    if (rec2(1:1) == '#' .or. len_trim(rec2) == 0) cycle
    if (has(rec2, 'units')) then
      if (len_trim(rec2(6:)) == 0) then
        engunit = 1.d0
      else if (has(rec2, 'ev')) then
        engunit = 9648.530821d0
      else if (has(rec2, 'kcal')) then
        engunit = 418.4d0
      else if (has(rec2, 'kj')) then
        engunit = 1.d2
      else if (has(rec2, 'k')) then
        engunit = boltz
      end if
    else if (has(rec2, 'vdw')) then
      ntpvdw = intstr(rec2)
      allocate(vdw_a(ntpvdw), vdw_b(ntpvdw), vdw_key(ntpvdw), vdw_prm(5 * ntpvdw))
      vdw_prm = 0.d0
      do k = 1, ntpvdw
        read(11, '(a)') rec
        a1 = getword(rec)
        a2 = getword(rec)
        call lowcase(rec)
        kw = getword(rec)
        select case (kw(1:4))
        case ('12-6'); vdw_key(k) = 1
        case ('lj  '); vdw_key(k) = 2
        case ('nm  '); vdw_key(k) = 3
        case ('buck'); vdw_key(k) = 4
        case ('bhm '); vdw_key(k) = 5
        case ('hbnd'); vdw_key(k) = 6
        case ('snm ', 'hcnm'); vdw_key(k) = 7
        case ('mors'); vdw_key(k) = 8
        case ('wca '); vdw_key(k) = 9
        case default
          call fail('unknown vdw potential '//kw)
        end select
        do j = 1, 5
          vdw_prm(5 * (k - 1) + j) = dblstr(rec)
        end do
        vdw_prm(5 * (k - 1) + 1) = vdw_prm(5 * (k - 1) + 1) * engunit
        if (vdw_key(k) == 1 .or. vdw_key(k) == 6) &
          vdw_prm(5 * (k - 1) + 2) = vdw_prm(5 * (k - 1) + 2) * engunit
        if (vdw_key(k) == 4) vdw_prm(5 * (k - 1) + 3) = vdw_prm(5 * (k - 1) + 3) * engunit
        if (vdw_key(k) == 5) then
          vdw_prm(5 * (k - 1) + 4) = vdw_prm(5 * (k - 1) + 4) * engunit
          vdw_prm(5 * (k - 1) + 5) = vdw_prm(5 * (k - 1) + 5) * engunit
        end if
        vdw_a(k) = 0; vdw_b(k) = 0
        do j = 1, ntpatm
          if (a1 == unqatm(j)) vdw_a(k) = j
          if (a2 == unqatm(j)) vdw_b(k) = j
        end do
      end do
    else if (has(rec2, 'molecu')) then
      ntpmls = intstr(rec2)
      allocate(mol_guest(ntpmls), mol_nummols(ntpmls), mol_nsite(ntpmls), mol_gi(ntpmls))
      allocate(mol_q(20000, ntpmls), mol_m(20000, ntpmls), mol_t(20000, ntpmls))
      allocate(mol_frz(20000, ntpmls), mol_name(20000, ntpmls))
      allocate(mol_x(maxs, ntpmls), mol_y(maxs, ntpmls), mol_z(maxs, ntpmls))
      ig = 0
      do n = 1, ntpmls
        read(11, '(a)') rec
        rec2 = rec
        call lowcase(rec2)
        mol_guest(n) = has(rec2, '&guest')
        mol_gi(n) = 0
        if (mol_guest(n)) then
          ig = ig + 1
          mol_gi(n) = ig
        end if
        mol_nummols(n) = 0
        do
          read(11, '(a)') rec2
          call lowcase(rec2)
          rec2 = adjustl(rec2)
          if (has(rec2, 'nummol')) then
            mol_nummols(n) = intstr(rec2)
          else if (has(rec2, 'atoms')) then
            numsit = intstr(rec2)
            ksite = 0
            do while (ksite < numsit)
              read(11, '(a)') rec
              a1 = getword(rec)
              ra = dblstr(rec)
              rb = dblstr(rec)
              if (mol_guest(n)) then
                nrept = 1; ifrz = 0
                mol_x(ksite + 1, n) = dblstr(rec)
                mol_y(ksite + 1, n) = dblstr(rec)
                mol_z(ksite + 1, n) = dblstr(rec)
              else
                nrept = intstr(rec)
                ifrz = intstr(rec)
              end if
              if (nrept == 0) nrept = 1
              k = 0
              do j = 1, ntpatm
                if (a1 == unqatm(j)) k = j
              end do
              if (k == 0) then
                ntpatm = ntpatm + 1
                unqatm(ntpatm) = a1
                k = ntpatm
              end if
              do j = 1, nrept
                ksite = ksite + 1
                mol_name(ksite, n) = a1
                mol_m(ksite, n) = ra
                mol_q(ksite, n) = rb
                mol_frz(ksite, n) = ifrz
                mol_t(ksite, n) = k
              end do
            end do
            mol_nsite(n) = numsit
          else if (has(rec2, 'finish')) then
            exit
          end if
        end do
      end do
    else if (has(rec2, 'close')) then
      exit
    end if
  end do
  close(11)
  if (ig /= ntpguest) call fail('guest count differs between CONTROL and FIELD (error 2312)')

  ! ------------------------------------------------------------------
  ! CONFIG (readinputs.f:309-513)
  ! ------------------------------------------------------------------
  open(12, file='CONFIG', status='old', action='read')
  read(12, '(a)') rec
  cfgname = rec(1:80)
  read(12, '(a)') rec
  levcfg = intstr(rec)
  imcon = intstr(rec)
  do k = 0, 2
    read(12, '(a)') rec
    do j = 1, 3
      cell(3 * k + j) = dblstr(rec)
    end do
  end do
  nfw = 0
  do n = 1, ntpmls
    if (.not. mol_guest(n)) nfw = nfw + mol_nummols(n) * mol_nsite(n)
  end do
  allocate(fx(nfw), fy(nfw), fz(nfw), fq(nfw), ft(nfw), ffz(nfw))
  gcap = maxguest
  allocate(gx(maxguest, maxg), gy(maxguest, maxg), gz(maxguest, maxg))
  nmol = 0; gns = 0
  i = 0
  do n = 1, ntpmls
    ig = mol_gi(n)
    if (ig > 0) then
      gns(ig) = mol_nsite(n)
      nmol(ig) = mol_nummols(n)
    end if
    k = 0
    do m = 1, mol_nummols(n)
      do s = 1, mol_nsite(n)
        read(12, '(a)') rec
        if (adjustl(rec(1:8)) /= adjustl(mol_name(s, n))) &
          call fail('unidentified atom label in CONFIG (error 25)')
        read(12, '(3f20.0)') ra, rb, rcc
        do j = 1, levcfg
          read(12, '(a)') rec
        end do
        if (ig > 0) then
          k = k + 1
          gx(k, ig) = ra; gy(k, ig) = rb; gz(k, ig) = rcc
        else
          i = i + 1
          fx(i) = ra; fy(i) = rb; fz(i) = rcc
          fq(i) = mol_q(s, n); ft(i) = mol_t(s, n); ffz(i) = mol_frz(s, n)
        end if
      end do
    end do
  end do
  close(12)

  ! restart: the REVIVE of a previous run replaces the molecule counts and
  ! coordinates (gcmc.f:349-410, revread utility_pack.f:1835-1862)
  chainstats = 0.d0
  have_rev = .false.; rev_prod = .false.
  if (lrestart) then
    inquire(file='branch01/REVIVE', exist=have_rev)
    if (have_rev) call read_revive()
  end if

  ! guest templates centred on their COM (readinputs.f:226-253)
  do n = 1, ntpmls
    ig = mol_gi(n)
    if (ig == 0) cycle
    ra = 0.d0; rb = 0.d0; rcc = 0.d0; r1 = 0.d0
    do s = 1, gns(ig)
      ra = ra + mol_x(s, n) * mol_m(s, n)
      rb = rb + mol_y(s, n) * mol_m(s, n)
      rcc = rcc + mol_z(s, n) * mol_m(s, n)
      r1 = r1 + mol_m(s, n)
    end do
    ra = ra / r1; rb = rb / r1; rcc = rcc / r1
    do s = 1, gns(ig)
      gtx(s, ig) = mol_x(s, n) - ra
      gty(s, ig) = mol_y(s, n) - rb
      gtz(s, ig) = mol_z(s, n) - rcc
      gq(s, ig) = mol_q(s, n)
      gm(s, ig) = mol_m(s, n)
      gtt(s, ig) = mol_t(s, n)
    end do
  end do

  ! ------------------------------------------------------------------
  ! build the device context
  ! ------------------------------------------------------------------
  call ck(fmc_ctx_create(ctx), 'fmc_ctx_create')
  call ck(fmc_set_cell(ctx, imcon, cell), 'fmc_set_cell')
  call ck(fmc_set_cutoffs(ctx, rcut, delr_in, rvdw), 'fmc_set_cutoffs')
  if (have_sum) then
    call ck(fmc_set_ewald(ctx, alpha, kmax(1), kmax(2), kmax(3), 1.d0), 'fmc_set_ewald')
  else if (have_prec) then
    call ck(fmc_set_ewald_precision(ctx, eps_prec, 1.d0), 'fmc_set_ewald_precision')
  else
    call fail('no ewald keyword in CONTROL (error 2311)')
  end if
  call ck(fmc_set_vdw(ctx, ntpatm, ntpvdw, vdw_a, vdw_b, vdw_key, vdw_prm), 'fmc_set_vdw')
  call ck(fmc_add_framework(ctx, nfw, fx, fy, fz, fq, ft, ffz), 'fmc_add_framework')
  do ig = 1, ntpguest
    call ck(fmc_add_guest_type(ctx, gns(ig), gtx(:, ig), gty(:, ig), gtz(:, ig), &
            gq(:, ig), gtt(:, ig), gm(:, ig), k), 'fmc_add_guest_type')
    do m = 1, nmol(ig)
      j = (m - 1) * gns(ig)
      call ck(fmc_add_guest_molecule(ctx, ig - 1, gx(j + 1:, ig), gy(j + 1:, ig), &
              gz(j + 1:, ig)), 'fmc_add_guest_molecule')
    end do
  end do
  call ck(fmc_set_capacity(ctx, maxguest), 'fmc_set_capacity')
  call ck(fmc_ctx_finalize(ctx), 'fmc_ctx_finalize')
  call ck(fmc_get_info(ctx, nk, volm, mxgrid, drewd, dlrpot), 'fmc_get_info')
  call invert(cell, rcell, det)
  statvolm = volm * 1d-30
  beta = 1.d0 / (kboltz * temp)
  do ig = 1, ntpguest
    call ck(fmc_guest_ewald3(ctx, ig - 1, e3(ig)), 'fmc_guest_ewald3')
  end do

  ! single_point (gcmc.f:500-527, 2518-2710)
  call ck(fmc_energy_full(ctx, e6), 'fmc_energy_full')
  ecoulg = e6(2) + e6(3) + e6(5)
  evdwg = e6(4) + e6(6)
  spenergy = (ecoulg + evdwg) / engunit
  dlpeng = (e6(1) + e6(2) + e6(3) + e6(4) + e6(5) + e6(6)) / engunit
  energy = spenergy       ! gcmc.f:555-560 (same start value per guest type)
  elrc = 0.d0             ! gcmc.f:138-139: main-program elrc starts at zero

  ! move frequencies (gcmc.f:274-331)
  do ig = 1, ntpguest
    r1 = sum(freq(:, ig))
    if (r1 == 0.d0) then
      freq(:, ig) = 0.d0
      freq(1:3, ig) = 1.d0 / 3.d0
      r1 = 1.d0
    end if
    cum(1, ig) = freq(1, ig) / r1
    do k = 2, 9
      cum(k, ig) = cum(k - 1, ig) + (freq(k, ig) / r1)
    end do
    if (freq(4, ig) > 0.d0 .or. freq(9, ig) > 0.d0) &
      call fail('jump / multi-jump moves are not on this path')
  end do

  ! RNG (gcmc.f:334-336, 2753-2758)
  if (.not. have_seed) then
    call date_and_time(values=values)
    seed(1) = mod(values(8) * 3, 177) + 1
    seed(2) = mod(values(7) * 23, 177) + 1
    seed(3) = mod(values(8) * 5, 177) + 1
    seed(4) = mod(values(7) * 3, 168)
  end if
  call duni_init(seed(1), seed(2), seed(3), seed(4))
  ra = duni()

  open(6, file='OUTPUT')
  write(6, "(/,20x,'FastMC MI355X driver (reference 1.4.0 move loop)',/)")
  write(6, "('simulation cell vectors:')")
  write(6, "(21x,3f12.6)") cell
  write(6, "(/,'system volume:',/,21x,1p,g22.12)") volm
  write(6, "(/,'reciprocal vectors kept: ',i8,'   kmax:',3i4)") nk, kmax
  write(6, "(/,'i,j,k,l values for duni()',4i5,/)") seed
  write(6, '(a35,f22.6)') 'Initial guest energy :', spenergy
  write(6, '(a35,f22.6)') 'van der Waals energy :', evdwg / engunit
  write(6, '(a35,f22.6)') 'Electrostatic energy :', ecoulg / engunit
  write(6, '(a35,f22.6)') 'Energy reported by DL_POLY:', dlpeng
  call flush(6)

  call execute_command_line('mkdir -p branch01')
  open(401, file='branch01/numguests.out')
  if (do_trace) open(402, file='branch01/trace.out')

  ! ------------------------------------------------------------------
  ! move loop (gcmc.f:589-1960)
  ! ------------------------------------------------------------------
  do ig = 1, ntpguest
    delrdisp(ig) = delr(ig); tran_delr(ig) = delr(ig)
  end do
  rotangle = pi / 3.d0; rota_rotangle = rotangle
  disp_count = 0; accept_disp = 0; tran_count = 0; accept_tran = 0
  rota_count = 0; accept_rota = 0; ins_count = 0; del_count = 0
  accept_ins = 0; accept_del = 0
  gcmccount = 0; prodcount = 0; lgchk = .true.; tick = .false.
  production = (eqsteps == 0)
  ! a restarted production run goes on averaging unless equilibration
  ! steps are asked for again (gcmc.f:398-410,573)
  if (have_rev) production = (rev_prod .and. eqsteps <= 0) .or. eqsteps == 0
  dele_last = 0.d0
  allocate(gasmol(maxsamp, ntpguest), gaseng(maxsamp, ntpguest))
  st_n = 0.d0; st_e = 0.d0; st_en = 0.d0; st_n2 = 0.d0; st_e2 = 0.d0
  avcount = 0; rollcount = 0; wsum = 0.d0; seltot = 0.d0; selwin = 0.d0
  call ck(fmc_ctx_timing(ctx, 1, tim), 'fmc_ctx_timing')
  t_rst = 0.d0
  call system_clock(c0, crate)
  do while (lgchk)
    gcmccount = gcmccount + 1
    if (mod(gcmccount, 400_8) == 0) then
      totalguests = sum(nmol(1:ntpguest))
      if (production) then
        if (mcsteps < 0) then
          tick = cshift(tick, -1)
          tick(1) = prodcount >= int(-mcsteps, 8) * (totalguests + 1)
          if (all(tick)) lgchk = .false.
        else
          if (prodcount >= mcsteps) lgchk = .false.
        end if
      else
        if (eqsteps > 0) then
          if (gcmccount >= eqsteps) production = .true.
        else if (eqsteps < 0) then
          tick = cshift(tick, -1)
          tick(1) = gcmccount >= int(-eqsteps, 8) * (totalguests + 1)
          if (all(tick)) then
            production = .true.
            tick = .false.
          end if
        end if
      end if
    end if

    ! restart files every 1000 moves (gcmc.f:737-740)
    if (mod(gcmccount, 1000_8) == 0) then
      call system_clock(cw0)
      call write_restart()
      call system_clock(cw1)
      t_rst = t_rst + dble(cw1 - cw0) / dble(crate)
    end if

    ig = 1
    if (ntpguest > 1) ig = floor(duni() * ntpguest) + 1
    randmov = duni()
    ins = .false.; del = .false.; dsp = .false.; tran = .false.; rota = .false.
    if (randmov < cum(1, ig)) then
      ins = .true.
    else if (randmov < cum(2, ig)) then
      del = .true.
    else if (randmov < cum(3, ig)) then
      dsp = .true.
    else if (randmov < cum(4, ig)) then
      continue
    else if (randmov < cum(5, ig)) then
      continue   ! flex: a no-op in the reference
    else if (randmov < cum(6, ig)) then
      continue   ! swap: a no-op in the reference
    else if (randmov < cum(7, ig)) then
      dsp = .true.; tran = .true.
    else if (randmov < cum(8, ig)) then
      dsp = .true.; rota = .true.
    else if (randmov < cum(9, ig)) then
      continue
    else
      dsp = .true.
    end if
    natms = gns(ig)
    nmols = nmol(ig)
    ! adaptive step sizes (gcmc.f:794-834)
    if (nmols >= 1 .and. disp_count(ig) >= 1) then
      if (mod(disp_count(ig), 100) == 0) then
        if ((accept_disp(ig) / dble(disp_count(ig))) > disp_ratio(ig)) then
          delrdisp(ig) = delrdisp(ig) * 1.05d0
        else
          if (delrdisp(ig) > delrmin) delrdisp(ig) = delrdisp(ig) * 0.95d0
        end if
      else if (mod(disp_count(ig), 50) == 0) then
        if ((accept_disp(ig) / dble(disp_count(ig))) > disp_ratio(ig)) then
          rotangle = rotangle * 1.05d0
        else
          if (rotangle > minangle) rotangle = rotangle * 0.95d0
        end if
      end if
    end if
    if (nmols >= 1 .and. tran_count(ig) >= 1 .and. mod(tran_count(ig), 100) == 0) then
      if ((accept_tran(ig) / dble(tran_count(ig))) > tran_ratio(ig)) then
        tran_delr(ig) = tran_delr(ig) * 1.05d0
      else
        if (tran_delr(ig) > delrmin) tran_delr(ig) = tran_delr(ig) * 0.95d0
      end if
    end if
    if (nmols >= 1 .and. rota_count >= 1 .and. mod(rota_count, 100) == 0) then
      if ((accept_rota / dble(rota_count)) > rota_ratio) then
        rota_rotangle = rota_rotangle * 1.05d0
      else
        if (rota_rotangle > minangle) rota_rotangle = rota_rotangle * 0.95d0
      end if
    end if
    if (nmols == 0) then
      ins = .true.; del = .false.; dsp = .false.; tran = .false.; rota = .false.
    end if

    kind = 0; accepted = .false.; estep = 0.d0; out = 0.d0; delrc = 0.d0
    if (ins) then
      kind = 1
      ins_count = ins_count + 1
      if (sum(nmol(1:ntpguest) * gns(1:ntpguest)) + natms >= maxguest) &
        call fail('too many guest atoms (reference error 75)')
      ra = duni(); rb = duni(); rcc = duni()
      r1 = duni(); r2 = duni(); r3 = duni(); r4 = duni()
      call random_ins(ig, ra, rb, rcc, r1, r2, r3, r4)
      call ck(fmc_trial(ctx, FMC_INSERT, ig - 1, -1, nx, ny, nz, out), 'fmc_trial')
      call ck(fmc_tail_delta(ctx, ig - 1, 1, elrc, delrc), 'fmc_tail_delta')
      estep = (out(1) + out(2) + e3(ig) + out(3) + delrc) / engunit
      rande = duni()
      accepted = energy_eval(estep, rande, press(ig), nmols, 1)
      if (accepted) then
        accept_ins = accept_ins + 1
        j = nmol(ig) * natms
        gx(j + 1:j + natms, ig) = nx(1:natms)
        gy(j + 1:j + natms, ig) = ny(1:natms)
        gz(j + 1:j + natms, ig) = nz(1:natms)
        call ck(fmc_commit(ctx), 'fmc_commit')
        elrc = elrc + delrc
        nmol(ig) = nmol(ig) + 1
        energy(ig) = energy(ig) + estep
      end if
    else if (del) then
      kind = 2
      del_count = del_count + 1
      randchoice = floor(duni() * nmols) + 1
      call ck(fmc_trial(ctx, FMC_DELETE, ig - 1, randchoice - 1, nx, ny, nz, out), &
              'fmc_trial')
      call ck(fmc_tail_delta(ctx, ig - 1, -1, elrc, delrc), 'fmc_tail_delta')
      estep = (out(1) - out(2) - e3(ig) - out(3) + delrc) / engunit
      rande = duni()
      accepted = energy_eval(estep, rande, press(ig), nmols, 2)
      if (accepted) then
        accept_del = accept_del + 1
        j = (randchoice - 1) * natms
        do i = j + 1, (nmols - 1) * natms
          gx(i, ig) = gx(i + natms, ig)
          gy(i, ig) = gy(i + natms, ig)
          gz(i, ig) = gz(i + natms, ig)
        end do
        call ck(fmc_commit(ctx), 'fmc_commit')
        elrc = elrc + delrc
        nmol(ig) = nmol(ig) - 1
        energy(ig) = energy(ig) + estep
      end if
    else if (dsp) then
      if (tran) then
        kind = 4
        tran_count(ig) = tran_count(ig) + 1
        ra = duni(); rb = duni(); rcc = duni()
        r1 = 0.d0; r2 = 0.d0; r3 = 0.d0; r4 = 0.d0
        dd = tran_delr(ig); da = 0.d0
      else if (rota) then
        kind = 5
        rota_count = rota_count + 1
        ra = 0.5d0; rb = 0.5d0; rcc = 0.5d0
        r1 = duni(); r2 = duni(); r3 = duni(); r4 = duni()
        dd = 0.d0; da = rota_rotangle
      else
        kind = 3
        disp_count(ig) = disp_count(ig) + 1
        ra = duni(); rb = duni(); rcc = duni()
        r1 = duni(); r2 = duni(); r3 = duni(); r4 = duni()
        dd = delrdisp(ig); da = rotangle
      end if
      randchoice = floor(duni() * nmols) + 1
      j = (randchoice - 1) * natms
      nx(1:natms) = gx(j + 1:j + natms, ig)
      ny(1:natms) = gy(j + 1:j + natms, ig)
      nz(1:natms) = gz(j + 1:j + natms, ig)
      call random_disp(ig, dd, da, ra, rb, rcc, r1, r2, r3, r4)
      call ck(fmc_trial(ctx, FMC_DISPLACE, ig - 1, randchoice - 1, nx, ny, nz, out), &
              'fmc_trial')
      estep = (out(1) + out(2) + out(3)) / engunit
      if (estep < 0.d0) then
        accepted = .true.
      else
        rande = duni()
        accepted = energy_eval(estep, rande, press(ig), nmols, 3)
      end if
      if (accepted) then
        if (tran) then
          accept_tran(ig) = accept_tran(ig) + 1
        else if (rota) then
          accept_rota = accept_rota + 1
        else
          accept_disp(ig) = accept_disp(ig) + 1
        end if
        gx(j + 1:j + natms, ig) = nx(1:natms)
        gy(j + 1:j + natms, ig) = ny(1:natms)
        gz(j + 1:j + natms, ig) = nz(1:natms)
        call ck(fmc_commit(ctx), 'fmc_commit')
        energy(ig) = energy(ig) + estep
      end if
    end if

    dele_last = merge(estep, 0.d0, accepted)   ! delE(iguest), gcmc.f:861,955,1124
    if (do_trace) write(402, '(i10,i3,i3,i3,i7,6es25.16e3)') gcmccount, ig, kind, &
      merge(1, 0, accepted), nmol(ig), estep, out(1), out(2), out(3), delrc, energy(ig)

    ! production statistics (gcmc.f:1814-1870)
    if (production) then
      prodcount = prodcount + 1
      do i = 1, ntpguest
        ra = dble(nmol(i))
        if (prodcount <= maxsamp) then
          gasmol(prodcount, i) = ra
          gaseng(prodcount, i) = energy(i)
        end if
        st_n(i) = st_n(i) + (ra - st_n(i)) / prodcount
        st_e(i) = st_e(i) + (energy(i) - st_e(i)) / prodcount
        st_en(i) = st_en(i) + (energy(i) * ra - st_en(i)) / prodcount
        st_n2(i) = st_n2(i) + (ra * ra - st_n2(i)) / prodcount
        st_e2(i) = st_e2(i) + (energy(i) ** 2 - st_e2(i)) / prodcount
      end do
      ! selectivity of guest j over guest i over windows of nwind samples,
      ! the last one partial at the end of the run (gcmc.f:1887-1916)
      rollcount = rollcount + 1
      do i = 1, ntpguest
        wsum(i) = wsum(i) + dble(nmol(i))
      end do
      if (mod(prodcount, int(nwind, 8)) == 0 .or. .not. lgchk) then
        avcount = avcount + 1
        weight = dble(rollcount) / dble(nwind)
        do i = 2, ntpguest
          aN = wsum(i) / dble(rollcount)
          do j = 1, i - 1
            if (aN == 0.d0) cycle
            ij_select = (wsum(j) / dble(rollcount)) / aN
            delta = weight * (ij_select - seltot(i, j))
            seltot(i, j) = seltot(i, j) + delta / avcount
            selwin(i, j) = selwin(i, j) + weight * (ij_select * ij_select - selwin(i, j)) / &
              (avcount - 1 + weight)
            seltot(j, i) = ((avcount - 1 + weight) * selwin(i, j) - &
              ((avcount - 1 + weight) * seltot(i, j) * seltot(i, j))) / (avcount - 1 + weight)
          end do
        end do
        wsum = 0.d0
        rollcount = 0
      end if
    end if
    if (mod(gcmccount, int(nnumg, 8)) == 0) then
      do i = 1, ntpguest
        write(401, '(i9,i7,2f20.6,6i4)') gcmccount, nmol(i), energy(i), &
          merge(estep, 0.d0, accepted .and. i == ig), merge(1, 0, kind == 1 .and. i == ig), &
          merge(1, 0, kind == 2 .and. i == ig), merge(kind - 2, 0, kind >= 3 .and. i == ig), 0, 0, 0
      end do
    end if
  end do
  call system_clock(c1)
  t1 = dble(c1 - c0) / dble(crate)

  ! ------------------------------------------------------------------
  ! final statistics (gcmc.f:1964-2332): 5 block means of the production
  ! samples, sigma = population std of the block means
  ! ------------------------------------------------------------------
  write(6, "(/,a17,i12,a15,f13.3,a8)") 'time elapsed for ', gcmccount, ' gcmc steps : ', t1, ' seconds'
  write(6, "(/,a30,f16.2)") 'trial moves per second : ', dble(gcmccount) / max(t1, 1d-9)
  ! where a move's time goes: the C ABI's trial and commit calls (host
  ! set-up, launch, wait for the sums) and the rest of the move loop
  call ck(fmc_ctx_timing(ctx, 0, tim), 'fmc_ctx_timing')
  if (tim(1) > 0.d0) then
    write(6, "(/,'microseconds per move:',f10.3,'  in fmc_trial: set-up',f8.3,'  launch',f8.3, &
      & '  wait',f8.3,'  fmc_commit',f8.3,'  restart files',f8.3,'  the rest',f8.3)") &
      1d6 * t1 / gcmccount, 1d6 * tim(2) / gcmccount, 1d6 * tim(3) / gcmccount, &
      1d6 * tim(4) / gcmccount, 1d6 * tim(6) / gcmccount, 1d6 * t_rst / gcmccount, &
      1d6 * (t1 - tim(2) - tim(3) - tim(4) - tim(6) - t_rst) / gcmccount
  end if
  write(6, "(/,a30,i12)") 'total accepted steps : ', accept_ins + accept_del + &
    sum(accept_disp) + sum(accept_tran) + accept_rota
  if (ins_count > 0) write(6, "(/,3x,a21,f15.9)") 'insertion ratio: ', dble(accept_ins) / dble(ins_count)
  if (del_count > 0) write(6, "(/,3x,a21,f15.9)") 'deletion ratio: ', dble(accept_del) / dble(del_count)
  if (sum(disp_count) > 0) write(6, "(/,3x,a21,f15.9)") 'displacement ratio: ', &
    dble(sum(accept_disp)) / dble(sum(disp_count))
  do i = 1, ntpguest
    call block_stats(i)
  end do
  ! gcmc.f:2315-2322: '+/-' is followed by seltot(j,i), a variance
  if (ntpguest > 1) then
    do i = 1, ntpguest
      do j = 1, i - 1
        write(6, "(/'selectivity stats for guest ',i2,'/'i2)") j, i
        write(6, "(a36,f15.6,a5,f12.3)") '<S>:', seltot(i, j), '+/-', seltot(j, i)
      end do
    end do
  end if
  close(6)
  close(401)
  if (do_trace) close(402)
  rc = fmc_ctx_destroy(ctx)

contains

  subroutine ck(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) then
      write(0, '(a,a,a,i4)') 'fmc_gcmc: ', what, ' failed with status ', code
      error stop 2
    end if
  end subroutine

  subroutine fail(msg)
    character(len=*), intent(in) :: msg
    write(0, '(a,a)') 'fmc_gcmc: ', msg
    error stop 3
  end subroutine

  subroutine set_move(r, f)
    character(len=*), intent(inout) :: r
    real(8), intent(inout) :: f(9)
    if (has(r, 'ins')) then
      f(1) = dblstr(r)
    else if (has(r, 'del')) then
      f(2) = dblstr(r)
    else if (has(r, 'dis')) then
      f(3) = dblstr(r)
    else if (has(r, 'jum')) then
      f(4) = dblstr(r)
    else if (has(r, 'fle')) then
      f(5) = dblstr(r)
    else if (has(r, 'swa')) then
      f(6) = dblstr(r)
    else if (has(r, 'tra')) then
      f(7) = dblstr(r)
    else if (has(r, 'rot')) then
      f(8) = dblstr(r)
    else if (has(r, 'mul')) then
      f(9) = dblstr(r)
    end if
  end subroutine

  ! utility_pack.f:1417-1467
  subroutine invert(a, b, d)
    real(8), intent(in) :: a(9)
    real(8), intent(out) :: b(9), d
    real(8) :: r
    b(1) = a(5) * a(9) - a(6) * a(8)
    b(2) = a(3) * a(8) - a(2) * a(9)
    b(3) = a(2) * a(6) - a(3) * a(5)
    b(4) = a(6) * a(7) - a(4) * a(9)
    b(5) = a(1) * a(9) - a(3) * a(7)
    b(6) = a(3) * a(4) - a(1) * a(6)
    b(7) = a(4) * a(8) - a(5) * a(7)
    b(8) = a(2) * a(7) - a(1) * a(8)
    b(9) = a(1) * a(5) - a(2) * a(4)
    d = a(1) * b(1) + a(4) * b(2) + a(7) * b(3)
    r = 0.d0
    if (abs(d) > 0.d0) r = 1.d0 / d
    b = r * b
  end subroutine

  ! utility_pack.f:3049-3095
  logical function energy_eval(eng, rnd, p, ngsts, knd)
    real(8), intent(in) :: eng, rnd, p
    integer, intent(in) :: ngsts, knd
    real(8) :: test
    test = 0.d0
    if (-1 * beta * eng > 700.d0) then
      test = exp(700.d0)
    else if (knd == 3) then
      test = exp(-1.d0 * beta * eng)
    else if (knd == 1) then
      test = statvolm * p / ((ngsts + 1) * boltz * temp) * exp(-1.d0 * beta * eng)
    else if (knd == 2) then
      if (abs(p) < 1.d-7) then
        test = 1.d0
      else
        test = ngsts * boltz * temp / (statvolm * p) * exp(-1.d0 * beta * eng)
      end if
    end if
    energy_eval = rnd < test
  end function

  ! utility_pack.f:2870-2906
  subroutine rotate(n, q1, q2, q3, q4)
    integer, intent(in) :: n
    real(8), intent(in) :: q1, q2, q3, q4
    real(8) :: qq(9), tx, ty, tz
    integer :: ii
    qq(1) = 2.d0 * q2 * q2 + 2.d0 * q1 * q1 - 1.d0
    qq(2) = 2.d0 * q2 * q3 - 2.d0 * q1 * q4
    qq(3) = 2.d0 * q2 * q4 + 2.d0 * q1 * q3
    qq(4) = 2.d0 * q2 * q3 + 2.d0 * q1 * q4
    qq(5) = 2.d0 * q1 * q1 + 2.d0 * q3 * q3 - 1.d0
    qq(6) = 2.d0 * q3 * q4 - 2.d0 * q1 * q2
    qq(7) = 2.d0 * q2 * q4 - 2.d0 * q1 * q3
    qq(8) = 2.d0 * q3 * q4 + 2.d0 * q1 * q2
    qq(9) = 2.d0 * q4 * q4 + 2.d0 * q1 * q1 - 1.d0
    do ii = 1, n
      tx = qq(1) * nx(ii) + qq(2) * ny(ii) + qq(3) * nz(ii)
      ty = qq(4) * nx(ii) + qq(5) * ny(ii) + qq(6) * nz(ii)
      tz = qq(7) * nx(ii) + qq(8) * ny(ii) + qq(9) * nz(ii)
      nx(ii) = tx; ny(ii) = ty; nz(ii) = tz
    end do
  end subroutine

  subroutine quat(theta, b2, b3, b4, q)
    real(8), intent(in) :: theta, b2, b3, b4
    real(8), intent(out) :: q(4)
    real(8) :: be1, be2, be3, st, nrm
    be1 = 2.d0 * b2 - 1.d0; be2 = 2.d0 * b3 - 1.d0; be3 = 2.d0 * b4 - 1.d0
    st = sin(theta / 2.d0)
    nrm = sqrt(be1 * be1 + be2 * be2 + be3 * be3)
    q(1) = cos(theta / 2.d0)
    q(2) = st * (be1 / nrm); q(3) = st * (be2 / nrm); q(4) = st * (be3 / nrm)
  end subroutine

  ! utility_pack.f:2545-2650 with cartesian (:3097-3111)
  subroutine random_ins(g, fa, fb, fc, q1, q2, q3, q4)
    integer, intent(in) :: g
    real(8), intent(in) :: fa, fb, fc, q1, q2, q3, q4
    real(8) :: xc, yc, zc, q(4)
    integer :: ii
    xc = fa * cell(1) + fb * cell(4) + fc * cell(7)
    yc = fa * cell(2) + fb * cell(5) + fc * cell(8)
    zc = fa * cell(3) + fb * cell(6) + fc * cell(9)
    do ii = 1, gns(g)
      nx(ii) = gtx(ii, g); ny(ii) = gty(ii, g); nz(ii) = gtz(ii, g)
    end do
    call quat(q1 * 2.d0 * pi, q2, q3, q4, q)
    call rotate(gns(g), q(1), q(2), q(3), q(4))
    do ii = 1, gns(g)
      nx(ii) = nx(ii) + xc; ny(ii) = ny(ii) + yc; nz(ii) = nz(ii) + zc
    end do
  end subroutine

  ! utility_pack.f:2679-2763 (random_disp) + 2929-3000 (random_rot)
  subroutine random_disp(g, dl, ang, fa, fb, fc, q1, q2, q3, q4)
    integer, intent(in) :: g
    real(8), intent(in) :: dl, ang, fa, fb, fc, q1, q2, q3, q4
    real(8) :: h, ta, tb, tc, cx, cy, cz, w, sx, sy, sz, xs, ys, zs, x, y, z, q(4)
    integer :: ii
    h = dl * 0.5d0
    ta = (2.d0 * fa - 1.d0) * h
    tb = (2.d0 * fb - 1.d0) * h
    tc = (2.d0 * fc - 1.d0) * h
    do ii = 1, gns(g)
      nx(ii) = ta + nx(ii); ny(ii) = tb + ny(ii); nz(ii) = tc + nz(ii)
    end do
    cx = 0.d0; cy = 0.d0; cz = 0.d0; w = 0.d0
    do ii = 1, gns(g)
      w = w + gm(ii, g)
      cx = cx + nx(ii) * gm(ii, g)
      cy = cy + ny(ii) * gm(ii, g)
      cz = cz + nz(ii) * gm(ii, g)
    end do
    cx = cx / w; cy = cy / w; cz = cz / w
    sx = (rcell(1) * cx + rcell(4) * cy + rcell(7) * cz)
    sy = (rcell(2) * cx + rcell(5) * cy + rcell(8) * cz)
    sz = (rcell(3) * cx + rcell(6) * cy + rcell(9) * cz)
    xs = sx - floor(sx); ys = sy - floor(sy); zs = sz - floor(sz)
    x = (cell(1) * xs + cell(4) * ys + cell(7) * zs)
    y = (cell(2) * xs + cell(5) * ys + cell(8) * zs)
    z = (cell(3) * xs + cell(6) * ys + cell(9) * zs)
    do ii = 1, gns(g)
      nx(ii) = nx(ii) + (x - cx); ny(ii) = ny(ii) + (y - cy); nz(ii) = nz(ii) + (z - cz)
    end do
    do ii = 1, gns(g)
      nx(ii) = nx(ii) - x; ny(ii) = ny(ii) - y; nz(ii) = nz(ii) - z
    end do
    call quat(q1 * ang, q2, q3, q4, q)
    call rotate(gns(g), q(1), q(2), q(3), q(4))
    do ii = 1, gns(g)
      nx(ii) = nx(ii) + x; ny(ii) = ny(ii) + y; nz(ii) = nz(ii) + z
    end do
  end subroutine

  subroutine read_revive()
    ! revread (utility_pack.f:1835-1862): production flag, molecule counts,
    ! rolling statistics, atom count, per molecule type x/y/z of every
    ! site, energies (replaced by the single_point energy, as gcmc.f:555-560
    ! does)
    integer :: u, nn, ng, kk, ns0, tot
    integer, allocatable :: num(:)
    real(8), allocatable :: bx(:), by(:), bz(:)
    logical :: lp
    real(8) :: erev(maxg)
    allocate(num(ntpmls))
    open(newunit=u, file='branch01/REVIVE', form='unformatted', status='old', action='read')
    read(u) lp
    read(u) num
    read(u) chainstats(1:14 * ntpguest + 1)
    read(u) tot
    kk = 0
    do nn = 1, ntpmls
      ns0 = num(nn) * mol_nsite(nn)
      allocate(bx(max(1, ns0)), by(max(1, ns0)), bz(max(1, ns0)))
      read(u) bx(1:ns0)
      read(u) by(1:ns0)
      read(u) bz(1:ns0)
      ng = mol_gi(nn)
      if (ng > 0) then
        if (ns0 > maxguest) call fail('REVIVE holds more guest atoms than max guest atoms (error 75)')
        nmol(ng) = num(nn)
        gx(1:ns0, ng) = bx(1:ns0)
        gy(1:ns0, ng) = by(1:ns0)
        gz(1:ns0, ng) = bz(1:ns0)
      else
        if (num(nn) /= mol_nummols(nn)) call fail('REVIVE framework differs from FIELD')
        fx(kk + 1:kk + ns0) = bx(1:ns0)
        fy(kk + 1:kk + ns0) = by(1:ns0)
        fz(kk + 1:kk + ns0) = bz(1:ns0)
        kk = kk + ns0
      end if
      deallocate(bx, by, bz)
    end do
    read(u) erev(1:ntpguest)
    close(u)
    rev_prod = lp
  end subroutine

  subroutine write_restart()
    ! revive (utility_pack.f:1772-1814): REVCON with every atom, guests
    ! first then the frameworks (the condense order, utility_pack.f:3337-
    ! 3380), and the unformatted REVIVE
    integer :: u, nn, ng, m2, s2, kk, a, ns0
    integer, allocatable :: num(:)
    character(len=80) :: title
    character(len=40) :: head
    character(len=60) :: vec(3)
    character(len=18) :: lab
    character(len=1), parameter :: nl = achar(10)
    ! REVCON's records are made by internal writes with the reference's
    ! edit descriptors and go out as bytes, each ended by a newline (the
    ! record a formatted sequential write makes); the framework block is
    ! one string of ready-made lines.  Both files are rewritten in place
    ! and cut at their new end (ENDFILE), which costs less than deleting
    ! and creating them every 1000 moves
    open(newunit=u, file='branch01/REVCON', form='unformatted', access='stream', &
         status='unknown', position='rewind')
    write(title, '(a80)') cfgname
    write(head, '(2i10,e20.10)') 0, imcon, dele_last
    write(u) title, nl, head, nl
    if (imcon > 0) then
      write(vec, '(3f20.12)') cell
      write(u) vec(1), nl, vec(2), nl, vec(3), nl
    end if
    a = 0
    do nn = 1, ntpmls
      ng = mol_gi(nn)
      if (ng == 0) cycle
      do m2 = 1, nmol(ng)
        do s2 = 1, gns(ng)
          a = a + 1
          kk = (m2 - 1) * gns(ng) + s2
          write(lab, '(a8,i10)') mol_name(s2, nn), a
          write(vec(1), '(3g20.10)') gx(kk, ng), gy(kk, ng), gz(kk, ng)
          write(u) lab, nl, vec(1), nl
        end do
      end do
    end do
    ! the framework never moves: its lines ('(a8,i10)' then '(3g20.10)',
    ! 80 bytes with the newlines) are formatted on the first write; later
    ! writes only renumber the atoms when the guest count has changed
    if (nfw > 0) then
      if (.not. allocated(fw_text)) call frame_text()
      if (fw_first /= a) then
        do kk = 1, nfw
          call put_i10(fw_text, 80 * (kk - 1) + 9, a + kk)
        end do
        fw_first = a
      end if
      write(u) fw_text
      a = a + nfw
    end if
    endfile(u)
    close(u)
    allocate(num(ntpmls))
    do nn = 1, ntpmls
      num(nn) = mol_nummols(nn)
      if (mol_gi(nn) > 0) num(nn) = nmol(mol_gi(nn))
    end do
    chainstats(1) = dble(prodcount)
    do ng = 1, ntpguest
      kk = 1 + 14 * (ng - 1)
      chainstats(kk + 1) = st_n(ng)
      chainstats(kk + 2) = st_e(ng)
      chainstats(kk + 3) = st_en(ng)
      chainstats(kk + 4) = st_n2(ng)
      chainstats(kk + 5) = st_e2(ng)
    end do
    open(newunit=u, file='branch01/REVIVE', form='unformatted', status='unknown', &
         position='rewind')
    write(u) production
    write(u) num
    write(u) chainstats(1:14 * ntpguest + 1)
    write(u) a
    kk = 0
    do nn = 1, ntpmls
      ng = mol_gi(nn)
      ns0 = num(nn) * mol_nsite(nn)
      if (ng > 0) then
        write(u) gx(1:ns0, ng)
        write(u) gy(1:ns0, ng)
        write(u) gz(1:ns0, ng)
      else
        write(u) fx(kk + 1:kk + ns0)
        write(u) fy(kk + 1:kk + ns0)
        write(u) fz(kk + 1:kk + ns0)
        kk = kk + ns0
      end if
    end do
    write(u) energy(1:ntpguest)
    endfile(u)
    close(u)
  end subroutine

  subroutine frame_text()
    ! the framework's REVCON lines in FIELD order, atom numbers zero
    character(len=18) :: lab
    character(len=60) :: xyz
    integer :: nn, m2, s2, kk, o
    integer :: j
    allocate(fw_text(80 * nfw))
    kk = 0
    do nn = 1, ntpmls
      if (mol_gi(nn) > 0) cycle
      do m2 = 1, mol_nummols(nn)
        do s2 = 1, mol_nsite(nn)
          kk = kk + 1
          write(lab, '(a8,i10)') mol_name(s2, nn), 0
          write(xyz, '(3g20.10)') fx(kk), fy(kk), fz(kk)
          o = 80 * (kk - 1)
          do j = 1, 18
            fw_text(o + j) = lab(j:j)
          end do
          fw_text(o + 19) = achar(10)
          do j = 1, 60
            fw_text(o + 19 + j) = xyz(j:j)
          end do
          fw_text(o + 80) = achar(10)
        end do
      end do
    end do
    fw_first = -1
  end subroutine

  subroutine put_i10(buf, pos, v)
    ! v >= 0 right-justified in buf(pos:pos+9), as the I10 edit writes it
    character(len=1), intent(inout) :: buf(:)
    integer, intent(in) :: pos, v
    integer :: p, x
    x = v
    do p = pos + 9, pos, -1
      if (x == 0 .and. p < pos + 9) then
        buf(p) = ' '
      else
        buf(p) = achar(48 + mod(x, 10))
        x = x / 10
      end if
    end do
  end subroutine

  subroutine block_stats(g)
    ! End-of-run statistics of guest g for this single branch, in the
    ! reference's arithmetic: five windows, window x summing the samples
    ! (x-1)*nwind ... (x-1)*nwind+nwind (nwind+1 of them, a sample past the
    ! last one reads zero) and dividing by nwind = int(samples/5); Welford
    ! mean and population deviation of the window <N>, <E>, <EN>, <N^2>,
    ! <E^2>, Q_st, C_v (gcmc.f:1969-2031); Q_st and C_v of the means
    ! (gcmc.f:2051-2100); the union over branches, here one, weighted by
    ! the sample count (gcmc.f:2366-2432); the report of gcmc.f:2280-2321.
    ! Beyond maxsamp samples the averages are the rolling ones and the
    ! errors (from the reference's rolling windows) are not computed: NaN.
    use ieee_arithmetic
    integer, intent(in) :: g
    integer :: nwind, x, z
    integer(8) :: y, k
    real(8) :: sw(5), w(7), mean(7), m2(7), delta, a, e, cs(6), sd(8), av(8)
    real(8) :: wt, ua(8), us(8)
    wt = dble(prodcount)
    if (prodcount <= int(maxsamp, 8)) then
      nwind = int(wt / 5)
      mean = 0.d0
      m2 = 0.d0
      do x = 1, 5
        y = int(x - 1, 8) * nwind
        sw = 0.d0
        do k = y + 1, y + nwind + 1
          a = 0.d0
          e = 0.d0
          if (k <= prodcount) then
            a = gasmol(k, g)
            e = gaseng(k, g)
          end if
          sw(1) = sw(1) + a
          sw(2) = sw(2) + e
          sw(3) = sw(3) + a * e
          sw(4) = sw(4) + a * a
          sw(5) = sw(5) + e * e
        end do
        w(1:5) = sw(1:5) / nwind
        w(6) = calc_qst(w(5), w(2), w(1), w(4), w(3))
        w(7) = calc_cv(w(5), w(2), w(1), w(4), w(3))
        do z = 1, 7
          delta = w(z) - mean(z)
          mean(z) = mean(z) + delta / x
          m2(z) = m2(z) + delta * (w(z) - mean(z))
        end do
      end do
      cs(1:6) = mean(1:6)
      sd(1:7) = sqrt(m2(1:7) / 5)
      sd(8) = 0.d0
    else
      cs = (/ st_n(g), st_e(g), st_en(g), st_n2(g), st_e2(g), 0.d0 /)
      sd = ieee_value(0.d0, ieee_quiet_nan)
    end if
    av(1:5) = cs(1:5)
    av(6) = calc_qst(cs(5), cs(2), cs(1), cs(4), cs(3))
    av(7) = calc_cv(cs(5), cs(2), cs(1), cs(4), cs(3))
    av(8) = cs(6)
    ua = (wt * av) / wt
    ua(6) = calc_qst(ua(5), ua(2), ua(1), ua(4), ua(3))
    ua(7) = calc_cv(ua(5), ua(2), ua(1), ua(4), ua(3))
    us = sqrt((wt * (sd ** 2 + av ** 2)) / wt - ua ** 2)
    write(6, "(/,'final stats for guest ',i2,/)") g
    write(6, "(/,a36,f15.6,a5,f12.3,/,a36,f15.6,a5,f12.3,/,a36,f15.6,a5,f12.3,/," // &
             "a36,f15.6,a5,f12.3,/,a36,f15.6,a5,f12.3,/)") &
      '<N>: ', ua(1), ' +/- ', us(1), '<E>: ', ua(2), ' +/- ', us(2), &
      '<E*N>: ', ua(3), ' +/- ', us(3), '<N*N>: ', ua(4), ' +/- ', us(4), &
      '<E*E>: ', ua(5), ' +/- ', us(5)
    write(6, "(/,a60,f15.6,/,a60,f15.6)") 'average number of guests: ', ua(1), &
      'standard error: ', us(1)
    write(6, "(a60,f15.6,/,a60,f15.6,/,a60,f15.6,/,a60,f15.6,/,a60,i15,/)") &
      'Isosteric heat of adsorption (kcal/mol): ', ua(6), 'Isosteric heat error: ', us(6), &
      'Heat capacity, Cv (kcal/mol/K): ', ua(7), 'Heat capacity error: ', us(7), &
      'Total steps counted: ', int(wt)
  end subroutine

  real(8) function calc_qst(e2, e, n, n2, en)
    ! utility_pack.f:2280-2290
    real(8), intent(in) :: e2, e, n, n2, en
    calc_qst = -1.d0 * (en - e * n) / (n2 - n * n) + rgas * temp
  end function

  real(8) function calc_cv(e2, e, n, n2, en)
    ! utility_pack.f:2305-2316
    real(8), intent(in) :: e2, e, n, n2, en
    calc_cv = (3.d0 / 2.d0 * rgas) + (e2 - (e * e) - (en - e * n) ** 2 &
              / (n2 - n * n)) / (n * kboltz * temp * temp)
  end function
end program fmc_gcmc
