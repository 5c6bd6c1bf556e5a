! blagpu: a Fortran host for the MI355X step library, taking the inputs of
! Simson's bla -- bla.i (version 20070716) and a .u velocity file -- and
! writing the end field in the same .u format.
!
! It keeps bla's driver logic on the host: the bla.i grammar of
! bla/rparambl.f:160-1860 (temporal channel subset), the .u reader/writer of
! bla/rdiscbl.f:93-411 and bla/wdiscbl.f:50-136, the RK weight bookkeeping of
! bla/bla.f:1465-1477 and 2384-2397, the CFL stop of bla/wcflbl.f:87-98 and
! the stop criteria of bla/bla.f:2779-2833, the amp.stat writer of
! bla/wampbl.f (channel), the x-y statistics schedule of bla/bla.f:1830,
! 2089-2142, 2907-2930 and file of bla/wxys.f, the sgs.i reader of
! bla/rparamles.f (ADM-RT and the Smagorinsky models) --
! and calls the library for prhs (bla.f:1562),
! the nonlinbl plane loop (bla.f:2251-2288), its plane statistics
! (boxamp.f, nonlinbl.f:430; boxxys.f) and the linearbl plane loop
! (bla.f:2433-2475)
! through ISO_C_BINDING.
!
! Run as `blagpu --pressure` it also writes the pressure file name.p beside
! every saved field, in the record layout of bla/wdiscp.f, as bla does when
! par.f sets pressure = 1 (bla.f:2660, 2968): the pressure of the saved
! state, from nonlinp + linearp (simson_gpu_pressure).
program blagpu
  use iso_c_binding
  use simson_gpu_mod
  implicit none

  integer, parameter :: dp = kind(1.d0)
  real(dp), parameter :: pi = 3.14159265358979323846_dp
  character(len=256) :: line, namnin, namnut, namamp, namxys
  character(len=256) :: nmsave(100)
  real(dp) :: tsave(101)
  integer :: version, maxit, nst, ibc, icfl, iamp, iext, ixys, msave
  integer :: mwave, npl, isfd, imhd, loctyp, wbci, i, it, isave, st, ixyss
  real(dp) :: tmax, cpumax, wallmax, dt, cflmaxin, xl, rot, retau, cdev
  logical :: write_inter, varsiz, cflux, pert, cim, gall, suction, spat
  logical :: sgs, tripf, longli, fileurms, pou, stopf, docfl, corrf, corrf_x, serf
  logical :: chkstop, tabfre, rbfl, osmod, streak, waves
  ! spatial simulations: fringe (ffun.f) and base flow [3][nyp][nxp]
  type(simson_fringe_t) :: fr
  type(simson_wall_t) :: wl
  type(simson_locf_t) :: lfc
  real(dp) :: cnrk(3)
  real(dp) :: ampob, amp2d, bstart, blength, rlam, spanv, du0upp
  real(dp), allocatable :: base(:,:,:), bjr(:,:), bji(:,:), tab(:,:), mforce(:,:)
  type(simson_baseflow_t) :: bf
  ! field and header
  integer :: nx, nyp, nzc, nfzsym, fltype
  real(dp) :: re, xlf, zl, t, xs, dstar
  real(dp), allocatable :: ur(:,:,:,:), ui(:,:,:,:), urx(:)
  ! time stepping (rkcoeff.f:21-32)
  real(dp) :: anrk(3), bnrk(3), dtn, dtnp1, an, bn, anp1, bnp1, cflp, cfl, px
  real(dp) :: dtmax, cflmax, tleft
  logical :: vart, do_pressure
  character(len=32) :: arg
  real(dp), allocatable :: prr(:,:,:), pri(:,:,:)
  real(dp), allocatable :: amp(:,:), wint(:)
  ! x-y statistics 1-15 of boxxys.f, xys(x, y, i), and the sum of weights
  integer, parameter :: nxys = 15
  real(dp), allocatable :: xys(:,:,:)
  real(dp) :: sumw, txys, zs, u0low, w0low
  type(simson_params_t) :: p
  type(simson_les_t) :: les
  integer :: sgsver
  real(dp) :: cutoff_inv
  type(c_ptr) :: g
  integer(8) :: c0, c1, crate

  do_pressure = .false.
  do i = 1, command_argument_count()
     call get_command_argument(i, arg)
     if (trim(arg) == '--pressure') do_pressure = .true.
  end do
  anrk = [8._dp/15._dp, 5._dp/12._dp, 3._dp/4._dp]
  cnrk = [0._dp, 8._dp/15._dp, 2._dp/3._dp]
  bnrk = [0._dp, -17._dp/60._dp, -5._dp/12._dp]

  ! ---------------- bla.i (rparambl.f) ----------------
  open(10, file='bla.i', status='old')
  call nxt(line); read(line, *) version
  if (version /= 20070716) call stopnow(345431)
  call nxt(line); namnin = word(line)
  call nxt(line); namnut = word(line)
  call nxt(line); read(line, *) tmax;     tmax = max(tmax, 0._dp)
  call nxt(line); read(line, *) maxit;    maxit = max(maxit, 0)
  call nxt(line); read(line, *) cpumax
  call nxt(line); read(line, *) wallmax
  call nxt(line); read(line, *) write_inter
  call nxt(line); read(line, *) dt
  call nxt(line); read(line, *) nst
  call nxt(line); read(line, *) cflmaxin
  call nxt(line); read(line, *) xl
  call nxt(line); read(line, *) varsiz
  call nxt(line); read(line, *) rot
  call nxt(line); read(line, *) cflux
  retau = 0._dp
  if (.not. cflux) then
     call nxt(line); read(line, *) retau
  end if
  call nxt(line); read(line, *) pert
  call nxt(line); read(line, *) ibc
  call nxt(line); read(line, *) cim
  call nxt(line); read(line, *) gall
  call nxt(line); read(line, *) suction
  if (suction) then
     ! the suction parameters follow in bla.i (rparambl.f); not supported
     write(*,*) 'blagpu: bla.i requests a feature outside the channel step library'
     call stopnow(13)
  end if
  call nxt(line); read(line, *) spat
  cdev = 0._dp
  if (spat) then
     ! rparambl.f:547-720: the fringe toward the computed base flow only
     call nxt(line); read(line, *) tabfre
     call nxt(line); read(line, *) rbfl
     call nxt(line); read(line, *) fr%fmax
     call nxt(line); read(line, *) fr%fstart
     call nxt(line); read(line, *) fr%fend
     call nxt(line); read(line, *) fr%frise
     call nxt(line); read(line, *) fr%ffall
     call nxt(line); read(line, *) ampob
     call nxt(line); read(line, *) amp2d
     call nxt(line); read(line, *) osmod
     call nxt(line); read(line, *) streak
     call nxt(line); read(line, *) waves
     fr%xsc = 0._dp
     if (tabfre .or. rbfl .or. ampob /= 0._dp .or. amp2d /= 0._dp .or. osmod .or. streak &
          .or. waves) then
        write(*,*) 'blagpu: only the fringe damping toward the computed base flow is supported'
        call stopnow(13)
     end if
  else if (.not. suction) then
     call nxt(line); read(line, *) cdev
  end if
  call nxt(line); read(line, *) sgs
  call nxt(line); read(line, *) isfd
  call nxt(line); read(line, *) imhd
  call nxt(line); read(line, *) loctyp
  if (loctyp == 3) then
     ! rparambl.f:908-935: amplitudes x, y, z, x scale, x position, y scale,
     ! x type, frequency of the force of locf.f:358-410
     call nxt(line); read(line, *) lfc%ampx
     call nxt(line); read(line, *) lfc%ampy
     call nxt(line); read(line, *) lfc%ampz
     call nxt(line); read(line, *) lfc%xscale
     call nxt(line); read(line, *) lfc%xloc0
     call nxt(line); read(line, *) lfc%yscale
     call nxt(line); read(line, *) lfc%xtype
     call nxt(line); read(line, *) lfc%tomeg
     lfc%xsc = 0._dp
  end if
  call nxt(line); read(line, *) tripf
  call nxt(line); read(line, *) wbci
  if (wbci == 1) then
     ! rparambl.f:1206-1232: the blowing/suction strip of cwallbc.f
     call nxt(line); read(line, *) wl%amp
     call nxt(line); read(line, *) wl%damp
     call nxt(line); read(line, *) wl%xstart
     call nxt(line); read(line, *) wl%xend
     call nxt(line); read(line, *) wl%xrise
     call nxt(line); read(line, *) wl%xfall
     call nxt(line); read(line, *) wl%zbet
     call nxt(line); read(line, *) wl%tomeg
     wl%xsc = 0._dp
     wl%zsc = 0._dp
  else if (wbci /= 0) then
     write(*,*) 'blagpu: bla.i requests a feature outside the channel step library'
     call stopnow(13)
  end if
  call nxt(line); read(line, *) icfl
  call nxt(line); read(line, *) iamp
  if (iamp > 0) then
     call nxt(line); namamp = word(line)
     call nxt(line); read(line, *) fileurms
  end if
  call nxt(line); read(line, *) longli
  call nxt(line); read(line, *) iext
  if (iext > 0) call nxt(line)
  call nxt(line); read(line, *) ixys
  ixys = max(0, ixys)
  txys = 0._dp
  ixyss = 0
  if (ixys > 0) then
     ! rparambl.f:1494-1660 with mxys = 1 (bla.f:200)
     ixys = max(ixys, nst) / nst * nst
     call nxt(line); namxys = word(line)
     call nxt(line); read(line, *) ixyss
     call nxt(line); read(line, *) txys
     ixyss = max(0, ixyss)
     if (ixyss > 0) ixyss = max(ixyss, nst) / nst * nst
     call nxt(line); read(line, *) corrf
     call nxt(line); read(line, *) corrf_x
     call nxt(line); read(line, *) serf
     if (corrf .or. corrf_x .or. serf) then
        write(*,*) 'blagpu: two-point correlations and time series are not supported'
        call stopnow(13)
     end if
  end if
  call nxt(line); read(line, *) msave
  if (msave < 0 .or. msave > 99) call stopnow(13)
  do i = 1, msave
     call nxt(line); read(line, *) tsave(i)
     call nxt(line); nmsave(i) = word(line)
  end do
  tsave(msave + 1) = 1.e30_dp
  call nxt(line); read(line, *) mwave
  call nxt(line); read(line, *) npl
  close(10)
  if (pert .or. cim .or. suction .or. isfd /= 0 .or. imhd /= 0 .or. &
       (loctyp /= 0 .and. loctyp /= 3) .or. tripf .or. (wbci /= 0 .and. wbci /= 1) .or. &
       (ibc /= 0 .and. ibc /= 100 .and. ibc /= 101) .or. nst /= 3 &
       .or. mwave > 0 .or. npl > 0) then
     write(*,*) 'blagpu: bla.i requests a feature outside the channel step library'
     call stopnow(13)
  end if
  ! rparambl.f:1860-1866, 2029-2031
  if (icfl > 0 .and. icfl < nst) icfl = nst
  if (icfl > 0) icfl = min(icfl / nst * nst, 3 * nst)
  maxit = maxit / nst * nst
  ! variable time step (rparambl.f:1996-2017): dt <= 0, with |dt| as the
  ! largest step when negative; the CFL target scaled by the RK3 stability
  ! bound sqrt(3)
  vart = dt <= 0._dp
  dtmax = 1.e30_dp
  if (dt < 0._dp) dtmax = -dt
  cflmax = cflmaxin * sqrt(3._dp)

  ! sgs.i (rparamles.f:23-165): ADM-RT (iles = 1), Smagorinsky (2), the
  ! Smagorinsky model on the high-pass filtered velocity (3) or no model
  if (sgs) then
     les%iles = 1; cutoff_inv = 1.5_dp; les%iord = 5; les%chi = 0.2_dp
     les%cs = 0._dp; les%ineg = 1
     open(10, file='sgs.i', status='old')
     call nxt(line); read(line, *) sgsver
     if (sgsver /= 20060909) call stopnow(345432)
     call nxt(line); read(line, *) les%iles
     if (les%iles == 1) then
        call nxt(line); read(line, *) cutoff_inv
        call nxt(line); read(line, *) les%iord
        call nxt(line); read(line, *) les%chi
     else if (les%iles == 2) then
        ! rparamles.f:72-81: cs, and ineg for the dynamic procedure (cs = 0)
        call nxt(line); read(line, *) les%cs
        if (les%cs == 0._dp) then
           call nxt(line); read(line, *) les%ineg
        end if
     else if (les%iles == 3) then
        ! rparamles.f:82-91: the filter, cs, and ineg for the dynamic procedure
        call nxt(line); read(line, *) cutoff_inv
        call nxt(line); read(line, *) les%iord
        call nxt(line); read(line, *) les%cs
        if (les%cs == 0._dp) then
           call nxt(line); read(line, *) les%ineg
        end if
     else if (les%iles /= 0) then
        write(*,*) 'blagpu: iles must be between 0 and 3: ', les%iles
        call stopnow(45543)
     end if
     close(10)
     les%cutoff = 1._dp / cutoff_inv
     if (les%cutoff < 0._dp .or. les%cutoff > 1._dp .or. les%iord < 0 .or. les%iord > 5) &
          call stopnow(45543)
     if (les%ineg < 0 .or. les%ineg > 2) call stopnow(45543)
     if (les%cs /= 0._dp .and. les%ineg /= 1) call stopnow(45543)
  end if

  ! ---------------- .u file (rdiscbl.f) ----------------
  open(12, file=trim(namnin), status='old', form='unformatted')
  read(12) re, pou, xlf, zl, t, xs
  read(12) nx, nyp, nzc, nfzsym
  read(12) fltype, dstar
  ! temporal (fltype 1) or spatial (fltype 4) Poiseuille flow, a spatial
  ! Blasius, Falkner-Skan or Falkner-Skan-Cooke boundary layer (6, 7, 8), or
  ! the temporal Blasius boundary layer (3)
  if (nfzsym /= 0 .or. .not. (fltype == 1 .or. (.not. spat .and. fltype == 3) .or. &
       (spat .and. (fltype == 4 .or. (fltype >= 6 .and. fltype <= 8))))) call stopnow(13)
  if (ibc >= 100 .and. fltype < 6 .and. fltype /= 3) call stopnow(13)
  if ((fltype >= 6 .or. fltype == 3) .and. (iamp > 0 .or. ixys > 0)) then
     write(*,*) 'blagpu: amplitude and x-y statistics files are written for channels only'
     call stopnow(13)
  end if
  bstart = 0._dp
  blength = 0._dp
  rlam = 0._dp
  spanv = 0._dp
  if (fltype >= 4 .and. fltype <= 6) read(12) bstart, blength
  if (fltype >= 7 .and. fltype <= 9) read(12) bstart, blength, rlam, spanv
  ! bla.i gives lengths and times in the units of dstar = 2/yl (1 for
  ! channels); scale them as rescale.f:33-110 does
  tmax = tmax * dstar
  dt = dt * dstar
  dtmax = dtmax * dstar
  tsave(1:msave) = tsave(1:msave) * dstar
  txys = txys * dstar
  rot = rot / dstar
  xl = xl * dstar
  fr%fstart = fr%fstart * dstar
  fr%fend = fr%fend * dstar
  fr%frise = fr%frise * dstar
  fr%ffall = fr%ffall * dstar
  fr%fmax = fr%fmax / dstar
  wl%xstart = wl%xstart * dstar
  wl%xend = wl%xend * dstar
  wl%xrise = wl%xrise * dstar
  wl%xfall = wl%xfall * dstar
  wl%zbet = wl%zbet / dstar
  wl%tomeg = wl%tomeg / dstar
  lfc%ampx = lfc%ampx / dstar
  lfc%ampy = lfc%ampy / dstar
  lfc%ampz = lfc%ampz / dstar
  lfc%tomeg = lfc%tomeg / dstar
  lfc%xscale = lfc%xscale * dstar
  lfc%xloc0 = lfc%xloc0 * dstar
  lfc%yscale = lfc%yscale * dstar
  allocate(ur(nx/2, nyp, nzc, 7), ui(nx/2, nyp, nzc, 7), urx(nx))
  ur = 0._dp
  ui = 0._dp
  do i = 1, 3
     do st = 1, nzc
        do it = 1, nyp
           read(12) urx
           ur(:, it, st, i) = urx(1:nx:2)
           ui(:, it, st, i) = urx(2:nx:2)
        end do
     end do
  end do
  close(12)
  ! lower-wall velocity of the mean mode (rdiscbl.f:316-323, bla.f:1041-1044)
  u0low = ur(1, nyp, 1, 1)
  w0low = ur(1, nyp, 1, 3)
  if (abs(u0low) < 1.e-8_dp) u0low = 0._dp
  if (abs(w0low) < 1.e-8_dp) w0low = 0._dp
  zs = 0._dp
  sumw = 0._dp
  allocate(xys(nx, nyp, nxys))
  xys = 0._dp

  ! ---------------- device state ----------------
  p%nx = nx; p%ny = nyp; p%nz = nzc
  p%precision = SIMSON_FP64; p%ibc = ibc
  ! no pressure gradient unless temporal Poiseuille flow (bla.f:1622-1662)
  if (fltype /= 1) then
     cflux = .false.
     retau = 0._dp
  end if
  p%cflux = merge(1, 0, cflux)
  p%re = re; p%xl = xl; p%zl = zl; p%rot = rot; p%retau = retau
  call chk(simson_gpu_init(p, g))
  if (sgs) then
     les%chi = les%chi / dstar                      ! chi_scaled, bla.f:2220
     call chk(simson_gpu_les(g, les))
  end if
  if (cflux) then
     px = -2._dp / re                             ! bla.f:1629-1630
  else
     px = -(retau / re)**2                        ! bla.f:1645 (0 for fltype 4)
  end if
  if (spat) then
     allocate(base(3 * nx / 2, nyp, 3))
     base = 0._dp
     if (fltype == 4) then
        ! the Poiseuille base flow of fltype 4 at the 3/2 x grid (cbflow.f:117-131)
        do it = 1, nyp
           base(:, it, 1) = 1._dp - cos(pi * real(it - 1, dp) / real(nyp - 1, dp))**2
        end do
     else
        ! boundary layer: the similarity table fsc.dat (rbla.f), the base flow
        ! of cbflow.f and its free-stream derivatives (bflow.f)
        call read_fsc()
        bf%nx = nx; bf%ny = nyp; bf%fltype = fltype
        bf%xl = xl; bf%re = re; bf%dstar = dstar; bf%bstart = bstart; bf%blength = blength
        bf%rlam = rlam; bf%spanv = spanv; bf%xsc = 0._dp; bf%u0low = 0._dp; bf%w0low = 0._dp
        allocate(bjr(nx / 2, 3), bji(nx / 2, 3))
        call chk(simson_base_flow(bf, tab, int(size(tab, 2), c_int), base, bjr, bji, du0upp))
        if (ibc == 100) call top_values()
        if (ibc >= 100) call chk(simson_gpu_free_stream(g, bjr, bji, du0upp))
     end if
     call chk(simson_gpu_fringe(g, fr, base))
  end if
  if (fltype == 3) then
     ! temporal boundary layer: the similarity table fsc.dat (rbla.f), the
     ! parallel profile's free-stream data (bflow.f) and the parallel-flow
     ! forcing of nonlinbl.f:771-795 (updated per substep when cdev /= 0)
     call read_fsc()
     bf%nx = nx; bf%ny = nyp; bf%fltype = fltype
     bf%xl = xl; bf%re = re; bf%dstar = dstar; bf%bstart = 0._dp; bf%blength = 0._dp
     bf%rlam = rlam; bf%spanv = spanv; bf%xsc = 0._dp; bf%u0low = 0._dp; bf%w0low = 0._dp
     allocate(base(3 * nx / 2, nyp, 3), bjr(nx / 2, 3), bji(nx / 2, 3), mforce(nyp, 3))
     call chk(simson_base_flow(bf, tab, int(size(tab, 2), c_int), base, bjr, bji, du0upp))
     if (ibc == 100) call top_values()
     if (ibc >= 100) call chk(simson_gpu_free_stream(g, bjr, bji, du0upp))
     call chk(simson_parallel_force(bf, tab, int(size(tab, 2), c_int), cdev, t, mforce))
     call chk(simson_gpu_mean_force(g, mforce))
  end if
  if (wbci == 1) call chk(simson_gpu_wall(g, wl))
  if (loctyp == 3) call chk(simson_gpu_local_force(g, lfc))
  call chk(simson_gpu_upload(g, ur, ui, 3_c_int, c_null_ptr, px))
  ! Clenshaw-Curtis weights (preprbl.f:78-87) for the amplitude file
  allocate(amp(nyp, 20), wint(nyp))
  do it = 1, nyp
     wint(it) = -.5_dp
     do i = 1, (nyp + 1) / 2 - 2
        wint(it) = wint(it) + cos(pi * real(2 * i * (it - 1), dp) / real(nyp - 1, dp)) / &
             real((2 * i)**2 - 1, dp)
     end do
     wint(it) = wint(it) + .5_dp * cos(pi * real(it - 1, dp)) / real((nyp - 1)**2 - 1, dp)
     wint(it) = -4._dp / real(nyp - 1, dp) * wint(it)
     if (it == 1 .or. it == nyp) wint(it) = wint(it) * .5_dp
  end do
  if (iamp > 0) open(15, file=trim(namamp), status='replace')

  ! ---------------- time stepping (bla.f:1465-1477, 1790-2885) ----------------
  dtn = dt
  if (vart) then
     ! first step from the CFL of the initial field (getdt.f, bla.f:1407-1415)
     call chk(simson_gpu_cfl(g, cflp))
     dtn = cflmax / (cflp * pi + 2._dp * abs(rot))
     write(*,*) 'First time step      : ', dtn / dstar
  end if
  ! the fringe's stability bound on the step (bla.f:1418-1432)
  if (spat .and. fr%fmax > 0._dp) dtmax = min(dtmax, 1.75_dp / fr%fmax)
  dtn = min(dtn, dtmax)
  an = dtn * anrk(1)
  dtnp1 = dtn
  anp1 = an
  bnp1 = 0._dp
  call chk(simson_gpu_prhs(g, anp1))
  isave = 1
  it = 1
  cfl = 0._dp
  call system_clock(c0, crate)
  do
     ! x-y statistics at the start of every ixys-th step from txys on
     ! (bla.f:1830, 2089-2106), written every ixyss-th (bla.f:2116-2142)
     if (ixys > 0) then
        if (mod(it - 1 - nst, ixys) == 0 .and. t >= txys) then
           call chk(simson_gpu_xys(g, dtn, xs, zs, int(nxys, c_int), xys))
           sumw = sumw + dtn
        end if
        if (ixyss > 0) then
           if (mod(it - 1, ixyss) == 0 .and. t >= txys .and. it > 1) call write_xys()
        end if
     end if
     ! plane statistics at the start of every iamp-th step (nonlinbl.f:430)
     ! and the amp.stat line (bla.f:2326-2331)
     ! (nested: Fortran's .and. need not short-circuit, and mod(., 0) traps)
     if (iamp > 0) then
        if (mod(it - 1 - nst, iamp) == 0) then
           call chk(simson_gpu_amp(g, amp))
           call chk(simson_gpu_download(g, ur, ui, 0_c_int, c_null_ptr, px))
           call write_amp(t, px)
        end if
     end if
     docfl = .false.
     if (icfl > 0) docfl = mod(it - 1, icfl) == 0
     if (docfl) then
        call chk(simson_gpu_nonlin(g, 1_c_int, cflp))
        cfl = (cflp * pi + 2._dp * abs(rot)) * dtn   ! wcflbl.f:87
        write(*, '(a,f22.16,a,f22.16)') 'CFL*dt ', cfl, ' dt ', dtn / dstar
        if (cfl > 10._dp) then
           write(*,*) 'CFL*dt too high. Stop'
           call stopnow(324234)
        end if
     else
        call chk(simson_gpu_nonlin(g, 0_c_int, cflp))
     end if
     ! the next step's dt from the latest CFL number (bla.f:2355-2380)
     if (vart .and. mod(it, nst) == 0) then
        if (tmax > 0._dp) then
           tleft = min(tmax, tsave(isave)) - t - dtn
        else
           tleft = tsave(isave) - t - dtn
        end if
        if (cfl / cflmax > 10._dp / 9._dp) then
           dtnp1 = dtn * cflmax / cfl
        else
           dtnp1 = dtn * (0.8_dp + 0.2_dp * cflmax / cfl)
        end if
        dtnp1 = min(dtnp1, dtmax)
        if (.5_dp * dtnp1 < tleft .and. tleft < 4._dp * dtnp1) then
           if (tmax > 0._dp) then
              write(*,*) 'Adjusting time step to ', min(tmax, tsave(isave)) / dstar
           else
              write(*,*) 'Adjusting time step to ', tsave(isave) / dstar
           end if
           dtnp1 = tleft / aint(tleft / dtnp1 + 1._dp)
        end if
     end if
     an = anp1
     bn = bnp1
     anp1 = dtnp1 * anrk(mod(it, nst) + 1)
     bnp1 = dtnp1 * bnrk(mod(it, nst) + 1)
     ! time of the start of this substep for the wall forcing (bla.f:1797)
     call chk(simson_gpu_set_time(g, t + dtn * cnrk(mod(it - 1, nst) + 1)))
     if (fltype == 3 .and. cdev /= 0._dp) then
        ! the boundary layer's position x0 + cdev tc of this substep (bla.f:1826)
        call chk(simson_parallel_force(bf, tab, int(size(tab, 2), c_int), cdev, &
             t + dtn * cnrk(mod(it - 1, nst) + 1), mforce))
        call chk(simson_gpu_mean_force(g, mforce))
     end if
     call chk(simson_gpu_linear(g, an, bn, anp1, bnp1, c_null_ptr))
     it = it + 1
     if (mod(it - 1, nst) == 0) then
        t = t + dtn
        xs = xs - dtn * u0low                        ! bla.f:2606-2609
        zs = zs - dtn * w0low
        if (vart) dtn = dtnp1
        if (t + .5_dp * dtn > tsave(isave)) then
           call save_field(trim(nmsave(isave)))
           isave = isave + 1
        end if
        ! with statistics on, the stop criteria wait for a statistics
        ! step (bla.f:2789)
        chkstop = .true.
        if (ixys > 0) then
           if (t >= txys .and. mod(it - 1, ixys) /= 0) chkstop = .false.
        end if
        if (.not. chkstop) cycle
        inquire(file='stop.now', exist=stopf)
        if (stopf) write(*,*) 'STOP DUE TO stop.now FILE'
        if (maxit > 0 .and. it - 1 >= maxit) then
           write(*,*) 'STOP DUE TO ITERATIONS'
           stopf = .true.
        end if
        if (tmax > 0._dp .and. t + .5_dp * dtn > tmax) then
           write(*,*) 'STOP DUE TO INTEGRATION TIME'
           stopf = .true.
        end if
        if (stopf) exit
     end if
  end do
  call system_clock(c1)
  if (iamp > 0) close(15)
  if (ixys > 0) then
     if (t >= txys) call write_xys()                  ! bla.f:2907-2930
  end if
  call save_field(trim(namnut))
  write(*,*) 'Average time step    : ', real(c1 - c0, dp) / real(crate, dp) / &
       real((it - 1) / nst, dp)
  call simson_gpu_finalize(g)
  write(*,*) 'blagpu terminated normally.'

contains

  subroutine chk(code)
    integer(c_int), intent(in) :: code
    if (code /= 0) then
       write(*,*) 'simson_gpu call failed with status ', code
       call stopnow(int(code))
    end if
  end subroutine chk

  ! stopnow.f:23-32: print the code and stop
  subroutine stopnow(code)
    integer, intent(in) :: code
    write(*,*) 'stopnow: ', code
    stop 1
  end subroutine stopnow

  ! next line that is not a comment (comment.f)
  subroutine nxt(l)
    character(len=*), intent(out) :: l
    do
       read(10, '(a)') l
       if (len_trim(l) == 0) cycle
       if (l(1:1) /= '#') exit
    end do
  end subroutine nxt

  function word(l) result(w)
    character(len=*), intent(in) :: l
    character(len=256) :: w
    integer :: k
    w = adjustl(l)
    k = index(trim(w), ' ')
    if (k > 0) w = w(1:k-1)
  end function word

  ! wampbl.f, channel flow (fltype 1): y integrals of the plane sums and one
  ! line of amp.stat.  amp(y, i) holds boxamp.f's amp(y, i).
  subroutine write_amp(tt, pxx)
    real(dp), intent(in) :: tt, pxx
    real(dp) :: sm(20), e0, hplus, tau1, tau2, wb, delta(3), gy
    integer :: y, j
    amp(:, 8) = -amp(:, 8)
    amp(:, 13) = amp(:, 9)
    sm = 0._dp
    do j = 1, 10
       sm(j) = sum(amp(:, j) * wint)
    end do
    sm(13) = sum(amp(:, 13) * wint)
    delta = 0._dp
    do y = 1, nyp
       gy = (1._dp + cos(pi * real(y - 1, dp) / real(nyp - 1, dp))) / dstar
       if (delta(1) == 0._dp .and. amp(y, 13) <= .99_dp) then
          if (amp(y, 13) == .99_dp .or. y == 1) then
             delta(1) = gy
          else
             delta(1) = gy + (.99_dp - amp(y, 13)) * &
                  ((1._dp + cos(pi * real(y - 2, dp) / real(nyp - 1, dp))) / dstar - gy) / &
                  (amp(y - 1, 13) - amp(y, 13))
          end if
       end if
       delta(2) = delta(2) + (1._dp - amp(y, 13)) * wint(y)
       delta(3) = delta(3) + amp(y, 13) * (1._dp - amp(y, 13)) * wint(y)
    end do
    delta(2:3) = delta(2:3) / dstar
    do j = 1, 7
       sm(j) = sqrt(sm(j) * .5_dp)
    end do
    wb = sm(10) * .5_dp
    e0 = sqrt((sm(9) + sm(10)) * .5_dp)
    tau1 = amp(1, 12)
    tau2 = amp(nyp, 12)
    hplus = sqrt((abs(tau1) + abs(tau2)) * re / 2._dp)
    write(15, '(19(e25.16e3,1x))') tt / dstar, (sm(j), j = 1, 3), (sm(j) * dstar, j = 4, 6), &
         sm(7), sm(8), e0, hplus, pxx, sm(13) * re / 2._dp, tau1, tau2, (delta(j), j = 1, 3), wb
  end subroutine write_amp

  ! wxys.f:83-135: the velocity statistics divided by the sum of weights
  ! (no scalars, correlations or time series)
  subroutine write_xys()
    integer :: j
    write(*,*) '** Write statistics ** ', t / dstar, sumw / dstar
    open(19, file=trim(namxys), form='unformatted', status='replace')
    write(19) re, .false., xl, zl, t, 0._dp
    write(19) 'A', 0._dp, 0._dp, 0._dp, 0._dp
    write(19) nx, nyp, nzc, nfzsym
    write(19) fltype, dstar
    write(19) sumw, nxys, 0, 0
    do j = 1, nxys
       write(19) xys(:, :, j) * (1._dp / sumw)
    end do
    close(19)
  end subroutine write_xys

  ! rbla.f:24-125: fsc.dat, header (0, ymax, n, m, dstar2, nscalar) and n
  ! rows of eta, f, df, d2f, d3f, g, dg, d2g; m must match the flow's
  ! ibc = 100: the base flow's values at y = +1 (the first grid row), as
  ! x Fourier coefficients over the 3/2 grid divided by nxp, kx < nx/2
  ! (bflow.f:80-84 and 103-111, bu1jr(:, i, 1)), in place of the
  ! derivatives in bjr, bji
  subroutine top_values()
    integer :: i, j, kx, nxpl
    real(dp) :: ph
    nxpl = size(base, 1)
    bjr = 0._dp
    bji = 0._dp
    do i = 1, 3
       do kx = 0, nx / 2 - 1
          do j = 0, nxpl - 1
             ph = -2._dp * pi * real(mod(kx * j, nxpl), dp) / real(nxpl, dp)
             bjr(kx + 1, i) = bjr(kx + 1, i) + base(j + 1, 1, i) * cos(ph)
             bji(kx + 1, i) = bji(kx + 1, i) + base(j + 1, 1, i) * sin(ph)
          end do
          bjr(kx + 1, i) = bjr(kx + 1, i) / real(nxpl, dp)
          bji(kx + 1, i) = bji(kx + 1, i) / real(nxpl, dp)
       end do
    end do
  end subroutine top_values

  subroutine read_fsc()
    real(dp) :: a, b, m
    integer :: n, nsc, r
    open(17, file='fsc.dat', status='old', form='unformatted')
    read(17) a, b, n, m, bf%dstar2, nsc
    if (a /= 0._dp .or. nsc /= 0 .or. m /= rlam .or. n < 4) then
       write(*,*) '** Please provide suitable base flow in fsc.dat'
       call stopnow(123554)
    end if
    allocate(tab(8, n))
    do r = 1, n
       read(17) tab(:, r)
    end do
    close(17)
  end subroutine read_fsc

  ! wdiscbl.f:50-136 (no shift)
  subroutine save_field(name)
    character(len=*), intent(in) :: name
    integer :: c, z, y
    real(dp) :: pxo
    call chk(simson_gpu_download(g, ur, ui, 3_c_int, c_null_ptr, pxo))
    open(11, file=name, form='unformatted', status='replace')
    write(11) re, .false., xl, zl, t, 0._dp
    write(11) nx, nyp, nzc, nfzsym
    write(11) fltype, dstar
    if (fltype >= 4 .and. fltype <= 6) write(11) bstart, blength
    if (fltype >= 7 .and. fltype <= 9) write(11) bstart, blength, rlam, spanv
    do c = 1, 3
       do z = 1, nzc
          do y = 1, nyp
             urx(1:nx:2) = ur(:, y, z, c)
             urx(2:nx:2) = ui(:, y, z, c)
             write(11) urx
          end do
       end do
    end do
    close(11)
    write(*,*) '** Write velocity field ** ', trim(name), t / dstar
    if (do_pressure) call save_pressure(name // '.p')
  end subroutine save_field

  ! wdiscp.f: the header of the velocity file, then slot 8 (the pressure at
  ! physical y) as one record per (z, y) of nx/2 complex values
  subroutine save_pressure(name)
    character(len=*), intent(in) :: name
    integer :: z, y
    if (.not. allocated(prr)) allocate(prr(nx / 2, nyp, nzc), pri(nx / 2, nyp, nzc))
    call chk(simson_gpu_pressure(g, prr, pri))
    open(11, file=name, form='unformatted', status='replace')
    write(11) re, .false., xl, zl, t, 0._dp
    write(11) nx, nyp, nzc, nfzsym
    write(11) fltype, dstar
    if (fltype >= 4 .and. fltype <= 6) write(11) bstart, blength
    if (fltype >= 7 .and. fltype <= 9) write(11) bstart, blength, rlam, spanv
    do z = 1, nzc
       do y = 1, nyp
          urx(1:nx:2) = prr(:, y, z)
          urx(2:nx:2) = pri(:, y, z)
          write(11) urx
       end do
    end do
    close(11)
  end subroutine save_pressure

end program blagpu
