! littleharsh_gpu.f90 -- Fortran driver of the GPU substep.  Mirrors the
! reference main program (littleharsh.f90:63-207) with the time loop and the
! RK loop restored: it reads input.in in the layout `start` reads
! (start.f90:62-227, formats 10/20/21/30 of start.f90:434-437, so a real
! written without a decimal point keeps the implied-decimal semantics),
! reads geom_circ_rough.txt, sets up the device solver, starts from field
! files (flag_init = 1 new, 2 continue: read_in, start.f90:1974-2009) or from
! the parabolic profile of flag_init = 3 (start.f90:2109-2249), calls the hot
! path through the C ABI in the reference order, writes the field files every
! nwrite steps at the start of a step (record_out, FOU3D.f90:58-70, 130-142)
! and once more at the end (finalize, littleharsh.f90:353-390), samples the
! statistics and spectra every nstat = min(20, nwrite) steps (start.f90:372;
! FOU3D.f90:56-61, 107-110, 185-221) and writes them with each record
! (FOU3D.f90:223-226), and appends a record to the history file hist_*.dat
! with each field record (start.f90:245-259; FOU3D.f90:2504-2534).  With
! CFD_LASTPLANE_NP=np in the environment it runs the nonlinear term of a run
! of the reference on np ranks (the last-plane replay).  The
! reference counts the sampling interval from
! iter0, which it also resets when it samples; a counter of its own, iterst,
! does that here so iter0 keeps the start of the run for maxiter.
!
! After the reference's last record (dstz) input.in may carry five more
! records in the same layout, each optional (an absent record keeps the
! default, so the reference's own input.in reads unchanged):
!   ygrid =   0   y map: 0 the reference polynomial (y_grid_canopy.f90), 1 tanh
!   ybeta =  2.d0 stretching of the tanh map
!   extr24=   1   1: the 4 fixed forcing points of the linked 24pt geometry
!   precis=   0   0 FP64; 1 mixed (FP32 transforms, products and nonlinear
!                 scratch; FP64 fields, solves, projection, flow rate)
!   fused =   1   1: one cfd_step per time step (fused column sweeps);
!                 0: the reference's call sequence, one entry point per
!                 subroutine (littleharsh.f90:165-196)
! Ngalx2/Ngalz2 below Ngalx1/Ngalz1 (multi-band resolution) run with
! fused = 1; their statistics sum the central planes on the Ngalx2 x Ngalz2
! grid and their spectra cover the long columns (stats.f90, spectra.f90).
! The time loop's own steps (those without a field record) are timed and the
! rate printed as " loop: <steps> steps <seconds> s <substeps/s> substeps/s".
program littleharsh_gpu
  use iso_c_binding
  use cfd_amd_mod
  implicit none
  type(cfd_params) :: p
  type(c_ptr) :: h
  integer(c_int) :: rc, d(16), kRK, iter0, dst(6), sflags
  integer :: i, Ngalx2, Ngalz2, flag_init, nwrite, iter, iwrite, maxiter, npeakx, npeakz, dstx, dstz
  integer :: nstat, iterst, fused, ios, ival, nloop
  integer(8) :: clk0, clk1, clkrate
  real(8) :: tloop, rval
  character(len=6) :: key
  integer(c_int) :: N20(20)
  real(8) :: hv(11)
  character(len=4) :: e1, e2, e3
  character(len=200) :: hname
  logical :: hexist
  integer :: physlim_bot, physlim_top, nyu, nyv, nyp, ncol, jlo, nn, j
  real(8) :: Kib, maxerr, maxt, maxA, hny, Lfracx, Lfracz, sc(8), Qx
  character(len=90) :: fnameimb, dirin, dirout
  character(len=32) :: arg, envnp
  integer :: envst, lastnp
  real(8), allocatable :: yu(:), yv(:), du(:), dv(:), d2u(:), d2v(:)
  complex(8), allocatable :: u1(:,:), u2(:,:), u3(:,:), pp(:,:)

10 format(7X,D10.1)
20 format(7X,I10)
21 format(12X,I5)
30 format(7X,A90)

  ! argument: number of time steps, or "parse" to print the parameters read
  ! from input.in and stop
  maxiter = huge(1)
  arg = ''
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    if (trim(arg) /= 'parse') read(arg, *) maxiter
  end if

  open(40, file='input.in', form='formatted', status='old')
  do i = 1, 9
    read(40,10)
  end do
  read(40,10) p%Re
  read(40,10)
  read(40,10)
  read(40,10) p%alp
  read(40,10) p%bet
  read(40,20) p%ngalx
  read(40,20) Ngalx2
  read(40,20) p%ngalz
  read(40,20) Ngalz2
  read(40,20) p%ny
  read(40,10) p%dyq
  read(40,20) p%ppp
  read(40,10)
  read(40,10)
  read(40,20) flag_init
  read(40,10)
  read(40,10)
  read(40,10)
  read(40,30) fnameimb
  read(40,10)
  read(40,30) dirin
  read(40,30) dirout
  read(40,10)
  read(40,10)
  read(40,10) p%CFL
  read(40,20) nwrite
  read(40,21) p%flag_ctpress
  read(40,10)
  p%mpgx = 0d0
  if (p%flag_ctpress == 1) then
    read(40,10) p%mpgx
  else if (p%flag_ctpress == 0) then
    read(40,10)
  else
    write(*,*) 'Wrong presure flag'
    error stop 1
  end if
  read(40,10)
  read(40,10)
  read(40,10)
  read(40,10) Kib
  read(40,10) maxerr
  read(40,10) maxt
  read(40,10) maxA
  read(40,20) physlim_bot
  read(40,20) physlim_top
  read(40,10)
  read(40,10)
  read(40,20) p%geometry_type
  read(40,10)
  read(40,10)
  read(40,10)
  read(40,10)
  read(40,20) p%ntilex
  read(40,20) p%ntilez
  read(40,10) p%posth
  read(40,10) p%posth1
  read(40,10) hny
  read(40,20) npeakx
  read(40,20) npeakz
  read(40,10) Lfracx
  read(40,10) Lfracz
  read(40,20) p%dsty
  read(40,20) p%dsty1
  read(40,20) dstx
  read(40,20) dstz
  ! optional extension records (defaults: the reference's polynomial grid,
  ! the 24pt extras, FP64, fused stepping)
  p%ygrid = 0
  p%ytanh_beta = 2d0
  p%extras24 = 1
  p%precision = 0
  fused = 1
  do
    read(40,'(A6)', iostat=ios, advance='no') key
    if (ios > 0) exit
    if (is_iostat_end(ios)) exit
    backspace(40)
    select case (key)
    case ('ygrid ')
      read(40,20) ival
      p%ygrid = ival
    case ('ybeta ')
      read(40,10) rval
      p%ytanh_beta = rval
    case ('extr24')
      read(40,20) ival
      p%extras24 = ival
    case ('precis')
      read(40,20) ival
      p%precision = ival
    case ('fused ')
      read(40,20) ival
      fused = ival
    case default
      read(40,*, iostat=ios)
      if (ios /= 0) exit
    end select
  end do
  close(40)
  ! the central band's resolution (multi-band when coarser than band 1):
  ! stepped by cfd_step only
  p%ngalx2 = Ngalx2
  p%ngalz2 = Ngalz2
  if ((Ngalx2 /= p%ngalx .or. Ngalz2 /= p%ngalz) .and. fused /= 1) then
    write(*,*) 'a coarser central band (Ngalx2, Ngalz2) runs with fused = 1 only'
    error stop 1
  end if
  if (p%geometry_type == 3) then
    open(69, file='geom_circ_rough.txt', form='formatted', status='old')
    read(69,10) p%r
    read(69,10) p%c_i
    read(69,10) p%c_k
    read(69,10) p%stag
    close(69)
  else
    p%r = 1d0; p%c_i = 2d0; p%c_k = 2d0; p%stag = 0d0
  end if
  p%device = 0

  if (trim(arg) == 'parse') then
    write(*,'(a,9i8)') ' ints', p%ngalx, p%ngalz, p%ny, p%dsty, p%dsty1, p%ppp, p%geometry_type, p%ntilex, p%ntilez
    write(*,'(a,i4)') ' flag_ctpress', p%flag_ctpress
    write(*,'(a,7es24.16)') ' reals', p%Re, p%alp, p%bet, p%dyq, p%posth, p%CFL, p%mpgx
    write(*,'(a,4es24.16)') ' post', p%r, p%c_i, p%c_k, p%stag
    write(*,'(a,4i4,es24.16)') ' ext', p%ygrid, p%extras24, p%precision, fused, p%ytanh_beta
    write(*,'(a,2i8)') ' band2', p%ngalx2, p%ngalz2
    stop
  end if
  rc = cfd_create(p, h);               call cfd_check(rc, h, 'cfd_create')
  ! CFD_LASTPLANE_NP=np: the as-shipped nonlinear term of a run of the
  ! reference on np ranks (ops_in_planes, FOU3D.f90:1553-1613), to compare
  ! with an MPI run; unset: the per-plane term
  call get_environment_variable('CFD_LASTPLANE_NP', envnp, status=envst)
  if (envst == 0 .and. len_trim(envnp) > 0) then
    read(envnp, *) lastnp
    rc = cfd_set_compat_lastplane(h, lastnp); call cfd_check(rc, h, 'cfd_set_compat_lastplane')
    write(*,'(a,i6,a)') ' nonlinear term of a run on', lastnp, ' ranks (last-plane replay)'
  end if
  rc = cfd_dims(h, d);                 call cfd_check(rc, h, 'cfd_dims')
  ncol = d(4); nyu = d(5); nyv = d(6); nyp = d(7); jlo = d(8)
  allocate(yu(nyu), yv(nyv), du(nyu), dv(nyv), d2u(3*nyu), d2v(3*nyv))
  rc = cfd_get_ygrid(h, yu, yv, du, dv, d2u, d2v)
  call cfd_check(rc, h, 'cfd_get_ygrid')

  allocate(u1(nyu, ncol), u2(nyv, ncol), u3(nyu, ncol), pp(nyp, ncol))
  iter0 = 0
  if (flag_init == 1 .or. flag_init == 2) then
    sc = 0d0
    sc(CFD_MPGX) = p%mpgx
    rc = cfd_set_scalars(h, sc);       call cfd_check(rc, h, 'set scalars')
    rc = cfd_read_in(h, trim(dirin)//c_null_char, trim(fnameimb)//c_null_char, flag_init, iter0)
    call cfd_check(rc, h, 'read_in')
    rc = cfd_get_scalars(h, sc);       call cfd_check(rc, h, 'get scalars')
    Qx = sc(CFD_QXT)
  else if (flag_init == 3) then
    ! parabolic mean profile and constant mean of v, w, p
    ! (mblock_ini_parabolic_profile, start.f90:2109-2249)
    u1 = 0d0; u2 = 0d0; u3 = 0d0; pp = 0d0
    do j = 1, nyu
      u1(j, 1) = 0.5d0 * p%Re * p%mpgx * (yu(j)**2 - 1d0)
      u3(j, 1) = 1d-5
    end do
    u2(:, 1) = 1d-6
    pp(:, 1) = 1d0
    rc = cfd_set_field(h, CFD_U1, u1);   call cfd_check(rc, h, 'set u1')
    rc = cfd_set_field(h, CFD_U2, u2);   call cfd_check(rc, h, 'set u2')
    rc = cfd_set_field(h, CFD_U3, u3);   call cfd_check(rc, h, 'set u3')
    rc = cfd_set_field(h, CFD_P, pp);    call cfd_check(rc, h, 'set p')
    ! flowrateIm tip to tip (littleharsh.f90:411-428) -> QxT
    nn = p%ny - 2
    Qx = 0d0
    do j = 1, nn + 1
      Qx = Qx + (yv(j - jlo + 1) - yv(j - jlo)) * real(u1(j - jlo + 1, 1), 8)
    end do
    sc = 0d0
    sc(CFD_MPGX) = p%mpgx
    sc(CFD_QXT) = Qx
    sc(CFD_QX) = Qx
    rc = cfd_set_scalars(h, sc);         call cfd_check(rc, h, 'set scalars')
  else
    write(*,*) 'INITIAL CONDITIONS NOT IMPLEMENTED YET'
    error stop 1
  end if

  write(*,'(a,3i6,a,i8,a,es12.5)') ' grid', p%ngalx, p%ny, p%ngalz, '  columns', ncol, '  QxT', Qx
  ! history file (start.f90:245-259): appended to when it exists, else
  ! created with the header records; t in the header is the value the
  ! reference's t holds before any initial condition is read, 0
  rc = cfd_band_table(p, N20);         call cfd_check(rc, h, 'band table')
  write(e1,'(i4.4)') d(1)
  write(e2,'(i4.4)') d(2)
  write(e3,'(i4.4)') p%ny
  hname = trim(dirout)//'hist_'//e1//'x'//e2//'x'//e3//'.dat'
  inquire(file=trim(hname), exist=hexist)
  if (hexist) then
    open(30, file=trim(hname), form='unformatted', status='old', position='append', action='write')
  else
    open(30, file=trim(hname), form='unformatted', status='new', action='write')
    write(30) 0d0, p%Re, p%alp, p%bet, p%mpgx, 3
    write(30) N20
  end if
  iter = iter0
  iwrite = iter
  iterst = iter0
  nstat = min(20, nwrite)
  ! the conditional statistics need wall bands (band limits in order)
  rc = cfd_stats_dims(h, dst)
  if (rc /= 0) then
    write(*,*) 'statistics off: no wall bands in the band table'
    nstat = huge(1)
  end if
  rc = cfd_get_scalars(h, sc);         call cfd_check(rc, h, 'get scalars')
  call system_clock(count_rate=clkrate)
  nloop = 0
  tloop = 0d0
  do while (sc(CFD_T) < maxt .and. iter - iter0 < maxiter)
    sflags = 0
    if (iter - iterst >= nstat) then
      sflags = 1
      iterst = iter
    end if
    if (iter >= iwrite) then
      rc = cfd_record_out(h, trim(dirout)//c_null_char, iter)
      call cfd_check(rc, h, 'record_out')
      call write_hist()
      iwrite = iwrite + nwrite
      if (nstat /= huge(1)) sflags = sflags + 2
    end if
    if (sflags /= 0) then
      rc = cfd_stats_request(h, sflags, trim(dirout)//c_null_char, iter)
      call cfd_check(rc, h, 'stats request')
    end if
    ! the previous step ended with a synchronous read of the scalars, so the
    ! device is idle here
    call system_clock(clk0)
    if (fused == 1) then
      ! three fused substeps and meanflow_ctP (flag_ctpress = 1)
      rc = cfd_step(h, 1);             call cfd_check(rc, h, 'step')
    else
      do kRK = 1, 3
        rc = cfd_rhs0(h, kRK);         call cfd_check(rc, h, 'RHS0')
        rc = cfd_nonlinear(h, kRK);    call cfd_check(rc, h, 'nonlinear')
        rc = cfd_solve_u(h, 1, kRK);   call cfd_check(rc, h, 'solveU(u1)')
        rc = cfd_solve_v(h, kRK);      call cfd_check(rc, h, 'solveV(u2)')
        rc = cfd_solve_u(h, 3, kRK);   call cfd_check(rc, h, 'solveU(u3)')
        if (p%flag_ctpress == 0) then
          rc = cfd_meanflow_ctU(h, kRK); call cfd_check(rc, h, 'meanflow_ctU')
        end if
        rc = cfd_solve_p(h, kRK);      call cfd_check(rc, h, 'solveP')
        rc = cfd_vcorr(h, kRK);        call cfd_check(rc, h, 'v_corr')
      end do
      if (p%flag_ctpress == 1) then
        rc = cfd_meanflow_ctP(h);      call cfd_check(rc, h, 'meanflow_ctP')
      end if
    end if
    iter = iter + 1
    rc = cfd_get_scalars(h, sc);       call cfd_check(rc, h, 'get scalars')
    call system_clock(clk1)
    ! steps that wrote statistics are left out of the rate
    if (sflags == 0) then
      nloop = nloop + 1
      tloop = tloop + real(clk1 - clk0, 8) / real(clkrate, 8)
    end if
    if (mod(iter, max(nwrite, 1)) == 0 .or. iter - iter0 == maxiter) then
      write(*,'(a,i8,a,es14.7,a,es12.5,a,es14.7)') ' iter', iter, '  t', sc(CFD_T), '  dt', sc(CFD_DT), &
                                                   '  mpgx', sc(CFD_MPGX)
    end if
  end do
  if (nloop > 0 .and. tloop > 0d0) then
    write(*,'(a,i8,a,es12.5,a,es12.5,a)') ' loop:', nloop, ' steps', tloop, ' s', 3d0 * nloop / tloop, ' substeps/s'
  end if
  ! finalize: one more record of the final state
  rc = cfd_record_out(h, trim(dirout)//c_null_char, iter)
  call cfd_check(rc, h, 'record_out')
  call write_hist()
  close(30)
  rc = cfd_get_field(h, CFD_U1, u1);   call cfd_check(rc, h, 'get u1')
  write(*,'(a,es22.15)') ' u1(0,0) at the centreline ', real(u1(nyu/2 + 1, 1), 8)
  rc = cfd_destroy(h)

contains

  ! the history record that record_out appends (FOU3D.f90:2504-2534)
  subroutine write_hist()
    rc = cfd_hist(h, hv);              call cfd_check(rc, h, 'hist')
    if (p%flag_ctpress == 0) then
      write(30) p%flag_ctpress, iter, hv(1), hv(2), hv(3), hv(4), hv(5), hv(6), hv(7), hv(8), hv(9), hv(10), hv(11)
    else
      write(30) p%flag_ctpress, iter, hv(1), hv(2), hv(3), hv(4), hv(5), hv(6), hv(8), hv(10), hv(11)
    end if
    flush(30)
  end subroutine

end program littleharsh_gpu
