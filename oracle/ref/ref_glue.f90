! ref_glue.f90 -- project-written C bindings around the MPI-free reference
! routines (FFTPACK wrappers cft/rft, y_grid_canopy, boundary_circ_rough,
! proc_lims_planes).
! Built only into oracle/_ref/ by oracle/Makefile.ref, from the reference
! sources where they lie.  Test infrastructure: used to pin the oracle.
module ref_glue
  use iso_c_binding
  use declaration
  implicit none
contains

  subroutine ref_cft(n, c, ise, isa, m, iopt) bind(C, name="ref_cft")
    integer(c_int), value :: n, ise, isa, m, iopt
    real(c_double) :: c(*)
    real(8), allocatable :: aux(:)
    allocate(aux(10*n+19))
    aux = 0d0
    call cfti(n, aux)
    call cft(c, ise, isa, m, iopt, aux)
    deallocate(aux)
  end subroutine

  subroutine ref_rft(n, c, isa, m, iopt) bind(C, name="ref_rft")
    integer(c_int), value :: n, isa, m, iopt
    real(c_double) :: c(*)
    real(8), allocatable :: aux(:)
    allocate(aux(2*n+18))
    aux = 0d0
    call rfti(n, aux)
    call rft(c, isa, m, iopt, aux)
    deallocate(aux)
  end subroutine

  ! y_grid_canopy with the module state it reads set from the arguments.
  ! Outputs are copied into caller arrays of length nyu = ny+2*dsty+1 and
  ! nyv = ny+2*dsty; the d2 coefficient arrays are (3, .) column-major.
  subroutine ref_ygrid(nyin, dstyin, dsty1in, dyqin, pppin, posthin, posth1in, outdir, &
                       yuo, yvo, dthuo, dthvo, d2uo, d2vo) bind(C, name="ref_ygrid")
    integer(c_int), value :: nyin, dstyin, dsty1in, pppin
    real(c_double), value :: dyqin, posthin, posth1in
    character(kind=c_char), dimension(*) :: outdir
    real(c_double) :: yuo(*), yvo(*), dthuo(*), dthvo(*), d2uo(*), d2vo(*)
    integer :: i, nyu, nyv
    if (associated(N)) deallocate(N)
    if (associated(Ngal)) deallocate(Ngal)
    if (associated(Ny)) deallocate(Ny)
    if (associated(yu)) deallocate(yu, yv, dthdyu, dthdyv, dyu2i, dyv2i)
    if (associated(h_ny)) deallocate(h_ny)
    allocate(N(4,0:nband+1), Ngal(4,0:nband+1), Ny(3,0:nband+1), h_ny(1))
    N = 0
    Ngal = 0
    Ny = 0
    N(3,nband) = nyin - 2
    Ngal(3,nband) = nyin - 2
    dsty = dstyin
    dsty1 = dsty1in
    dyq = dyqin
    ppp = pppin
    posth = posthin
    posth1 = posth1in
    h_ny(1) = 0.1d0
    dirout = ''
    i = 1
    do while (outdir(i) /= c_null_char)
      dirout(i:i) = outdir(i)
      i = i + 1
    end do
    call y_grid_canopy
    nyu = nyin + 2*dstyin + 1
    nyv = nyin + 2*dstyin
    do i = 1, nyu
      yuo(i) = yu(-dstyin + i - 1)
      dthuo(i) = dthdyu(-dstyin + i - 1)
      d2uo(3*(i-1)+1) = dyu2i(1, -dstyin + i - 1)
      d2uo(3*(i-1)+2) = dyu2i(2, -dstyin + i - 1)
      d2uo(3*(i-1)+3) = dyu2i(3, -dstyin + i - 1)
    end do
    do i = 1, nyv
      yvo(i) = yv(-dstyin + i - 1)
      dthvo(i) = dthdyv(-dstyin + i - 1)
      d2vo(3*(i-1)+1) = dyv2i(1, -dstyin + i - 1)
      d2vo(3*(i-1)+2) = dyv2i(2, -dstyin + i - 1)
      d2vo(3*(i-1)+3) = dyv2i(3, -dstyin + i - 1)
    end do
  end subroutine

  ! boundary_circ_rough: reads geom_circ_rough.txt from the working directory
  ! and writes outdir//'boundary_'//ext1//'x'//ext2//'x'//ext3//'.dat'.
  subroutine ref_geom(ngx, ngz, nyin, dstyin, ntx, ntz, outdir) bind(C, name="ref_geom")
    integer(c_int), value :: ngx, ngz, nyin, dstyin, ntx, ntz
    character(kind=c_char), dimension(*) :: outdir
    integer :: i
    if (associated(Ngal)) deallocate(Ngal)
    allocate(Ngal(4,0:nband+1))
    Ngal = 0
    Ngal(1,1:nband) = ngx
    Ngal(2,1:nband) = ngz
    Ngal(3,3) = nyin - 2 + dstyin
    Ngal(4,3) = nyin - 1 + dstyin
    dsty = dstyin
    ntilex = ntx
    ntilez = ntz
    Lx = 1d0
    Ly = 2d0
    Lz = 1d0
    ext1 = '0000'
    ext2 = '0000'
    ext3 = '0000'
    dirout = ''
    i = 1
    do while (outdir(i) /= c_null_char)
      dirout(i:i) = outdir(i)
      i = i + 1
    end do
    call boundary_circ_rough
  end subroutine

  ! y_grid_canopy with h_ny(1) given: the band table N(4, 0:nband+1) it
  ! computes (rows 3 and 4; rows 1 and 2 come from start and stay 0 here),
  ! copied column-major into nout(20).
  subroutine ref_ygrid_bands(nyin, dstyin, dsty1in, dyqin, pppin, posthin, posth1in, hnyin, outdir, nout) &
      bind(C, name="ref_ygrid_bands")
    integer(c_int), value :: nyin, dstyin, dsty1in, pppin
    real(c_double), value :: dyqin, posthin, posth1in, hnyin
    character(kind=c_char), dimension(*) :: outdir
    integer(c_int) :: nout(*)
    integer :: i, b, r
    if (associated(N)) deallocate(N)
    if (associated(Ngal)) deallocate(Ngal)
    if (associated(Ny)) deallocate(Ny)
    if (associated(yu)) deallocate(yu, yv, dthdyu, dthdyv, dyu2i, dyv2i)
    if (associated(h_ny)) deallocate(h_ny)
    allocate(N(4,0:nband+1), Ngal(4,0:nband+1), Ny(3,0:nband+1), h_ny(1))
    N = 0
    Ngal = 0
    Ny = 0
    N(3,nband) = nyin - 2
    Ngal(3,nband) = nyin - 2
    dsty = dstyin
    dsty1 = dsty1in
    dyq = dyqin
    ppp = pppin
    posth = posthin
    posth1 = posth1in
    h_ny(1) = hnyin
    dirout = ''
    i = 1
    do while (outdir(i) /= c_null_char)
      dirout(i:i) = outdir(i)
      i = i + 1
    end do
    call y_grid_canopy
    do b = 0, nband + 1
      do r = 1, 4
        nout(4*b + r) = N(r, b)
      end do
    end do
  end subroutine

  ! proc_lims_planes (start.f90) with the module state it reads set from the
  ! arguments: the band table N(1:4,0:nband+1) (column-major, 20 values), one
  ! x/z resolution for every band (Ngal rows 3 and 4 are N's, as y_grid_canopy
  ! sets them), np ranks (np >= 3: the balance loop needs a central rank).
  ! Called for a rank other than 0, so it prints nothing.  Out: per rank the
  ! first and last planes of the u grid, then of the v grid (lim(4*np)).
  subroutine ref_plane_partition(n20, ngalx, ngalz, npin, lim) bind(C, name="ref_plane_partition")
    integer(c_int) :: n20(20)
    integer(c_int), value :: ngalx, ngalz, npin
    integer(c_int) :: lim(*)
    integer :: r, b
    if (associated(N)) deallocate(N)
    if (associated(Ngal)) deallocate(Ngal)
    if (associated(jlim)) deallocate(jlim)
    if (associated(procs)) deallocate(procs)
    if (associated(procs_b)) deallocate(procs_b)
    if (associated(procs_c)) deallocate(procs_c)
    if (associated(planelim)) deallocate(planelim)
    if (associated(limPL_incw)) deallocate(limPL_incw)
    if (associated(limPL_excw)) deallocate(limPL_excw)
    if (associated(limPL_FFT)) deallocate(limPL_FFT)
    if (associated(bandPL)) deallocate(bandPL)
    if (associated(bandPL_FFT)) deallocate(bandPL_FFT)
    allocate(N(4,0:nband+1), Ngal(4,0:nband+1), jlim(2,3,nband))
    N = reshape(n20, [4, nband+2])
    Ngal = 0
    do b = 0, nband+1
      Ngal(3,b) = N(3,b)
      Ngal(4,b) = N(4,b)
    end do
    do b = 1, nband
      Ngal(1,b) = ngalx
      Ngal(2,b) = ngalz
    end do
    ! only the FFT-plane limits (limPL_FFT) read these; they are not compared
    jlim = 0
    jlim(1,ugrid,2) = N(4,1)
    jlim(2,ugrid,2) = N(4,2)
    physlim_bot = jlim(1,ugrid,2) + npin
    physlim_top = jlim(2,ugrid,2) - npin
    np = npin
    call proc_lims_planes(1)
    do r = 0, np-1
      lim(1 + r) = planelim(ugrid,1,r)
      lim(1 + np + r) = planelim(ugrid,2,r)
      lim(1 + 2*np + r) = planelim(vgrid,1,r)
      lim(1 + 3*np + r) = planelim(vgrid,2,r)
    end do
  end subroutine

end module ref_glue
