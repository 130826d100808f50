! cfd_amd_mod.f90 -- ISO_C_BINDING interface to libcfd_amd.so (include/cfd_amd.h).
! A Fortran driver of the reference's shape calls these in place of RHS0_u1/u2/u3,
! nonlinear, solveU, solveV, meanflow_ctU, solveP and v_corr
! (littleharsh.f90:163-189).  Nonzero status -> error stop via cfd_check.
module cfd_amd_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: CFD_U1 = 0, CFD_U2 = 1, CFD_U3 = 2, CFD_P = 3
  integer(c_int), parameter :: CFD_NU1 = 4, CFD_NU2 = 5, CFD_NU3 = 6
  integer(c_int), parameter :: CFD_DU1 = 7, CFD_DU2 = 8, CFD_DU3 = 9, CFD_PSI = 10, CFD_DIV = 11
  integer, parameter :: CFD_T = 1, CFD_DT = 2, CFD_DTI = 3, CFD_MPGX = 4, CFD_DGX = 5, CFD_QXT = 6, CFD_QX = 7

  type, bind(C) :: cfd_params
    integer(c_int) :: ngalx, ngalz, ny, dsty, dsty1, ppp, ygrid, geometry_type
    integer(c_int) :: ntilex, ntilez, extras24, flag_ctpress, precision, device
    integer(c_int) :: ngalx2, ngalz2
    real(c_double) :: Re, alp, bet, dyq, posth, posth1, ytanh_beta, CFL, mpgx
    real(c_double) :: r, c_i, c_k, stag
  end type cfd_params

  interface
    integer(c_int) function cfd_create(p, h) bind(C, name="cfd_create")
      import :: c_int, c_ptr, cfd_params
      type(cfd_params), intent(in) :: p
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function cfd_destroy(h) bind(C, name="cfd_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function cfd_last_error(h) bind(C, name="cfd_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cfd_dims(h, d) bind(C, name="cfd_dims")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: d(16)
    end function
    integer(c_int) function cfd_get_ygrid(h, yu, yv, du, dv, d2u, d2v) bind(C, name="cfd_get_ygrid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: yu(*), yv(*), du(*), dv(*), d2u(*), d2v(*)
    end function
    integer(c_int) function cfd_set_field(h, which, f) bind(C, name="cfd_set_field")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      integer(c_int), value :: which
      complex(c_double_complex), intent(in) :: f(*)
    end function
    integer(c_int) function cfd_get_field(h, which, f) bind(C, name="cfd_get_field")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: h
      integer(c_int), value :: which
      complex(c_double_complex) :: f(*)
    end function
    integer(c_int) function cfd_set_scalars(h, v) bind(C, name="cfd_set_scalars")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: v(8)
    end function
    integer(c_int) function cfd_get_scalars(h, v) bind(C, name="cfd_get_scalars")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: v(8)
    end function
    integer(c_int) function cfd_rhs0(h, kRK) bind(C, name="cfd_rhs0")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function cfd_nonlinear(h, kRK) bind(C, name="cfd_nonlinear")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function cfd_solve_u(h, comp, kRK) bind(C, name="cfd_solve_u")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: comp, kRK
    end function
    integer(c_int) function cfd_solve_v(h, kRK) bind(C, name="cfd_solve_v")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function cfd_meanflow_ctU(h, kRK) bind(C, name="cfd_meanflow_ctU")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function cfd_solve_p(h, kRK) bind(C, name="cfd_solve_p")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function cfd_vcorr(h, kRK) bind(C, name="cfd_vcorr")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function cfd_substep(h, kRK) bind(C, name="cfd_substep")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function cfd_step(h, n) bind(C, name="cfd_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: n
    end function
    integer(c_int) function cfd_sync(h) bind(C, name="cfd_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    ! meanflow_ctP (littleharsh.f90:504-520): flow rate Qx of the (0,0) column
    integer(c_int) function cfd_meanflow_ctP(h) bind(C, name="cfd_meanflow_ctP")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    ! the history record of record_out (FOU3D.f90:2504-2534): t, dtv, dtc,
    ! dt, err, Qx, QxT, mpgx, dgx, Umax, Uslip
    integer(c_int) function cfd_hist(h, v) bind(C, name="cfd_hist")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: v(11)
    end function
    ! the as-shipped last-plane nonlinear term of a run on np ranks
    ! (ops_in_planes, FOU3D.f90:1553-1613); np = 0: per plane
    integer(c_int) function cfd_set_compat_lastplane(h, np) bind(C, name="cfd_set_compat_lastplane")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: np
    end function
    ! the band table N(1:4, 0:4) (start.f90:89-111; y_grid_canopy.f90:250-275)
    integer(c_int) function cfd_band_table(p, N20) bind(C, name="cfd_band_table")
      import :: c_int, cfd_params
      type(cfd_params), intent(in) :: p
      integer(c_int) :: N20(20)
    end function
    ! record_out (FOU3D.f90:2172-2345): u1, u2, u3, p files in dirout
    integer(c_int) function cfd_record_out(h, dirout, iter) bind(C, name="cfd_record_out")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char) :: dirout(*)
      integer(c_int), value :: iter
    end function
    ! mblock_ini + read_in (start.f90:2061-2097, 2251-2700), flag_init 1 or 2
    integer(c_int) function cfd_read_in(h, dirin, fnameimb, flag_init, iter0) bind(C, name="cfd_read_in")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char) :: dirin(*), fnameimb(*)
      integer(c_int), value :: flag_init
      integer(c_int) :: iter0
    end function
    ! statistics and spectra (stats.f90, spectra.f90): one sample of the
    ! current state; write the stats_ / sp*_ files and zero the sums; serve
    ! flags (bit 0 sample, bit 1 write) in the next nonlinear term with
    ! kRK = 1 (FOU3D.f90:56-69, 107-110, 177-226); d(6) = dnx, dnz, wall-band
    ! planes of the u, v, p grids, istat
    integer(c_int) function cfd_stats(h) bind(C, name="cfd_stats")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function cfd_write_stats(h, dirout, iter) bind(C, name="cfd_write_stats")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char) :: dirout(*)
      integer(c_int), value :: iter
    end function
    integer(c_int) function cfd_stats_request(h, flags, dirout, iter) bind(C, name="cfd_stats_request")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      integer(c_int), value :: flags
      character(kind=c_char) :: dirout(*)
      integer(c_int), value :: iter
    end function
    integer(c_int) function cfd_stats_dims(h, d) bind(C, name="cfd_stats_dims")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: d(6)
    end function
  end interface

contains

  ! stop on a nonzero status with the library's message (replaces `stop`)
  subroutine cfd_check(rc, h, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: h
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=256) :: text
    integer :: i
    if (rc == 0) return
    text = ''
    call c_f_pointer(cfd_last_error(h), msg, [256])
    do i = 1, 256
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write(*,*) 'cfd_amd: ', what, ' failed: ', trim(text)
    error stop 1
  end subroutine

end module cfd_amd_mod
