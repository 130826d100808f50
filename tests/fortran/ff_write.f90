! ff_write.f90 -- test program: writes a field file with the write statements
! of record_out (FOU3D.f90:2231-2241): header, N, y/dthetavi/dthdy, then one
! record per plane j = j0..j1 with buff(i,k) = i + 1000 k + 1e6 (j - j0), all
! exact in binary, so a reader can check every value.
program ff_write
  implicit none
  character(len=1024) :: fn
  integer, parameter :: nband = 3, nx = 10, nz = 8, j0 = -2, j1 = 4
  integer :: N(4, 0:nband+1), dummint(88), j, i, k, iter
  real(8) :: t, Re, alp, bet, mpgx, dthetavi, yu(j0:j1), dthdyu(j0:j1), buffSR(nx, nz)
  call get_command_argument(1, fn)
  t = 12.5d0
  Re = 2800d0
  alp = 1d0
  bet = 2d0
  mpgx = -0.0625d0
  iter = 1234
  dthetavi = 0d0
  dummint = 0
  do k = 0, nband + 1
    do i = 1, 4
      N(i, k) = 10 * k + i
    end do
  end do
  do j = j0, j1
    yu(j) = 0.25d0 * j
    dthdyu(j) = 4d0 + j
  end do
  open(10, file=trim(fn), form='unformatted')
  write(10) t, Re, alp, bet, mpgx, nband, iter, dummint
  write(10) N
  write(10) yu, dthetavi, dthdyu
  do j = j0, j1
    do k = 1, nz
      do i = 1, nx
        buffSR(i, k) = dble(i) + 1000d0 * k + 1d6 * (j - j0)
      end do
    end do
    write(10) j, 1, nx, nz, yu(j), buffSR
  end do
  close(10)
end program
