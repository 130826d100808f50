! ff_read.f90 -- test program: reads a field file with the record structure
! of read_in (start.f90:2277-2300, 2354-2366): the header record, the band
! table N(4,0:nband+1), the grid record, then <nplanes> plane records
! j, id, nx, nz, y(j), buff(nx,nz); prints what it read.
program ff_read
  implicit none
  character(len=1024) :: fn, arg
  integer :: nband2, iter0, jin, id, nxin, nzin, nplanes, q, i, k
  real(8) :: t, Re2, alp2, bet2, dummRe, yj, cks
  integer :: dummint(88)
  integer, allocatable :: N2(:,:)
  real(8), allocatable :: buff(:,:)
  call get_command_argument(1, fn)
  call get_command_argument(2, arg)
  read(arg, *) nplanes
  open(10, file=trim(fn), form='unformatted', status='old')
  read(10) t, Re2, alp2, bet2, dummRe, nband2, iter0, dummint
  allocate(N2(4, 0:nband2+1))
  read(10) N2
  read(10)
  write(*, '(5ES25.17)') t, Re2, alp2, bet2, dummRe
  write(*, '(3I12)') nband2, iter0, sum(abs(dummint))
  write(*, '(20I8)') N2
  do q = 1, nplanes
    read(10) jin, id, nxin, nzin
    backspace(10)
    allocate(buff(nxin, nzin))
    read(10) jin, id, nxin, nzin, yj, buff
    cks = 0d0
    do k = 1, nzin
      do i = 1, nxin
        cks = cks + buff(i, k) * dble(i + 7 * k)
      end do
    end do
    write(*, '(4I8, 2ES25.17)') jin, id, nxin, nzin, yj, cks
    deallocate(buff)
  end do
  close(10)
end program
