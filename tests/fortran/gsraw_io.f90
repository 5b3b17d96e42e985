! Test helper: list-directed I/O in the item order of the GS.raw eigen-sector list.
!   gsraw_io write <file>  writes two sectors with known contents
!   gsraw_io read  <file>  reads a file and prints, per state: npart_up npart_dn val n sum(v) sum(i*v(i))
program gsraw_io
  implicit none
  character(len=256) :: mode, fname
  call get_command_argument(1, mode)
  call get_command_argument(2, fname)
  if (trim(mode) == 'write') then
    call write_file(trim(fname))
  else
    call read_file(trim(fname))
  end if
contains
  subroutine write_file(fn)
    character(len=*), intent(in) :: fn
    integer :: u, i, s, k
    real(8) :: v(7)
    open(newunit=u, file=fn, status='replace', action='write')
    do s = 1, 2
      write(u,*) .true.
      write(u,*) 'updo'
      write(u,*) 4
      write(u,*) s + 1
      write(u,*) 4
      write(u,*) 2
      do k = 1, s
        do i = 1, 7
          v(i) = dble(i) / dble(s + k) - 0.125d0 * i * i
        end do
        write(u,*) .true.
        write(u,*) -1.5d0 * s - 0.25d0 * k
        write(u,*) .true.
        write(u,*) k
        write(u,*) 7
        write(u,*) v
      end do
      write(u,*) .false.
    end do
    write(u,*) .false.
    close(u)
  end subroutine
  subroutine read_file(fn)
    character(len=*), intent(in) :: fn
    integer :: u, i, no, nu, nd, rk, n
    logical :: more, conv
    character(len=16) :: kind
    real(8) :: val
    real(8), allocatable :: v(:)
    open(newunit=u, file=fn, status='old', action='read')
    do
      read(u,*) more
      if (.not. more) exit
      read(u,*) kind
      read(u,*) no
      read(u,*) nu
      read(u,*) no
      read(u,*) nd
      do
        read(u,*) more
        if (.not. more) exit
        read(u,*) val
        read(u,*) conv
        read(u,*) rk
        read(u,*) n
        allocate(v(n))
        read(u,*) v
        write(*,'(2I4,3ES26.17E3,I6)') nu, nd, val, sum(v), sum([(i * v(i), i = 1, n)]), n
        deallocate(v)
      end do
    end do
    close(u)
  end subroutine
end program
