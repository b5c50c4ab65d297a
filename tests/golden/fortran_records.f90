! Test helper: writes and reads the restart and analysis record layouts of
! restar.f90 with flang's own unformatted sequential I/O, so that the tests
! can compare the library's C writer and reader with the Fortran runtime
! byte for byte.  Arrays travel to and from Python as raw stream files.
!
!   fortran_records w ni nj nk state.bin restart.outdat   (store1 layout)
!   fortran_records r ni nj nk restart.outdat state.bin   (restar2 layout)
!   fortran_records o ni nj nk dump.bin structr.outdat    (output1 layout)
!
! state.bin: int32 myrank, nd, it0, ih0; real64 time; uu, u0, uri, uri0
! (each 9*ni*nj*nk, column-major (n,i,j,k)).
! dump.bin: real64 time; int32 myrank, npe; x1(ni), x2(nj), x3(nk),
! uri(9,ni,nj,nk), detg(ni,nj,nk).
program fortran_records
  implicit none
  character(len=512) :: mode, arg, fin, fout
  integer :: ni, nj, nk, i, j, k
  integer(4) :: myrank, nd, it0, ih0, npe
  real(8) :: time
  real(8), allocatable :: a(:,:,:,:,:), x1(:), x2(:), x3(:), w(:,:,:,:), g(:,:,:)

  call get_command_argument(1, mode)
  call get_command_argument(2, arg); read(arg, *) ni
  call get_command_argument(3, arg); read(arg, *) nj
  call get_command_argument(4, arg); read(arg, *) nk
  call get_command_argument(5, fin)
  call get_command_argument(6, fout)

  select case (trim(mode))
  case ('w')
    allocate(a(9, ni, nj, nk, 4))
    open(20, file=trim(fin), access='stream', form='unformatted', status='old')
    read(20) myrank, nd, it0, ih0, time, a
    close(20)
    open(9, file=trim(fout), form='unformatted', access='sequential', status='replace')
    write(9) myrank, nd
    write(9) it0, ih0
    write(9) time
    do k = 1, nk
      do j = 1, nj
        do i = 1, ni
          write(9) a(:, i, j, k, 1)
          write(9) a(:, i, j, k, 2)
          write(9) a(:, i, j, k, 3)
          write(9) a(:, i, j, k, 4)
        end do
      end do
    end do
    close(9)
  case ('r')
    allocate(a(9, ni, nj, nk, 4))
    open(9, file=trim(fin), form='unformatted', access='sequential', status='old')
    read(9) myrank, nd
    read(9) it0, ih0
    read(9) time
    do k = 1, nk
      do j = 1, nj
        do i = 1, ni
          read(9) a(:, i, j, k, 1)
          read(9) a(:, i, j, k, 2)
          read(9) a(:, i, j, k, 3)
          read(9) a(:, i, j, k, 4)
        end do
      end do
    end do
    close(9)
    open(20, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(20) myrank, nd, it0, ih0, time, a
    close(20)
  case ('o')
    allocate(x1(ni), x2(nj), x3(nk), w(9, ni, nj, nk), g(ni, nj, nk))
    open(20, file=trim(fin), access='stream', form='unformatted', status='old')
    read(20) time, myrank, npe, x1, x2, x3, w, g
    close(20)
    open(8, file=trim(fout), form='unformatted', access='sequential', status='replace')
    write(8) time, myrank, npe
    write(8) 1, ni, 1, nj, 1, nk
    do i = 1, ni
      do j = 1, nj
        do k = 1, nk
          write(8) x1(i), x2(j), x3(k), w(:, i, j, k), g(i, j, k)
        end do
      end do
    end do
    close(8)
  case default
    stop 2
  end select
end program fortran_records
