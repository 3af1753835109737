! Writes Plot3D grid / solution files and microphone text with Fortran's own
! unformatted-sequential and F12.5 editing, from inputs in a stream file, so
! that the Python restatement of the output formats can be checked against a
! Fortran runtime byte for byte.  The record shapes are those of the
! reference's plot3d_output.f95:130-131,196-200,256-305 and mic_write.f95:
! 186-200 / mic_initialize.f95:38-61; the code here is written for this check.
!
! usage: plot3d_layout <input.bin> <output directory>
program plot3d_layout
  implicit none
  integer :: nb, i, a, b, c, k, nval, ts, ios
  integer, allocatable :: dims(:, :), blanks(:, :, :), codes(:, :, :)
  real(kind=8), allocatable :: start(:, :), dx(:, :), vals(:), fld(:, :, :, :)
  real, allocatable :: x(:, :, :), y(:, :, :), z(:, :, :)
  real(kind=8) :: hv(9)
  character(len=512) :: inp, dir

  call get_command_argument(1, inp)
  call get_command_argument(2, dir)
  open(unit=10, file=trim(inp), access='stream', form='unformatted', status='old', iostat=ios)
  if (ios /= 0) stop 2
  read(10) nb
  allocate(dims(3, nb), start(3, nb), dx(3, nb))
  read(10) dims, start, dx, hv, ts

  open(unit=17, file=trim(dir)//'/grid.x', status='replace', form='unformatted', action='write')
  open(unit=18, file=trim(dir)//'/data.q', status='replace', form='unformatted', action='write')
  write(17) nb
  write(18) nb
  write(17) (dims(1, i), dims(2, i), dims(3, i), i = 1, nb)
  write(18) (dims(1, i), dims(2, i), dims(3, i), i = 1, nb), 6, 0
  do i = 1, nb
    allocate(x(dims(1, i), dims(2, i), dims(3, i)), y(dims(1, i), dims(2, i), dims(3, i)))
    allocate(z(dims(1, i), dims(2, i), dims(3, i)), blanks(dims(1, i), dims(2, i), dims(3, i)))
    allocate(codes(dims(1, i), dims(2, i), dims(3, i)))
    allocate(fld(dims(1, i), dims(2, i), dims(3, i), 6))
    read(10) codes, fld
    do c = 1, dims(3, i)
      do b = 1, dims(2, i)
        do a = 1, dims(1, i)
          x(a, b, c) = start(1, i) + dx(1, i) * (a - 1)
          y(a, b, c) = start(2, i) + dx(2, i) * (b - 1)
          z(a, b, c) = start(3, i) + dx(3, i) * (c - 1)
          if (codes(a, b, c) >= 0) then
            blanks(a, b, c) = 1
          else if (codes(a, b, c) == -666) then
            blanks(a, b, c) = -1
          else
            blanks(a, b, c) = 0
          end if
        end do
      end do
    end do
    write(17) x, y, z, blanks
    write(18) hv(1), 0.0d0, hv(2), hv(3), hv(4), 0.0d0, hv(5), 0, hv(6), hv(7), hv(8), hv(9), &
      hv(9), hv(1), hv(3), ts
    write(18) (fld(:, :, :, k), k = 1, 6)
    deallocate(x, y, z, blanks, codes, fld)
  end do
  close(17)
  close(18)

  read(10) nval
  allocate(vals(nval))
  read(10) vals
  close(10)
  open(unit=19, file=trim(dir)//'/mic.txt', status='replace', form='formatted', action='write')
  write(19, *) trim('P,          ;Temp,       ;Density,    ;U-vel,      ;V-vel,      ;W-vel,      ;')
  write(19, *) trim('P,          ;Density,    ;')
  do i = 1, nval
    write(19, 100) vals(i)
  end do
  do i = 1, nval - 5, 6
    write(19, 105) vals(i), vals(i + 1), vals(i + 2), vals(i + 3), vals(i + 4), vals(i + 5)
  end do
  close(19)
100 format (F12.5)
105 format (F12.5, ' ', F12.5, ' ', F12.5, ' ', F12.5, ' ', F12.5, ' ', F12.5)
end program
