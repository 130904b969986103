!> A Fortran caller of module w90_library (the host library of this build)
!> that uses the entry points beside the wannierisation step: the input
!> reader, w90_set_constant_bohr_to_ang, w90_get_fortran_file and
!> w90_get_proj.
!>
!>   library_demo seedname [bohr_to_angstrom]
!>
!> Writes seedname.proj.txt: ierr and n, the lattice (rows), then per
!> projection l, m, s, rad, site(3), z(3), x(3), zona.
program library_demo
  use w90_library
  implicit none
  type(lib_common_type) :: lib
  integer, parameter :: cap = 64
  integer :: n, ierr, unit_out, unit_err, i, unit
  integer :: l(cap), m(cap), s(cap), rad(cap)
  real(8) :: site(3, cap), z(3, cap), x(3, cap), zona(cap), bohr
  character(len=256) :: arg

  call get_command_argument(1, arg)
  lib%seedname = trim(arg)
  call get_command_argument(2, arg)
  if (len_trim(arg) > 0) then
    read (arg, *) bohr
    call w90_set_constant_bohr_to_ang(lib, bohr)
  end if
  call w90_get_fortran_stdout(unit_out)
  call w90_get_fortran_stderr(unit_err)
  call w90_input_reader(lib, unit_out, unit_err, ierr)
  if (ierr /= 0) stop 1
  n = 0
  call w90_get_proj(lib, n, site, l, m, s, unit_out, unit_err, ierr, z=z, x=x, rad=rad, zona=zona)
  call w90_get_fortran_file(unit, trim(lib%seedname)//'.proj.txt')
  write (unit, *) ierr, n
  do i = 1, 3
    write (unit, '(3es26.17)') lib%real_lattice(i, :)
  end do
  do i = 1, n
    write (unit, '(4i5,10es26.17)') l(i), m(i), s(i), rad(i), site(:, i), z(:, i), x(:, i), zona(i)
  end do
  close (unit)
end program library_demo
