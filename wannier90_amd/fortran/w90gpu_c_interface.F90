!> C entry points of the host library.  The exported names and C
!> signatures are those of wannier90's C interface (src/c_interface.F90,
!> prototypes in src/wannier90.hh), declared for this build in
!> include/w90_library_c.h, so a C or C++ caller written against the
!> reference links unchanged.  Every entry point turns the opaque handle
!> into the library state (state_of), views caller memory through
!> c_f_pointer with the shape the state knows (the caller's array is
!> associated, never copied), and forwards to module w90_library.  The
!> steps (reading the input, projection, disentanglement, minimisation) go
!> through one dispatcher that fetches the Fortran output units.
!>
!> Strings (keywords, seedname) are F2018 assumed-length character dummies:
!> the C side passes CFI descriptors.  Added here: cproject_overlap and
!> cget_omega, which the reference does not export to C.
module w90_library_c
  use iso_c_binding
  use w90_library
  implicit none
  private

  public :: w90_create, w90_delete
  public :: cinput_setopt, cinput_reader
  public :: cset_eigval, cset_m_local, cset_u_matrix, cset_u_opt
  public :: cget_nn, cget_nnkp
  public :: cdisentangle, cproject_overlap, cwannierise
  public :: cget_centres, cget_spreads, cget_omega
  public :: cset_option_int, cset_option_int3, cset_option_intx
  public :: cset_option_float, cset_option_float33, cset_option_floatxy

  ! the steps of a calculation, run by do_step
  enum, bind(c)
    enumerator :: read_input = 1, project, disentangle, wannierise
  end enum

contains

  ! ------------------------------------------------------------- helpers

  !> The library state behind a handle made by w90_create.
  function state_of(handle) result(lib)
    type(c_ptr), intent(in) :: handle
    type(lib_common_type), pointer :: lib
    call c_f_pointer(handle, lib)
  end function state_of

  !> Run one step on the state behind `handle`; its status goes to ierr.
  subroutine do_step(handle, which, ierr)
    type(c_ptr), intent(in) :: handle
    integer, intent(in) :: which
    integer(c_int), intent(out) :: ierr
    type(lib_common_type), pointer :: lib
    integer :: unit_out, unit_err, status
    lib => state_of(handle)
    call w90_get_fortran_stdout(unit_out)
    call w90_get_fortran_stderr(unit_err)
    status = 0
    select case (which)
    case (read_input)
      call w90_input_reader(lib, unit_out, unit_err, status)
    case (project)
      call w90_project_overlap(lib, unit_out, unit_err, status)
    case (disentangle)
      call w90_disentangle(lib, unit_out, unit_err, status)
    case (wannierise)
      call w90_wannierise(lib, unit_out, unit_err, status)
    end select
    ierr = int(status, c_int)
  end subroutine do_step

  !> The k-mesh tables exist once the options are applied; the array
  !> setters and getters that need nntot build them on first use.
  subroutine need_kmesh(lib)
    type(lib_common_type), intent(inout) :: lib
    integer :: unit_out, unit_err, status
    if (lib%setup_complete) return
    call w90_get_fortran_stdout(unit_out)
    call w90_get_fortran_stderr(unit_err)
    call w90_create_kmesh(lib, unit_out, unit_err, status)
  end subroutine need_kmesh

  ! ----------------------------------------------------- state and steps

  type(c_ptr) function w90_create() bind(c)
    type(lib_common_type), pointer :: lib
    allocate (lib)
    w90_create = c_loc(lib)
  end function w90_create

  subroutine w90_delete(handle) bind(c)
    type(c_ptr), value :: handle
    type(lib_common_type), pointer :: lib
    lib => state_of(handle)
    if (associated(lib)) deallocate (lib)
  end subroutine w90_delete

  subroutine cinput_setopt(handle, seedname, ierr, comm) bind(c)
    type(c_ptr), value :: handle
    character(*, kind=c_char) :: seedname
    integer(c_int) :: ierr
    integer(c_int), value :: comm
    type(lib_common_type), pointer :: lib
    integer :: unit_out, unit_err, status
    lib => state_of(handle)
    call w90_get_fortran_stdout(unit_out)
    call w90_get_fortran_stderr(unit_err)
    call w90_input_setopt(lib, seedname, int(comm), unit_out, unit_err, status)
    ierr = int(status, c_int)
  end subroutine cinput_setopt

  subroutine cinput_reader(handle, ierr) bind(c)
    type(c_ptr), value :: handle
    integer(c_int) :: ierr
    call do_step(handle, read_input, ierr)
  end subroutine cinput_reader

  subroutine cproject_overlap(handle, ierr) bind(c)
    type(c_ptr), value :: handle
    integer(c_int) :: ierr
    call do_step(handle, project, ierr)
  end subroutine cproject_overlap

  subroutine cdisentangle(handle, ierr) bind(c)
    type(c_ptr), value :: handle
    integer(c_int) :: ierr
    call do_step(handle, disentangle, ierr)
  end subroutine cdisentangle

  subroutine cwannierise(handle, ierr) bind(c)
    type(c_ptr), value :: handle
    integer(c_int) :: ierr
    call do_step(handle, wannierise, ierr)
  end subroutine cwannierise

  ! ------------------------------------------------ caller-owned arrays

  !> eigval(num_bands, num_kpts)
  subroutine cset_eigval(handle, data) bind(c)
    type(c_ptr), value :: handle, data
    type(lib_common_type), pointer :: lib
    real(c_double), pointer :: eig(:, :)
    lib => state_of(handle)
    call c_f_pointer(data, eig, [lib%num_bands, lib%num_kpts])
    call w90_set_eigval(lib, eig)
  end subroutine cset_eigval

  !> m(num_bands, num_bands, nntot, num_kpts)
  subroutine cset_m_local(handle, data) bind(c)
    type(c_ptr), value :: handle, data
    type(lib_common_type), pointer :: lib
    complex(c_double_complex), pointer :: overlaps(:, :, :, :)
    lib => state_of(handle)
    call need_kmesh(lib)
    call c_f_pointer(data, overlaps, [lib%num_bands, lib%num_bands, lib%nntot, lib%num_kpts])
    call w90_set_m_local(lib, overlaps)
  end subroutine cset_m_local

  !> u(num_wann, num_wann, num_kpts)
  subroutine cset_u_matrix(handle, data) bind(c)
    type(c_ptr), value :: handle, data
    type(lib_common_type), pointer :: lib
    complex(c_double_complex), pointer :: rot(:, :, :)
    lib => state_of(handle)
    call c_f_pointer(data, rot, [lib%num_wann, lib%num_wann, lib%num_kpts])
    call w90_set_u_matrix(lib, rot)
  end subroutine cset_u_matrix

  !> u_opt(num_bands, num_wann, num_kpts)
  subroutine cset_u_opt(handle, data) bind(c)
    type(c_ptr), value :: handle, data
    type(lib_common_type), pointer :: lib
    complex(c_double_complex), pointer :: rot(:, :, :)
    lib => state_of(handle)
    call c_f_pointer(data, rot, [lib%num_bands, lib%num_wann, lib%num_kpts])
    call w90_set_u_opt(lib, rot)
  end subroutine cset_u_opt

  ! ------------------------------------------------------------- k-mesh

  subroutine cget_nn(handle, nntot) bind(c)
    type(c_ptr), value :: handle
    integer(c_int), intent(out) :: nntot
    type(lib_common_type), pointer :: lib
    integer :: unit_out, unit_err, status, n
    lib => state_of(handle)
    call w90_get_fortran_stdout(unit_out)
    call w90_get_fortran_stderr(unit_err)
    call w90_get_nn(lib, n, unit_out, unit_err, status)
    nntot = int(n, c_int)
  end subroutine cget_nn

  !> nnlist(num_kpts, nntot), 1-based
  subroutine cget_nnkp(handle, data) bind(c)
    type(c_ptr), value :: handle, data
    type(lib_common_type), pointer :: lib
    integer(c_int), pointer :: table(:, :)
    integer, allocatable :: nnlist(:, :)
    integer :: unit_out, unit_err, status
    lib => state_of(handle)
    call need_kmesh(lib)
    call w90_get_fortran_stdout(unit_out)
    call w90_get_fortran_stderr(unit_err)
    allocate (nnlist(lib%num_kpts, lib%nntot))
    call w90_get_nnkp(lib, nnlist, unit_out, unit_err, status)
    call c_f_pointer(data, table, [lib%num_kpts, lib%nntot])
    table = int(nnlist, c_int)
  end subroutine cget_nnkp

  ! ------------------------------------------------------------ results

  !> centres(3, num_wann), Ang
  subroutine cget_centres(handle, data) bind(c)
    type(c_ptr), value :: handle, data
    type(lib_common_type), pointer :: lib
    real(c_double), pointer :: xyz(:, :)
    lib => state_of(handle)
    call c_f_pointer(data, xyz, [3, lib%num_wann])
    call w90_get_centres(lib, xyz)
  end subroutine cget_centres

  !> spreads(num_wann), Ang^2
  subroutine cget_spreads(handle, data) bind(c)
    type(c_ptr), value :: handle, data
    type(lib_common_type), pointer :: lib
    real(c_double), pointer :: sprd(:)
    lib => state_of(handle)
    call c_f_pointer(data, sprd, [lib%num_wann])
    call w90_get_spreads(lib, sprd)
  end subroutine cget_spreads

  !> omega(4) = Omega_I, Omega_D, Omega_OD, Omega (Ang^2)
  subroutine cget_omega(handle, data) bind(c)
    type(c_ptr), value :: handle, data
    type(lib_common_type), pointer :: lib
    real(c_double), pointer :: parts(:)
    lib => state_of(handle)
    call c_f_pointer(data, parts, [4])
    call w90_get_omega(lib, parts)
  end subroutine cget_omega

  ! ------------------------------------------------------------ options

  subroutine cset_option_int(handle, keyword, value) bind(c)
    type(c_ptr), value :: handle
    character(*, kind=c_char) :: keyword
    integer(c_int), value :: value
    type(lib_common_type), pointer :: lib
    lib => state_of(handle)
    call w90_set_option(lib, keyword, int(value))
  end subroutine cset_option_int

  !> `where` holds the address of three ints (the caller passes int**)
  subroutine cset_option_int3(handle, keyword, where) bind(c)
    type(c_ptr), value :: handle
    character(*, kind=c_char) :: keyword
    type(c_ptr), intent(in) :: where
    integer(c_int), pointer :: triple(:)
    type(lib_common_type), pointer :: lib
    lib => state_of(handle)
    call c_f_pointer(where, triple, [3])
    call w90_set_option(lib, keyword, int(triple))
  end subroutine cset_option_int3

  subroutine cset_option_intx(handle, keyword, values, count) bind(c)
    type(c_ptr), value :: handle, values
    character(*, kind=c_char) :: keyword
    integer(c_int), value :: count
    integer(c_int), pointer :: list(:)
    type(lib_common_type), pointer :: lib
    lib => state_of(handle)
    call c_f_pointer(values, list, [count])
    call w90_set_option(lib, keyword, int(list))
  end subroutine cset_option_intx

  subroutine cset_option_float(handle, keyword, value) bind(c)
    type(c_ptr), value :: handle
    character(*, kind=c_char) :: keyword
    real(c_double), value :: value
    type(lib_common_type), pointer :: lib
    lib => state_of(handle)
    call w90_set_option(lib, keyword, real(value, kind(1.0d0)))
  end subroutine cset_option_float

  !> `where` holds the address of a C double[3][3] (row i = vector i),
  !> which Fortran sees column by column as (3,3)
  subroutine cset_option_float33(handle, keyword, where) bind(c)
    type(c_ptr), value :: handle
    character(*, kind=c_char) :: keyword
    type(c_ptr), intent(in) :: where
    real(c_double), pointer :: block(:, :)
    type(lib_common_type), pointer :: lib
    lib => state_of(handle)
    call c_f_pointer(where, block, [3, 3])
    call w90_set_option(lib, keyword, block)
  end subroutine cset_option_float33

  !> a C double[x][y]: y values per row, seen by Fortran as (y, x)
  subroutine cset_option_floatxy(handle, keyword, values, x, y) bind(c)
    type(c_ptr), value :: handle, values
    character(*, kind=c_char) :: keyword
    integer(c_int), value :: x, y
    real(c_double), pointer :: rows(:, :)
    type(lib_common_type), pointer :: lib
    lib => state_of(handle)
    call c_f_pointer(values, rows, [y, x])
    call w90_set_option(lib, keyword, rows)
  end subroutine cset_option_floatxy

end module w90_library_c
