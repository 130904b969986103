!> C interface of the wannierisation library with the names and signatures
!> of wannier90's src/c_interface.F90 (module w90_library_c), so that C and
!> C++ callers written against src/wannier90.hh link unchanged.  Strings
!> are F2018 assumed-length character dummies (CFI descriptors from C).
!> Added here: cproject_overlap (w90_project_overlap, which the reference
!> does not export to C) and cget_omega (Omega decomposition).
module w90_library_c
  use iso_c_binding
  use w90_library
  implicit none
  public
contains
  type(c_ptr) function w90_create() bind(c)
    type(lib_common_type), pointer :: common_data
    allocate (common_data)
    w90_create = c_loc(common_data)
  end function w90_create

  subroutine w90_delete(common_cptr) bind(c)
    type(c_ptr), value :: common_cptr
    type(lib_common_type), pointer :: common_fptr
    call c_f_pointer(common_cptr, common_fptr)
    deallocate (common_fptr)
  end subroutine w90_delete

  subroutine cdisentangle(common_cptr, ierr) bind(c)
    type(c_ptr), value :: common_cptr
    integer(kind=c_int) :: ierr
    type(lib_common_type), pointer :: common_fptr
    integer :: istdout, istderr, ie
    call w90_get_fortran_stderr(istderr)
    call w90_get_fortran_stdout(istdout)
    call c_f_pointer(common_cptr, common_fptr)
    call w90_disentangle(common_fptr, istdout, istderr, ie)
    ierr = ie
  end subroutine cdisentangle

  subroutine cproject_overlap(common_cptr, ierr) bind(c)
    type(c_ptr), value :: common_cptr
    integer(kind=c_int) :: ierr
    type(lib_common_type), pointer :: common_fptr
    integer :: istdout, istderr, ie
    call w90_get_fortran_stderr(istderr)
    call w90_get_fortran_stdout(istdout)
    call c_f_pointer(common_cptr, common_fptr)
    call w90_project_overlap(common_fptr, istdout, istderr, ie)
    ierr = ie
  end subroutine cproject_overlap

  subroutine cwannierise(common_cptr, ierr) bind(c)
    type(c_ptr), value :: common_cptr
    integer(kind=c_int) :: ierr
    type(lib_common_type), pointer :: common_fptr
    integer :: istdout, istderr, ie
    call w90_get_fortran_stderr(istderr)
    call w90_get_fortran_stdout(istdout)
    call c_f_pointer(common_cptr, common_fptr)
    call w90_wannierise(common_fptr, istdout, istderr, ie)
    ierr = ie
  end subroutine cwannierise

  subroutine cinput_setopt(common_cptr, seedname, ierr, comm) bind(c)
    type(c_ptr), value :: common_cptr
    character(*, kind=c_char) :: seedname
    integer(kind=c_int) :: ierr
    integer(kind=c_int), value :: comm
    type(lib_common_type), pointer :: common_fptr
    integer :: istdout, istderr, ie
    call w90_get_fortran_stderr(istderr)
    call w90_get_fortran_stdout(istdout)
    call c_f_pointer(common_cptr, common_fptr)
    call w90_input_setopt(common_fptr, seedname, int(comm), istdout, istderr, ie)
    ierr = ie
  end subroutine cinput_setopt

  subroutine cinput_reader(common_cptr, ierr) bind(c)
    type(c_ptr), value :: common_cptr
    integer(kind=c_int) :: ierr
    type(lib_common_type), pointer :: common_fptr
    integer :: istdout, istderr, ie
    call w90_get_fortran_stderr(istderr)
    call w90_get_fortran_stdout(istdout)
    call c_f_pointer(common_cptr, common_fptr)
    call w90_input_reader(common_fptr, istdout, istderr, ie)
    ierr = ie
  end subroutine cinput_reader

  subroutine cset_eigval(common_cptr, eigval_cptr) bind(c)
    type(c_ptr), value :: common_cptr, eigval_cptr
    real(8), pointer :: eigval_fptr(:, :)
    type(lib_common_type), pointer :: common_fptr
    call c_f_pointer(common_cptr, common_fptr)
    call c_f_pointer(eigval_cptr, eigval_fptr, [common_fptr%num_bands, common_fptr%num_kpts])
    call w90_set_eigval(common_fptr, eigval_fptr)
  end subroutine cset_eigval

  subroutine cset_m_local(common_cptr, m_cptr) bind(c)
    type(c_ptr), value :: common_cptr, m_cptr
    complex(8), pointer :: m_fptr(:, :, :, :)
    type(lib_common_type), pointer :: common_fptr
    integer :: istdout, istderr, ie
    call c_f_pointer(common_cptr, common_fptr)
    if (.not. common_fptr%setup_complete) then
      call w90_get_fortran_stderr(istderr)
      call w90_get_fortran_stdout(istdout)
      call w90_create_kmesh(common_fptr, istdout, istderr, ie)
    end if
    call c_f_pointer(m_cptr, m_fptr, [common_fptr%num_bands, common_fptr%num_bands, &
                                      common_fptr%nntot, common_fptr%num_kpts])
    call w90_set_m_local(common_fptr, m_fptr)
  end subroutine cset_m_local

  subroutine cset_u_matrix(common_cptr, a_cptr) bind(c)
    type(c_ptr), value :: common_cptr, a_cptr
    complex(8), pointer :: a_fptr(:, :, :)
    type(lib_common_type), pointer :: common_fptr
    call c_f_pointer(common_cptr, common_fptr)
    call c_f_pointer(a_cptr, a_fptr, [common_fptr%num_wann, common_fptr%num_wann, common_fptr%num_kpts])
    call w90_set_u_matrix(common_fptr, a_fptr)
  end subroutine cset_u_matrix

  subroutine cset_u_opt(common_cptr, a_cptr) bind(c)
    type(c_ptr), value :: common_cptr, a_cptr
    complex(8), pointer :: a_fptr(:, :, :)
    type(lib_common_type), pointer :: common_fptr
    call c_f_pointer(common_cptr, common_fptr)
    call c_f_pointer(a_cptr, a_fptr, [common_fptr%num_bands, common_fptr%num_wann, common_fptr%num_kpts])
    call w90_set_u_opt(common_fptr, a_fptr)
  end subroutine cset_u_opt

  subroutine cget_nn(common_cptr, n) bind(c)
    type(c_ptr), value :: common_cptr, n
    type(lib_common_type), pointer :: common_fptr
    integer(kind=c_int), pointer :: ndat
    integer :: istdout, istderr, ierr, nn
    call w90_get_fortran_stderr(istderr)
    call w90_get_fortran_stdout(istdout)
    call c_f_pointer(common_cptr, common_fptr)
    call c_f_pointer(n, ndat)
    call w90_get_nn(common_fptr, nn, istdout, istderr, ierr)
    ndat = nn
  end subroutine cget_nn

  subroutine cget_nnkp(common_cptr, nnkp) bind(c)
    type(c_ptr), value :: common_cptr, nnkp
    type(lib_common_type), pointer :: common_fptr
    integer(kind=c_int), pointer :: nfptr(:, :)
    integer :: istdout, istderr, ierr
    integer, allocatable :: tmp(:, :)
    call w90_get_fortran_stderr(istderr)
    call w90_get_fortran_stdout(istdout)
    call c_f_pointer(common_cptr, common_fptr)
    if (.not. common_fptr%setup_complete) call w90_create_kmesh(common_fptr, istdout, istderr, ierr)
    call c_f_pointer(nnkp, nfptr, [common_fptr%num_kpts, common_fptr%nntot])
    allocate (tmp(common_fptr%num_kpts, common_fptr%nntot))
    call w90_get_nnkp(common_fptr, tmp, istdout, istderr, ierr)
    nfptr = tmp
  end subroutine cget_nnkp

  subroutine cget_spreads(common_cptr, spreads) bind(c)
    type(c_ptr), value :: common_cptr, spreads
    type(lib_common_type), pointer :: common_fptr
    real(kind=8), pointer :: fspreads(:)
    call c_f_pointer(common_cptr, common_fptr)
    call c_f_pointer(spreads, fspreads, [common_fptr%num_wann])
    call w90_get_spreads(common_fptr, fspreads)
  end subroutine cget_spreads

  subroutine cget_centres(common_cptr, centres) bind(c)
    type(c_ptr), value :: common_cptr, centres
    type(lib_common_type), pointer :: common_fptr
    real(kind=8), pointer :: fcentres(:, :)
    call c_f_pointer(common_cptr, common_fptr)
    call c_f_pointer(centres, fcentres, [3, common_fptr%num_wann])
    call w90_get_centres(common_fptr, fcentres)
  end subroutine cget_centres

  !> omega(4) = Omega_I, Omega_D, Omega_OD, Omega (Ang^2)
  subroutine cget_omega(common_cptr, omega) bind(c)
    type(c_ptr), value :: common_cptr, omega
    type(lib_common_type), pointer :: common_fptr
    real(kind=8), pointer :: fomega(:)
    call c_f_pointer(common_cptr, common_fptr)
    call c_f_pointer(omega, fomega, [4])
    call w90_get_omega(common_fptr, fomega)
  end subroutine cget_omega

  subroutine cset_option_int(common_cptr, keyword, ival) bind(c)
    type(c_ptr), value :: common_cptr
    character(*, kind=c_char) :: keyword
    integer(kind=c_int), value :: ival
    type(lib_common_type), pointer :: common_fptr
    call c_f_pointer(common_cptr, common_fptr)
    call w90_set_option(common_fptr, keyword, int(ival))
  end subroutine cset_option_int

  subroutine cset_option_int3(common_cptr, keyword, ival) bind(c)
    type(c_ptr), value :: common_cptr
    type(c_ptr) :: ival
    character(*, kind=c_char) :: keyword
    integer(kind=c_int), pointer :: m_fptr(:)
    type(lib_common_type), pointer :: common_fptr
    call c_f_pointer(common_cptr, common_fptr)
    call c_f_pointer(ival, m_fptr, [3])
    call w90_set_option(common_fptr, keyword, int(m_fptr))
  end subroutine cset_option_int3

  subroutine cset_option_intx(common_cptr, keyword, ival, ict) bind(c)
    type(c_ptr), value :: common_cptr
    type(c_ptr), value :: ival
    character(*, kind=c_char) :: keyword
    integer(kind=c_int), value :: ict
    integer(kind=c_int), pointer :: m_fptr(:)
    type(lib_common_type), pointer :: common_fptr
    call c_f_pointer(common_cptr, common_fptr)
    call c_f_pointer(ival, m_fptr, [ict])
    call w90_set_option(common_fptr, keyword, int(m_fptr))
  end subroutine cset_option_intx

  subroutine cset_option_float(common_cptr, keyword, ival) bind(c)
    type(c_ptr), value :: common_cptr
    character(*, kind=c_char) :: keyword
    real(kind=c_double), value :: ival
    type(lib_common_type), pointer :: common_fptr
    call c_f_pointer(common_cptr, common_fptr)
    call w90_set_option(common_fptr, keyword, real(ival, 8))
  end subroutine cset_option_float

  subroutine cset_option_float33(common_cptr, keyword, ival) bind(c)
    type(c_ptr), value :: common_cptr
    character(*, kind=c_char) :: keyword
    type(c_ptr) :: ival
    type(lib_common_type), pointer :: common_fptr
    real(kind=8), pointer :: m_fptr(:, :)
    call c_f_pointer(ival, m_fptr, [3, 3])
    call c_f_pointer(common_cptr, common_fptr)
    call w90_set_option(common_fptr, keyword, m_fptr)
  end subroutine cset_option_float33

  subroutine cset_option_floatxy(common_cptr, keyword, ival, x, y) bind(c)
    type(c_ptr), value :: common_cptr
    character(*, kind=c_char) :: keyword
    type(c_ptr), value :: ival
    integer(kind=c_int), value :: x, y
    type(lib_common_type), pointer :: common_fptr
    real(kind=8), pointer :: m_fptr(:, :)
    call c_f_pointer(ival, m_fptr, [y, x])
    call c_f_pointer(common_cptr, common_fptr)
    call w90_set_option(common_fptr, keyword, m_fptr)
  end subroutine cset_option_floatxy

end module w90_library_c
