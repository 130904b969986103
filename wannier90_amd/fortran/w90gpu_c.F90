!> ISO_C_BINDING interface to libw90gpu.so (include/w90gpu.h): the MI355X
!> library of the wannierisation step.  Complex arrays are passed as the
!> caller's complex(c_double_complex) Fortran arrays without copies.
module w90gpu_c
  use iso_c_binding
  implicit none
  public

  integer(c_int), parameter :: W90GPU_OK = 0
  integer(c_int), parameter :: W90GPU_ERR_FATAL = 1
  integer(c_int), parameter :: W90GPU_ERR_ALLOC = 2
  integer(c_int), parameter :: W90GPU_ERR_INPUT = 5
  integer(c_int), parameter :: W90GPU_ERR_FILE = 6
  integer(c_int), parameter :: W90GPU_NUM_NNMAX = 12
  integer(c_int), parameter :: W90GPU_PREC_FP64 = 0
  integer(c_int), parameter :: W90GPU_PREC_MIXED = 1
  integer(c_int), parameter :: W90GPU_DIS_MAX_BANDS = 256

  !> wann_control_type fields on this path (w90gpu_control)
  type, bind(c) :: w90gpu_control
    integer(c_int) :: num_iter
    integer(c_int) :: num_cg_steps
    integer(c_int) :: conv_window
    integer(c_int) :: lfixstep
    real(c_double) :: conv_tol
    real(c_double) :: trial_step
    real(c_double) :: fixed_step
    real(c_double) :: conv_noise_amp
    integer(c_int) :: precond
    integer(c_int) :: guiding_centres
    integer(c_int) :: selective_loc
    integer(c_int) :: precision
  end type w90gpu_control

  !> sizes and k-mesh tables of one problem (w90gpu_problem)
  type, bind(c) :: w90gpu_problem
    integer(c_int) :: num_wann
    integer(c_int) :: num_kpts
    integer(c_int) :: nntot
    type(c_ptr) :: nnlist
    type(c_ptr) :: bk
    type(c_ptr) :: wb
    real(c_double) :: wbtot
  end type w90gpu_problem

  !> spread decomposition and loop statistics (w90gpu_result)
  type, bind(c) :: w90gpu_result
    real(c_double) :: omega_total
    real(c_double) :: omega_invariant
    real(c_double) :: omega_d
    real(c_double) :: omega_od
    real(c_double) :: omega_tilde
    real(c_double) :: gcnorm1
    real(c_double) :: delta_omega
    real(c_double) :: solve_seconds
    integer(c_int) :: iterations
    integer(c_int) :: converged
    integer(c_int) :: unitarity_ok
    integer(c_int) :: status
    integer(c_int) :: bulk_iterations
    integer(c_int) :: held_second_steps
  end type w90gpu_result

  !> dis_control_type and the windows of dis_manifold_type (w90gpu_dis_control)
  type, bind(c) :: w90gpu_dis_control
    real(c_double) :: win_min
    real(c_double) :: win_max
    real(c_double) :: froz_min
    real(c_double) :: froz_max
    integer(c_int) :: frozen_states
    integer(c_int) :: num_iter
    real(c_double) :: mix_ratio
    real(c_double) :: conv_tol
    integer(c_int) :: conv_window
  end type w90gpu_dis_control

  !> Omega_I and loop statistics of a disentanglement (w90gpu_dis_result)
  type, bind(c) :: w90gpu_dis_result
    real(c_double) :: omega_invariant
    real(c_double) :: seconds
    real(c_double) :: sweeps_mean
    integer(c_int) :: iterations
    integer(c_int) :: converged
    integer(c_int) :: status
    integer(c_int) :: sweeps_max
  end type w90gpu_dis_result

  interface
    subroutine w90gpu_control_defaults(ctl) bind(c, name='w90gpu_control_defaults')
      import :: w90gpu_control
      type(w90gpu_control), intent(out) :: ctl
    end subroutine

    function w90gpu_create(device, ctx) bind(c, name='w90gpu_create') result(rc)
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function

    function w90gpu_destroy(ctx) bind(c, name='w90gpu_destroy') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function w90gpu_last_error(ctx) bind(c, name='w90gpu_last_error') result(msg)
      import :: c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr) :: msg
    end function

    function w90gpu_kmesh_get(real_lattice, num_kpts, kpt_latt, search_shells, tol, num_shells, &
                              shell_list, skip_b1_tests, nntot, nnlist, nncell, bk, wb, wbtot, &
                              errmsg, errlen) bind(c, name='w90gpu_kmesh_get') result(rc)
      import :: c_int, c_double, c_char
      real(c_double), intent(in) :: real_lattice(*)
      integer(c_int), value :: num_kpts
      real(c_double), intent(in) :: kpt_latt(*)
      integer(c_int), value :: search_shells
      real(c_double), value :: tol
      integer(c_int), value :: num_shells
      integer(c_int), intent(in) :: shell_list(*)
      integer(c_int), value :: skip_b1_tests
      integer(c_int), intent(out) :: nntot
      integer(c_int), intent(out) :: nnlist(*)
      integer(c_int), intent(out) :: nncell(*)
      real(c_double), intent(out) :: bk(*)
      real(c_double), intent(out) :: wb(*)
      real(c_double), intent(out) :: wbtot
      character(kind=c_char), intent(out) :: errmsg(*)
      integer(c_int), value :: errlen
      integer(c_int) :: rc
    end function

    function w90gpu_read_mmn(path, num_bands, num_kpts, nntot, nnlist, nncell, m_matrix, nthreads, &
                             errmsg, errlen) bind(c, name='w90gpu_read_mmn') result(rc)
      import :: c_int, c_double_complex, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: num_bands, num_kpts, nntot
      integer(c_int), intent(in) :: nnlist(*), nncell(*)
      complex(c_double_complex), intent(out) :: m_matrix(*)
      integer(c_int), value :: nthreads
      character(kind=c_char), intent(out) :: errmsg(*)
      integer(c_int), value :: errlen
      integer(c_int) :: rc
    end function

    function w90gpu_read_amn(path, num_bands, num_kpts, num_proj, a_matrix, nthreads, errmsg, errlen) &
      bind(c, name='w90gpu_read_amn') result(rc)
      import :: c_int, c_double_complex, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: num_bands, num_kpts, num_proj
      complex(c_double_complex), intent(out) :: a_matrix(*)
      integer(c_int), value :: nthreads
      character(kind=c_char), intent(out) :: errmsg(*)
      integer(c_int), value :: errlen
      integer(c_int) :: rc
    end function

    function w90gpu_wann_main(ctx, prob, ctl, m_matrix, u_matrix, centres, spreads, res) &
      bind(c, name='w90gpu_wann_main') result(rc)
      import :: c_int, c_ptr, c_double, c_double_complex, w90gpu_problem, w90gpu_control, w90gpu_result
      type(c_ptr), value :: ctx
      type(w90gpu_problem), intent(in) :: prob
      type(w90gpu_control), intent(in) :: ctl
      complex(c_double_complex), intent(inout) :: m_matrix(*)
      complex(c_double_complex), intent(inout) :: u_matrix(*)
      real(c_double), intent(out) :: centres(*)
      real(c_double), intent(out) :: spreads(*)
      type(w90gpu_result), intent(out) :: res
      integer(c_int) :: rc
    end function

    function w90gpu_project(ctx, prob, a_matrix, u_matrix, m_matrix) &
      bind(c, name='w90gpu_project') result(rc)
      import :: c_int, c_ptr, c_double_complex, w90gpu_problem
      type(c_ptr), value :: ctx
      type(w90gpu_problem), intent(in) :: prob
      complex(c_double_complex), intent(in) :: a_matrix(*)
      complex(c_double_complex), intent(out) :: u_matrix(*)
      complex(c_double_complex), intent(inout) :: m_matrix(*)
      integer(c_int) :: rc
    end function

    function w90gpu_session_create(ctx, prob, ctl, sess) bind(c, name='w90gpu_session_create') result(rc)
      import :: c_int, c_ptr, w90gpu_problem, w90gpu_control
      type(c_ptr), value :: ctx
      type(w90gpu_problem), intent(in) :: prob
      type(w90gpu_control), intent(in) :: ctl
      type(c_ptr), intent(out) :: sess
      integer(c_int) :: rc
    end function

    function w90gpu_session_destroy(sess) bind(c, name='w90gpu_session_destroy') result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: sess
      integer(c_int) :: rc
    end function

    function w90gpu_session_load(sess, m_matrix, u_matrix) bind(c, name='w90gpu_session_load') result(rc)
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: sess
      complex(c_double_complex), intent(in) :: m_matrix(*)
      complex(c_double_complex), intent(in) :: u_matrix(*)
      integer(c_int) :: rc
    end function

    function w90gpu_session_solve(sess, res) bind(c, name='w90gpu_session_solve') result(rc)
      import :: c_int, c_ptr, w90gpu_result
      type(c_ptr), value :: sess
      type(w90gpu_result), intent(out) :: res
      integer(c_int) :: rc
    end function

    function w90gpu_session_fetch(sess, m_matrix, u_matrix, centres, spreads) &
      bind(c, name='w90gpu_session_fetch') result(rc)
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: sess
      complex(c_double_complex), intent(out) :: m_matrix(*)
      complex(c_double_complex), intent(out) :: u_matrix(*)
      real(c_double), intent(out) :: centres(*)
      real(c_double), intent(out) :: spreads(*)
      integer(c_int) :: rc
    end function

    function w90gpu_session_history(sess, out, rows) bind(c, name='w90gpu_session_history') result(n)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sess
      real(c_double), intent(out) :: out(*)
      integer(c_int), value :: rows
      integer(c_int) :: n
    end function

    subroutine w90gpu_dis_control_defaults(ctl) bind(c, name='w90gpu_dis_control_defaults')
      import :: w90gpu_dis_control
      type(w90gpu_dis_control), intent(out) :: ctl
    end subroutine

    function w90gpu_disentangle(ctx, prob, num_bands, ctl, eigval, a_matrix, m_matrix_orig, u_matrix_opt, &
                                u_matrix, m_matrix, eigval_opt, nfirstwin, ndimwin, history, history_rows, res) &
      bind(c, name='w90gpu_disentangle') result(rc)
      import :: c_int, c_ptr, c_double, c_double_complex, w90gpu_problem, w90gpu_dis_control, w90gpu_dis_result
      type(c_ptr), value :: ctx
      type(w90gpu_problem), intent(in) :: prob
      integer(c_int), value :: num_bands
      type(w90gpu_dis_control), intent(in) :: ctl
      real(c_double), intent(in) :: eigval(*)
      complex(c_double_complex), intent(in) :: a_matrix(*)
      complex(c_double_complex), intent(in) :: m_matrix_orig(*)
      complex(c_double_complex), intent(out) :: u_matrix_opt(*)
      complex(c_double_complex), intent(out) :: u_matrix(*)
      complex(c_double_complex), intent(out) :: m_matrix(*)
      real(c_double), intent(out) :: eigval_opt(*)
      integer(c_int), intent(out) :: nfirstwin(*)
      integer(c_int), intent(out) :: ndimwin(*)
      real(c_double), intent(out) :: history(*)
      integer(c_int), value :: history_rows
      type(w90gpu_dis_result), intent(out) :: res
      integer(c_int) :: rc
    end function

    function c_strlen(s) bind(c, name='strlen') result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  !> Message of the last error of a context (or of context creation).
  function w90gpu_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = w90gpu_last_error(ctx)
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [max(n, 1)])
    allocate (character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function w90gpu_error_message

end module w90gpu_c
