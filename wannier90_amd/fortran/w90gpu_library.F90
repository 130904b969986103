!> Library interface of the wannierisation step, MI355X build.
!>
!> Keeps the procedure names, arguments and error behaviour of wannier90's
!> module w90_library (src/library_interface.F90) for the wannierisation
!> step: options through w90_set_option + w90_input_setopt or a .win file
!> through w90_input_reader, caller-owned M/U/A/eigenvalue arrays
!> associated by w90_set_*, k-mesh through w90_create_kmesh/w90_get_nn/
!> w90_get_nnkp, the Lowdin projection in w90_project_overlap, the
!> disentanglement in w90_disentangle and the minimisation in
!> w90_wannierise.  The arithmetic runs in libw90gpu.so through
!> ISO_C_BINDING (module w90gpu_c); nothing falls back to the CPU.
!> Status: ierr = 0 ok, > 0 error (codes of src/error.F90:27-36), < 0 warning.
module w90_library
  use iso_c_binding
  use w90gpu_c
  use w90gpu_io
  implicit none
  private

  public :: lib_common_type
  public :: w90_get_fortran_stdout, w90_get_fortran_stderr, w90_get_fortran_file
  public :: w90_set_constant_bohr_to_ang, w90_get_proj
  public :: w90_set_option, w90_input_setopt, w90_input_reader
  public :: w90_create_kmesh, w90_get_nn, w90_get_nnkp, w90_get_gkpb
  public :: w90_set_m_local, w90_set_u_matrix, w90_set_u_opt, w90_set_eigval
  public :: w90_project_overlap, w90_wannierise, w90_disentangle, w90_plot, w90_transport
  public :: w90_get_centres, w90_get_spreads, w90_get_omega, w90_write_chkpt, w90_read_chkpt
  public :: w90_write_nnkp, w90_print_conv

  type lib_common_type
    character(len=128) :: seedname = 'wannier'
    complex(dp), pointer :: a_matrix(:, :, :) => null()
    complex(dp), pointer :: m_matrix_local(:, :, :, :) => null()
    complex(dp), pointer :: u_matrix(:, :, :) => null()
    complex(dp), pointer :: u_opt(:, :, :) => null()
    real(dp), pointer :: eigval(:, :) => null()
    integer :: mp_grid(3) = 0
    integer :: num_bands = 0
    integer :: num_kpts = 0
    integer :: num_wann = 0
    integer :: num_proj = 0
    logical :: gamma_only = .false.
    logical :: have_disentangled = .false.
    logical :: setup_complete = .false.
    real(dp), allocatable :: kpt_latt(:, :)
    real(dp) :: real_lattice(3, 3) = 0.0_dp
    ! Angstrom per Bohr for inputs in Bohr (w90_set_constant_bohr_to_ang)
    real(dp) :: bohr = bohr_ang
    ! kmesh_info_type / kmesh_input_type
    integer :: nntot = 0
    integer(c_int), allocatable :: nnlist(:, :)
    integer(c_int), allocatable :: nncell(:, :, :)
    real(dp), allocatable :: bk(:, :, :)
    real(dp), allocatable :: wb(:)
    real(dp) :: wbtot = 0.0_dp
    integer :: search_shells = 36
    integer :: num_shells = 0
    integer :: shell_list(6) = 0
    real(dp) :: kmesh_tol = 1.0e-6_dp
    logical :: skip_b1_tests = .false.
    ! wann_control_type and print settings
    type(w90gpu_control) :: control
    logical :: control_set = .false.
    integer :: iprint = 1
    integer :: num_print_cycles = 1
    integer :: device = 0
    logical :: write_xyz = .false.
    logical :: write_u_matrices = .false.
    logical :: translate_home_cell = .false.
    character(len=16) :: restart = ''
    ! atoms (written to _centres.xyz)
    integer :: num_atoms = 0
    character(len=20), allocatable :: atom_symbol(:)
    real(dp), allocatable :: atom_cart(:, :)
    ! wannier_data / wann_omega_type
    real(dp), allocatable :: centres(:, :)
    real(dp), allocatable :: spreads(:)
    type(w90gpu_result) :: result
    real(dp), allocatable :: history(:, :)
    ! dis_control_type / dis_manifold_type, and what w90_disentangle leaves
    type(w90gpu_dis_control) :: dis_control
    logical :: dis_set = .false.
    type(w90gpu_dis_result) :: dis_result
    real(dp) :: omega_invariant = 0.0_dp
    integer, allocatable :: ndimwin(:), nfirstwin(:)
    logical, allocatable :: lwindow(:, :)
    real(dp), allocatable :: eigval_opt(:, :)
    real(dp), allocatable :: dis_history(:, :)
    ! options (set_option entries or .win data)
    type(w90_settings) :: settings
    logical :: have_win = .false.
    logical :: setopt_done = .false.  ! input_setopt applied the interface's options
    type(c_ptr) :: ctx = c_null_ptr
  end type lib_common_type

  interface w90_set_option
    module procedure w90_set_option_text
    module procedure w90_set_option_logical
    module procedure w90_set_option_int
    module procedure w90_set_option_i1d
    module procedure w90_set_option_i2d
    module procedure w90_set_option_real
    module procedure w90_set_option_r1d
    module procedure w90_set_option_r2d
  end interface w90_set_option

contains

  subroutine w90_get_fortran_stdout(istdout)
    use iso_fortran_env, only: output_unit
    integer, intent(out) :: istdout
    istdout = output_unit
  end subroutine w90_get_fortran_stdout

  subroutine w90_get_fortran_stderr(istderr)
    use iso_fortran_env, only: error_unit
    integer, intent(out) :: istderr
    istderr = error_unit
  end subroutine w90_get_fortran_stderr

  !> A new formatted unit on file `name` (w90_get_fortran_file,
  !> src/library_interface.F90:154): callers route the library's output to
  !> it in place of the standard units.
  subroutine w90_get_fortran_file(output, name)
    integer, intent(out) :: output
    character(len=*), intent(in) :: name
    open (newunit=output, file=name, form='formatted', status='unknown')
  end subroutine w90_get_fortran_file

  !> The length of the Bohr radius in Angstrom of the calling code
  !> (w90_set_constant_bohr_to_ang, src/library_interface.F90:655): used
  !> for blocks given in bohr by the options applied after this call.
  subroutine w90_set_constant_bohr_to_ang(common_data, bohr_to_angstrom)
    type(lib_common_type), intent(inout) :: common_data
    real(dp), intent(in) :: bohr_to_angstrom
    common_data%bohr = bohr_to_angstrom
  end subroutine w90_set_constant_bohr_to_ang

  !> report an error like prterr (src/io.F90:395-452): message on stderr
  subroutine report(istdout, istderr, code, msg, ierr)
    integer, intent(in) :: istdout, istderr, code
    character(len=*), intent(in) :: msg
    integer, intent(out) :: ierr
    ierr = code
    if (code > 0) then
      write (istderr, '(1x,a)') 'Exiting.......'
      write (istderr, '(1x,a)') trim(msg)
      write (istdout, '(1x,a)') 'Exiting.......'
      write (istdout, '(1x,a)') trim(msg)
    else if (code < 0) then
      write (istdout, '(1x,a)') 'Warning: '//trim(msg)
    end if
  end subroutine report

  ! ------------------------------------------------------------ options

  subroutine w90_set_option_text(common_data, keyword, text)
    type(lib_common_type), intent(inout) :: common_data
    character(*), intent(in) :: keyword, text
    type(w90_setting) :: e
    e%keyword = keyword
    e%kind = K_TEXT
    e%text = text
    call settings_add(common_data%settings, e)
  end subroutine w90_set_option_text

  subroutine w90_set_option_logical(common_data, keyword, bool)
    type(lib_common_type), intent(inout) :: common_data
    character(*), intent(in) :: keyword
    logical, intent(in) :: bool
    type(w90_setting) :: e
    e%keyword = keyword
    e%kind = K_LOGICAL
    e%lval = bool
    call settings_add(common_data%settings, e)
  end subroutine w90_set_option_logical

  subroutine w90_set_option_int(common_data, keyword, ival)
    type(lib_common_type), intent(inout) :: common_data
    character(*), intent(in) :: keyword
    integer, intent(in) :: ival
    type(w90_setting) :: e
    e%keyword = keyword
    e%kind = K_INT
    e%ival = ival
    call settings_add(common_data%settings, e)
  end subroutine w90_set_option_int

  subroutine w90_set_option_i1d(common_data, keyword, arr)
    type(lib_common_type), intent(inout) :: common_data
    character(*), intent(in) :: keyword
    integer, intent(in) :: arr(:)
    type(w90_setting) :: e
    e%keyword = keyword
    e%kind = K_I1D
    e%i1d = arr
    call settings_add(common_data%settings, e)
  end subroutine w90_set_option_i1d

  subroutine w90_set_option_i2d(common_data, keyword, arr)
    type(lib_common_type), intent(inout) :: common_data
    character(*), intent(in) :: keyword
    integer, intent(in) :: arr(:, :)
    type(w90_setting) :: e
    e%keyword = keyword
    e%kind = K_I2D
    e%i2d = arr
    call settings_add(common_data%settings, e)
  end subroutine w90_set_option_i2d

  subroutine w90_set_option_real(common_data, keyword, rval)
    type(lib_common_type), intent(inout) :: common_data
    character(*), intent(in) :: keyword
    real(dp), intent(in) :: rval
    type(w90_setting) :: e
    e%keyword = keyword
    e%kind = K_REAL
    e%rval = rval
    call settings_add(common_data%settings, e)
  end subroutine w90_set_option_real

  subroutine w90_set_option_r1d(common_data, keyword, arr)
    type(lib_common_type), intent(inout) :: common_data
    character(*), intent(in) :: keyword
    real(dp), intent(in) :: arr(:)
    type(w90_setting) :: e
    e%keyword = keyword
    e%kind = K_R1D
    e%r1d = arr
    call settings_add(common_data%settings, e)
  end subroutine w90_set_option_r1d

  !> blocks (unit_cell_cart, kpoints) are given as r(columns, rows), e.g.
  !> kpoints(3, num_kpts), as the reference stores them
  subroutine w90_set_option_r2d(common_data, keyword, arr)
    type(lib_common_type), intent(inout) :: common_data
    character(*), intent(in) :: keyword
    real(dp), intent(in) :: arr(:, :)
    type(w90_setting) :: e
    e%keyword = keyword
    e%kind = K_R2D
    e%r2d = arr
    call settings_add(common_data%settings, e)
  end subroutine w90_set_option_r2d

  !> Apply the option entries (w90_input_setopt, src/library_interface.F90:161)
  subroutine w90_input_setopt(common_data, seedname, comm, istdout, istderr, ierr)
    type(lib_common_type), intent(inout) :: common_data
    character(len=*), intent(in) :: seedname
    integer, intent(in) :: comm
    integer, intent(in) :: istdout, istderr
    integer, intent(out) :: ierr
    character(len=256) :: msg
    ierr = 0
    if (common_data%have_win) then
      call report(istdout, istderr, 1, ' readinput and setopt clash at input_setopt call', ierr)
      return
    end if
    if (common_data%settings%n == 0) then
      call report(istdout, istderr, 1, ' input_setopt called with no input', ierr)
      return
    end if
    common_data%seedname = seedname
    call apply_settings(common_data, ierr, msg)
    if (ierr > 0) then
      call report(istdout, istderr, ierr, msg, ierr)
      return
    end if
    common_data%setopt_done = .true.
    if (comm < -huge(comm)) ierr = ierr  ! single device: the communicator is not used
  end subroutine w90_input_setopt

  !> Read seedname.win (w90_input_reader, src/library_interface.F90:294).
  !> After input_setopt (which sets the seedname) the options given through
  !> the interface stay and take precedence: the first occurrence of a
  !> keyword is the one used, and the .win's entries follow them.
  subroutine w90_input_reader(common_data, istdout, istderr, ierr)
    type(lib_common_type), intent(inout) :: common_data
    integer, intent(in) :: istdout, istderr
    integer, intent(out) :: ierr
    character(len=256) :: msg
    ierr = 0
    if (common_data%settings%n > 0 .and. .not. common_data%have_win .and. .not. common_data%setopt_done) then
      call report(istdout, istderr, 1, &
                  'Error: input reader called when unspent options present (setopt must be called first)', ierr)
      return
    end if
    if (.not. common_data%setopt_done) call settings_clear(common_data%settings)
    call read_win(trim(common_data%seedname)//'.win', common_data%settings, ierr, msg)
    if (ierr > 0) then
      call report(istdout, istderr, ierr, msg, ierr)
      return
    end if
    common_data%have_win = .true.
    call apply_settings(common_data, ierr, msg)
    if (ierr > 0) call report(istdout, istderr, ierr, msg, ierr)
  end subroutine w90_input_reader

  !> Keywords of the wannierise path (src/wannier90_readwrite.F90:157-384,
  !> :628-804, src/readwrite.F90 lattice/k-point/atom blocks).  Keywords of
  !> the branches this build does not carry are rejected with an input error.
  subroutine apply_settings(cd, ierr, msg)
    type(lib_common_type), intent(inout) :: cd
    integer, intent(out) :: ierr
    character(len=*), intent(out) :: msg
    logical :: found, lval, found_trial
    integer :: nk, i, n, ios, nsh
    real(dp) :: rval, uc(3, 3), pos(3)
    character(len=64) :: text
    character(len=512) :: line
    integer :: sl(6)
    ierr = 0
    msg = ''
    call w90gpu_control_defaults(cd%control)
    cd%control_set = .true.
    call get_int(cd%settings, 'num_wann', cd%num_wann, found, ierr)
    if (.not. found .or. cd%num_wann <= 0) then
      ierr = 5
      msg = 'Error: num_wann must be provided and positive'
      return
    end if
    cd%num_bands = cd%num_wann
    call get_int(cd%settings, 'num_bands', cd%num_bands, found, ierr)
    if (cd%num_bands < cd%num_wann) then
      ierr = 5
      msg = 'Error: num_bands must be greater than or equal to num_wann'
      return
    else if (cd%num_bands > W90GPU_DIS_MAX_BANDS .and. cd%num_bands > cd%num_wann) then
      ierr = 5
      write (msg, '(a,i0,a)') 'Error: disentanglement in this build takes at most ', W90GPU_DIS_MAX_BANDS, &
        ' bands (num_bands)'
      return
    end if
    call read_dis_settings(cd, ierr, msg)
    if (ierr /= 0) return
    call get_i1d(cd%settings, 'mp_grid', cd%mp_grid, 3, found, ierr)
    if (.not. found .or. ierr /= 0) then
      ierr = 5
      msg = 'Error: Did not find the kpoint grid (mp_grid) in input file'
      return
    end if
    call get_r2d(cd%settings, 'unit_cell_cart', uc, 3, 3, .true., found, ierr, cd%bohr)
    if (.not. found .or. ierr /= 0) then
      ierr = 5
      msg = 'Error: Did not find the cell information (unit_cell_cart) in the input file'
      return
    end if
    cd%real_lattice = transpose(uc)  ! real_lattice(i,:) is lattice vector i
    nk = product(cd%mp_grid)
    call get_int(cd%settings, 'num_kpts', nk, found, ierr)
    cd%num_kpts = nk
    if (allocated(cd%kpt_latt)) deallocate (cd%kpt_latt)
    allocate (cd%kpt_latt(3, nk))
    call get_r2d(cd%settings, 'kpoints', cd%kpt_latt, 3, nk, .false., found, ierr)
    if (.not. found .or. ierr /= 0) then
      ierr = 5
      msg = 'Error: Did not find the kpoint information in the input file'
      return
    end if
    ! wann_control (defaults of src/wannier90_types.F90:170-192)
    call get_int(cd%settings, 'num_iter', cd%control%num_iter, found, ierr)
    call get_int(cd%settings, 'num_cg_steps', cd%control%num_cg_steps, found, ierr)
    call get_int(cd%settings, 'conv_window', cd%control%conv_window, found, ierr)
    call get_real(cd%settings, 'conv_tol', cd%control%conv_tol, found, ierr)
    call get_real(cd%settings, 'trial_step', cd%control%trial_step, found_trial, ierr)
    rval = -999.0_dp
    call get_real(cd%settings, 'fixed_step', rval, found, ierr)
    if (found) then
      if (found_trial) then
        ierr = 5
        msg = 'Error: cannot specify both fixed_step and trial_step'
        return
      end if
      cd%control%fixed_step = rval
      cd%control%lfixstep = 1
      if (rval <= 0.0_dp) then
        ierr = 5
        msg = 'Error: fixed_step must be greater than zero'
        return
      end if
    end if
    rval = -1.0_dp
    call get_real(cd%settings, 'conv_noise_amp', rval, found, ierr)
    cd%control%conv_noise_amp = rval
    if (ierr /= 0) then
      msg = 'Error: malformed wannierise keyword'
      return
    end if
    ! branches of wann_main outside this path
    do i = 1, 5
      lval = .false.
      select case (i)
      case (1)
        text = 'precond'
      case (2)
        text = 'guiding_centres'
      case (3)
        text = 'slwf_constrain'
      case (4)
        text = 'gamma_only'
      case (5)
        text = 'site_symmetry'
      end select
      call get_logical(cd%settings, trim(text), lval, found, ierr)
      if (lval) then
        ierr = 5
        msg = 'Error: '//trim(text)//' is not supported by this build (wannierise path without it)'
        return
      end if
    end do
    n = 0
    call get_int(cd%settings, 'slwf_num', n, found, ierr)
    if (found .and. n /= cd%num_wann) then
      ierr = 5
      msg = 'Error: selective localisation (slwf_num) is not supported by this build'
      return
    end if
    if (settings_find(cd%settings, 'exclude_bands') > 0) then
      ierr = 5
      msg = 'Error: exclude_bands is not supported by this build'
      return
    end if
    if (cd%control%conv_noise_amp > 0.0_dp) then
      ierr = 5
      msg = 'Error: conv_noise_amp > 0 is not supported by this build'
      return
    end if
    ! print and output options
    call get_int(cd%settings, 'iprint', cd%iprint, found, ierr)
    call get_int(cd%settings, 'num_print_cycles', cd%num_print_cycles, found, ierr)
    if (cd%num_print_cycles <= 0) cd%num_print_cycles = 1
    call get_logical(cd%settings, 'write_xyz', cd%write_xyz, found, ierr)
    call get_logical(cd%settings, 'write_u_matrices', cd%write_u_matrices, found, ierr)
    call get_logical(cd%settings, 'translate_home_cell', cd%translate_home_cell, found, ierr)
    call get_text(cd%settings, 'restart', cd%restart, found)
    ! k-mesh input
    call get_int(cd%settings, 'search_shells', cd%search_shells, found, ierr)
    call get_real(cd%settings, 'kmesh_tol', cd%kmesh_tol, found, ierr)
    call get_logical(cd%settings, 'skip_b1_tests', cd%skip_b1_tests, found, ierr)
    sl = 0
    i = settings_find(cd%settings, 'shell_list')
    if (i > 0) then
      if (cd%settings%e(i)%kind == K_I1D) then
        nsh = min(6, size(cd%settings%e(i)%i1d))
        sl(1:nsh) = cd%settings%e(i)%i1d(1:nsh)
      else
        nsh = 0
        do n = 6, 1, -1
          read (cd%settings%e(i)%text, *, iostat=ios) sl(1:n)
          if (ios == 0) then
            nsh = n
            exit
          end if
        end do
      end if
      cd%num_shells = nsh
      cd%shell_list = sl
    end if
    ! MI355X-specific options
    call get_int(cd%settings, 'gpu_device', cd%device, found, ierr)
    text = 'fp64'
    call get_text(cd%settings, 'precision', text, found)
    if (trim(text) == 'mixed') then
      cd%control%precision = W90GPU_PREC_MIXED
    else
      cd%control%precision = W90GPU_PREC_FP64
    end if
    ! atoms for _centres.xyz
    cd%num_atoms = 0
    if (allocated(cd%atom_symbol)) deallocate (cd%atom_symbol)
    if (allocated(cd%atom_cart)) deallocate (cd%atom_cart)
    i = settings_find(cd%settings, 'atoms_frac')
    if (i == 0) i = settings_find(cd%settings, 'atoms_cart')
    if (i > 0) then
      if (cd%settings%e(i)%kind == K_BLOCK) then
        n = size(cd%settings%e(i)%lines)
        allocate (cd%atom_symbol(n), cd%atom_cart(3, n))
        rval = 1.0_dp
        do nk = 1, n
          line = cd%settings%e(i)%lines(nk)
          if (lower(adjustl(line(1:4))) == 'bohr') then
            rval = cd%bohr
            cycle
          end if
          if (lower(adjustl(line(1:3))) == 'ang') cycle
          read (line, *, iostat=ios) text, pos
          if (ios /= 0) cycle
          cd%num_atoms = cd%num_atoms + 1
          cd%atom_symbol(cd%num_atoms) = text
          if (trim(cd%settings%e(i)%keyword) == 'atoms_frac') then
            cd%atom_cart(:, cd%num_atoms) = matmul(pos, cd%real_lattice)
          else
            cd%atom_cart(:, cd%num_atoms) = pos * rval
          end if
        end do
      end if
    end if
    ierr = 0
  end subroutine apply_settings

  ! ------------------------------------------------------------ k-mesh

  !> The disentanglement keywords (w90_readwrite_read_dis_manifold,
  !> src/readwrite.F90:636-684, and the dis_control part of
  !> src/wannier90_readwrite.F90:815-856), with the reference's checks.
  subroutine read_dis_settings(cd, ierr, msg)
    type(lib_common_type), intent(inout) :: cd
    integer, intent(out) :: ierr
    character(len=*), intent(out) :: msg
    logical :: found, found2
    integer :: n
    ierr = 0
    msg = ''
    call w90gpu_dis_control_defaults(cd%dis_control)
    cd%dis_set = .true.
    call get_real(cd%settings, 'dis_win_min', cd%dis_control%win_min, found, ierr)
    call get_real(cd%settings, 'dis_win_max', cd%dis_control%win_max, found, ierr)
    if (cd%dis_control%win_max < cd%dis_control%win_min) then
      ierr = 5
      msg = 'Error: w90_readwrite_read_dis_manifold: check disentanglement windows (win_max < win_min !)'
      return
    end if
    call get_real(cd%settings, 'dis_froz_max', cd%dis_control%froz_max, found, ierr)
    if (found) cd%dis_control%frozen_states = 1
    call get_real(cd%settings, 'dis_froz_min', cd%dis_control%froz_min, found2, ierr)
    if (cd%dis_control%froz_max < cd%dis_control%froz_min) then
      ierr = 5
      msg = 'Error: w90_readwrite_read_dis_manifold: check disentanglement frozen windows'
      return
    end if
    if (found2 .and. .not. found) then
      ierr = 5
      msg = 'Error: w90_readwrite_read_dis_manifold: found dis_froz_min but not dis_froz_max'
      return
    end if
    call get_int(cd%settings, 'dis_num_iter', cd%dis_control%num_iter, found, ierr)
    if (cd%dis_control%num_iter < 0) then
      ierr = 5
      msg = 'Error: dis_num_iter must be positive'
      return
    end if
    call get_real(cd%settings, 'dis_mix_ratio', cd%dis_control%mix_ratio, found, ierr)
    if (cd%dis_control%mix_ratio <= 0.0_dp .or. cd%dis_control%mix_ratio > 1.0_dp) then
      ierr = 5
      msg = 'Error: dis_mix_ratio must be greater than 0.0 but not greater than 1.0'
      return
    end if
    call get_real(cd%settings, 'dis_conv_tol', cd%dis_control%conv_tol, found, ierr)
    if (cd%dis_control%conv_tol < 0.0_dp) then
      ierr = 5
      msg = 'Error: dis_conv_tol must be positive'
      return
    end if
    call get_int(cd%settings, 'dis_conv_window', cd%dis_control%conv_window, found, ierr)
    if (cd%dis_control%conv_window < 0) then
      ierr = 5
      msg = 'Error: dis_conv_window must be positive'
      return
    end if
    n = 0
    call get_int(cd%settings, 'dis_spheres_num', n, found, ierr)
    if (n > 0) then
      ierr = 5
      msg = 'Error: dis_spheres is not supported by this build'
      return
    end if
    if (ierr /= 0) msg = 'Error: malformed disentanglement keyword'
  end subroutine read_dis_settings

  subroutine ensure_context(cd, istdout, istderr, ierr)
    type(lib_common_type), intent(inout) :: cd
    integer, intent(in) :: istdout, istderr
    integer, intent(out) :: ierr
    integer(c_int) :: rc
    ierr = 0
    if (c_associated(cd%ctx)) return
    rc = w90gpu_create(int(cd%device, c_int), cd%ctx)
    if (rc /= W90GPU_OK) then
      call report(istdout, istderr, int(rc), 'Error: no MI355X device for the wannierisation library: '// &
                  w90gpu_error_message(c_null_ptr), ierr)
      cd%ctx = c_null_ptr
    end if
  end subroutine ensure_context

  !> w90_create_kmesh (src/library_interface.F90:810): shells, weights and
  !> neighbour lists, computed by the library's k-mesh (kmesh_get).
  subroutine w90_create_kmesh(common_data, istdout, istderr, ierr)
    type(lib_common_type), intent(inout) :: common_data
    integer, intent(in) :: istdout, istderr
    integer, intent(out) :: ierr
    integer(c_int) :: rc, nntot, nk
    integer(c_int), allocatable :: nnl(:), nnc(:), sl(:)
    real(c_double), allocatable :: bk(:), wb(:)
    real(c_double) :: wbtot
    character(kind=c_char) :: err(256)
    character(len=256) :: msg
    integer :: i
    ierr = 0
    nk = int(common_data%num_kpts, c_int)
    if (nk <= 0 .or. .not. allocated(common_data%kpt_latt)) then
      call report(istdout, istderr, 1, 'Error: kmesh requested before the k-points are set', ierr)
      return
    end if
    allocate (nnl(nk * W90GPU_NUM_NNMAX), nnc(3 * nk * W90GPU_NUM_NNMAX), bk(3 * W90GPU_NUM_NNMAX * nk), &
              wb(W90GPU_NUM_NNMAX), sl(6))
    sl = common_data%shell_list
    rc = w90gpu_kmesh_get(common_data%real_lattice, nk, common_data%kpt_latt, &
                          int(common_data%search_shells, c_int), real(common_data%kmesh_tol, c_double), &
                          int(common_data%num_shells, c_int), sl, &
                          merge(1_c_int, 0_c_int, common_data%skip_b1_tests), nntot, nnl, nnc, bk, wb, &
                          wbtot, err, 256_c_int)
    if (rc /= W90GPU_OK) then
      msg = ''
      do i = 1, 256
        if (err(i) == c_null_char) exit
        msg(i:i) = err(i)
      end do
      call report(istdout, istderr, int(rc), 'Error: '//trim(msg), ierr)
      return
    end if
    common_data%nntot = nntot
    if (allocated(common_data%nnlist)) deallocate (common_data%nnlist, common_data%nncell, common_data%bk, &
                                                   common_data%wb)
    allocate (common_data%nnlist(nk, nntot), common_data%nncell(3, nk, nntot), &
              common_data%bk(3, nntot, nk), common_data%wb(nntot))
    common_data%nnlist = reshape(nnl(1:nk * nntot), [nk, nntot])
    common_data%nncell = reshape(nnc(1:3 * nk * nntot), [3, nk, nntot])
    common_data%bk = reshape(bk(1:3 * nntot * nk), [3, nntot, nk])
    common_data%wb = wb(1:nntot)
    common_data%wbtot = wbtot
    common_data%setup_complete = .true.
  end subroutine w90_create_kmesh

  subroutine w90_get_nn(common_data, nn, istdout, istderr, ierr)
    type(lib_common_type), intent(inout) :: common_data
    integer, intent(out) :: nn, ierr
    integer, intent(in) :: istdout, istderr
    ierr = 0
    if (.not. common_data%setup_complete) then
      call w90_create_kmesh(common_data, istdout, istderr, ierr)
      if (ierr > 0) return
    end if
    nn = common_data%nntot
  end subroutine w90_get_nn

  subroutine w90_get_nnkp(common_data, nnkp, istdout, istderr, ierr)
    type(lib_common_type), intent(inout) :: common_data
    integer, intent(out) :: nnkp(:, :), ierr
    integer, intent(in) :: istdout, istderr
    ierr = 0
    if (.not. common_data%setup_complete) then
      call w90_create_kmesh(common_data, istdout, istderr, ierr)
      if (ierr > 0) return
    end if
    nnkp = common_data%nnlist
  end subroutine w90_get_nnkp

  subroutine w90_get_gkpb(common_data, gkpb, istdout, istderr, ierr)
    type(lib_common_type), intent(inout) :: common_data
    integer, intent(out) :: gkpb(:, :, :), ierr
    integer, intent(in) :: istdout, istderr
    ierr = 0
    if (.not. common_data%setup_complete) then
      call w90_create_kmesh(common_data, istdout, istderr, ierr)
      if (ierr > 0) return
    end if
    gkpb = common_data%nncell
  end subroutine w90_get_gkpb

  ! ------------------------------------------------------------ arrays

  subroutine w90_set_m_local(common_data, m_orig)
    type(lib_common_type), intent(inout) :: common_data
    complex(dp), intent(inout), target :: m_orig(:, :, :, :)
    common_data%m_matrix_local => m_orig
  end subroutine w90_set_m_local

  subroutine w90_set_u_matrix(common_data, u_matrix)
    type(lib_common_type), intent(inout) :: common_data
    complex(dp), intent(inout), target :: u_matrix(:, :, :)
    common_data%u_matrix => u_matrix
  end subroutine w90_set_u_matrix

  subroutine w90_set_u_opt(common_data, u_opt)
    type(lib_common_type), intent(inout) :: common_data
    complex(dp), intent(inout), target :: u_opt(:, :, :)
    common_data%u_opt => u_opt
  end subroutine w90_set_u_opt

  subroutine w90_set_eigval(common_data, eigval)
    type(lib_common_type), intent(inout) :: common_data
    real(dp), intent(inout), target :: eigval(:, :)
    common_data%eigval => eigval
  end subroutine w90_set_eigval

  subroutine make_problem(cd, prob)
    type(lib_common_type), intent(inout), target :: cd
    type(w90gpu_problem), intent(out) :: prob
    prob%num_wann = int(cd%num_wann, c_int)
    prob%num_kpts = int(cd%num_kpts, c_int)
    prob%nntot = int(cd%nntot, c_int)
    prob%nnlist = c_loc(cd%nnlist)
    prob%bk = c_loc(cd%bk)
    prob%wb = c_loc(cd%wb)
    prob%wbtot = real(cd%wbtot, c_double)
  end subroutine make_problem

  !> w90_project_overlap (src/library_interface.F90:422): on entry u_matrix
  !> holds A, on exit the Lowdin U; m_matrix_local is rotated with it.
  subroutine w90_project_overlap(common_data, istdout, istderr, ierr)
    type(lib_common_type), intent(inout), target :: common_data
    integer, intent(in) :: istdout, istderr
    integer, intent(out) :: ierr
    type(w90gpu_problem) :: prob
    complex(dp), allocatable :: a(:, :, :)
    integer(c_int) :: rc
    ierr = 0
    if (.not. associated(common_data%m_matrix_local)) then
      call report(istdout, istderr, 1, 'm_matrix_local not set for projovlp call', ierr)
      return
    else if (.not. associated(common_data%u_matrix)) then
      call report(istdout, istderr, 1, 'u_matrix not set for projovlp call', ierr)
      return
    end if
    if (common_data%have_disentangled) return
    if (.not. common_data%setup_complete) then
      call w90_create_kmesh(common_data, istdout, istderr, ierr)
      if (ierr > 0) return
    end if
    call ensure_context(common_data, istdout, istderr, ierr)
    if (ierr > 0) return
    call make_problem(common_data, prob)
    a = common_data%u_matrix
    rc = w90gpu_project(common_data%ctx, prob, a, common_data%u_matrix, common_data%m_matrix_local)
    if (rc /= W90GPU_OK) call report(istdout, istderr, int(rc), w90gpu_error_message(common_data%ctx), ierr)
  end subroutine w90_project_overlap

  !> w90_wannierise (src/library_interface.F90:464): the minimisation of
  !> wann_main on the GPU, in place on the caller's u_matrix and
  !> m_matrix_local; centres, spreads and Omega kept in common_data.
  subroutine w90_wannierise(common_data, istdout, istderr, ierr)
    type(lib_common_type), intent(inout), target :: common_data
    integer, intent(in) :: istdout, istderr
    integer, intent(out) :: ierr
    type(w90gpu_problem) :: prob
    type(c_ptr) :: sess
    integer(c_int) :: rc, n
    real(c_double), allocatable :: hist(:)
    complex(dp), allocatable :: mw(:, :, :, :)
    ierr = 0
    if (.not. associated(common_data%m_matrix_local)) then
      call report(istdout, istderr, 1, 'm_matrix_local not set for wannierise', ierr)
      return
    else if (.not. associated(common_data%u_opt)) then
      call report(istdout, istderr, 1, 'u_opt not set for wannierise', ierr)
      return
    else if (.not. associated(common_data%u_matrix)) then
      call report(istdout, istderr, 1, 'u_matrix not set for wannierise', ierr)
      return
    end if
    if (common_data%gamma_only) then
      call report(istdout, istderr, 5, 'Error: gamma_only wannierisation is not supported by this build', ierr)
      return
    end if
    if (.not. common_data%setup_complete) then
      call w90_create_kmesh(common_data, istdout, istderr, ierr)
      if (ierr > 0) return
    end if
    if (.not. common_data%control_set) then
      call w90gpu_control_defaults(common_data%control)
      common_data%control_set = .true.
    end if
    call ensure_context(common_data, istdout, istderr, ierr)
    if (ierr > 0) return
    call make_problem(common_data, prob)
    rc = w90gpu_session_create(common_data%ctx, prob, common_data%control, sess)
    if (rc /= W90GPU_OK) then
      call report(istdout, istderr, int(rc), w90gpu_error_message(common_data%ctx), ierr)
      return
    end if
    ! after a disentanglement m_matrix_local is (num_bands, num_bands, ...)
    ! with the overlaps in its num_wann x num_wann corner (setup_m_loc)
    call pack_m(common_data, mw)
    rc = w90gpu_session_load(sess, mw, common_data%u_matrix)
    if (rc == W90GPU_OK) rc = w90gpu_session_solve(sess, common_data%result)
    if (allocated(common_data%centres)) deallocate (common_data%centres, common_data%spreads)
    allocate (common_data%centres(3, common_data%num_wann), common_data%spreads(common_data%num_wann))
    if (rc <= 0) rc = w90gpu_session_fetch(sess, mw, common_data%u_matrix, common_data%centres, &
                                          common_data%spreads)
    if (rc <= 0) call unpack_m(common_data, mw)
    if (rc <= 0) then
      allocate (hist(4 * (common_data%control%num_iter + 1)))
      n = w90gpu_session_history(sess, hist, int(common_data%control%num_iter + 1, c_int))
      if (allocated(common_data%history)) deallocate (common_data%history)
      allocate (common_data%history(4, 0:max(n - 1, 0)))
      if (n > 0) common_data%history = reshape(hist(1:4 * n), [4, n])
    end if
    if (rc /= W90GPU_OK) call report(istdout, istderr, int(rc), w90gpu_error_message(common_data%ctx), ierr)
    rc = w90gpu_session_destroy(sess)
  end subroutine w90_wannierise

  !> The num_wann x num_wann overlaps out of m_matrix_local (its corner when
  !> it is num_bands x num_bands) into a contiguous array, and back.
  subroutine pack_m(cd, mw)
    type(lib_common_type), intent(in) :: cd
    complex(dp), allocatable, intent(out) :: mw(:, :, :, :)
    integer :: nw
    nw = cd%num_wann
    allocate (mw(nw, nw, cd%nntot, cd%num_kpts))
    mw = cd%m_matrix_local(1:nw, 1:nw, 1:cd%nntot, 1:cd%num_kpts)
  end subroutine pack_m

  subroutine unpack_m(cd, mw)
    type(lib_common_type), intent(inout) :: cd
    complex(dp), intent(in) :: mw(:, :, :, :)
    integer :: nw
    nw = cd%num_wann
    cd%m_matrix_local(1:nw, 1:nw, 1:cd%nntot, 1:cd%num_kpts) = mw
  end subroutine unpack_m

  !> w90_disentangle (src/library_interface.F90:362-420): dis_main and
  !> setup_m_loc on the GPU.  On entry u_opt holds the projections A
  !> (num_bands, num_wann, num_kpts), m_matrix_local the overlaps of all
  !> bands and eigval the band energies; on exit u_opt is U_opt, u_matrix
  !> the initial U of wannierise, and the num_wann x num_wann corner of each
  !> block of m_matrix_local the overlaps in their gauge.  The DIS table of
  !> dis_extract goes to istdout.
  subroutine w90_disentangle(common_data, istdout, istderr, ierr)
    type(lib_common_type), intent(inout), target :: common_data
    integer, intent(in) :: istdout, istderr
    integer, intent(out) :: ierr
    type(w90gpu_problem) :: prob
    integer(c_int) :: rc
    integer :: nb, nw, nk, nn, k, j, rows
    complex(dp), allocatable :: a(:, :, :), mw(:, :, :, :)
    real(dp), allocatable :: hist(:, :)
    integer(c_int), allocatable :: nfirst(:), ndim(:)
    ierr = 0
    if (.not. associated(common_data%m_matrix_local)) then
      call report(istdout, istderr, 1, 'm_matrix_local not associated for disentangle call', ierr)
      return
    else if (.not. associated(common_data%u_matrix)) then
      call report(istdout, istderr, 1, 'u_matrix not associated for disentangle call', ierr)
      return
    else if (.not. associated(common_data%u_opt)) then
      call report(istdout, istderr, 1, 'u_opt not associated for disentangle call', ierr)
      return
    else if (.not. associated(common_data%eigval)) then
      call report(istdout, istderr, 1, 'eigval not associated for disentangle call', ierr)
      return
    end if
    if (common_data%gamma_only) then
      call report(istdout, istderr, 5, 'Error: gamma_only disentanglement is not supported by this build', ierr)
      return
    end if
    if (.not. common_data%setup_complete) then
      call w90_create_kmesh(common_data, istdout, istderr, ierr)
      if (ierr > 0) return
    end if
    if (.not. common_data%dis_set) then
      call w90gpu_dis_control_defaults(common_data%dis_control)
      common_data%dis_set = .true.
    end if
    call ensure_context(common_data, istdout, istderr, ierr)
    if (ierr > 0) return
    nb = common_data%num_bands
    nw = common_data%num_wann
    nk = common_data%num_kpts
    nn = common_data%nntot
    call make_problem(common_data, prob)
    allocate (a(nb, nw, nk), mw(nw, nw, nn, nk), nfirst(nk), ndim(nk))
    a = common_data%u_opt(1:nb, 1:nw, 1:nk)
    rows = max(common_data%dis_control%num_iter, 1)
    allocate (hist(3, rows))
    if (allocated(common_data%eigval_opt)) deallocate (common_data%eigval_opt)
    allocate (common_data%eigval_opt(nb, nk))
    rc = w90gpu_disentangle(common_data%ctx, prob, int(nb, c_int), common_data%dis_control, common_data%eigval, a, &
                            common_data%m_matrix_local, common_data%u_opt, common_data%u_matrix, mw, &
                            common_data%eigval_opt, nfirst, ndim, hist, int(rows, c_int), common_data%dis_result)
    if (rc /= W90GPU_OK) then
      call report(istdout, istderr, int(rc), w90gpu_error_message(common_data%ctx), ierr)
      return
    end if
    call unpack_m(common_data, mw)
    if (allocated(common_data%ndimwin)) deallocate (common_data%ndimwin, common_data%nfirstwin, &
                                                    common_data%lwindow)
    allocate (common_data%ndimwin(nk), common_data%nfirstwin(nk), common_data%lwindow(nb, nk))
    common_data%ndimwin = ndim
    common_data%nfirstwin = nfirst
    common_data%lwindow = .false.
    do k = 1, nk
      do j = nfirst(k), nfirst(k) + ndim(k) - 1
        common_data%lwindow(j, k) = .true.
      end do
    end do
    if (allocated(common_data%dis_history)) deallocate (common_data%dis_history)
    allocate (common_data%dis_history(3, common_data%dis_result%iterations))
    if (common_data%dis_result%iterations > 0) &
      common_data%dis_history = hist(:, 1:common_data%dis_result%iterations)
    common_data%omega_invariant = common_data%dis_result%omega_invariant
    common_data%have_disentangled = .true.
    if (common_data%iprint > 0) call print_dis(common_data, istdout)
  end subroutine w90_disentangle

  !> The DIS table and the closing lines of dis_extract
  !> (src/disentangle.F90:2017-2023, :2209-2212, :2545-2571), with the
  !> device time spread evenly over the iterations in the time column.
  subroutine print_dis(cd, istdout)
    type(lib_common_type), intent(in) :: cd
    integer, intent(in) :: istdout
    integer :: it, nit
    nit = cd%dis_result%iterations
    write (istdout, '(/1x,a)') '                  Extraction of optimally-connected subspace                  '
    write (istdout, '(1x,a)') '                  ------------------------------------------                  '
    write (istdout, '(1x,a)') '+---------------------------------------------------------------------+<-- DIS'
    write (istdout, '(1x,a)') '|  Iter     Omega_I(i-1)      Omega_I(i)      Delta (frac.)    Time   |<-- DIS'
    write (istdout, '(1x,a)') '+---------------------------------------------------------------------+<-- DIS'
    do it = 1, nit
      write (istdout, '(2x,i6,3x,f14.8,3x,f14.8,6x,es10.3,2x,f8.2,4x,a)') it, cd%dis_history(1, it), &
        cd%dis_history(2, it), cd%dis_history(3, it), cd%dis_result%seconds * it / max(nit, 1), '<-- DIS'
    end do
    if (cd%dis_result%converged == 1) then
      write (istdout, '(/13x,a,es10.3,a,i2,a)') '<<<      Delta <', cd%dis_control%conv_tol, '  over ', &
        cd%dis_control%conv_window, ' iterations     >>>'
      write (istdout, '(13x,a)') '<<< Disentanglement convergence criteria satisfied >>>'
    else
      write (istdout, '(/5x,a)') '<<< Warning: Maximum number of disentanglement iterations reached >>>'
      write (istdout, '(10x,a)') '<<< Disentanglement convergence criteria not satisfied >>>'
    end if
    write (istdout, '(/8x,a,f14.8,a/)') 'Final Omega_I ', cd%omega_invariant, ' (Ang^2)'
    write (istdout, '(1x,a/)') '+----------------------------------------------------------------------------+'
  end subroutine print_dis

  !> The output writers of plot_main on this path: _centres.xyz (write_xyz)
  !> and _u.mat, with _u_dis.mat after a disentanglement (write_u_matrices).
  !> Band/Fermi-surface/WF plotting is not carried by this build.
  subroutine w90_plot(common_data, istdout, istderr, ierr)
    type(lib_common_type), intent(inout) :: common_data
    integer, intent(in) :: istdout, istderr
    integer, intent(out) :: ierr
    character(len=20), allocatable :: sym(:)
    real(dp), allocatable :: pos(:, :)
    ierr = 0
    if (common_data%write_xyz .and. allocated(common_data%centres)) then
      if (common_data%num_atoms > 0) then
        sym = common_data%atom_symbol(1:common_data%num_atoms)
        pos = common_data%atom_cart(:, 1:common_data%num_atoms)
      else
        allocate (sym(0), pos(3, 0))
      end if
      call write_xyz(trim(common_data%seedname)//'_centres.xyz', common_data%num_wann, &
                     common_data%centres, common_data%real_lattice, common_data%translate_home_cell, &
                     common_data%num_atoms, sym, pos)
    end if
    if (common_data%write_u_matrices .and. associated(common_data%u_matrix)) then
      call write_umat(trim(common_data%seedname)//'_u.mat', common_data%num_wann, common_data%num_kpts, &
                      common_data%kpt_latt, common_data%u_matrix)
      if (common_data%have_disentangled .and. associated(common_data%u_opt) .and. &
          allocated(common_data%ndimwin)) then
        call write_umat_dis(trim(common_data%seedname)//'_u_dis.mat', common_data%num_wann, &
                            common_data%num_bands, common_data%num_kpts, common_data%kpt_latt, &
                            common_data%u_opt(1:common_data%num_bands, 1:common_data%num_wann, &
                                              1:common_data%num_kpts), common_data%ndimwin, &
                            common_data%nfirstwin)
      end if
    end if
    if (istdout < 0 .and. istderr < 0) ierr = 0
  end subroutine w90_plot

  subroutine w90_transport(common_data, istdout, istderr, ierr)
    type(lib_common_type), intent(inout) :: common_data
    integer, intent(in) :: istdout, istderr
    integer, intent(out) :: ierr
    call report(istdout, istderr, 5, 'Error: transport is not supported by this build', ierr)
    if (common_data%num_wann < 0) ierr = 5
  end subroutine w90_transport

  ! ------------------------------------------------------------ results

  subroutine w90_get_centres(common_data, centres)
    type(lib_common_type), intent(in) :: common_data
    real(dp), intent(out) :: centres(:, :)
    if (allocated(common_data%centres)) then
      centres = common_data%centres
    else
      centres = 0.0_dp
    end if
  end subroutine w90_get_centres

  subroutine w90_get_spreads(common_data, spreads)
    type(lib_common_type), intent(in) :: common_data
    real(dp), intent(out) :: spreads(:)
    if (allocated(common_data%spreads)) then
      spreads = common_data%spreads
    else
      spreads = 0.0_dp
    end if
  end subroutine w90_get_spreads

  !> The projections of the input (w90_get_proj,
  !> src/library_interface.F90:902): n functions with their site(3,:) in
  !> fractional coordinates, l, m (the mr index) and s, and where asked
  !> the spin axis sqa, the axes z and x, the radial index rad and zona.
  !> The projections block is expanded as the reference reads it
  !> (src/readwrite.F90:3390-4076) for inputs without spinors: a site is an
  !> atom label of the atoms block (every atom of it), f= fractional or c=
  !> Cartesian (Ang, or Bohr when the block starts with bohr); the angular
  !> part is a ;-list of names (s, p, pz, ..., sp3d2-6) or l=[,mr=]; the
  !> options z=, x=, r=, zona= follow after further colons.  Per line the
  !> functions go out by site, then l from -5 to 3, then mr.  Errors as
  !> the reference's: no projections block, arrays too short.
  subroutine w90_get_proj(common_data, n, site, l, m, s, istdout, istderr, ierr, sqa, z, x, rad, zona)
    type(lib_common_type), intent(in) :: common_data
    integer, intent(inout) :: n, l(:), m(:), s(:)
    real(dp), intent(inout) :: site(:, :)
    integer, intent(in) :: istdout, istderr
    integer, intent(out) :: ierr
    integer, intent(inout), optional :: rad(:)
    real(dp), intent(inout), optional :: sqa(:, :), z(:, :), x(:, :), zona(:)
    integer, allocatable :: pl(:), pm(:), prad(:)
    real(dp), allocatable :: psite(:, :), pz(:, :), px(:, :), pzona(:)
    character(len=256) :: msg
    integer :: np, i
    ierr = 0
    call expand_projections(common_data, np, pl, pm, psite, pz, px, prad, pzona, ierr, msg)
    if (ierr /= 0) then
      call report(istdout, istderr, ierr, msg, ierr)
      return
    end if
    n = np
    if (size(l) < n .or. size(m) < n .or. size(s) < n .or. size(site, 2) < n) then
      call report(istdout, istderr, 1, 'Array argument in get_proj() call is insufficiently sized', ierr)
      return
    end if
    if (present(sqa)) then
      if (size(sqa, 2) < n) ierr = 1
    end if
    if (present(z)) then
      if (size(z, 2) < n) ierr = 1
    end if
    if (present(x)) then
      if (size(x, 2) < n) ierr = 1
    end if
    if (present(rad)) then
      if (size(rad) < n) ierr = 1
    end if
    if (present(zona)) then
      if (size(zona) < n) ierr = 1
    end if
    if (ierr /= 0) then
      call report(istdout, istderr, 1, 'Optional array argument in get_proj() call is insufficiently sized', ierr)
      return
    end if
    do i = 1, n
      l(i) = pl(i)
      m(i) = pm(i)
      s(i) = 0
      site(:, i) = psite(:, i)
      if (present(sqa)) sqa(:, i) = [0.0_dp, 0.0_dp, 1.0_dp]
      if (present(z)) z(:, i) = pz(:, i)
      if (present(x)) x(:, i) = px(:, i)
      if (present(rad)) rad(i) = prad(i)
      if (present(zona)) zona(i) = pzona(i)
    end do
  end subroutine w90_get_proj

  !> The (l, first mr, last mr) of a projection name, or l = -99.
  subroutine proj_name(name, pl, m0, m1)
    character(len=*), intent(in) :: name
    integer, intent(out) :: pl, m0, m1
    character(len=12), parameter :: one(0:3, 7) = reshape([character(len=12) :: &
      's', 'pz', 'dz2', 'fz3', &
      '', 'px', 'dxz', 'fxz2', &
      '', 'py', 'dyz', 'fyz2', &
      '', '', 'dx2-y2', 'fxyz', &
      '', '', 'dxy', 'fz(x2-y2)', &
      '', '', '', 'fx(x2-3y2)', &
      '', '', '', 'fy(3x2-y2)'], [4, 7])
    character(len=5), parameter :: hyb(5) = [character(len=5) :: 'sp', 'sp2', 'sp3', 'sp3d', 'sp3d2']
    character(len=1), parameter :: whole(0:3) = ['s', 'p', 'd', 'f']
    integer :: i, j, dash, k, ios
    pl = -99
    do i = 0, 3
      if (trim(name) == whole(i)) then
        pl = i
        m0 = 1
        m1 = 2 * i + 1
        return
      end if
      do j = 1, 7
        if (len_trim(one(i, j)) > 0 .and. trim(name) == trim(one(i, j))) then
          pl = i
          m0 = j
          m1 = j
          return
        end if
      end do
    end do
    ! hybrids: sp..sp3d2 (l = -1..-5, mr 1..|l|+1), or one of them as name-k
    dash = index(name, '-')
    do i = 1, 5
      if (dash == 0) then
        if (trim(name) /= trim(hyb(i))) cycle
        pl = -i
        m0 = 1
        m1 = i + 1
        return
      else if (name(1:dash - 1) == trim(hyb(i))) then
        read (name(dash + 1:), *, iostat=ios) k
        if (ios /= 0 .or. k < 1 .or. k > i + 1) return
        pl = -i
        m0 = k
        m1 = k
        return
      end if
    end do
  end subroutine proj_name

  !> The functions of the projections block (see w90_get_proj).
  subroutine expand_projections(cd, np, pl, pm, psite, pz, px, prad, pzona, ierr, msg)
    type(lib_common_type), intent(in) :: cd
    integer, intent(out) :: np
    integer, allocatable, intent(out) :: pl(:), pm(:), prad(:)
    real(dp), allocatable, intent(out) :: psite(:, :), pz(:, :), px(:, :), pzona(:)
    integer, intent(out) :: ierr
    character(len=*), intent(out) :: msg
    integer, parameter :: cap = 4096
    character(len=512) :: line, field, part, rest
    logical :: states(1:7, -5:3)
    real(dp) :: inv(3, 3), frac(3), zax(3), xax(3), zn, scale
    real(dp), allocatable :: sites(:, :)
    integer :: ib, nl, j, c, nsite, ia, ll, mm, m0, m1, ios, rad, first, q
    ierr = 0
    msg = ''
    np = 0
    ib = settings_find(cd%settings, 'projections')
    if (ib == 0) then
      ierr = 1
      msg = 'Projectors are not setup in Wannier90 library when requested via get_proj()'
      return
    end if
    if (cd%settings%e(ib)%kind /= K_BLOCK) then
      ierr = 1
      msg = 'Projectors are not setup in Wannier90 library when requested via get_proj()'
      return
    end if
    allocate (pl(cap), pm(cap), prad(cap), psite(3, cap), pz(3, cap), px(3, cap), pzona(cap))
    inv = inverse3(cd%real_lattice)
    nl = size(cd%settings%e(ib)%lines)
    scale = 1.0_dp
    first = 1
    if (nl > 0) then
      line = lower(adjustl(cd%settings%e(ib)%lines(1)))
      if (line(1:4) == 'bohr') then
        scale = cd%bohr
        first = 2
      else if (line(1:3) == 'ang') then
        first = 2
      else if (line(1:6) == 'random') then
        ierr = 5
        msg = 'Error: random projections are not supported by this build'
        return
      end if
    end if
    do j = first, nl
      line = strip_spaces(lower(cd%settings%e(ib)%lines(j)))
      if (len_trim(line) == 0) cycle
      c = index(line, ':')
      if (c == 0) then
        ierr = 5
        msg = 'Error: malformed projection definition: '//trim(line)
        return
      end if
      field = line(:c - 1)
      rest = line(c + 1:)
      ! the site(s)
      if (field(1:2) == 'f=' .or. field(1:2) == 'c=') then
        read (field(3:), *, iostat=ios) frac
        if (ios /= 0) then
          ierr = 5
          msg = 'Error: malformed projection site: '//trim(field)
          return
        end if
        if (field(1:1) == 'c') frac = matmul(frac * scale, inv)
        nsite = 1
        allocate (sites(3, 1))
        sites(:, 1) = frac
      else
        nsite = 0
        do ia = 1, cd%num_atoms
          if (lower(trim(cd%atom_symbol(ia))) == trim(field)) nsite = nsite + 1
        end do
        if (nsite == 0) then
          ierr = 5
          msg = 'Error: Atom site not recognised '//trim(field)
          return
        end if
        allocate (sites(3, nsite))
        nsite = 0
        do ia = 1, cd%num_atoms
          if (lower(trim(cd%atom_symbol(ia))) /= trim(field)) cycle
          nsite = nsite + 1
          sites(:, nsite) = matmul(cd%atom_cart(:, ia), inv)
        end do
      end if
      if (index(rest, '[') > 0 .or. index(rest, '(u') > 0 .or. index(rest, '(d') > 0) then
        ierr = 5
        msg = 'Error: spinor projections are not supported by this build'
        return
      end if
      ! angular part: up to the next colon, ;-separated
      c = index(rest, ':')
      if (c > 0) then
        part = rest(:c - 1)
        rest = rest(c + 1:)
      else
        part = rest
        rest = ''
      end if
      states = .false.
      do while (len_trim(part) > 0)
        q = index(part, ';')
        if (q > 0) then
          field = part(:q - 1)
          part = part(q + 1:)
        else
          field = part
          part = ''
        end if
        if (field(1:2) == 'l=') then
          q = index(field, ',mr=')
          if (q > 0) then
            read (field(3:q - 1), *, iostat=ios) ll
          else
            read (field(3:), *, iostat=ios) ll
          end if
          if (ios /= 0 .or. ll < -5 .or. ll > 3) then
            ierr = 5
            msg = 'Error: malformed angular momentum: '//trim(field)
            return
          end if
          m1 = merge(2 * ll + 1, 1 - ll, ll >= 0)
          if (q == 0) then
            states(1:m1, ll) = .true.
          else
            field = field(q + 4:)
            do while (len_trim(field) > 0)
              ia = index(field, ',')
              if (ia == 0) ia = len_trim(field) + 1
              read (field(:ia - 1), *, iostat=ios) mm
              if (ios /= 0 .or. mm < 1 .or. mm > m1) then
                ierr = 5
                msg = 'Error: malformed mr in projection: '//trim(line)
                return
              end if
              states(mm, ll) = .true.
              field = field(ia + 1:)
            end do
          end if
        else
          call proj_name(trim(field), ll, m0, m1)
          if (ll == -99) then
            ierr = 5
            msg = 'Error: unknown projection '//trim(field)
            return
          end if
          states(m0:m1, ll) = .true.
        end if
      end do
      ! options z=, x=, r=, zona=
      zax = [0.0_dp, 0.0_dp, 1.0_dp]
      xax = [1.0_dp, 0.0_dp, 0.0_dp]
      zn = 1.0_dp
      rad = 1
      do while (len_trim(rest) > 0)
        c = index(rest, ':')
        if (c > 0) then
          field = rest(:c - 1)
          rest = rest(c + 1:)
        else
          field = rest
          rest = ''
        end if
        ios = 0
        if (field(1:2) == 'z=') then
          read (field(3:), *, iostat=ios) zax
        else if (field(1:2) == 'x=') then
          read (field(3:), *, iostat=ios) xax
        else if (field(1:5) == 'zona=') then
          read (field(6:), *, iostat=ios) zn
        else if (field(1:2) == 'r=') then
          read (field(3:), *, iostat=ios) rad
        else
          ios = 1
        end if
        if (ios /= 0) then
          ierr = 5
          msg = 'Error: malformed projection option: '//trim(field)
          return
        end if
      end do
      do ia = 1, nsite
        do ll = -5, 3
          do mm = 1, 7
            if (.not. states(mm, ll)) cycle
            if (np == cap) then
              ierr = 5
              msg = 'Error: too many projection functions'
              return
            end if
            np = np + 1
            pl(np) = ll
            pm(np) = mm
            psite(:, np) = sites(:, ia)
            pz(:, np) = zax
            px(:, np) = xax
            prad(np) = rad
            pzona(np) = zn
          end do
        end do
      end do
      deallocate (sites)
    end do
    if (np < cd%num_wann) then
      ierr = 5
      msg = 'Error: too few projection functions defined'
    end if
  end subroutine expand_projections

  !> `s` without blanks
  pure function strip_spaces(s) result(r)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: r
    integer :: i, k
    r = ''
    k = 0
    do i = 1, len_trim(s)
      if (s(i:i) == ' ' .or. s(i:i) == achar(9)) cycle
      k = k + 1
      r(k:k) = s(i:i)
    end do
  end function strip_spaces

  pure function inverse3(a) result(b)
    real(dp), intent(in) :: a(3, 3)
    real(dp) :: b(3, 3), det
    b(1, 1) = a(2, 2) * a(3, 3) - a(2, 3) * a(3, 2)
    b(1, 2) = a(1, 3) * a(3, 2) - a(1, 2) * a(3, 3)
    b(1, 3) = a(1, 2) * a(2, 3) - a(1, 3) * a(2, 2)
    b(2, 1) = a(2, 3) * a(3, 1) - a(2, 1) * a(3, 3)
    b(2, 2) = a(1, 1) * a(3, 3) - a(1, 3) * a(3, 1)
    b(2, 3) = a(1, 3) * a(2, 1) - a(1, 1) * a(2, 3)
    b(3, 1) = a(2, 1) * a(3, 2) - a(2, 2) * a(3, 1)
    b(3, 2) = a(1, 2) * a(3, 1) - a(1, 1) * a(3, 2)
    b(3, 3) = a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)
    det = a(1, 1) * b(1, 1) + a(1, 2) * b(2, 1) + a(1, 3) * b(3, 1)
    b = b / det
  end function inverse3

  !> Omega decomposition: omega(1:4) = Omega_I, Omega_D, Omega_OD, Omega;
  !> the reference keeps these in common_data%omega without a getter.
  subroutine w90_get_omega(common_data, omega)
    type(lib_common_type), intent(in) :: common_data
    real(dp), intent(out) :: omega(4)
    omega = [common_data%result%omega_invariant, common_data%result%omega_d, &
             common_data%result%omega_od, common_data%result%omega_total]
  end subroutine w90_get_omega

  !> .chk (w90_wannier90_readwrite_write_chkpt); with the disentanglement
  !> records after w90_disentangle.  Centres and spreads not computed yet
  !> (a 'postdis' checkpoint) are written as zeros.
  subroutine w90_write_chkpt(common_data, chkpt)
    type(lib_common_type), intent(in) :: common_data
    character(len=*), intent(in) :: chkpt
    complex(dp), allocatable :: mw(:, :, :, :)
    real(dp), allocatable :: cen(:, :), spr(:)
    integer :: nw, nb
    nw = common_data%num_wann
    nb = common_data%num_bands
    call pack_m(common_data, mw)
    allocate (cen(3, nw), spr(nw))
    cen = 0.0_dp
    spr = 0.0_dp
    if (allocated(common_data%centres)) then
      cen = common_data%centres
      spr = common_data%spreads
    end if
    if (common_data%have_disentangled) then
      call write_chk(trim(common_data%seedname)//'.chk', chkpt, nb, common_data%real_lattice, &
                     common_data%num_kpts, common_data%mp_grid, common_data%kpt_latt, common_data%nntot, nw, &
                     common_data%u_matrix, mw, cen, spr, common_data%omega_invariant, common_data%lwindow, &
                     common_data%ndimwin, common_data%u_opt(1:nb, 1:nw, 1:common_data%num_kpts))
    else
      call write_chk(trim(common_data%seedname)//'.chk', chkpt, nb, common_data%real_lattice, &
                     common_data%num_kpts, common_data%mp_grid, common_data%kpt_latt, common_data%nntot, nw, &
                     common_data%u_matrix, mw, cen, spr)
    end if
  end subroutine w90_write_chkpt

  !> U and M from the .chk, and after a disentanglement U_opt, lwindow,
  !> ndimwin and Omega_I (the restart from 'postdis' or 'postwann').
  subroutine w90_read_chkpt(common_data, istdout, istderr, ierr)
    type(lib_common_type), intent(inout) :: common_data
    integer, intent(in) :: istdout, istderr
    integer, intent(out) :: ierr
    character(len=256) :: msg
    complex(dp), allocatable :: mw(:, :, :, :)
    integer :: nb, nw, nk, k, j
    logical :: dis
    nb = common_data%num_bands
    nw = common_data%num_wann
    nk = common_data%num_kpts
    allocate (mw(nw, nw, common_data%nntot, nk))
    if (allocated(common_data%ndimwin)) deallocate (common_data%ndimwin, common_data%nfirstwin, &
                                                    common_data%lwindow)
    allocate (common_data%ndimwin(nk), common_data%nfirstwin(nk), common_data%lwindow(nb, nk))
    if (nb > nw .and. .not. associated(common_data%u_opt)) then
      call report(istdout, istderr, 1, 'u_opt not associated for reading a disentangled checkpoint', ierr)
      return
    end if
    if (nb > nw) then
      call read_chk(trim(common_data%seedname)//'.chk', nw, nk, common_data%nntot, common_data%u_matrix, mw, &
                    ierr, msg, nb, dis, common_data%omega_invariant, common_data%lwindow, common_data%ndimwin, &
                    common_data%u_opt)
    else
      call read_chk(trim(common_data%seedname)//'.chk', nw, nk, common_data%nntot, common_data%u_matrix, mw, &
                    ierr, msg)
      dis = .false.
    end if
    if (ierr > 0) then
      call report(istdout, istderr, ierr, msg, ierr)
      return
    end if
    call unpack_m(common_data, mw)
    common_data%have_disentangled = dis
    if (dis) then
      do k = 1, nk
        common_data%nfirstwin(k) = 0
        do j = nb, 1, -1
          if (common_data%lwindow(j, k)) common_data%nfirstwin(k) = j
        end do
      end do
    end if
  end subroutine w90_read_chkpt

  subroutine w90_write_nnkp(common_data)
    type(lib_common_type), intent(in) :: common_data
    call write_nnkp(trim(common_data%seedname)//'.nnkp', common_data%real_lattice, common_data%num_kpts, &
                    common_data%kpt_latt, common_data%nntot, common_data%nnlist, common_data%nncell)
  end subroutine w90_write_nnkp

  !> The CONV/SPRD lines and final state of wann_main
  !> (src/wannierise.F90:486-523, :703-756, :838-880) from the per-iteration
  !> record of the solve.
  subroutine w90_print_conv(common_data, istdout, wall)
    type(lib_common_type), intent(in) :: common_data
    integer, intent(in) :: istdout
    real(dp), intent(in) :: wall
    integer :: it, iw, nit
    real(dp) :: delta
    if (.not. allocated(common_data%history)) return
    nit = ubound(common_data%history, 2)
    write (istdout, *)
    write (istdout, '(1x,a)') '*------------------------------- WANNIERISE ---------------------------------*'
    write (istdout, '(1x,a)') '+--------------------------------------------------------------------+<-- CONV'
    write (istdout, '(1x,a)') '| Iter  Delta Spread     RMS Gradient      Spread (Ang^2)      Time  |<-- CONV'
    write (istdout, '(1x,a)') '+--------------------------------------------------------------------+<-- CONV'
    write (istdout, *)
    do it = 0, nit
      if (.not. (it == 0 .or. it == 1 .or. it == nit .or. mod(it, common_data%num_print_cycles) == 0)) cycle
      if (it == 0) then
        delta = common_data%history(1, 0)
      else
        delta = common_data%history(1, it) - common_data%history(1, it - 1)
      end if
      write (istdout, '(1x,i6,2x,E12.3,2x,F15.10,2x,F18.10,3x,F8.2,2x,a)') it, delta, &
        sqrt(abs(common_data%history(4, it))), common_data%history(1, it), wall, '<-- CONV'
      write (istdout, '(8x,a,F15.7,a,F15.7,a,F15.7,a)') 'O_D=', common_data%history(2, it), ' O_OD=', &
        common_data%history(3, it), ' O_TOT=', common_data%history(1, it), ' <-- SPRD'
    end do
    if (common_data%result%converged == 1) then
      write (istdout, '(/13x,a,es10.3,a,i2,a)') '<<<     Delta <', common_data%control%conv_tol, &
        '  over ', common_data%control%conv_window, ' iterations     >>>'
      write (istdout, '(13x,a/)') '<<< Wannierisation convergence criteria satisfied >>>'
    end if
    write (istdout, '(1x,a)') 'Final State'
    do iw = 1, common_data%num_wann
      write (istdout, '(2x,"WF centre and spread",i5,2x,"(",f10.6,",",f10.6,",",f10.6," )",f15.8)') iw, &
        common_data%centres(:, iw), common_data%spreads(iw)
    end do
    write (istdout, '(2x,"Sum of centres and spreads",1x,"(",f10.6,",",f10.6,",",f10.6," )",f15.8)') &
      sum(common_data%centres(1, :)), sum(common_data%centres(2, :)), sum(common_data%centres(3, :)), &
      sum(common_data%spreads)
    write (istdout, *)
    write (istdout, '(3x,a21,a,f15.9)') '     Spreads (Ang^2)', '       Omega I      = ', &
      common_data%result%omega_invariant
    write (istdout, '(3x,a,f15.9)') '     ================       Omega D      = ', common_data%result%omega_d
    write (istdout, '(3x,a,f15.9)') '                            Omega OD     = ', common_data%result%omega_od
    write (istdout, '(3x,a21,a,f15.9)') 'Final Spread (Ang^2)', '       Omega Total  = ', &
      common_data%result%omega_total
    write (istdout, '(1x,a78)') repeat('-', 78)
  end subroutine w90_print_conv

end module w90_library
