!> wannier90_gpu.x [-pp] seedname
!>
!> The wannier90.x driver (src/wannier_prog.F90) for the wannierisation
!> step, with the disentanglement and the minimisation on an MI355X: reads
!> seedname.win, builds the k-mesh (-pp: writes seedname.nnkp and stops),
!> reads seedname.mmn/.amn (and seedname.eig when num_bands > num_wann), or
!> with restart = wannierise U, M (and U_opt) from seedname.chk; projects
!> (Lowdin) or disentangles (then writes seedname.chk 'postdis'), minimises
!> the spread, and writes seedname.wout, seedname.chk ('postwann'), and,
!> when asked for, seedname_centres.xyz and seedname_u.mat.
program wannier90_gpu
  use iso_c_binding
  use w90_library
  use w90gpu_io, only: dp, read_mmn, read_amn, read_eig, datetime
  implicit none
  type(lib_common_type), target :: cd
  character(len=256) :: arg, seed, msg
  character(len=9) :: cdate, ctime
  logical :: pp
  integer :: i, nargs, stdout, stderr, ierr, nk, nw, nn, iw, nb
  integer(8) :: c0, c1, crate
  real(dp) :: wall
  complex(dp), allocatable, target :: m(:, :, :, :), u(:, :, :), uopt(:, :, :), a(:, :, :)
  real(dp), allocatable, target :: eig(:, :)

  call system_clock(c0, crate)
  pp = .false.
  seed = 'wannier'
  nargs = command_argument_count()
  do i = 1, nargs
    call get_command_argument(i, arg)
    if (trim(arg) == '-pp') then
      pp = .true.
    else
      seed = arg
    end if
  end do
  i = index(seed, '.win', back=.true.)
  if (i > 0 .and. i == len_trim(seed) - 3) seed = seed(1:i - 1)
  cd%seedname = seed

  open (newunit=stdout, file=trim(seed)//'.wout', form='formatted', status='replace')
  open (newunit=stderr, file=trim(seed)//'.werr', form='formatted', status='replace')
  call datetime(cdate, ctime)
  write (stdout, '(1x,a)') '+---------------------------------------------------+'
  write (stdout, '(1x,a)') '|   wannier90_gpu: wannierisation on AMD MI355X     |'
  write (stdout, '(1x,a)') '|   (disentanglement and spread minimisation)       |'
  write (stdout, '(1x,a)') '+---------------------------------------------------+'
  write (stdout, '(1x,a)') 'Execution started on '//cdate//' at '//ctime
  write (stdout, *)

  call w90_input_reader(cd, stdout, stderr, ierr)
  if (ierr > 0) call fail()
  call w90_create_kmesh(cd, stdout, stderr, ierr)
  if (ierr > 0) call fail()
  nk = cd%num_kpts
  nw = cd%num_wann
  nn = cd%nntot
  write (stdout, '(1x,a,i8,a,i4,a,i4)') 'num_kpts =', nk, '   nntot =', nn, '   num_wann =', nw
  write (stdout, '(1x,a)') 'b_k Vectors (Ang^-1) and Weights (Ang^2)'
  do i = 1, nn
    write (stdout, '(1x,i6,3x,3f12.6,3x,f12.6)') i, cd%bk(:, i, 1), cd%wb(i)
  end do
  if (pp) then
    call w90_write_nnkp(cd)
    write (stdout, '(/1x,a)') 'Exiting... '//trim(seed)//'.nnkp written.'
    close (stderr, status='delete')
    close (stdout)
    stop
  end if

  nb = cd%num_bands
  allocate (m(nb, nb, nn, nk), u(nw, nw, nk), uopt(nb, nw, nk))
  uopt = (0.0_dp, 0.0_dp)
  do i = 1, nk
    do iw = 1, nw
      uopt(iw, iw, i) = (1.0_dp, 0.0_dp)
    end do
  end do
  call w90_set_m_local(cd, m)
  call w90_set_u_matrix(cd, u)
  call w90_set_u_opt(cd, uopt)

  if (trim(cd%restart) == 'wannierise') then
    write (stdout, '(1x,a)') 'Restarting the wannierisation from '//trim(seed)//'.chk'
    call w90_read_chkpt(cd, stdout, stderr, ierr)
    if (ierr > 0) call fail()
  else if (nb > nw) then
    ! disentanglement: the projections go in u_opt, the band energies from .eig
    allocate (eig(nb, nk))
    call read_eig(trim(seed)//'.eig', nb, nk, eig, ierr, msg)
    if (ierr > 0) call fail_msg(msg)
    call w90_set_eigval(cd, eig)
    call read_mmn(trim(seed)//'.mmn', nb, nk, nn, cd%nnlist, cd%nncell, m, ierr, msg)
    if (ierr > 0) call fail_msg(msg)
    call read_amn(trim(seed)//'.amn', nb, nk, nw, uopt, ierr, msg)
    if (ierr > 0) call fail_msg(msg)
    call w90_disentangle(cd, stdout, stderr, ierr)
    if (ierr > 0) call fail()
    write (stdout, '(1x,a,f12.3,a,i6,a)') 'Time for disentangle (device) ', cd%dis_result%seconds, ' (sec), ', &
      cd%dis_result%iterations, ' iterations'
    call w90_write_chkpt(cd, 'postdis')
  else
    call read_mmn(trim(seed)//'.mmn', nw, nk, nn, cd%nnlist, cd%nncell, m, ierr, msg)
    if (ierr > 0) call fail_msg(msg)
    allocate (a(nw, nw, nk))
    call read_amn(trim(seed)//'.amn', nw, nk, nw, a, ierr, msg)
    if (ierr > 0) call fail_msg(msg)
    u = a
    call w90_project_overlap(cd, stdout, stderr, ierr)
    if (ierr > 0) call fail()
  end if

  call w90_wannierise(cd, stdout, stderr, ierr)
  if (ierr > 0) call fail()
  call system_clock(c1)
  wall = real(c1 - c0, dp) / real(crate, dp)
  call w90_print_conv(cd, stdout, wall)
  write (stdout, '(1x,a,f12.3,a,i6,a)') 'Time for wannierise (device) ', cd%result%solve_seconds, &
    ' (sec), ', cd%result%iterations, ' iterations'

  call w90_write_chkpt(cd, 'postwann')
  call w90_plot(cd, stdout, stderr, ierr)
  if (ierr > 0) call fail()

  call system_clock(c1)
  write (stdout, '(/1x,a,f12.3,a)') 'Total Execution Time ', real(c1 - c0, dp) / real(crate, dp), ' (sec)'
  write (stdout, '(1x,a)') 'All done: wannier90 exiting'
  close (stderr, status='delete')
  close (stdout)

contains

  subroutine fail()
    write (stdout, '(1x,a)') 'Error encountered; check '//trim(seed)//'.werr'
    close (stdout)
    close (stderr)
    error stop 1
  end subroutine fail

  subroutine fail_msg(text)
    character(len=*), intent(in) :: text
    write (stderr, '(1x,a)') 'Exiting.......'
    write (stderr, '(1x,a)') trim(text)
    call fail()
  end subroutine fail_msg

end program wannier90_gpu
