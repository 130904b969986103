!> Runs the reference's wannierisation step -- module w90_library of the
!> wannier90 sources, compiled by oracle/Makefile.ref -- on an initial state
!> handed over in binary, serially or (built with -DMPI) with the k-points
!> distributed over MPI ranks in contiguous blocks, the distribution
!> wannier90.x uses (src/wannier_prog.F90:127-137).  TEST INFRASTRUCTURE
!> AND CPU BASELINE ONLY: it writes fixtures and times the reference; the
!> product never runs it.
!>
!> The options go in through w90_set_option + w90_input_setopt
!> (src/library_interface.F90:161), the arrays through w90_set_m_local,
!> w90_set_u_matrix and w90_set_u_opt, the step is w90_wannierise (:464).
!> No .mmn/.amn text and no checkpoint are read or written, so a large
!> problem costs the minimisation only.
!>
!>   ref_driver.x input.bin output.bin seedname
!>
!> input.bin (little endian): int32 nw, nk, nn, mp(3), num_iter,
!> num_cg_steps, conv_window, num_print_cycles, nb (num_bands),
!> dis_num_iter, dis_conv_window, frozen_states; float64 conv_tol,
!> trial_step, dis_win_min, dis_win_max, dis_froz_min, dis_froz_max,
!> dis_mix_ratio, dis_conv_tol, cell(3,3) (column i = lattice vector i,
!> Ang), kpt_latt(3,nk); int32 nnlist(nk,nn) (1-based, checked against the
!> reference's own k-mesh); complex128 M(nb,nb,nn,nk); then for nb = nw
!> U(nw,nw,nk), for nb > nw the projections A(nb,nw,nk) and float64
!> eigval(nb,nk) (w90_disentangle runs first; windows beyond +-1e300 are
!> left at the reference's defaults).
!> output.bin (rank 0): float64 seconds (wall time of w90_wannierise, the
!> slowest rank), float64 omega(3) = (invariant, tilde, total), float64
!> centres(3,nw), spreads(nw); complex128 U(nw,nw,nk); for nb > nw also
!> float64 seconds of w90_disentangle and complex128 U_opt(nb,nw,nk).  The
!> CONV and DIS lines go to seedname.wout.
program ref_driver
  use w90_library
  implicit none
#ifdef MPI
  include 'mpif.h'
#endif
  integer, parameter :: dp = kind(1.0d0)
  type(lib_common_type), target :: lib
  integer :: nw, nk, nn, mp(3), num_iter, num_cg_steps, conv_window, num_print_cycles
  integer :: nb, dis_num_iter, dis_conv_window, frozen_states
  real(dp) :: conv_tol, trial_step, cell(3, 3), t0, t1, tmax, tdis
  real(dp) :: win_min, win_max, froz_min, froz_max, dis_mix_ratio, dis_conv_tol
  real(dp), allocatable :: kpt(:, :), eig(:, :)
  integer, allocatable :: nnlist(:, :), nnref(:, :), distk(:)
  complex(dp), allocatable :: mall(:, :, :, :), mloc(:, :, :, :), u(:, :, :), uopt(:, :, :)
  integer :: rank, nranks, ierr, unit_in, unit_out, unit_wout, unit_err, nkl, k, k0, k1, i, nn_ref, comm
  character(len=512) :: fin, fout, seed

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, seed)
#ifdef MPI
  call mpi_init(ierr)
  call mpi_comm_rank(MPI_COMM_WORLD, rank, ierr)
  call mpi_comm_size(MPI_COMM_WORLD, nranks, ierr)
  comm = MPI_COMM_WORLD
#else
  rank = 0
  nranks = 1
  comm = 0
#endif

  open (newunit=unit_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read (unit_in) nw, nk, nn, mp, num_iter, num_cg_steps, conv_window, num_print_cycles, nb, dis_num_iter, &
    dis_conv_window, frozen_states
  read (unit_in) conv_tol, trial_step, win_min, win_max, froz_min, froz_max, dis_mix_ratio, dis_conv_tol, cell
  allocate (kpt(3, nk), nnlist(nk, nn), mall(nb, nb, nn, nk), u(nw, nw, nk), uopt(nb, nw, nk))
  read (unit_in) kpt, nnlist, mall
  if (nb == nw) then
    read (unit_in) u
  else
    allocate (eig(nb, nk))
    read (unit_in) uopt, eig
    u = (0.0_dp, 0.0_dp)
  end if
  close (unit_in)

  ! contiguous blocks of k-points per rank
  allocate (distk(nk))
  k = 0
  do i = 0, nranks - 1
    nkl = nk / nranks
    if (mod(nk, nranks) > i) nkl = nkl + 1
    distk(k + 1:k + nkl) = i
    k = k + nkl
  end do

  call w90_set_option(lib, 'num_wann', nw)
  call w90_set_option(lib, 'num_bands', nb)
  if (nb > nw) then
    call w90_set_option(lib, 'dis_num_iter', dis_num_iter)
    call w90_set_option(lib, 'dis_conv_window', dis_conv_window)
    call w90_set_option(lib, 'dis_mix_ratio', dis_mix_ratio)
    call w90_set_option(lib, 'dis_conv_tol', dis_conv_tol)
    if (abs(win_min) < 1.0e300_dp) call w90_set_option(lib, 'dis_win_min', win_min)
    if (abs(win_max) < 1.0e300_dp) call w90_set_option(lib, 'dis_win_max', win_max)
    if (frozen_states /= 0) then
      call w90_set_option(lib, 'dis_froz_max', froz_max)
      if (abs(froz_min) < 1.0e300_dp) call w90_set_option(lib, 'dis_froz_min', froz_min)
    end if
  end if
  call w90_set_option(lib, 'unit_cell_cart', cell)
  call w90_set_option(lib, 'mp_grid', mp)
  call w90_set_option(lib, 'kpoints', kpt)
  call w90_set_option(lib, 'num_iter', num_iter)
  call w90_set_option(lib, 'num_cg_steps', num_cg_steps)
  call w90_set_option(lib, 'conv_window', conv_window)
  call w90_set_option(lib, 'conv_tol', conv_tol)
  call w90_set_option(lib, 'trial_step', trial_step)
  call w90_set_option(lib, 'num_print_cycles', num_print_cycles)
  call w90_set_option(lib, 'search_shells', 36)
  call w90_set_option(lib, 'distk', distk)

  call w90_get_fortran_stderr(unit_err)
  if (rank == 0) then
    open (newunit=unit_wout, file=trim(seed)//'.wout', status='replace')
  else
    open (newunit=unit_wout, file='/dev/null', status='old')
  end if
  call w90_input_setopt(lib, trim(seed), comm, unit_wout, unit_err, ierr)
  if (ierr /= 0) call fail('w90_input_setopt')
  call w90_create_kmesh(lib, unit_wout, unit_err, ierr)
  if (ierr /= 0) call fail('w90_create_kmesh')
  call w90_get_nn(lib, nn_ref, unit_wout, unit_err, ierr)
  if (nn_ref /= nn) call fail('nntot differs from the reference k-mesh')
  allocate (nnref(nk, nn))
  call w90_get_nnkp(lib, nnref, unit_wout, unit_err, ierr)
  if (any(nnref /= nnlist)) call fail('nnlist differs from the reference k-mesh')

  ! this rank's overlaps, U for all k (wann_main keeps the global U)
  k0 = 0
  do while (distk(k0 + 1) /= rank)
    k0 = k0 + 1
  end do
  k1 = k0 + count(distk == rank)
  allocate (mloc(nb, nb, nn, k1 - k0))
  mloc = mall(:, :, :, k0 + 1:k1)
  deallocate (mall)
  if (nb == nw) then
    uopt = (0.0_dp, 0.0_dp)
    do k = 1, nk
      do i = 1, nw
        uopt(i, i, k) = (1.0_dp, 0.0_dp)
      end do
    end do
  end if
  call w90_set_m_local(lib, mloc)
  call w90_set_u_matrix(lib, u)
  call w90_set_u_opt(lib, uopt)

  tdis = 0.0_dp
  if (nb > nw) then
    call w90_set_eigval(lib, eig)
#ifdef MPI
    call mpi_barrier(MPI_COMM_WORLD, ierr)
    t0 = mpi_wtime()
#else
    call wall(t0)
#endif
    call w90_disentangle(lib, unit_wout, unit_err, ierr)
#ifdef MPI
    t1 = mpi_wtime() - t0
    call mpi_reduce(t1, tdis, 1, MPI_DOUBLE_PRECISION, MPI_MAX, 0, MPI_COMM_WORLD, ierr)
#else
    call wall(t1)
    tdis = t1 - t0
#endif
    if (ierr /= 0) call fail('w90_disentangle')
  end if

#ifdef MPI
  call mpi_barrier(MPI_COMM_WORLD, ierr)
  t0 = mpi_wtime()
#else
  call wall(t0)
#endif
  call w90_wannierise(lib, unit_wout, unit_err, ierr)
#ifdef MPI
  t1 = mpi_wtime() - t0
  call mpi_reduce(t1, tmax, 1, MPI_DOUBLE_PRECISION, MPI_MAX, 0, MPI_COMM_WORLD, ierr)
#else
  call wall(t1)
  tmax = t1 - t0
#endif
  if (ierr /= 0) call fail('w90_wannierise')
  close (unit_wout)

  if (rank == 0) then
    open (newunit=unit_out, file=trim(fout), access='stream', form='unformatted', status='replace')
    write (unit_out) tmax, lib%omega%invariant, lib%omega%tilde, lib%omega%total
    write (unit_out) lib%wannier_data%centres, lib%wannier_data%spreads, u
    if (nb > nw) write (unit_out) tdis, uopt
    close (unit_out)
  end if
#ifdef MPI
  call mpi_finalize(ierr)
#endif

contains

  subroutine fail(what)
    character(len=*), intent(in) :: what
    write (*, '(a,a)') 'ref_driver: failed in ', what
#ifdef MPI
    call mpi_abort(MPI_COMM_WORLD, 1, ierr)
#endif
    stop 1
  end subroutine fail

  subroutine wall(t)
    real(dp), intent(out) :: t
    integer(8) :: c, r
    call system_clock(c, r)
    t = real(c, dp) / real(r, dp)
  end subroutine wall

end program ref_driver
