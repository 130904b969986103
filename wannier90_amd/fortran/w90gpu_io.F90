!> Input/output of the wannierisation step: the keyword store filled from a
!> .win file or from set_option calls, the .mmn/.amn readers, and the
!> .chk/.nnkp/_centres.xyz/_u.mat writers.  File formats are wannier90's:
!> .win (src/readwrite.F90:2151-2712), .mmn/.amn (src/overlap.F90:152-430),
!> .eig (src/readwrite.F90:588-630),
!> .chk (src/wannier90_readwrite.F90:2278-2373), .nnkp (src/kmesh.F90:743),
!> _u.mat (src/plot.F90:2224-2288), _centres.xyz (src/plot.F90:2790-2862).
module w90gpu_io
  use iso_fortran_env, only: int32, real64
  use iso_c_binding, only: c_int, c_char, c_null_char
  use w90gpu_c, only: w90gpu_read_mmn, w90gpu_read_amn
  implicit none
  private

  integer, parameter, public :: dp = real64
  !> Bohr radius in Angstrom (CODATA 2006, the reference's default)
  real(dp), parameter, public :: bohr_ang = 0.52917720859_dp
  integer, parameter :: linelen = 512

  integer, parameter, public :: K_TEXT = 0, K_INT = 1, K_REAL = 2, K_LOGICAL = 3, K_I1D = 4, &
                                K_R1D = 5, K_R2D = 6, K_I2D = 7, K_BLOCK = 8

  !> one keyword: a scalar text value (.win), a typed value (set_option) or
  !> the lines of a begin/end block
  type, public :: w90_setting
    character(len=64) :: keyword = ''
    integer :: kind = K_TEXT
    character(len=linelen) :: text = ''
    integer :: ival = 0
    real(dp) :: rval = 0.0_dp
    logical :: lval = .false.
    integer, allocatable :: i1d(:), i2d(:, :)
    real(dp), allocatable :: r1d(:), r2d(:, :)
    character(len=linelen), allocatable :: lines(:)
  end type w90_setting

  type, public :: w90_settings
    type(w90_setting), allocatable :: e(:)
    integer :: n = 0
  end type w90_settings

  public :: settings_add, settings_find, settings_clear, read_win, lower
  public :: get_int, get_real, get_logical, get_text, get_i1d, get_r2d, get_block
  public :: read_mmn, read_amn, read_eig, write_chk, read_chk, write_nnkp, write_xyz, write_umat, write_umat_dis
  public :: datetime

contains

  pure function lower(s) result(r)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: r
    integer :: i, c
    r = s
    do i = 1, len(s)
      c = iachar(s(i:i))
      if (c >= iachar('A') .and. c <= iachar('Z')) r(i:i) = achar(c + 32)
    end do
  end function lower

  subroutine datetime(cdate, ctime)
    character(len=9), intent(out) :: cdate, ctime
    character(len=3), parameter :: months(12) = ['Jan', 'Feb', 'Mar', 'Apr', 'May', 'Jun', &
                                                 'Jul', 'Aug', 'Sep', 'Oct', 'Nov', 'Dec']
    integer :: v(8)
    call date_and_time(values=v)
    write (cdate, '(i2,a3,i4)') v(3), months(v(2)), v(1)
    write (ctime, '(i2.2,":",i2.2,":",i2.2," ")') v(5), v(6), v(7)
  end subroutine datetime

  subroutine settings_add(s, entry)
    type(w90_settings), intent(inout) :: s
    type(w90_setting), intent(in) :: entry
    type(w90_setting), allocatable :: tmp(:)
    if (.not. allocated(s%e)) allocate (s%e(32))
    if (s%n == size(s%e)) then
      allocate (tmp(2 * size(s%e)))
      tmp(1:s%n) = s%e(1:s%n)
      call move_alloc(tmp, s%e)
    end if
    s%n = s%n + 1
    s%e(s%n) = entry
    s%e(s%n)%keyword = lower(adjustl(entry%keyword))
  end subroutine settings_add

  subroutine settings_clear(s)
    type(w90_settings), intent(inout) :: s
    if (allocated(s%e)) deallocate (s%e)
    s%n = 0
  end subroutine settings_clear

  !> index of the first entry for keyword, 0 if absent (the reference uses
  !> the first occurrence, src/readwrite.F90:2585)
  integer function settings_find(s, keyword)
    type(w90_settings), intent(in) :: s
    character(len=*), intent(in) :: keyword
    integer :: i
    settings_find = 0
    do i = 1, s%n
      if (trim(s%e(i)%keyword) == lower(trim(keyword))) then
        settings_find = i
        return
      end if
    end do
  end function settings_find

  !> Tokenise a .win file: "keyword = value", "keyword : value" or
  !> "keyword value" lines and "begin name ... end name" blocks; text after
  !> ! or # is a comment; keywords are case-insensitive.
  subroutine read_win(path, s, ierr, msg)
    character(len=*), intent(in) :: path
    type(w90_settings), intent(inout) :: s
    integer, intent(out) :: ierr
    character(len=*), intent(out) :: msg
    integer :: u, ios, p, nb
    character(len=linelen) :: line, key, rest
    character(len=linelen), allocatable :: blk(:), tmp(:)
    type(w90_setting) :: e
    logical :: inblock
    character(len=64) :: bname
    ierr = 0
    msg = ''
    open (newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      ierr = 6
      msg = 'Error: Problem opening input file '//trim(path)
      return
    end if
    inblock = .false.
    allocate (blk(64))
    nb = 0
    do
      read (u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      p = scan(line, '!#')
      if (p > 0) line(p:) = ''
      line = adjustl(line)
      if (len_trim(line) == 0) cycle
      if (inblock) then
        if (lower(line(1:4)) == 'end ' .or. lower(trim(line)) == 'end') then
          e = w90_setting()
          e%keyword = bname
          e%kind = K_BLOCK
          allocate (e%lines(nb))
          e%lines = blk(1:nb)
          call settings_add(s, e)
          inblock = .false.
          nb = 0
        else
          if (nb == size(blk)) then
            allocate (tmp(2 * size(blk)))
            tmp(1:nb) = blk(1:nb)
            call move_alloc(tmp, blk)
          end if
          nb = nb + 1
          blk(nb) = line
        end if
        cycle
      end if
      if (lower(line(1:6)) == 'begin ') then
        bname = lower(adjustl(line(7:)))
        inblock = .true.
        nb = 0
        cycle
      end if
      p = scan(line, '=: ')
      if (p == 0) then
        key = line
        rest = ''
      else
        key = line(1:p - 1)
        rest = adjustl(line(p + 1:))
        if (len_trim(rest) > 0) then
          if (rest(1:1) == '=' .or. rest(1:1) == ':') rest = adjustl(rest(2:))
        end if
      end if
      e = w90_setting()
      e%keyword = key
      e%kind = K_TEXT
      e%text = rest
      call settings_add(s, e)
    end do
    close (u)
    if (inblock) then
      ierr = 5
      msg = 'Error: block '//trim(bname)//' has no end'
    end if
  end subroutine read_win

  subroutine get_int(s, key, val, found, ierr)
    type(w90_settings), intent(in) :: s
    character(len=*), intent(in) :: key
    integer, intent(inout) :: val
    logical, intent(out) :: found
    integer, intent(inout) :: ierr
    integer :: i, ios
    i = settings_find(s, key)
    found = i > 0
    if (.not. found) return
    select case (s%e(i)%kind)
    case (K_INT)
      val = s%e(i)%ival
    case (K_TEXT)
      read (s%e(i)%text, *, iostat=ios) val
      if (ios /= 0) ierr = 5
    case default
      ierr = 5
    end select
  end subroutine get_int

  subroutine get_real(s, key, val, found, ierr)
    type(w90_settings), intent(in) :: s
    character(len=*), intent(in) :: key
    real(dp), intent(inout) :: val
    logical, intent(out) :: found
    integer, intent(inout) :: ierr
    integer :: i, ios
    character(len=linelen) :: t
    i = settings_find(s, key)
    found = i > 0
    if (.not. found) return
    select case (s%e(i)%kind)
    case (K_REAL)
      val = s%e(i)%rval
    case (K_INT)
      val = real(s%e(i)%ival, dp)
    case (K_TEXT)
      t = s%e(i)%text
      call fortran_exponent(t)
      read (t, *, iostat=ios) val
      if (ios /= 0) ierr = 5
    case default
      ierr = 5
    end select
  end subroutine get_real

  !> accept Fortran double-precision exponents (1.0d-10)
  subroutine fortran_exponent(t)
    character(len=*), intent(inout) :: t
    integer :: i
    do i = 1, len_trim(t)
      if (t(i:i) == 'd' .or. t(i:i) == 'D') t(i:i) = 'e'
    end do
  end subroutine fortran_exponent

  subroutine get_logical(s, key, val, found, ierr)
    type(w90_settings), intent(in) :: s
    character(len=*), intent(in) :: key
    logical, intent(inout) :: val
    logical, intent(out) :: found
    integer, intent(inout) :: ierr
    integer :: i
    character(len=linelen) :: t
    i = settings_find(s, key)
    found = i > 0
    if (.not. found) return
    select case (s%e(i)%kind)
    case (K_LOGICAL)
      val = s%e(i)%lval
    case (K_INT)
      val = s%e(i)%ival /= 0
    case (K_TEXT)
      t = lower(adjustl(s%e(i)%text))
      if (t(1:1) == '.') t = t(2:)
      if (t(1:1) == 't') then
        val = .true.
      else if (t(1:1) == 'f') then
        val = .false.
      else
        ierr = 5
      end if
    case default
      ierr = 5
    end select
  end subroutine get_logical

  subroutine get_text(s, key, val, found)
    type(w90_settings), intent(in) :: s
    character(len=*), intent(in) :: key
    character(len=*), intent(inout) :: val
    logical, intent(out) :: found
    integer :: i
    i = settings_find(s, key)
    found = i > 0
    if (found) val = lower(trim(adjustl(s%e(i)%text)))
  end subroutine get_text

  !> integer vector: an i1d entry or a text value with n integers
  subroutine get_i1d(s, key, val, n, found, ierr)
    type(w90_settings), intent(in) :: s
    character(len=*), intent(in) :: key
    integer, intent(in) :: n
    integer, intent(inout) :: val(n)
    logical, intent(out) :: found
    integer, intent(inout) :: ierr
    integer :: i, ios
    i = settings_find(s, key)
    found = i > 0
    if (.not. found) return
    select case (s%e(i)%kind)
    case (K_I1D)
      if (size(s%e(i)%i1d) < n) then
        ierr = 5
      else
        val = s%e(i)%i1d(1:n)
      end if
    case (K_TEXT)
      read (s%e(i)%text, *, iostat=ios) val
      if (ios /= 0) ierr = 5
    case default
      ierr = 5
    end select
  end subroutine get_i1d

  !> real block r(columns, rows): a r2d entry, or the numeric lines of a
  !> block whose optional first line gives the unit (bohr/ang); values in
  !> Bohr are converted to Angstrom when length_block is set
  subroutine get_r2d(s, key, r, cols, rows, length_block, found, ierr, bohr)
    type(w90_settings), intent(in) :: s
    character(len=*), intent(in) :: key
    integer, intent(in) :: cols, rows
    real(dp), intent(out) :: r(cols, rows)
    logical, intent(in) :: length_block
    logical, intent(out) :: found
    integer, intent(inout) :: ierr
    real(dp), intent(in), optional :: bohr  ! Angstrom per Bohr (default bohr_ang)
    integer :: i, j, first, ios
    real(dp) :: scale
    character(len=linelen) :: t
    i = settings_find(s, key)
    found = i > 0
    if (.not. found) return
    select case (s%e(i)%kind)
    case (K_R2D)
      if (size(s%e(i)%r2d, 1) /= cols .or. size(s%e(i)%r2d, 2) < rows) then
        ierr = 5
        return
      end if
      r = s%e(i)%r2d(:, 1:rows)
    case (K_BLOCK)
      first = 1
      scale = 1.0_dp
      if (size(s%e(i)%lines) > 0) then
        t = lower(adjustl(s%e(i)%lines(1)))
        if (t(1:4) == 'bohr') then
          first = 2
          if (length_block) then
            scale = bohr_ang
            if (present(bohr)) scale = bohr
          end if
        else if (t(1:3) == 'ang') then
          first = 2
        end if
      end if
      if (size(s%e(i)%lines) - first + 1 < rows) then
        ierr = 5
        return
      end if
      do j = 1, rows
        t = s%e(i)%lines(first + j - 1)
        call fortran_exponent(t)
        read (t, *, iostat=ios) r(:, j)
        if (ios /= 0) then
          ierr = 5
          return
        end if
      end do
      r = r * scale
    case default
      ierr = 5
    end select
  end subroutine get_r2d

  !> number of lines of a block (0 if absent)
  integer function get_block(s, key)
    type(w90_settings), intent(in) :: s
    character(len=*), intent(in) :: key
    integer :: i
    get_block = 0
    i = settings_find(s, key)
    if (i > 0) then
      if (s%e(i)%kind == K_BLOCK) get_block = size(s%e(i)%lines)
      if (s%e(i)%kind == K_R2D) get_block = size(s%e(i)%r2d, 2)
    end if
  end function get_block

  !> .mmn: each (k, k', G) block is stored at the b of the k-mesh with
  !> nnlist(k,b) = k' and nncell(:,k,b) = G (overlap_read), by the
  !> library's parallel reader (w90gpu_read_mmn)
  subroutine read_mmn(path, nb, nk, nntot, nnlist, nncell, m, ierr, msg)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nb, nk, nntot
    integer, intent(in) :: nnlist(nk, nntot), nncell(3, nk, nntot)
    complex(dp), intent(out) :: m(nb, nb, nntot, nk)
    integer, intent(out) :: ierr
    character(len=*), intent(out) :: msg
    character(kind=c_char) :: err(512)
    integer(c_int), allocatable :: nnl(:, :), nnc(:, :, :)
    allocate (nnl(nk, nntot), nnc(3, nk, nntot))
    nnl = int(nnlist, c_int)
    nnc = int(nncell, c_int)
    ierr = int(w90gpu_read_mmn(c_path(path), int(nb, c_int), int(nk, c_int), int(nntot, c_int), nnl, nnc, &
                               m, 0_c_int, err, 512_c_int))
    msg = c_message(err)
  end subroutine read_mmn

  !> .amn (overlap_read), by the library's parallel reader (w90gpu_read_amn)
  subroutine read_amn(path, nb, nk, nproj, a, ierr, msg)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nb, nk, nproj
    complex(dp), intent(out) :: a(nb, nproj, nk)
    integer, intent(out) :: ierr
    character(len=*), intent(out) :: msg
    character(kind=c_char) :: err(512)
    ierr = int(w90gpu_read_amn(c_path(path), int(nb, c_int), int(nk, c_int), int(nproj, c_int), a, 0_c_int, &
                               err, 512_c_int))
    msg = c_message(err)
  end subroutine read_amn

  !> .eig: lines "band k-point energy" in band-fastest order, each checked
  !> against the expected indices (w90_readwrite_read_eigvals)
  subroutine read_eig(path, nb, nk, eig, ierr, msg)
    character(len=*), intent(in) :: path
    integer, intent(in) :: nb, nk
    real(dp), intent(out) :: eig(nb, nk)
    integer, intent(out) :: ierr
    character(len=*), intent(out) :: msg
    integer :: unit, ios, i, j, n, k
    logical :: there
    ierr = 0
    msg = ''
    inquire (file=path, exist=there)
    if (.not. there) then
      ierr = 6
      msg = 'No '//trim(path)//' file found. Needed for disentanglement'
      return
    end if
    open (newunit=unit, file=path, form='formatted', status='old', iostat=ios)
    if (ios /= 0) then
      ierr = 6
      msg = 'Error: Problem opening eigenvalue file '//trim(path)
      return
    end if
    do k = 1, nk
      do n = 1, nb
        read (unit, *, iostat=ios) i, j, eig(n, k)
        if (ios /= 0) then
          ierr = 6
          msg = 'Error: Problem reading eigenvalue file '//trim(path)
          close (unit)
          return
        end if
        if (i /= n .or. j /= k) then
          ierr = 6
          write (msg, '(a,i0,a,i0,a,i0,a,i0,a)') 'mismatch in '//trim(path)//' (wanted band ', n, &
            ' k-point ', k, ', found band ', i, ' k-point ', j, ')'
          close (unit)
          return
        end if
      end do
    end do
    close (unit)
  end subroutine read_eig

  !> NUL-terminated copy of a Fortran file name
  function c_path(path) result(cp)
    character(len=*), intent(in) :: path
    character(kind=c_char) :: cp(len_trim(path) + 1)
    integer :: i
    do i = 1, len_trim(path)
      cp(i) = path(i:i)
    end do
    cp(len_trim(path) + 1) = c_null_char
  end function c_path

  !> Fortran text of a NUL-terminated message
  function c_message(err) result(msg)
    character(kind=c_char), intent(in) :: err(:)
    character(len=size(err)) :: msg
    integer :: i
    msg = ''
    do i = 1, size(err)
      if (err(i) == c_null_char) exit
      msg(i:i) = err(i)
    end do
  end function c_message

  !> unformatted checkpoint, record by record as the reference writes it;
  !> after a disentanglement (u_opt present) with Omega_I, lwindow, ndimwin
  !> and U_opt after have_disentangled
  subroutine write_chk(path, chkpt, num_bands, real_lattice, num_kpts, mp_grid, kpt_latt, nntot, &
                       num_wann, u, m, centres, spreads, omega_invariant, lwindow, ndimwin, u_opt)
    character(len=*), intent(in) :: path, chkpt
    integer, intent(in) :: num_bands, num_kpts, mp_grid(3), nntot, num_wann
    real(dp), intent(in) :: real_lattice(3, 3), kpt_latt(3, num_kpts)
    complex(dp), intent(in) :: u(num_wann, num_wann, num_kpts)
    complex(dp), intent(in) :: m(num_wann, num_wann, nntot, num_kpts)
    real(dp), intent(in) :: centres(3, num_wann), spreads(num_wann)
    real(dp), intent(in), optional :: omega_invariant
    logical, intent(in), optional :: lwindow(num_bands, num_kpts)
    integer, intent(in), optional :: ndimwin(num_kpts)
    complex(dp), intent(in), optional :: u_opt(num_bands, num_wann, num_kpts)
    integer :: unit, i, j
    character(len=9) :: cdate, ctime
    character(len=33) :: header
    character(len=20) :: chkpt1
    real(dp) :: recip(3, 3)
    call datetime(cdate, ctime)
    header = 'written on '//cdate//' at '//ctime
    call recip_lattice(real_lattice, recip)
    open (newunit=unit, file=path, form='unformatted')
    write (unit) header
    write (unit) num_bands
    write (unit) 0
    write (unit) (i, i=1, 0)
    write (unit) ((real_lattice(i, j), i=1, 3), j=1, 3)
    write (unit) ((recip(i, j), i=1, 3), j=1, 3)
    write (unit) num_kpts
    write (unit) (mp_grid(i), i=1, 3)
    write (unit) ((kpt_latt(i, j), i=1, 3), j=1, num_kpts)
    write (unit) nntot
    write (unit) num_wann
    chkpt1 = adjustl(trim(chkpt))
    write (unit) chkpt1
    if (present(u_opt)) then
      write (unit) .true.
      write (unit) omega_invariant
      write (unit) lwindow
      write (unit) ndimwin
      write (unit) u_opt
    else
      write (unit) .false.
    end if
    write (unit) u
    write (unit) m
    write (unit) centres
    write (unit) spreads
    close (unit)
  end subroutine write_chk

  !> read U and M of a checkpoint, and the disentanglement records into
  !> the optional arguments (have_dis says whether there were any)
  subroutine read_chk(path, num_wann, num_kpts, nntot, u, m, ierr, msg, num_bands, have_dis, omega_invariant, &
                      lwindow, ndimwin, u_opt)
    character(len=*), intent(in) :: path
    integer, intent(in) :: num_wann, num_kpts, nntot
    complex(dp), intent(out) :: u(num_wann, num_wann, num_kpts)
    complex(dp), intent(out) :: m(num_wann, num_wann, nntot, num_kpts)
    integer, intent(out) :: ierr
    character(len=*), intent(out) :: msg
    integer, intent(in), optional :: num_bands
    logical, intent(out), optional :: have_dis
    real(dp), intent(out), optional :: omega_invariant
    logical, intent(out), optional :: lwindow(:, :)
    integer, intent(out), optional :: ndimwin(:)
    complex(dp), intent(out), optional :: u_opt(:, :, :)
    integer :: unit, ios, nb, nex, nk, mp(3), nn, nw
    character(len=33) :: header
    character(len=20) :: chkpt1
    logical :: dis
    real(dp) :: rl(3, 3), rc(3, 3)
    real(dp), allocatable :: kp(:, :)
    integer, allocatable :: ex(:)
    ierr = 0
    msg = ''
    open (newunit=unit, file=path, form='unformatted', status='old', iostat=ios)
    if (ios /= 0) then
      ierr = 6
      msg = 'Error: Problem opening checkpoint file '//trim(path)
      return
    end if
    read (unit, iostat=ios) header
    read (unit, iostat=ios) nb
    read (unit, iostat=ios) nex
    allocate (ex(max(nex, 0)))
    read (unit, iostat=ios) ex
    read (unit, iostat=ios) rl
    read (unit, iostat=ios) rc
    read (unit, iostat=ios) nk
    read (unit, iostat=ios) mp
    allocate (kp(3, max(nk, 1)))
    read (unit, iostat=ios) kp(:, 1:nk)
    read (unit, iostat=ios) nn
    read (unit, iostat=ios) nw
    read (unit, iostat=ios) chkpt1
    read (unit, iostat=ios) dis
    if (ios /= 0 .or. nk /= num_kpts .or. nn /= nntot .or. nw /= num_wann) then
      ierr = 6
      msg = 'Error: checkpoint '//trim(path)//' does not match the input'
      close (unit)
      return
    end if
    if (present(have_dis)) have_dis = dis
    if (dis) then
      if (.not. present(u_opt)) then
        ierr = 5
        msg = 'Error: checkpoint '//trim(path)//' holds a disentanglement this caller does not read'
        close (unit)
        return
      end if
      if (nb /= num_bands) then
        ierr = 6
        msg = 'Error: checkpoint '//trim(path)//' does not match the input (num_bands)'
        close (unit)
        return
      end if
      read (unit, iostat=ios) omega_invariant
      read (unit, iostat=ios) lwindow
      read (unit, iostat=ios) ndimwin
      read (unit, iostat=ios) u_opt
    end if
    read (unit, iostat=ios) u
    read (unit, iostat=ios) m
    if (ios /= 0) then
      ierr = 6
      msg = 'Error: Problem reading checkpoint file '//trim(path)
    end if
    close (unit)
  end subroutine read_chk

  subroutine recip_lattice(real_lattice, recip)
    real(dp), intent(in) :: real_lattice(3, 3)
    real(dp), intent(out) :: recip(3, 3)
    real(dp) :: vol
    real(dp), parameter :: twopi = 6.283185307179586476925286766559_dp
    recip(1, 1) = real_lattice(2, 2) * real_lattice(3, 3) - real_lattice(3, 2) * real_lattice(2, 3)
    recip(1, 2) = real_lattice(2, 3) * real_lattice(3, 1) - real_lattice(3, 3) * real_lattice(2, 1)
    recip(1, 3) = real_lattice(2, 1) * real_lattice(3, 2) - real_lattice(3, 1) * real_lattice(2, 2)
    recip(2, 1) = real_lattice(3, 2) * real_lattice(1, 3) - real_lattice(1, 2) * real_lattice(3, 3)
    recip(2, 2) = real_lattice(3, 3) * real_lattice(1, 1) - real_lattice(1, 3) * real_lattice(3, 1)
    recip(2, 3) = real_lattice(3, 1) * real_lattice(1, 2) - real_lattice(1, 1) * real_lattice(3, 2)
    recip(3, 1) = real_lattice(1, 2) * real_lattice(2, 3) - real_lattice(2, 2) * real_lattice(1, 3)
    recip(3, 2) = real_lattice(1, 3) * real_lattice(2, 1) - real_lattice(2, 3) * real_lattice(1, 1)
    recip(3, 3) = real_lattice(1, 1) * real_lattice(2, 2) - real_lattice(2, 1) * real_lattice(1, 2)
    vol = real_lattice(1, 1) * recip(1, 1) + real_lattice(1, 2) * recip(1, 2) + real_lattice(1, 3) * recip(1, 3)
    recip = twopi * recip / vol
  end subroutine recip_lattice

  subroutine write_nnkp(path, real_lattice, num_kpts, kpt_latt, nntot, nnlist, nncell)
    character(len=*), intent(in) :: path
    integer, intent(in) :: num_kpts, nntot
    real(dp), intent(in) :: real_lattice(3, 3), kpt_latt(3, num_kpts)
    integer, intent(in) :: nnlist(num_kpts, nntot), nncell(3, num_kpts, nntot)
    integer :: u, i, k, b
    character(len=9) :: cdate, ctime
    real(dp) :: recip(3, 3)
    call datetime(cdate, ctime)
    call recip_lattice(real_lattice, recip)
    open (newunit=u, file=path, form='formatted')
    write (u, '(4(a),/)') '# File written on ', cdate, ' at ', ctime
    write (u, '(a,l2,/)') 'calc_only_A  : ', .false.
    write (u, '(a)') 'begin real_lattice'
    do i = 1, 3
      write (u, '(3(f12.7))') real_lattice(i, :)
    end do
    write (u, '(a/)') 'end real_lattice'
    write (u, '(a)') 'begin recip_lattice'
    do i = 1, 3
      write (u, '(3f12.7)') recip(i, :)
    end do
    write (u, '(a/)') 'end recip_lattice'
    write (u, '(a)') 'begin kpoints'
    write (u, '(i6)') num_kpts
    do k = 1, num_kpts
      write (u, '(3f14.8)') kpt_latt(:, k)
    end do
    write (u, '(a/)') 'end kpoints'
    write (u, '(a)') 'begin projections'
    write (u, '(i6)') 0
    write (u, '(a/)') 'end projections'
    write (u, '(a)') 'begin nnkpts'
    write (u, '(i4)') nntot
    do k = 1, num_kpts
      do b = 1, nntot
        write (u, '(2i6,3x,3i4)') k, nnlist(k, b), nncell(:, k, b)
      end do
    end do
    write (u, '(a/)') 'end nnkpts'
    write (u, '(a)') 'begin exclude_bands'
    write (u, '(i4)') 0
    write (u, '(a)') 'end exclude_bands'
    close (u)
  end subroutine write_nnkp

  subroutine write_xyz(path, num_wann, centres, real_lattice, translate_home_cell, nat, symbols, pos_cart)
    character(len=*), intent(in) :: path
    integer, intent(in) :: num_wann, nat
    real(dp), intent(in) :: centres(3, num_wann), real_lattice(3, 3)
    logical, intent(in) :: translate_home_cell
    character(len=*), intent(in) :: symbols(:)
    real(dp), intent(in) :: pos_cart(:, :)
    integer :: u, iw, i
    character(len=9) :: cdate, ctime
    real(dp) :: wc(3), frac(3), recip(3, 3)
    real(dp), parameter :: twopi = 6.283185307179586476925286766559_dp
    call datetime(cdate, ctime)
    call recip_lattice(real_lattice, recip)
    open (newunit=u, file=path, form='formatted')
    write (u, '(i6)') num_wann + nat
    write (u, *) 'Wannier centres, written by Wannier90 on'//cdate//' at '//ctime
    do iw = 1, num_wann
      wc = centres(:, iw)
      if (translate_home_cell) then
        frac = matmul(recip, wc) / twopi
        frac = frac - floor(frac)
        wc = matmul(frac, real_lattice)
      end if
      write (u, '("X",6x,3(f14.8,3x))') wc
    end do
    do i = 1, nat
      write (u, '(a2,5x,3(f14.8,3x))') symbols(i), pos_cart(:, i)
    end do
    close (u)
  end subroutine write_xyz

  subroutine write_umat(path, num_wann, num_kpts, kpt_latt, u_matrix)
    character(len=*), intent(in) :: path
    integer, intent(in) :: num_wann, num_kpts
    real(dp), intent(in) :: kpt_latt(3, num_kpts)
    complex(dp), intent(in) :: u_matrix(num_wann, num_wann, num_kpts)
    integer :: u, k, i, j
    character(len=9) :: cdate, ctime
    call datetime(cdate, ctime)
    open (newunit=u, file=path, form='formatted')
    write (u, *) 'written on '//cdate//' at '//ctime
    write (u, '(3i12)') num_kpts, num_wann, num_wann
    do k = 1, num_kpts
      write (u, *)
      write (u, '(f15.10,sp,f15.10,sp,f15.10)') kpt_latt(:, k)
      write (u, '(f15.10,sp,f15.10)') ((u_matrix(i, j, k), i=1, num_wann), j=1, num_wann)
    end do
    close (u)
  end subroutine write_umat

  !> _u_dis.mat (plot_u_matrices, src/plot.F90:2266-2285): U_opt with its
  !> window rows at their band positions (zero outside the window), in the
  !> layout of _u.mat; the header gives num_kpts, num_wann, num_bands.
  subroutine write_umat_dis(path, num_wann, num_bands, num_kpts, kpt_latt, u_opt, ndimwin, nfirstwin)
    character(len=*), intent(in) :: path
    integer, intent(in) :: num_wann, num_bands, num_kpts
    real(dp), intent(in) :: kpt_latt(3, num_kpts)
    complex(dp), intent(in) :: u_opt(num_bands, num_wann, num_kpts)
    integer, intent(in) :: ndimwin(num_kpts), nfirstwin(num_kpts)
    complex(dp) :: col(num_bands)
    integer :: u, k, i, j, f, nd
    character(len=9) :: cdate, ctime
    call datetime(cdate, ctime)
    open (newunit=u, file=path, form='formatted')
    write (u, *) 'written on '//cdate//' at '//ctime
    write (u, '(3i12)') num_kpts, num_wann, num_bands
    do k = 1, num_kpts
      write (u, *)
      write (u, '(f15.10,sp,f15.10,sp,f15.10)') kpt_latt(:, k)
      f = nfirstwin(k)
      nd = ndimwin(k)
      do j = 1, num_wann
        col = (0.0_dp, 0.0_dp)
        col(f:f + nd - 1) = u_opt(1:nd, j, k)
        write (u, '(f15.10,sp,f15.10)') (col(i), i=1, num_bands)
      end do
    end do
    close (u)
  end subroutine write_umat_dis

end module w90gpu_io
