﻿!mod$ v1 sum:e34ef4e13c24167b
!need$ ad019ea2035d982d n w90gpu_c
!need$ 7d55b91d8c4c39ed n w90gpu_io
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module w90_library
use,intrinsic::__fortran_builtins,only:__builtin_c_ptr
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
use,intrinsic::iso_c_binding,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use,intrinsic::iso_c_binding,only:c_int8_t
use,intrinsic::iso_c_binding,only:c_int16_t
use,intrinsic::iso_c_binding,only:c_int32_t
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_int128_t
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_short
use,intrinsic::iso_c_binding,only:c_long
use,intrinsic::iso_c_binding,only:c_long_long
use,intrinsic::iso_c_binding,only:c_signed_char
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_intmax_t
use,intrinsic::iso_c_binding,only:c_intptr_t
use,intrinsic::iso_c_binding,only:c_ptrdiff_t
use,intrinsic::iso_c_binding,only:c_int_least8_t
use,intrinsic::iso_c_binding,only:c_int_fast8_t
use,intrinsic::iso_c_binding,only:c_int_least16_t
use,intrinsic::iso_c_binding,only:c_int_fast16_t
use,intrinsic::iso_c_binding,only:c_int_least32_t
use,intrinsic::iso_c_binding,only:c_int_fast32_t
use,intrinsic::iso_c_binding,only:c_int_least64_t
use,intrinsic::iso_c_binding,only:c_int_fast64_t
use,intrinsic::iso_c_binding,only:c_int_least128_t
use,intrinsic::iso_c_binding,only:c_int_fast128_t
use,intrinsic::iso_c_binding,only:c_float
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_long_double
use,intrinsic::iso_c_binding,only:c_float_complex
use,intrinsic::iso_c_binding,only:c_double_complex
use,intrinsic::iso_c_binding,only:c_long_double_complex
use,intrinsic::iso_c_binding,only:c_bool
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_char
use,intrinsic::iso_c_binding,only:c_alert
use,intrinsic::iso_c_binding,only:c_backspace
use,intrinsic::iso_c_binding,only:c_form_feed
use,intrinsic::iso_c_binding,only:c_new_line
use,intrinsic::iso_c_binding,only:c_carriage_return
use,intrinsic::iso_c_binding,only:c_horizontal_tab
use,intrinsic::iso_c_binding,only:c_vertical_tab
use,intrinsic::iso_c_binding,only:c_float128
use,intrinsic::iso_c_binding,only:c_float128_complex
use,intrinsic::iso_c_binding,only:c_uint8_t
use,intrinsic::iso_c_binding,only:c_uint16_t
use,intrinsic::iso_c_binding,only:c_uint32_t
use,intrinsic::iso_c_binding,only:c_uint64_t
use,intrinsic::iso_c_binding,only:c_uint128_t
use,intrinsic::iso_c_binding,only:c_unsigned_char
use,intrinsic::iso_c_binding,only:c_unsigned_short
use,intrinsic::iso_c_binding,only:c_unsigned
use,intrinsic::iso_c_binding,only:c_unsigned_long
use,intrinsic::iso_c_binding,only:c_unsigned_long_long
use,intrinsic::iso_c_binding,only:c_uintmax_t
use,intrinsic::iso_c_binding,only:c_uint_fast8_t
use,intrinsic::iso_c_binding,only:c_uint_fast16_t
use,intrinsic::iso_c_binding,only:c_uint_fast32_t
use,intrinsic::iso_c_binding,only:c_uint_fast64_t
use,intrinsic::iso_c_binding,only:c_uint_fast128_t
use,intrinsic::iso_c_binding,only:c_uint_least8_t
use,intrinsic::iso_c_binding,only:c_uint_least16_t
use,intrinsic::iso_c_binding,only:c_uint_least32_t
use,intrinsic::iso_c_binding,only:c_uint_least64_t
use,intrinsic::iso_c_binding,only:c_uint_least128_t
use,intrinsic::iso_c_binding,only:c_f_procpointer
use w90gpu_c,only:w90gpu_ok
use w90gpu_c,only:w90gpu_err_fatal
use w90gpu_c,only:w90gpu_err_alloc
use w90gpu_c,only:w90gpu_err_input
use w90gpu_c,only:w90gpu_err_file
use w90gpu_c,only:w90gpu_num_nnmax
use w90gpu_c,only:w90gpu_prec_fp64
use w90gpu_c,only:w90gpu_prec_mixed
use w90gpu_c,only:w90gpu_control
use w90gpu_c,only:w90gpu_problem
use w90gpu_c,only:w90gpu_result
use w90gpu_c,only:w90gpu_control_defaults
use w90gpu_c,only:w90gpu_create
use w90gpu_c,only:w90gpu_destroy
use w90gpu_c,only:w90gpu_last_error
use w90gpu_c,only:w90gpu_kmesh_get
use w90gpu_c,only:w90gpu_wann_main
use w90gpu_c,only:w90gpu_project
use w90gpu_c,only:w90gpu_session_create
use w90gpu_c,only:w90gpu_session_destroy
use w90gpu_c,only:w90gpu_session_load
use w90gpu_c,only:w90gpu_session_solve
use w90gpu_c,only:w90gpu_session_fetch
use w90gpu_c,only:w90gpu_session_history
use w90gpu_c,only:c_strlen
use w90gpu_c,only:w90gpu_error_message
use w90gpu_io,only:dp
use w90gpu_io,only:bohr_ang
use w90gpu_io,only:k_text
use w90gpu_io,only:k_int
use w90gpu_io,only:k_real
use w90gpu_io,only:k_logical
use w90gpu_io,only:k_i1d
use w90gpu_io,only:k_r1d
use w90gpu_io,only:k_r2d
use w90gpu_io,only:k_i2d
use w90gpu_io,only:k_block
use w90gpu_io,only:w90_setting
use w90gpu_io,only:w90_settings
use w90gpu_io,only:lower
use w90gpu_io,only:datetime
use w90gpu_io,only:settings_add
use w90gpu_io,only:settings_clear
use w90gpu_io,only:settings_find
use w90gpu_io,only:read_win
use w90gpu_io,only:get_int
use w90gpu_io,only:get_real
use w90gpu_io,only:get_logical
use w90gpu_io,only:get_text
use w90gpu_io,only:get_i1d
use w90gpu_io,only:get_r2d
use w90gpu_io,only:get_block
use w90gpu_io,only:read_mmn
use w90gpu_io,only:read_amn
use w90gpu_io,only:write_chk
use w90gpu_io,only:read_chk
use w90gpu_io,only:write_nnkp
use w90gpu_io,only:write_xyz
use w90gpu_io,only:write_umat
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::__builtin_c_ptr
private::c_associated
private::c_funloc
private::c_funptr
private::c_f_pointer
private::c_loc
private::c_null_funptr
private::c_null_ptr
private::c_ptr
private::c_sizeof
private::operator(==)
private::operator(/=)
private::c_int8_t
private::c_int16_t
private::c_int32_t
private::c_int64_t
private::c_int128_t
private::c_int
private::c_short
private::c_long
private::c_long_long
private::c_signed_char
private::c_size_t
private::c_intmax_t
private::c_intptr_t
private::c_ptrdiff_t
private::c_int_least8_t
private::c_int_fast8_t
private::c_int_least16_t
private::c_int_fast16_t
private::c_int_least32_t
private::c_int_fast32_t
private::c_int_least64_t
private::c_int_fast64_t
private::c_int_least128_t
private::c_int_fast128_t
private::c_float
private::c_double
private::c_long_double
private::c_float_complex
private::c_double_complex
private::c_long_double_complex
private::c_bool
private::c_char
private::c_null_char
private::c_alert
private::c_backspace
private::c_form_feed
private::c_new_line
private::c_carriage_return
private::c_horizontal_tab
private::c_vertical_tab
private::c_float128
private::c_float128_complex
private::c_uint8_t
private::c_uint16_t
private::c_uint32_t
private::c_uint64_t
private::c_uint128_t
private::c_unsigned_char
private::c_unsigned_short
private::c_unsigned
private::c_unsigned_long
private::c_unsigned_long_long
private::c_uintmax_t
private::c_uint_fast8_t
private::c_uint_fast16_t
private::c_uint_fast32_t
private::c_uint_fast64_t
private::c_uint_fast128_t
private::c_uint_least8_t
private::c_uint_least16_t
private::c_uint_least32_t
private::c_uint_least64_t
private::c_uint_least128_t
private::c_f_procpointer
private::w90gpu_ok
private::w90gpu_err_fatal
private::w90gpu_err_alloc
private::w90gpu_err_input
private::w90gpu_err_file
private::w90gpu_num_nnmax
private::w90gpu_prec_fp64
private::w90gpu_prec_mixed
private::w90gpu_control
private::w90gpu_problem
private::w90gpu_result
private::w90gpu_control_defaults
private::w90gpu_create
private::w90gpu_destroy
private::w90gpu_last_error
private::w90gpu_kmesh_get
private::w90gpu_wann_main
private::w90gpu_project
private::w90gpu_session_create
private::w90gpu_session_destroy
private::w90gpu_session_load
private::w90gpu_session_solve
private::w90gpu_session_fetch
private::w90gpu_session_history
private::c_strlen
private::w90gpu_error_message
private::dp
private::bohr_ang
private::k_text
private::k_int
private::k_real
private::k_logical
private::k_i1d
private::k_r1d
private::k_r2d
private::k_i2d
private::k_block
private::w90_setting
private::w90_settings
private::lower
private::datetime
private::settings_add
private::settings_clear
private::settings_find
private::read_win
private::get_int
private::get_real
private::get_logical
private::get_text
private::get_i1d
private::get_r2d
private::get_block
private::read_mmn
private::read_amn
private::write_chk
private::read_chk
private::write_nnkp
private::write_xyz
private::write_umat
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
type::lib_common_type
character(128_4,1)::seedname="wannier                                                                                                                         "
complex(8),pointer::a_matrix(:,:,:)=>NULL()
complex(8),pointer::m_matrix_local(:,:,:,:)=>NULL()
complex(8),pointer::u_matrix(:,:,:)=>NULL()
complex(8),pointer::u_opt(:,:,:)=>NULL()
real(8),pointer::eigval(:,:)=>NULL()
integer(4)::mp_grid(1_8:3_8)=[INTEGER(4)::0_4,0_4,0_4]
integer(4)::num_bands=0_4
integer(4)::num_kpts=0_4
integer(4)::num_wann=0_4
integer(4)::num_proj=0_4
logical(4)::gamma_only=.false._4
logical(4)::have_disentangled=.false._4
logical(4)::setup_complete=.false._4
real(8),allocatable::kpt_latt(:,:)
real(8)::real_lattice(1_8:3_8,1_8:3_8)=reshape([REAL(8)::0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8,0._8],shape=[3,3])
integer(4)::nntot=0_4
integer(4),allocatable::nnlist(:,:)
integer(4),allocatable::nncell(:,:,:)
real(8),allocatable::bk(:,:,:)
real(8),allocatable::wb(:)
real(8)::wbtot=0._8
integer(4)::search_shells=36_4
integer(4)::num_shells=0_4
integer(4)::shell_list(1_8:6_8)=[INTEGER(4)::0_4,0_4,0_4,0_4,0_4,0_4]
real(8)::kmesh_tol=9.99999999999999954748111825886258685613938723690807819366455078125e-7_8
logical(4)::skip_b1_tests=.false._4
type(w90gpu_control)::control
logical(4)::control_set=.false._4
integer(4)::iprint=1_4
integer(4)::num_print_cycles=1_4
integer(4)::device=0_4
logical(4)::write_xyz=.false._4
logical(4)::write_u_matrices=.false._4
logical(4)::translate_home_cell=.false._4
character(16_4,1)::restart="                "
integer(4)::num_atoms=0_4
character(20_4,1),allocatable::atom_symbol(:)
real(8),allocatable::atom_cart(:,:)
real(8),allocatable::centres(:,:)
real(8),allocatable::spreads(:)
type(w90gpu_result)::result
real(8),allocatable::history(:,:)
type(w90_settings)::settings
logical(4)::have_win=.false._4
type(c_ptr)::ctx=__builtin_c_ptr(__address=0_8)
end type
intrinsic::null
private::null
private::report
private::w90_set_option_text
private::w90_set_option_logical
private::w90_set_option_int
private::w90_set_option_i1d
private::w90_set_option_i2d
private::w90_set_option_real
private::w90_set_option_r1d
private::w90_set_option_r2d
private::apply_settings
private::ensure_context
private::make_problem
interface w90_set_option
procedure::w90_set_option_text
procedure::w90_set_option_logical
procedure::w90_set_option_int
procedure::w90_set_option_i1d
procedure::w90_set_option_i2d
procedure::w90_set_option_real
procedure::w90_set_option_r1d
procedure::w90_set_option_r2d
end interface
contains
subroutine w90_get_fortran_stdout(istdout)
integer(4),intent(out)::istdout
end
subroutine w90_get_fortran_stderr(istderr)
integer(4),intent(out)::istderr
end
subroutine report(istdout,istderr,code,msg,ierr)
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(in)::code
character(*,1),intent(in)::msg
integer(4),intent(out)::ierr
end
subroutine w90_set_option_text(common_data,keyword,text)
type(lib_common_type),intent(inout)::common_data
character(*,1),intent(in)::keyword
character(*,1),intent(in)::text
end
subroutine w90_set_option_logical(common_data,keyword,bool)
type(lib_common_type),intent(inout)::common_data
character(*,1),intent(in)::keyword
logical(4),intent(in)::bool
end
subroutine w90_set_option_int(common_data,keyword,ival)
type(lib_common_type),intent(inout)::common_data
character(*,1),intent(in)::keyword
integer(4),intent(in)::ival
end
subroutine w90_set_option_i1d(common_data,keyword,arr)
type(lib_common_type),intent(inout)::common_data
character(*,1),intent(in)::keyword
integer(4),intent(in)::arr(:)
end
subroutine w90_set_option_i2d(common_data,keyword,arr)
type(lib_common_type),intent(inout)::common_data
character(*,1),intent(in)::keyword
integer(4),intent(in)::arr(:,:)
end
subroutine w90_set_option_real(common_data,keyword,rval)
type(lib_common_type),intent(inout)::common_data
character(*,1),intent(in)::keyword
real(8),intent(in)::rval
end
subroutine w90_set_option_r1d(common_data,keyword,arr)
type(lib_common_type),intent(inout)::common_data
character(*,1),intent(in)::keyword
real(8),intent(in)::arr(:)
end
subroutine w90_set_option_r2d(common_data,keyword,arr)
type(lib_common_type),intent(inout)::common_data
character(*,1),intent(in)::keyword
real(8),intent(in)::arr(:,:)
end
subroutine w90_input_setopt(common_data,seedname,comm,istdout,istderr,ierr)
type(lib_common_type),intent(inout)::common_data
character(*,1),intent(in)::seedname
integer(4),intent(in)::comm
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine w90_input_reader(common_data,istdout,istderr,ierr)
type(lib_common_type),intent(inout)::common_data
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine apply_settings(cd,ierr,msg)
type(lib_common_type),intent(inout)::cd
integer(4),intent(out)::ierr
character(*,1),intent(out)::msg
end
subroutine ensure_context(cd,istdout,istderr,ierr)
type(lib_common_type),intent(inout)::cd
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine w90_create_kmesh(common_data,istdout,istderr,ierr)
type(lib_common_type),intent(inout)::common_data
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine w90_get_nn(common_data,nn,istdout,istderr,ierr)
type(lib_common_type),intent(inout)::common_data
integer(4),intent(out)::nn
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine w90_get_nnkp(common_data,nnkp,istdout,istderr,ierr)
type(lib_common_type),intent(inout)::common_data
integer(4),intent(out)::nnkp(:,:)
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine w90_get_gkpb(common_data,gkpb,istdout,istderr,ierr)
type(lib_common_type),intent(inout)::common_data
integer(4),intent(out)::gkpb(:,:,:)
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine w90_set_m_local(common_data,m_orig)
type(lib_common_type),intent(inout)::common_data
complex(8),intent(inout),target::m_orig(:,:,:,:)
end
subroutine w90_set_u_matrix(common_data,u_matrix)
type(lib_common_type),intent(inout)::common_data
complex(8),intent(inout),target::u_matrix(:,:,:)
end
subroutine w90_set_u_opt(common_data,u_opt)
type(lib_common_type),intent(inout)::common_data
complex(8),intent(inout),target::u_opt(:,:,:)
end
subroutine w90_set_eigval(common_data,eigval)
type(lib_common_type),intent(inout)::common_data
real(8),intent(inout),target::eigval(:,:)
end
subroutine make_problem(cd,prob)
type(lib_common_type),intent(inout),target::cd
type(w90gpu_problem),intent(out)::prob
end
subroutine w90_project_overlap(common_data,istdout,istderr,ierr)
type(lib_common_type),intent(inout),target::common_data
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine w90_wannierise(common_data,istdout,istderr,ierr)
type(lib_common_type),intent(inout),target::common_data
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine w90_disentangle(common_data,istdout,istderr,ierr)
type(lib_common_type),intent(inout)::common_data
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine w90_plot(common_data,istdout,istderr,ierr)
type(lib_common_type),intent(inout)::common_data
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine w90_transport(common_data,istdout,istderr,ierr)
type(lib_common_type),intent(inout)::common_data
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine w90_get_centres(common_data,centres)
type(lib_common_type),intent(in)::common_data
real(8),intent(out)::centres(:,:)
end
subroutine w90_get_spreads(common_data,spreads)
type(lib_common_type),intent(in)::common_data
real(8),intent(out)::spreads(:)
end
subroutine w90_get_omega(common_data,omega)
type(lib_common_type),intent(in)::common_data
real(8),intent(out)::omega(1_8:4_8)
end
subroutine w90_write_chkpt(common_data,chkpt)
type(lib_common_type),intent(in)::common_data
character(*,1),intent(in)::chkpt
end
subroutine w90_read_chkpt(common_data,istdout,istderr,ierr)
type(lib_common_type),intent(inout)::common_data
integer(4),intent(in)::istdout
integer(4),intent(in)::istderr
integer(4),intent(out)::ierr
end
subroutine w90_write_nnkp(common_data)
type(lib_common_type),intent(in)::common_data
end
subroutine w90_print_conv(common_data,istdout,wall)
type(lib_common_type),intent(in)::common_data
integer(4),intent(in)::istdout
real(8),intent(in)::wall
end
end
