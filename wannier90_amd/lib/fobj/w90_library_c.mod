﻿!mod$ v1 sum:b92451e51e165643
!need$ e34ef4e13c24167b n w90_library
!need$ 0bde2ac47243ead2 i iso_c_binding
module w90_library_c
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
use w90_library,only:lib_common_type
use w90_library,only:w90_set_option
use w90_library,only:w90_get_fortran_stdout
use w90_library,only:w90_get_fortran_stderr
use w90_library,only:w90_input_setopt
use w90_library,only:w90_input_reader
use w90_library,only:w90_create_kmesh
use w90_library,only:w90_get_nn
use w90_library,only:w90_get_nnkp
use w90_library,only:w90_get_gkpb
use w90_library,only:w90_set_m_local
use w90_library,only:w90_set_u_matrix
use w90_library,only:w90_set_u_opt
use w90_library,only:w90_set_eigval
use w90_library,only:w90_project_overlap
use w90_library,only:w90_wannierise
use w90_library,only:w90_disentangle
use w90_library,only:w90_plot
use w90_library,only:w90_transport
use w90_library,only:w90_get_centres
use w90_library,only:w90_get_spreads
use w90_library,only:w90_get_omega
use w90_library,only:w90_write_chkpt
use w90_library,only:w90_read_chkpt
use w90_library,only:w90_write_nnkp
use w90_library,only:w90_print_conv
use w90_library,only:w90_library$w90_library$w90_set_option_i1d=>w90_set_option_i1d
use w90_library,only:w90_library$w90_library$w90_set_option_int=>w90_set_option_int
use w90_library,only:w90_library$w90_library$w90_set_option_r2d=>w90_set_option_r2d
use w90_library,only:w90_library$w90_library$w90_set_option_real=>w90_set_option_real
private::w90_library$w90_library$w90_set_option_i1d
private::w90_library$w90_library$w90_set_option_int
private::w90_library$w90_library$w90_set_option_r2d
private::w90_library$w90_library$w90_set_option_real
contains
function w90_create() bind(c)
type(c_ptr)::w90_create
end
subroutine w90_delete(common_cptr) bind(c)
type(c_ptr),value::common_cptr
end
subroutine cdisentangle(common_cptr,ierr) bind(c)
type(c_ptr),value::common_cptr
integer(4)::ierr
end
subroutine cproject_overlap(common_cptr,ierr) bind(c)
type(c_ptr),value::common_cptr
integer(4)::ierr
end
subroutine cwannierise(common_cptr,ierr) bind(c)
type(c_ptr),value::common_cptr
integer(4)::ierr
end
subroutine cinput_setopt(common_cptr,seedname,ierr,comm) bind(c)
type(c_ptr),value::common_cptr
character(*,1)::seedname
integer(4)::ierr
integer(4),value::comm
end
subroutine cinput_reader(common_cptr,ierr) bind(c)
type(c_ptr),value::common_cptr
integer(4)::ierr
end
subroutine cset_eigval(common_cptr,eigval_cptr) bind(c)
type(c_ptr),value::common_cptr
type(c_ptr),value::eigval_cptr
end
subroutine cset_m_local(common_cptr,m_cptr) bind(c)
type(c_ptr),value::common_cptr
type(c_ptr),value::m_cptr
end
subroutine cset_u_matrix(common_cptr,a_cptr) bind(c)
type(c_ptr),value::common_cptr
type(c_ptr),value::a_cptr
end
subroutine cset_u_opt(common_cptr,a_cptr) bind(c)
type(c_ptr),value::common_cptr
type(c_ptr),value::a_cptr
end
subroutine cget_nn(common_cptr,n) bind(c)
type(c_ptr),value::common_cptr
type(c_ptr),value::n
end
subroutine cget_nnkp(common_cptr,nnkp) bind(c)
type(c_ptr),value::common_cptr
type(c_ptr),value::nnkp
end
subroutine cget_spreads(common_cptr,spreads) bind(c)
type(c_ptr),value::common_cptr
type(c_ptr),value::spreads
end
subroutine cget_centres(common_cptr,centres) bind(c)
type(c_ptr),value::common_cptr
type(c_ptr),value::centres
end
subroutine cget_omega(common_cptr,omega) bind(c)
type(c_ptr),value::common_cptr
type(c_ptr),value::omega
end
subroutine cset_option_int(common_cptr,keyword,ival) bind(c)
type(c_ptr),value::common_cptr
character(*,1)::keyword
integer(4),value::ival
end
subroutine cset_option_int3(common_cptr,keyword,ival) bind(c)
type(c_ptr),value::common_cptr
character(*,1)::keyword
type(c_ptr)::ival
end
subroutine cset_option_intx(common_cptr,keyword,ival,ict) bind(c)
type(c_ptr),value::common_cptr
character(*,1)::keyword
type(c_ptr),value::ival
integer(4),value::ict
end
subroutine cset_option_float(common_cptr,keyword,ival) bind(c)
type(c_ptr),value::common_cptr
character(*,1)::keyword
real(8),value::ival
end
subroutine cset_option_float33(common_cptr,keyword,ival) bind(c)
type(c_ptr),value::common_cptr
character(*,1)::keyword
type(c_ptr)::ival
end
subroutine cset_option_floatxy(common_cptr,keyword,ival,x,y) bind(c)
type(c_ptr),value::common_cptr
character(*,1)::keyword
type(c_ptr),value::ival
integer(4),value::x
integer(4),value::y
end
end
