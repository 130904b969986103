﻿!mod$ v1 sum:ad019ea2035d982d
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module w90gpu_c
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
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
integer(4),parameter::w90gpu_ok=0_4
integer(4),parameter::w90gpu_err_fatal=1_4
integer(4),parameter::w90gpu_err_alloc=2_4
integer(4),parameter::w90gpu_err_input=5_4
integer(4),parameter::w90gpu_err_file=6_4
integer(4),parameter::w90gpu_num_nnmax=12_4
integer(4),parameter::w90gpu_prec_fp64=0_4
integer(4),parameter::w90gpu_prec_mixed=1_4
type,bind(c)::w90gpu_control
integer(4)::num_iter
integer(4)::num_cg_steps
integer(4)::conv_window
integer(4)::lfixstep
real(8)::conv_tol
real(8)::trial_step
real(8)::fixed_step
real(8)::conv_noise_amp
integer(4)::precond
integer(4)::guiding_centres
integer(4)::selective_loc
integer(4)::precision
end type
type,bind(c)::w90gpu_problem
integer(4)::num_wann
integer(4)::num_kpts
integer(4)::nntot
type(c_ptr)::nnlist
type(c_ptr)::bk
type(c_ptr)::wb
real(8)::wbtot
end type
type,bind(c)::w90gpu_result
real(8)::omega_total
real(8)::omega_invariant
real(8)::omega_d
real(8)::omega_od
real(8)::omega_tilde
real(8)::gcnorm1
real(8)::delta_omega
real(8)::solve_seconds
integer(4)::iterations
integer(4)::converged
integer(4)::unitarity_ok
integer(4)::status
integer(4)::bulk_iterations
integer(4)::reserved
end type
interface
subroutine w90gpu_control_defaults(ctl) bind(c,name="w90gpu_control_defaults")
import::w90gpu_control
type(w90gpu_control),intent(out)::ctl
end
end interface
interface
function w90gpu_create(device,ctx) bind(c,name="w90gpu_create") result(rc)
import::c_ptr
integer(4),value::device
type(c_ptr),intent(out)::ctx
integer(4)::rc
end
end interface
interface
function w90gpu_destroy(ctx) bind(c,name="w90gpu_destroy") result(rc)
import::c_ptr
type(c_ptr),value::ctx
integer(4)::rc
end
end interface
interface
function w90gpu_last_error(ctx) bind(c,name="w90gpu_last_error") result(msg)
import::c_ptr
type(c_ptr),value::ctx
type(c_ptr)::msg
end
end interface
interface
function w90gpu_kmesh_get(real_lattice,num_kpts,kpt_latt,search_shells,tol,num_shells,shell_list,skip_b1_tests,nntot,nnlist,nncell,bk,wb,wbtot,errmsg,errlen) bind(c,name="w90gpu_kmesh_get") result(rc)
real(8),intent(in)::real_lattice(1_8:*)
integer(4),value::num_kpts
real(8),intent(in)::kpt_latt(1_8:*)
integer(4),value::search_shells
real(8),value::tol
integer(4),value::num_shells
integer(4),intent(in)::shell_list(1_8:*)
integer(4),value::skip_b1_tests
integer(4),intent(out)::nntot
integer(4),intent(out)::nnlist(1_8:*)
integer(4),intent(out)::nncell(1_8:*)
real(8),intent(out)::bk(1_8:*)
real(8),intent(out)::wb(1_8:*)
real(8),intent(out)::wbtot
character(1_8,1),intent(out)::errmsg(1_8:*)
integer(4),value::errlen
integer(4)::rc
end
end interface
interface
function w90gpu_wann_main(ctx,prob,ctl,m_matrix,u_matrix,centres,spreads,res) bind(c,name="w90gpu_wann_main") result(rc)
import::c_ptr
import::w90gpu_control
import::w90gpu_problem
import::w90gpu_result
type(c_ptr),value::ctx
type(w90gpu_problem),intent(in)::prob
type(w90gpu_control),intent(in)::ctl
complex(8),intent(inout)::m_matrix(1_8:*)
complex(8),intent(inout)::u_matrix(1_8:*)
real(8),intent(out)::centres(1_8:*)
real(8),intent(out)::spreads(1_8:*)
type(w90gpu_result),intent(out)::res
integer(4)::rc
end
end interface
interface
function w90gpu_project(ctx,prob,a_matrix,u_matrix,m_matrix) bind(c,name="w90gpu_project") result(rc)
import::c_ptr
import::w90gpu_problem
type(c_ptr),value::ctx
type(w90gpu_problem),intent(in)::prob
complex(8),intent(in)::a_matrix(1_8:*)
complex(8),intent(out)::u_matrix(1_8:*)
complex(8),intent(inout)::m_matrix(1_8:*)
integer(4)::rc
end
end interface
interface
function w90gpu_session_create(ctx,prob,ctl,sess) bind(c,name="w90gpu_session_create") result(rc)
import::c_ptr
import::w90gpu_control
import::w90gpu_problem
type(c_ptr),value::ctx
type(w90gpu_problem),intent(in)::prob
type(w90gpu_control),intent(in)::ctl
type(c_ptr),intent(out)::sess
integer(4)::rc
end
end interface
interface
function w90gpu_session_destroy(sess) bind(c,name="w90gpu_session_destroy") result(rc)
import::c_ptr
type(c_ptr),value::sess
integer(4)::rc
end
end interface
interface
function w90gpu_session_load(sess,m_matrix,u_matrix) bind(c,name="w90gpu_session_load") result(rc)
import::c_ptr
type(c_ptr),value::sess
complex(8),intent(in)::m_matrix(1_8:*)
complex(8),intent(in)::u_matrix(1_8:*)
integer(4)::rc
end
end interface
interface
function w90gpu_session_solve(sess,res) bind(c,name="w90gpu_session_solve") result(rc)
import::c_ptr
import::w90gpu_result
type(c_ptr),value::sess
type(w90gpu_result),intent(out)::res
integer(4)::rc
end
end interface
interface
function w90gpu_session_fetch(sess,m_matrix,u_matrix,centres,spreads) bind(c,name="w90gpu_session_fetch") result(rc)
import::c_ptr
type(c_ptr),value::sess
complex(8),intent(out)::m_matrix(1_8:*)
complex(8),intent(out)::u_matrix(1_8:*)
real(8),intent(out)::centres(1_8:*)
real(8),intent(out)::spreads(1_8:*)
integer(4)::rc
end
end interface
interface
function w90gpu_session_history(sess,out,rows) bind(c,name="w90gpu_session_history") result(n)
import::c_ptr
type(c_ptr),value::sess
real(8),intent(out)::out(1_8:*)
integer(4),value::rows
integer(4)::n
end
end interface
interface
function c_strlen(s) bind(c,name="strlen") result(n)
import::c_ptr
type(c_ptr),value::s
integer(8)::n
end
end interface
contains
function w90gpu_error_message(ctx) result(msg)
type(c_ptr),intent(in)::ctx
character(:,1),allocatable::msg
end
end
