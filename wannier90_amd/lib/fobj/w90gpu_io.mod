﻿!mod$ v1 sum:7d55b91d8c4c39ed
!need$ f1de5abe9bfe2168 i iso_fortran_env
module w90gpu_io
use,intrinsic::iso_fortran_env,only:int32
use,intrinsic::iso_fortran_env,only:real64
private::int32
private::real64
integer(4),parameter::dp=8_4
real(8),parameter::bohr_ang=5.291772085899999478186828127945773303508758544921875e-1_8
integer(4),parameter,private::linelen=512_4
integer(4),parameter::k_text=0_4
integer(4),parameter::k_int=1_4
integer(4),parameter::k_real=2_4
integer(4),parameter::k_logical=3_4
integer(4),parameter::k_i1d=4_4
integer(4),parameter::k_r1d=5_4
integer(4),parameter::k_r2d=6_4
integer(4),parameter::k_i2d=7_4
integer(4),parameter::k_block=8_4
type::w90_setting
character(64_4,1)::keyword="                                                                "
integer(4)::kind=0_4
character(512_4,1)::text="                                                                                                                                                                                                                                                                                                                                                                                                                                                                                                                                "
integer(4)::ival=0_4
real(8)::rval=0._8
logical(4)::lval=.false._4
integer(4),allocatable::i1d(:)
integer(4),allocatable::i2d(:,:)
real(8),allocatable::r1d(:)
real(8),allocatable::r2d(:,:)
character(512_4,1),allocatable::lines(:)
end type
type::w90_settings
type(w90_setting),allocatable::e(:)
integer(4)::n=0_4
end type
private::fortran_exponent
private::recip_lattice
contains
pure function lower(s) result(r)
character(*,1),intent(in)::s
character(int(int(s%len,kind=8),kind=4),1)::r
end
subroutine datetime(cdate,ctime)
character(9_4,1),intent(out)::cdate
character(9_4,1),intent(out)::ctime
end
subroutine settings_add(s,entry)
type(w90_settings),intent(inout)::s
type(w90_setting),intent(in)::entry
end
subroutine settings_clear(s)
type(w90_settings),intent(inout)::s
end
function settings_find(s,keyword)
type(w90_settings),intent(in)::s
character(*,1),intent(in)::keyword
integer(4)::settings_find
end
subroutine read_win(path,s,ierr,msg)
character(*,1),intent(in)::path
type(w90_settings),intent(inout)::s
integer(4),intent(out)::ierr
character(*,1),intent(out)::msg
end
subroutine get_int(s,key,val,found,ierr)
type(w90_settings),intent(in)::s
character(*,1),intent(in)::key
integer(4),intent(inout)::val
logical(4),intent(out)::found
integer(4),intent(inout)::ierr
end
subroutine get_real(s,key,val,found,ierr)
type(w90_settings),intent(in)::s
character(*,1),intent(in)::key
real(8),intent(inout)::val
logical(4),intent(out)::found
integer(4),intent(inout)::ierr
end
subroutine fortran_exponent(t)
character(*,1),intent(inout)::t
end
subroutine get_logical(s,key,val,found,ierr)
type(w90_settings),intent(in)::s
character(*,1),intent(in)::key
logical(4),intent(inout)::val
logical(4),intent(out)::found
integer(4),intent(inout)::ierr
end
subroutine get_text(s,key,val,found)
type(w90_settings),intent(in)::s
character(*,1),intent(in)::key
character(*,1),intent(inout)::val
logical(4),intent(out)::found
end
subroutine get_i1d(s,key,val,n,found,ierr)
type(w90_settings),intent(in)::s
character(*,1),intent(in)::key
integer(4),intent(in)::n
integer(4),intent(inout)::val(1_8:int(n,kind=8))
logical(4),intent(out)::found
integer(4),intent(inout)::ierr
end
subroutine get_r2d(s,key,r,cols,rows,length_block,found,ierr)
type(w90_settings),intent(in)::s
character(*,1),intent(in)::key
integer(4),intent(in)::cols
integer(4),intent(in)::rows
real(8),intent(out)::r(1_8:int(cols,kind=8),1_8:int(rows,kind=8))
logical(4),intent(in)::length_block
logical(4),intent(out)::found
integer(4),intent(inout)::ierr
end
function get_block(s,key)
type(w90_settings),intent(in)::s
character(*,1),intent(in)::key
integer(4)::get_block
end
subroutine read_mmn(path,nb,nk,nntot,nnlist,nncell,m,ierr,msg)
character(*,1),intent(in)::path
integer(4),intent(in)::nb
integer(4),intent(in)::nk
integer(4),intent(in)::nntot
integer(4),intent(in)::nnlist(1_8:int(nk,kind=8),1_8:int(nntot,kind=8))
integer(4),intent(in)::nncell(1_8:3_8,1_8:int(nk,kind=8),1_8:int(nntot,kind=8))
complex(8),intent(out)::m(1_8:int(nb,kind=8),1_8:int(nb,kind=8),1_8:int(nntot,kind=8),1_8:int(nk,kind=8))
integer(4),intent(out)::ierr
character(*,1),intent(out)::msg
end
subroutine read_amn(path,nb,nk,nproj,a,ierr,msg)
character(*,1),intent(in)::path
integer(4),intent(in)::nb
integer(4),intent(in)::nk
integer(4),intent(in)::nproj
complex(8),intent(out)::a(1_8:int(nb,kind=8),1_8:int(nproj,kind=8),1_8:int(nk,kind=8))
integer(4),intent(out)::ierr
character(*,1),intent(out)::msg
end
subroutine write_chk(path,chkpt,num_bands,real_lattice,num_kpts,mp_grid,kpt_latt,nntot,num_wann,u,m,centres,spreads)
character(*,1),intent(in)::path
character(*,1),intent(in)::chkpt
integer(4),intent(in)::num_bands
real(8),intent(in)::real_lattice(1_8:3_8,1_8:3_8)
integer(4),intent(in)::num_kpts
integer(4),intent(in)::mp_grid(1_8:3_8)
real(8),intent(in)::kpt_latt(1_8:3_8,1_8:int(num_kpts,kind=8))
integer(4),intent(in)::nntot
integer(4),intent(in)::num_wann
complex(8),intent(in)::u(1_8:int(num_wann,kind=8),1_8:int(num_wann,kind=8),1_8:int(num_kpts,kind=8))
complex(8),intent(in)::m(1_8:int(num_wann,kind=8),1_8:int(num_wann,kind=8),1_8:int(nntot,kind=8),1_8:int(num_kpts,kind=8))
real(8),intent(in)::centres(1_8:3_8,1_8:int(num_wann,kind=8))
real(8),intent(in)::spreads(1_8:int(num_wann,kind=8))
end
subroutine read_chk(path,num_wann,num_kpts,nntot,u,m,ierr,msg)
character(*,1),intent(in)::path
integer(4),intent(in)::num_wann
integer(4),intent(in)::num_kpts
integer(4),intent(in)::nntot
complex(8),intent(out)::u(1_8:int(num_wann,kind=8),1_8:int(num_wann,kind=8),1_8:int(num_kpts,kind=8))
complex(8),intent(out)::m(1_8:int(num_wann,kind=8),1_8:int(num_wann,kind=8),1_8:int(nntot,kind=8),1_8:int(num_kpts,kind=8))
integer(4),intent(out)::ierr
character(*,1),intent(out)::msg
end
subroutine recip_lattice(real_lattice,recip)
real(8),intent(in)::real_lattice(1_8:3_8,1_8:3_8)
real(8),intent(out)::recip(1_8:3_8,1_8:3_8)
end
subroutine write_nnkp(path,real_lattice,num_kpts,kpt_latt,nntot,nnlist,nncell)
character(*,1),intent(in)::path
real(8),intent(in)::real_lattice(1_8:3_8,1_8:3_8)
integer(4),intent(in)::num_kpts
real(8),intent(in)::kpt_latt(1_8:3_8,1_8:int(num_kpts,kind=8))
integer(4),intent(in)::nntot
integer(4),intent(in)::nnlist(1_8:int(num_kpts,kind=8),1_8:int(nntot,kind=8))
integer(4),intent(in)::nncell(1_8:3_8,1_8:int(num_kpts,kind=8),1_8:int(nntot,kind=8))
end
subroutine write_xyz(path,num_wann,centres,real_lattice,translate_home_cell,nat,symbols,pos_cart)
character(*,1),intent(in)::path
integer(4),intent(in)::num_wann
real(8),intent(in)::centres(1_8:3_8,1_8:int(num_wann,kind=8))
real(8),intent(in)::real_lattice(1_8:3_8,1_8:3_8)
logical(4),intent(in)::translate_home_cell
integer(4),intent(in)::nat
character(*,1),intent(in)::symbols(:)
real(8),intent(in)::pos_cart(:,:)
end
subroutine write_umat(path,num_wann,num_kpts,kpt_latt,u_matrix)
character(*,1),intent(in)::path
integer(4),intent(in)::num_wann
integer(4),intent(in)::num_kpts
real(8),intent(in)::kpt_latt(1_8:3_8,1_8:int(num_kpts,kind=8))
complex(8),intent(in)::u_matrix(1_8:int(num_wann,kind=8),1_8:int(num_wann,kind=8),1_8:int(num_kpts,kind=8))
end
end
