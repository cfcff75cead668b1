﻿!mod$ v1 sum:ed38b9e32766ea62
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module bgk_hip
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
integer(4),parameter::bgk_bc_driven=0_4
integer(4),parameter::bgk_bc_periodic=1_4
integer(4),parameter::bgk_bc_channel=2_4
integer(4),parameter::bgk_bc_inflow=3_4
integer(4),parameter::bgk_bc_couette=4_4
integer(4),parameter::bgk_bc_step=5_4
integer(4),parameter::bgk_bc_open=6_4
integer(4),parameter::bgk_prec_single=0_4
integer(4),parameter::bgk_prec_double=1_4
integer(4),parameter::bgk_prec_fp16=2_4
integer(4),parameter::bgk_prec_mixed2=3_4
integer(4),parameter::bgk_prec_bf16=4_4
integer(4),parameter::bgk_impl_original=0_4
integer(4),parameter::bgk_impl_fused=1_4
type,bind(c)::bgk_params_t
integer(4)::lx
integer(4)::ly
integer(4)::ipad
integer(4)::jpad
integer(4)::bc
integer(4)::obstacle
integer(4)::impl
integer(4)::prec
integer(4)::radius
integer(4)::device
integer(4)::mixed1_quirk
integer(4)::form
integer(4)::rounding
integer(4)::tblock
integer(4)::les
integer(4)::trick1
integer(4)::forcing_y
integer(4)::pad0
real(8)::svisc
real(8)::u0
real(8)::scale
real(8)::cte_s
end type
interface
function bgk_create(p,ctx) bind(c,name="bgk_create")
import::bgk_params_t
import::c_ptr
type(bgk_params_t),intent(in)::p
type(c_ptr),intent(out)::ctx
integer(4)::bgk_create
end
end interface
interface
function bgk_destroy(ctx) bind(c,name="bgk_destroy")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::bgk_destroy
end
end interface
interface
function bgk_last_error(ctx) bind(c,name="bgk_last_error")
import::c_ptr
type(c_ptr),value::ctx
type(c_ptr)::bgk_last_error
end
end interface
interface
function bgk_constants(ctx,out) bind(c,name="bgk_constants")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::out(1_8:4_8)
integer(4)::bgk_constants
end
end interface
interface
function bgk_build_obs(ctx) bind(c,name="bgk_build_obs")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::bgk_build_obs
end
end interface
interface
function bgk_set_obstacle(ctx,obs) bind(c,name="bgk_set_obstacle")
import::c_ptr
type(c_ptr),value::ctx
integer(4),intent(in)::obs(1_8:*)
integer(4)::bgk_set_obstacle
end
end interface
interface
function bgk_draglift(ctx,i0,i1,j0,j1,out) bind(c,name="bgk_draglift")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::i0
integer(4),value::i1
integer(4),value::j0
integer(4),value::j1
real(8),intent(out)::out(1_8:2_8)
integer(4)::bgk_draglift
end
end interface
interface
function bgk_fields(ctx,u,v,rho) bind(c,name="bgk_fields")
import::c_ptr
type(c_ptr),value::ctx
real(4),intent(out)::u(1_8:*)
real(4),intent(out)::v(1_8:*)
real(4),intent(out)::rho(1_8:*)
integer(4)::bgk_fields
end
end interface
interface
function bgk_init(ctx,init_v,amplitude) bind(c,name="bgk_init")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::init_v
real(8),value::amplitude
integer(4)::bgk_init
end
end interface
interface
function bgk_boundaries(ctx) bind(c,name="bgk_boundaries")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::bgk_boundaries
end
end interface
interface
function bgk_propagation(ctx) bind(c,name="bgk_propagation")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::bgk_propagation
end
end interface
interface
function bgk_collision(ctx,itime) bind(c,name="bgk_collision")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::itime
integer(4)::bgk_collision
end
end interface
interface
function bgk_run(ctx,nsteps) bind(c,name="bgk_run")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::nsteps
integer(4)::bgk_run
end
end interface
interface
function bgk_sync(ctx) bind(c,name="bgk_sync")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::bgk_sync
end
end interface
interface
function bgk_diagno(ctx,out) bind(c,name="bgk_diagno")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::out(1_8:4_8)
integer(4)::bgk_diagno
end
end interface
interface
function bgk_varm(ctx,u,v,rho) bind(c,name="bgk_varm")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::u(1_8:*)
real(8),intent(out)::v(1_8:*)
real(8),intent(out)::rho(1_8:*)
integer(4)::bgk_varm
end
end interface
interface
function bgk_prof_i(ctx,jc,u,v,rho) bind(c,name="bgk_prof_i")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::jc
real(8),intent(out)::u(1_8:*)
real(8),intent(out)::v(1_8:*)
real(8),intent(out)::rho(1_8:*)
integer(4)::bgk_prof_i
end
end interface
interface
function bgk_prof_j(ctx,ic,u,v,rho) bind(c,name="bgk_prof_j")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::ic
real(8),intent(out)::u(1_8:*)
real(8),intent(out)::v(1_8:*)
real(8),intent(out)::rho(1_8:*)
integer(4)::bgk_prof_j
end
end interface
interface
function bgk_probe(ctx,i0,j0,out) bind(c,name="bgk_probe")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::i0
integer(4),value::j0
real(8),intent(out)::out(1_8:3_8)
integer(4)::bgk_probe
end
end interface
interface
function bgk_check(ctx,umax) bind(c,name="bgk_check")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::umax
integer(4)::bgk_check
end
end interface
interface
function bgk_download(ctx,k,host) bind(c,name="bgk_download")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::k
type(c_ptr),value::host
integer(4)::bgk_download
end
end interface
interface
function bgk_upload(ctx,k,host) bind(c,name="bgk_upload")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::k
type(c_ptr),value::host
integer(4)::bgk_upload
end
end interface
interface
function bgk_storage_bytes(ctx) bind(c,name="bgk_storage_bytes")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::bgk_storage_bytes
end
end interface
contains
subroutine bgk_ok(ctx,rc,what)
type(c_ptr),intent(in)::ctx
integer(4),intent(in)::rc
character(*,1),intent(in)::what
end
end
