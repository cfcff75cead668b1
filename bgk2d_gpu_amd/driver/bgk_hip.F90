! bgk_hip.F90 -- ISO_C_BINDING interface to libbgk2d_hip.so (include/bgk2d.h).
!
! The reference's only existing C binding (SRC/get_mem_mod.F90:7) is the
! model: one `interface ... bind(c)` block per entry point, plain C types.
! The helpers at the bottom turn a nonzero status into `stop` with the
! library's message, as the reference stops on fatal input errors
! (SRC/input.F90:83-86, SRC/init.F90:68-71).
module bgk_hip
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: BGK_BC_DRIVEN = 0, BGK_BC_PERIODIC = 1, &
       BGK_BC_CHANNEL = 2, BGK_BC_INFLOW = 3, BGK_BC_COUETTE = 4, BGK_BC_STEP = 5, &
       BGK_BC_OPEN = 6
  integer(c_int32_t), parameter :: BGK_PREC_SINGLE = 0, BGK_PREC_DOUBLE = 1, &
       BGK_PREC_FP16 = 2, BGK_PREC_MIXED2 = 3, BGK_PREC_BF16 = 4
  integer(c_int32_t), parameter :: BGK_IMPL_ORIGINAL = 0, BGK_IMPL_FUSED = 1

  type, bind(c) :: bgk_params_t
     integer(c_int32_t) :: lx, ly, ipad, jpad, bc, obstacle, impl, prec
     integer(c_int32_t) :: radius, device, mixed1_quirk, form, rounding, tblock
     integer(c_int32_t) :: les, trick1, forcing_y, noshift
     real(c_double)     :: svisc, u0, scale, cte_s
  end type bgk_params_t

  interface
     integer(c_int) function bgk_create(p, ctx) bind(c, name='bgk_create')
       import :: c_int, c_ptr, bgk_params_t
       type(bgk_params_t), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
     end function
     integer(c_int) function bgk_destroy(ctx) bind(c, name='bgk_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     type(c_ptr) function bgk_last_error(ctx) bind(c, name='bgk_last_error')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function bgk_constants(ctx, out) bind(c, name='bgk_constants')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out(4)
     end function
     integer(c_int) function bgk_build_obs(ctx) bind(c, name='bgk_build_obs')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function bgk_set_obstacle(ctx, obs) bind(c, name='bgk_set_obstacle')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: obs(*)
     end function
     integer(c_int) function bgk_draglift(ctx, i0, i1, j0, j1, out) bind(c, name='bgk_draglift')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: i0, i1, j0, j1
       real(c_double), intent(out) :: out(2)
     end function
     integer(c_int) function bgk_fields(ctx, u, v, rho) bind(c, name='bgk_fields')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(out) :: u(*), v(*), rho(*)
     end function
     integer(c_int) function bgk_init(ctx, init_v, amplitude) bind(c, name='bgk_init')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: init_v
       real(c_double), value :: amplitude
     end function
     integer(c_int) function bgk_boundaries(ctx) bind(c, name='bgk_boundaries')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function bgk_propagation(ctx) bind(c, name='bgk_propagation')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function bgk_collision(ctx, itime) bind(c, name='bgk_collision')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: itime
     end function
     integer(c_int) function bgk_run(ctx, nsteps) bind(c, name='bgk_run')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsteps
     end function
     integer(c_int) function bgk_sync(ctx) bind(c, name='bgk_sync')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function bgk_diagno(ctx, out) bind(c, name='bgk_diagno')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out(4)
     end function
     integer(c_int) function bgk_varm(ctx, u, v, rho) bind(c, name='bgk_varm')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: u(*), v(*), rho(*)
     end function
     integer(c_int) function bgk_prof_i(ctx, jc, u, v, rho) bind(c, name='bgk_prof_i')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: jc
       real(c_double), intent(out) :: u(*), v(*), rho(*)
     end function
     integer(c_int) function bgk_prof_j(ctx, ic, u, v, rho) bind(c, name='bgk_prof_j')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ic
       real(c_double), intent(out) :: u(*), v(*), rho(*)
     end function
     integer(c_int) function bgk_probe(ctx, i0, j0, out) bind(c, name='bgk_probe')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: i0, j0
       real(c_double), intent(out) :: out(3)
     end function
     integer(c_int) function bgk_check(ctx, umax) bind(c, name='bgk_check')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: umax
     end function
     integer(c_int) function bgk_download(ctx, k, host) bind(c, name='bgk_download')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: k
       type(c_ptr), value :: host
     end function
     integer(c_int) function bgk_upload(ctx, k, host) bind(c, name='bgk_upload')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: k
       type(c_ptr), value :: host
     end function
     integer(c_int) function bgk_storage_bytes(ctx) bind(c, name='bgk_storage_bytes')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
  end interface

contains

  ! stop with the library's message when a call fails
  subroutine bgk_ok(ctx, rc, what)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=256) :: text
    integer :: n
    if (rc == 0) return
    text = ''
    if (c_associated(ctx)) then
       call c_f_pointer(bgk_last_error(ctx), msg, [256])
       do n = 1, 256
          if (msg(n) == c_null_char) exit
          text(n:n) = msg(n)
       end do
    end if
    write(6,*) 'ERROR: ', what, ' failed, status', rc, ': ', trim(text)
    error stop 1
  end subroutine bgk_ok

end module bgk_hip
