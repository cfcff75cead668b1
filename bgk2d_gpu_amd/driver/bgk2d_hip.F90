! bgk2d_hip.F90 -- BGK2D driver on the MI355X library.
!
! Keeps the reference driver's structure and inputs (SRC/bgk2d.F90:43-143):
! reads `bgk.input` namelist &parameters with the same names and defaults
! (SRC/input.F90:44-75), then setup -> initialize -> time loop
! [boundaries, propagation, collision(itime), diagnostic] -> final
! diagnostics -> finalize, writing diagno.dat, probe.dat, u_med.dat,
! prof_i.dat, prof_j.dat, save.XXXXXXXX.bin and bgk.perf in the reference's
! formats.  The reference's compile-time choices (-DLDC/-DTGV/..., SINGLE/
! DOUBLE/MIXED1/MIXED2, ORIG/FUSED) are read at run time from an optional
! second namelist &gpu in the same file:
!     case_name      'LDC' 'TGV' 'POF' 'VKS' 'COU' 'BFS', or 'CYL': the
!                    open channel past a cylinder (parabolic inflow of
!                    peak u0, zero-gradient outflow, free-slip walls)
!                    (default 'LDC')
!     precision      'SINGLE' 'DOUBLE' 'MIXED1' 'MIXED2' 'FP16' 'BF16'
!     implementation 'ORIGINAL' 'FUSED'                    (default 'FUSED')
!     scale          FP16/BF16 storage scale, a power of two (default 1)
!     amplitude      Taylor-Green amplitude (default 0.1, as SRC/init.F90)
!     phases         1: call boundaries/propagation/collision every step;
!                    0: advance between diagnostic steps with bgk_run
!     device         HIP device ordinal
!     form           0: reference operation order; 1: deviation form
!     rounding       FP16/BF16 stores: 0 nearest even; 1 stochastic
!     obs_x, obs_y   cylinder centre of VKS/CYL (default: the reference's
!                    2l/5, m/2, SRC/build.obs.F90:174-175); radius from
!                    &parameters
!     drag           1: write drag.lift.dat at every icheck step, as the
!                    DRAG build does (SRC/diagnostic.F90:80-82)
!     les            1: Smagorinsky LES collision, as the LES build does,
!                    with &parameters cteS
!     trick1         1: driven-cavity corner order of the TRICK1 build
!     forcing_y      1: POF body force along y, as the FORCING_Y build does
!     shape          obstacle of VKS/CYL: 'CYLINDER' (default), 'SQUARE',
!                    'HALF_CIRCLE', 'POROUS' (the SQUARE / HALF_CIRCLE /
!                    POROUS builds; POROUS draws from the Park-Miller
!                    minimal standard generator, see include/bgk2d.h)
!     noshift        1: whole populations f stored instead of f - w, as
!                    the NOSHIFT build does (form 0 only)
program bgk2d_hip
  use iso_c_binding
  use bgk_hip
  implicit none

  integer :: lx, ly, itfin, ivtim, isignal, itsave, icheck, irestart, init_v
  integer :: flag1, flag2, flag3, ipad, jpad, radius
  real(8) :: svisc, u0, cteS
  character(len=16) :: case_name, precision, implementation
  real(8) :: scale, amplitude
  integer :: phases, device, form, rounding, obs_x, obs_y, drag, les, trick1, forcing_y, noshift
  character(len=16) :: shape
  namelist /gpu/ case_name, precision, implementation, scale, amplitude, phases, device, &
       form, rounding, obs_x, obs_y, drag, les, trick1, forcing_y, shape, noshift

  type(bgk_params_t) :: p
  type(c_ptr) :: ctx
  integer :: itime, itstart, l, m, nextev, nrun, ios
  integer(8) :: c0, c1, crate
  real(8) :: time_loop, time_init
  real(8) :: consts(4)

  call system_clock(c0, crate)
  ! ------------------------------------------------------------- setup
  open(16, file='bgk.log', status='unknown')
  open(61, file='prof_i.dat', status='unknown')
  open(68, file='probe.dat', status='unknown')
  open(64, file='prof_j.dat', status='unknown')
  open(62, file='u_med.dat', status='unknown')
  open(63, file='diagno.dat', status='unknown')
  open(66, file='drag.lift.dat', status='unknown')

  case_name = 'LDC'
  precision = 'SINGLE'
  implementation = 'FUSED'
  scale = 1.0d0
  amplitude = 0.1d0
  phases = 0
  device = 0
  form = 0
  rounding = 0
  obs_x = -1
  obs_y = -1
  drag = 0
  les = 0
  trick1 = 0
  forcing_y = 0
  noshift = 0
  shape = 'CYLINDER' 
  open(15, file='bgk.input', status='old', iostat=ios)
  if (ios /= 0) error stop 'bgk.input not found'
  read(15, gpu, iostat=ios)
  close(15)

  call read_parameters()
  l = lx
  m = ly

  p%lx = lx
  p%ly = ly
  p%ipad = ipad
  p%jpad = jpad
  p%radius = radius
  p%device = device
  p%mixed1_quirk = 0
  p%form = form
  p%rounding = rounding
  p%tblock = 0
  p%les = les
  p%trick1 = trick1
  p%forcing_y = forcing_y
  p%noshift = noshift
  p%obstacle = 0
  p%scale = scale
  select case (trim(case_name))
  case ('LDC'); p%bc = BGK_BC_DRIVEN
  case ('TGV'); p%bc = BGK_BC_PERIODIC
  case ('POF'); p%bc = BGK_BC_CHANNEL
  case ('VKS'); p%bc = BGK_BC_INFLOW; p%obstacle = 1
  case ('COU'); p%bc = BGK_BC_COUETTE
  case ('BFS'); p%bc = BGK_BC_STEP
  case ('CYL'); p%bc = BGK_BC_OPEN; p%obstacle = 1
  case default; error stop 'unknown case_name'
  end select
  if (p%obstacle == 1) then
     select case (trim(shape))
     case ('CYLINDER')
     case ('SQUARE'); p%obstacle = 2
     case ('HALF_CIRCLE'); p%obstacle = 3
     case default; error stop 'unknown shape'
     end select
  end if
  select case (trim(precision))
  case ('SINGLE'); p%prec = BGK_PREC_SINGLE
  case ('DOUBLE'); p%prec = BGK_PREC_DOUBLE
  case ('MIXED1'); p%prec = BGK_PREC_FP16; p%mixed1_quirk = 1
  case ('FP16');   p%prec = BGK_PREC_FP16
  case ('MIXED2'); p%prec = BGK_PREC_MIXED2
  case ('BF16');   p%prec = BGK_PREC_BF16
  case default; error stop 'unknown precision'
  end select
  select case (trim(implementation))
  case ('ORIGINAL', 'ORIG'); p%impl = BGK_IMPL_ORIGINAL
  case ('FUSED');            p%impl = BGK_IMPL_FUSED
  case default; error stop 'unknown implementation'
  end select
  call read_parameters_kind()
  ! save.*.bin / restore.bin hold the storage kind's raw values in the
  ! layout of SRC/save_raw.F90, which has no field for a scale: with 16-bit
  ! storage they are written or read at scale 1 only
  if ((p%prec == BGK_PREC_FP16 .or. p%prec == BGK_PREC_BF16) .and. scale /= 1.0d0 .and. &
      (irestart == 1 .or. (itsave > 0 .and. itsave <= itfin))) &
     error stop 'scale /= 1: save/restore files hold 16-bit populations at scale 1 only'
  p%svisc = svisc
  p%u0 = u0
  p%cte_s = cteS

  ctx = c_null_ptr
  call bgk_ok(ctx, bgk_create(p, ctx), 'bgk_create')
  write(6,*)  'INFO: lattice', l, m, ' case ', trim(case_name), ' ', trim(precision), ' ', &
       trim(implementation)
  write(16,*) 'INFO: lattice', l, m, ' case ', trim(case_name), ' ', trim(precision), ' ', &
       trim(implementation)

  ! -------------------------------------------------------- initialize
  if (p%obstacle >= 1) then
     if (p%obstacle == 1 .and. obs_x > 0 .and. obs_y > 0) then
        call set_cylinder()
     else
        call bgk_ok(ctx, bgk_build_obs(ctx), 'bgk_build_obs')
        obs_x = 2 * l / 5
        obs_y = m / 2
     end if
  end if
  if (irestart == 1) then
     call restore_raw(itstart)
     write(6,*) 'INFO: restoring at timestep ', itstart
  else
     itstart = 0
     call bgk_ok(ctx, bgk_init(ctx, int(init_v, c_int32_t), amplitude), 'bgk_init')
     call varm_out(itstart)
     call prof_i_out(itstart, m/2)
     call prof_j_out(itstart, l/2)
  end if
  call bgk_ok(ctx, bgk_constants(ctx, consts), 'bgk_constants')
  write(16,*) 'INFO: omega, fgrad, u00', consts(1), consts(2), consts(3)
  call system_clock(c1)
  time_init = real(c1 - c0, 8) / real(crate, 8)

  ! --------------------------------------------------------- time loop
  call system_clock(c0)
  itime = itstart
  do while (itime < itfin)
     if (phases == 1) then
        itime = itime + 1
        call bgk_ok(ctx, bgk_boundaries(ctx), 'bgk_boundaries')
        call bgk_ok(ctx, bgk_propagation(ctx), 'bgk_propagation')
        call bgk_ok(ctx, bgk_collision(ctx, int(itime, c_int32_t)), 'bgk_collision')
     else
        ! advance to the next step at which a diagnostic or a save is due
        nextev = min(itfin, next_multiple(itime, ivtim), next_multiple(itime, icheck), &
             next_multiple(itime, itsave), next_multiple(itime, isignal))
        nrun = nextev - itime
        call bgk_ok(ctx, bgk_run(ctx, int(nrun, c_int32_t)), 'bgk_run')
        itime = nextev
     end if
     call diagnostic(itime)
     if (mod(itime, isignal) == 0) then
        call bgk_ok(ctx, bgk_sync(ctx), 'bgk_sync')
        call system_clock(c1)
        write(6,'(a,i9,a,f12.3)') ' INFO: step ', itime, '  elapsed (s) ', &
             real(c1 - c0, 8) / real(crate, 8)
     end if
  end do
  call bgk_ok(ctx, bgk_sync(ctx), 'bgk_sync')
  call system_clock(c1)
  time_loop = real(c1 - c0, 8) / real(crate, 8)

  ! final diagnostic (SRC/bgk2d.F90:122-125)
  call diagno_out(itime)
  call varm_out(itime)
  call prof_i_out(itime, m/2)
  call prof_j_out(itime, l/2)
  call finalize()
  call bgk_ok(ctx, bgk_destroy(ctx), 'bgk_destroy')
  write(6,*) "That's all folks!!!!"

contains

  integer function next_multiple(t, every)
    integer, intent(in) :: t, every
    if (every <= 0) then
       next_multiple = huge(1)
    else
       next_multiple = (t / every + 1) * every
    end if
  end function next_multiple

  ! &parameters with the reference's defaults (SRC/input.F90:51-61); values
  ! read into real(8) here, and again in the compute kind below
  subroutine read_parameters()
    namelist /parameters/ svisc, u0, itfin, ivtim, isignal, itsave, icheck, irestart, &
         init_v, lx, ly, flag1, flag2, flag3, ipad, jpad, radius, cteS
    flag1 = 0
    flag2 = 0
    flag3 = 0
    init_v = 0
    cteS = 0.1   ! a default-real literal, as SRC/input.F90:55
    ipad = 0
    jpad = 0
    radius = 16
    open(15, file='bgk.input', status='old')
    read(15, parameters)
    close(15)
  end subroutine read_parameters

  ! svisc and u0 are converted from their text directly into the compute
  ! kind, as the reference's namelist read into real(mykind) does; for an
  ! fp32 compute kind the single value is passed on exactly as a double.
  subroutine read_parameters_kind()
    real(4) :: svisc4, u04, ctes4
    if (p%prec == BGK_PREC_DOUBLE .or. p%prec == BGK_PREC_MIXED2) return
    call read_sp(svisc4, u04, ctes4)
    svisc = real(svisc4, 8)
    u0 = real(u04, 8)
    cteS = real(ctes4, 8)
  end subroutine read_parameters_kind

  subroutine read_sp(svisc4, u04, ctes4)
    real(4), intent(out) :: svisc4, u04, ctes4
    real(4) :: svisc, u0, cteS
    integer :: itfin, ivtim, isignal, itsave, icheck, irestart, init_v, lx, ly
    integer :: flag1, flag2, flag3, ipad, jpad, radius
    namelist /parameters/ svisc, u0, itfin, ivtim, isignal, itsave, icheck, irestart, &
         init_v, lx, ly, flag1, flag2, flag3, ipad, jpad, radius, cteS
    cteS = 0.1
    open(15, file='bgk.input', status='old')
    read(15, parameters)
    close(15)
    svisc4 = svisc
    u04 = u0
    ctes4 = cteS
  end subroutine read_sp

  ! diagnostic(itime,ivtim,icheck,itsave) (SRC/diagnostic.F90:32-103)
  subroutine diagnostic(it)
    integer, intent(in) :: it
    integer(c_int) :: rc
    if (mod(it, ivtim) == 0) then
       call vtk_xy_out(it)
       call vtk_om_out(it)
    end if
    if (mod(it, icheck) == 0) then
       call diagno_out(it)
       call probe_out(it, l/2, m/2)
       call varm_out(it)
       call prof_i_out(it, m/2)
       call prof_j_out(it, l/2)
       rc = bgk_check(ctx, 0.5d0)
       if (rc /= 0) write(16,*) 'WARNING: non-finite or |u| > 0.5 at step ', it
       if (drag == 1 .and. p%obstacle >= 1) call draglift_out(it)
    end if
    if (mod(it, itsave) == 0) call save_raw(it)
  end subroutine diagnostic

  ! vtk_xy_bin (SRC/vtk_xy.binary.F90:46-150): legacy VTK, text header,
  ! big-endian FP32 blocks: velocity vectors (u, w, 0), then w, u, rho;
  ! the fields come from bgk_fields (u/rho, w/rho, rho in FP32)
  subroutine vtk_xy_out(it)
    integer, intent(in) :: it
    real(c_float), allocatable :: u(:), w(:), den(:)
    character(len=24) :: fname
    integer :: q
    allocate(u(l * m), w(l * m), den(l * m))
    call bgk_ok(ctx, bgk_fields(ctx, u, w, den), 'bgk_fields')
    fname = 'tec_xy.xxxxxxxx.vtk'
    write(fname(8:15), '(i8.8)') it
    open(52, file=fname, status='replace')
    write(52,'(A26)') '# vtk DataFile Version 2.0'
    write(52,'(A5)') 'Campo'
    write(52,'(A6)') 'BINARY'
    write(52,'(A25)') 'DATASET STRUCTURED_POINTS'
    write(52,'(A11,I10,A1,I10,A1,I10)') 'DIMENSIONS ', l, ' ', m, ' ', 1
    write(52,'(A7,I10,A1,I10,A1,I10)') 'ORIGIN ', 1, ' ', 1, ' ', 1
    write(52,'(A8,I10,A1,I10,A1,I10)') 'SPACING ', 1, ' ', 1, ' ', 1
    write(52,'(A10,I10)') 'POINT_DATA ', l * m
    write(52,'(A23)') 'VECTORS velocity float'
    close(52)
    call vtk_append(fname)
    write(52) (u(q), w(q), 0.0_c_float, q = 1, l * m)
    close(52)
    call vtk_label(fname, 'SCALARS w float', 21)
    call vtk_append(fname)
    write(52) w
    close(52)
    call vtk_label(fname, 'SCALARS u float', 21)
    call vtk_append(fname)
    write(52) u
    close(52)
    call vtk_label(fname, 'SCALARS rho float', 21)
    call vtk_append(fname)
    write(52) den
    close(52)
    deallocate(u, w, den)
    write(16,*) 'I/O: plane xy (vtk,binary) done'
  end subroutine vtk_xy_out

  ! vtk_om_bin (SRC/vtk_om.binary.F90:50-160): vorticity
  ! 0.5 (u(i,j+1) - u(i,j-1)) - 0.5 (w(i+1,j) - w(i-1,j)) and the stream
  ! function of nodes 2..l-1 x 2..m-1.  The stream function starts its
  ! column-1 prefix at row 1 (the reference adds the halo node (1,0), a
  ! boundary value this library does not keep between steps).
  subroutine vtk_om_out(it)
    integer, intent(in) :: it
    real(c_float), allocatable :: u(:), w(:), den(:), phi(:)
    character(len=24) :: fname
    integer :: i, j
    allocate(u(l * m), w(l * m), den(l * m), phi(0:m))
    call bgk_ok(ctx, bgk_fields(ctx, u, w, den), 'bgk_fields')
    phi(0) = 0.0
    do j = 1, m
       phi(j) = phi(j - 1) + u(1 + (j - 1) * l)
    end do
    fname = 'tec_om.xxxxxxxx.vtk'
    write(fname(8:15), '(i8.8)') it
    open(55, file=fname, status='replace')
    write(55,'(A26)') '# vtk DataFile Version 2.0'
    write(55,'(A5)') 'Campo'
    write(55,'(A6)') 'BINARY'
    write(55,'(A25)') 'DATASET STRUCTURED_POINTS'
    write(55,'(A11,I10,A1,I10,A1,I10)') 'DIMENSIONS ', l - 2, ' ', m - 2, ' ', 1
    write(55,'(A7,I10,A1,I10,A1,I10)') 'ORIGIN ', 2, ' ', 2, ' ', 1
    write(55,'(A8,I10,A1,I10,A1,I10)') 'SPACING ', 1, ' ', 1, ' ', 1
    write(55,'(A10,I10)') 'POINT_DATA ', (l - 2) * (m - 2)
    write(55,'(A24)') 'SCALARS vorticity float'
    write(55,'(A20)') 'LOOKUP_TABLE default'
    close(55)
    open(55, file=fname, status='old', position='append', form='unformatted', &
         access='STREAM', convert='BIG_ENDIAN')
    do j = 2, m - 1
       do i = 2, l - 1
          write(55) 0.5 * (u(i + j * l) - u(i + (j - 2) * l)) - &
               0.5 * (w(i + 1 + (j - 1) * l) - w(i - 1 + (j - 1) * l))
       end do
    end do
    close(55)
    open(55, file=fname, status='old', position='append')
    write(55,'(A25)') 'SCALARS stream float'
    write(55,'(A20)') 'LOOKUP_TABLE default'
    close(55)
    open(55, file=fname, status='old', position='append', form='unformatted', &
         access='STREAM', convert='BIG_ENDIAN')
    do j = 2, m - 1
       do i = 2, l - 1
          phi(j) = phi(j) - w(i + (j - 1) * l)
          write(55) phi(j)
       end do
    end do
    close(55)
    deallocate(u, w, den, phi)
    write(16,*) 'I/O: vorticity xy (vtk) done'
  end subroutine vtk_om_out

  subroutine vtk_append(fname)
    character(len=*), intent(in) :: fname
    open(52, file=fname, status='old', position='append', form='unformatted', &
         access='STREAM', convert='BIG_ENDIAN')
  end subroutine vtk_append

  ! a label line written with the reference's (Aw) edit descriptor, which
  ! right-justifies the text in w columns
  subroutine vtk_label(fname, label, width)
    character(len=*), intent(in) :: fname, label
    integer, intent(in) :: width
    character(len=16) :: fmt
    write(fmt, '(a,i0,a)') '(A', width, ')'
    open(52, file=fname, status='old', position='append')
    write(52, fmt) label
    write(52,'(A20)') 'LOOKUP_TABLE default'
    close(52)
  end subroutine vtk_label

  ! cylinder (i-obs_x)^2 + (j-obs_y)^2 <= radius^2 as build_obs's
  ! (SRC/build.obs.F90:170-205), at a chosen centre
  subroutine set_cylinder()
    integer(c_int32_t), allocatable :: mask(:)
    integer :: i, j
    allocate(mask(l * m))
    do j = 1, m
       do i = 1, l
          mask(i + (j - 1) * l) = 0
          if ((i - obs_x)**2 + (j - obs_y)**2 <= radius * radius) mask(i + (j - 1) * l) = 1
       end do
    end do
    call bgk_ok(ctx, bgk_set_obstacle(ctx, mask), 'bgk_set_obstacle')
    deallocate(mask)
  end subroutine set_cylinder

  ! draglift (SRC/draglift.F90:41-101): box = centre +- (radius + 3),
  ! coefficients F / (U^2 radius), U the inflow speed (VKS: u_inflow = 0.1)
  subroutine draglift_out(it)
    integer, intent(in) :: it
    real(8) :: f(2), uin, norm
    integer, parameter :: delta = 3
    uin = 0.1d0
    if (p%bc == BGK_BC_OPEN) uin = u0
    norm = 1.0d0 / (uin * uin * real(radius, 8))
    call bgk_ok(ctx, bgk_draglift(ctx, int(obs_x - radius - delta, c_int32_t), &
         int(obs_x + radius + delta, c_int32_t), int(obs_y - radius - delta, c_int32_t), &
         int(obs_y + radius + delta, c_int32_t), f), 'bgk_draglift')
    write(66,*) it, f(1) * norm, f(2) * norm
    flush(66)
  end subroutine draglift_out

  subroutine diagno_out(it)
    integer, intent(in) :: it
    real(8) :: d(4)
    call bgk_ok(ctx, bgk_diagno(ctx, d), 'bgk_diagno')
    write(16,1001) it
    write(16,1002) d(3)
    write(16,1003) d(1), d(2), d(4)
    flush(16)
    write(63,1004) it, d(1), d(2), d(3), d(4)
    flush(63)
1001 format(" Timestep ",i8)
1002 format("       mean rho ",1(e14.6,1x))
1003 format("       mean vel ",3(e14.6,1x))
1004 format(i8,4(e14.6,1x))
  end subroutine diagno_out

  subroutine probe_out(it, i0, j0)
    integer, intent(in) :: it, i0, j0
    real(8) :: d(3)
    call bgk_ok(ctx, bgk_probe(ctx, int(i0, c_int32_t), int(j0, c_int32_t), d), 'bgk_probe')
    write(68,1002) it, d(1), d(2), d(3)
1002 format(i8,3(e14.6,1x))
  end subroutine probe_out

  subroutine varm_out(it)
    integer, intent(in) :: it
    real(8), allocatable :: u(:), v(:), r(:)
    integer :: j
    allocate(u(m), v(m), r(m))
    call bgk_ok(ctx, bgk_varm(ctx, u, v, r), 'bgk_varm')
    write(62,1005) it
    do j = 1, m
       write(62,1003) j, u(j), v(j), r(j)
    end do
    write(62,'(a1)')
    write(62,'(a1)')
1003 format(i5,3(e14.6,1x))
1005 format("# t=",i7)
  end subroutine varm_out

  subroutine prof_i_out(it, jc)
    integer, intent(in) :: it, jc
    real(8), allocatable :: u(:), v(:), r(:)
    integer :: i
    allocate(u(l), v(l), r(l))
    call bgk_ok(ctx, bgk_prof_i(ctx, int(jc, c_int32_t), u, v, r), 'bgk_prof_i')
    write(61,1005) it
    do i = 1, l
       write(61,1002) (i-0.5), u(i), v(i), r(i)
    end do
    write(61,'(a1)')
    write(61,'(a1)')
1002 format(4(e14.6,1x))
1005 format("# t=",i7)
  end subroutine prof_i_out

  subroutine prof_j_out(it, ic)
    integer, intent(in) :: it, ic
    real(8), allocatable :: u(:), v(:), r(:)
    integer :: j
    allocate(u(m), v(m), r(m))
    call bgk_ok(ctx, bgk_prof_j(ctx, int(ic, c_int32_t), u, v, r), 'bgk_prof_j')
    write(64,1005) it
    do j = 1, m
       write(64,1002) (j-0.5), u(j), v(j), r(j)
    end do
    write(64,'(a1)')
    write(64,'(a1)')
1002 format(4(e14.6,1x))
1005 format("# t=",i7)
  end subroutine prof_j_out

  ! save_raw (SRC/save_raw.F90:37-57): unformatted sequential, record 1 the
  ! step, then a01..a19 over (0:l+1, 0:m+1) in the storage kind
  subroutine save_raw(it)
    integer, intent(in) :: it
    character(len=18) :: file_name
    integer :: k, es
    real(4), allocatable, target :: b4(:)
    real(8), allocatable, target :: b8(:)
    integer(2), allocatable, target :: b2(:)
    file_name = 'save.xxxxxxxx.bin'
    write(file_name(6:13), '(i8.8)') it
    es = bgk_storage_bytes(ctx)
    open(20, file=file_name, form='unformatted', status='unknown')
    write(20) it
    select case (es)
    case (8)
       allocate(b8((l+2)*(m+2)))
       do k = 0, 8
          call bgk_ok(ctx, bgk_download(ctx, int(k, c_int32_t), c_loc(b8)), 'bgk_download')
          write(20) b8
       end do
    case (2)
       allocate(b2((l+2)*(m+2)))
       do k = 0, 8
          call bgk_ok(ctx, bgk_download(ctx, int(k, c_int32_t), c_loc(b2)), 'bgk_download')
          write(20) b2
       end do
    case default
       allocate(b4((l+2)*(m+2)))
       do k = 0, 8
          call bgk_ok(ctx, bgk_download(ctx, int(k, c_int32_t), c_loc(b4)), 'bgk_download')
          write(20) b4
       end do
    end select
    close(20)
    write(16,*) 'task ', 0, 'saving t=', it, file_name
  end subroutine save_raw

  ! restore_raw (SRC/restore_raw.F90:41-51) from restore.bin
  subroutine restore_raw(it)
    integer, intent(out) :: it
    integer :: k, es
    real(4), allocatable, target :: b4(:)
    real(8), allocatable, target :: b8(:)
    integer(2), allocatable, target :: b2(:)
    es = bgk_storage_bytes(ctx)
    open(20, file='restore.bin', form='unformatted', status='old')
    read(20) it
    select case (es)
    case (8)
       allocate(b8((l+2)*(m+2)))
       do k = 0, 8
          read(20) b8
          call bgk_ok(ctx, bgk_upload(ctx, int(k, c_int32_t), c_loc(b8)), 'bgk_upload')
       end do
    case (2)
       allocate(b2((l+2)*(m+2)))
       do k = 0, 8
          read(20) b2
          call bgk_ok(ctx, bgk_upload(ctx, int(k, c_int32_t), c_loc(b2)), 'bgk_upload')
       end do
    case default
       allocate(b4((l+2)*(m+2)))
       do k = 0, 8
          read(20) b4
          call bgk_ok(ctx, bgk_upload(ctx, int(k, c_int32_t), c_loc(b4)), 'bgk_upload')
       end do
    end select
    close(20)
  end subroutine restore_raw

  ! finalize (SRC/finalize.F90:34-81): MLUPS = l*m*steps/loop time
  subroutine finalize()
    real(8) :: mlups
    mlups = real(l,8) * real(m,8) * real(itfin - itstart, 8) / time_loop / 1.0d6
    open(69, file='bgk.perf', status='unknown')
    write(69,9999)
    write(69,*) "# Run info "
    write(69,*) l, m
    write(69,9999)
    write(69,*) "# Time for section "
    write(69,1100) time_init
    write(69,1101) time_loop
    write(69,9999)
    write(69,*) "# Derived (global) metrics "
    write(69,1106) mlups
    write(69,1108) 72.0d0 * real(l,8) * real(m,8) * real(itfin - itstart, 8) / time_loop / 1.0d9
    write(69,9999)
    close(69)
    write(6,1106) mlups
9999 format(" #--------------------------------")
1100 format(" # init   time",1(e14.6,1x))
1101 format(" # loop   time",1(e14.6,1x))
1106 format(" # Mlups      ",1(f14.6,1x))
1108 format(" # fp32 BW    ",1(e14.6,1x), "GB/s at 72 B per update")
  end subroutine finalize

end program bgk2d_hip
