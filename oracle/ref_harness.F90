! ref_harness.F90 -- drives the reference's own hot-path routines.
!
! TEST INFRASTRUCTURE ONLY.  Linked (by oracle/Makefile) against the
! reference's SRC/*.F90 files for the path, compiled unchanged from where
! they lie; this program replaces only SRC/bgk2d.F90's driver role so that
! setup/check_case (which need build-system-generated git_info/make_info)
! are not part of the build.  It reproduces the driver's call order
! (SRC/bgk2d.F90:59-113, SRC/initialize.F90:70-128):
!     alloca -> build_bcond -> [build_obs] -> init -> hencol
!     do itime: boundaries -> propagation -> collision(itime)
! and dumps the current `a` populations (with halos) as raw stream bytes.
! A dump file name of "-" skips the dump (timing runs).
!
! usage: ref_harness lx ly svisc u0 nsteps radius dumpfile
!   dump layout: int32 l, m, storage bytes, nsteps;
!                then a01,a03,a05,a08,a10,a12,a14,a17,a19 each
!                (0:l+1, 0:m+1), x fastest, in the storage kind;
!                then real(8) omega, fgrad, u00;
!                with OBSTACLE, then int32 obs(1:l,1:m), x fastest.
! Text diagnostics diagno/varm/prof_i/prof_j/probe of the last step land
! in fort.63 / fort.62 / fort.61 / fort.64 / fort.68 as in the reference.
program ref_harness
  use storage
  use timing
  use real_kinds
  implicit none
  integer :: itime, nsteps, ios, esize
  integer(8) :: c0, c1, crate
  character(len=256) :: arg, dumpfile

  call get_command_argument(1, arg); read(arg, *) lx
  call get_command_argument(2, arg); read(arg, *) ly
  call get_command_argument(3, arg); read(arg, *) svisc
  call get_command_argument(4, arg); read(arg, *) u0
  call get_command_argument(5, arg); read(arg, *) nsteps
  call get_command_argument(6, arg); read(arg, *) radius
  call get_command_argument(7, dumpfile)

  myrank = 0
  ipad = 0
  jpad = 0
  flag1 = 0
  flag2 = 0
  flag3 = 0
  cteS = 0.1
  l = lx
  m = ly
  l1 = l + 1
  m1 = m + 1
  stepy = m/2
  stepx = l/20
  u0x = zero
  u0y = zero

  call alloca()
  call build_bcond
#ifdef OBSTACLE
  call build_obs
#endif
  call init(0)
  call hencol

  call system_clock(c0, crate)
  do itime = 1, nsteps
     call boundaries
     call propagation
     call collision(itime)
  end do
  call system_clock(c1)
#ifdef DRAG
  ! drag and lift of the last step (SRC/draglift.F90), written to fort.66
  call draglift(nsteps)
#endif
  ! loop wall time and MLUPS as SRC/finalize.F90:60-61 defines it
  write(6,'(a,f14.6,a,f14.3)') 'HARNESS loop_s=', real(c1-c0,8)/real(crate,8), &
       ' mlups=', real(l,8)*real(m,8)*real(nsteps,8)/(real(c1-c0,8)/real(crate,8))/1.0d6

  esize = storage_size(a01(0,0)) / 8
  if (trim(dumpfile) == '-') stop
  open(30, file=trim(dumpfile), access='stream', form='unformatted', &
       status='replace', iostat=ios)
  if (ios /= 0) stop 'cannot open dump file'
  write(30) int(l,4), int(m,4), int(esize,4), int(nsteps,4)
  write(30) a01(0:l1,0:m1)
  write(30) a03(0:l1,0:m1)
  write(30) a05(0:l1,0:m1)
  write(30) a08(0:l1,0:m1)
  write(30) a10(0:l1,0:m1)
  write(30) a12(0:l1,0:m1)
  write(30) a14(0:l1,0:m1)
  write(30) a17(0:l1,0:m1)
  write(30) a19(0:l1,0:m1)
  write(30) real(omega,8), real(fgrad,8), real(u00,8)
#ifdef OBSTACLE
  ! the mask build_obs made (SRC/build.obs.F90), 1 = solid
  write(30) int(obs(1:l,1:m),4)
#endif
  close(30)

  call diagno(nsteps)
  call varm(nsteps)
  call prof_i(nsteps, m/2)
  call prof_j(nsteps, l/2)
  call probe(nsteps, l/2, m/2)
end program ref_harness
