! sparc_timestepping -- a Fortran host driving the GPU engine through the
! ISO_C_BINDING module: the forward loop of TIMESTEPPING (driver.f90:398-547)
! with `call step()` (driver.f90:424) replaced by sparc_step.
!
!   sparc_timestepping BUNDLE OUT
!
! BUNDLE (stream, little endian) holds what mp_Initialize_RHS, INIT_DAMPING
! and DETERMINE_STATUS leave in module globals, written by
! sparc_amd.driver.write_bundle:
!   "SPARCBN1", sparc_config, int32 maxtime, int32 steps,
!   int32 nbg  x (int32 id, int64 n, real64(n)),
!   int32 nprof x (int32 id, int32 n, real64(n)),
!   int64 n, real64(n)            damping rates,
!   int64 n, real64(n) x 2        decay_coeff_x, decay_coeff_y,
!   int32 kind, int64 n, real64(n), int32 nt, real64(nt), int32 active_until.
! OUT is a direct-access file with one record of v_z(:,:,o_rad) (nx*ny real64)
! after every step i with mod(i, steps) == 0, the records vz_cc.bin holds
! (driver.f90:460-476).  The L2 norm of v_z is printed at each record, as the
! reference prints norm(a(:,:,:,6)) per iteration.
program sparc_timestepping
  use, intrinsic :: iso_c_binding
  use sparc_gpu
  implicit none
  character(len=4096) :: bundle, outfile
  character(len=8) :: magic
  type(sparc_config) :: cfg
  type(c_ptr) :: ctx
  integer(c_int32_t) :: maxtime, steps, nitems, id, n32, kind, nt, active_until
  integer(c_int64_t) :: n64
  real(c_double), allocatable :: buf(:), buf2(:), mask(:), series(:), plane(:)
  real(c_double) :: nrm
  integer :: u, v, i, q, ierr, reclen, rec

  call get_command_argument(1, bundle)
  call get_command_argument(2, outfile)
  open(newunit=u, file=trim(bundle), access='stream', form='unformatted', status='old', action='read')
  read(u) magic
  if (magic /= 'SPARCBN1') stop 'not a set-up bundle'
  read(u) cfg
  read(u) maxtime, steps

  ierr = sparc_create(cfg, ctx)
  call check(ierr, 'sparc_create')

  read(u) nitems
  do q = 1, nitems
    read(u) id, n64
    allocate(buf(n64))
    read(u) buf
    call check(sparc_set_background(ctx, id, buf, int(n64, c_size_t)), 'sparc_set_background')
    deallocate(buf)
  end do
  read(u) nitems
  do q = 1, nitems
    read(u) id, n32
    allocate(buf(n32))
    read(u) buf
    call check(sparc_set_profile(ctx, id, buf, n32), 'sparc_set_profile')
    deallocate(buf)
  end do
  read(u) n64
  if (n64 > 0) then
    allocate(buf(n64))
    read(u) buf
    call check(sparc_set_damping(ctx, buf, int(n64, c_size_t)), 'sparc_set_damping')
    deallocate(buf)
  end if
  read(u) n64
  allocate(buf(n64))
  read(u) buf
  read(u) n64
  allocate(buf2(n64))
  read(u) buf2
  call check(sparc_set_filter(ctx, buf, buf2), 'sparc_set_filter')
  read(u) kind, n64
  allocate(mask(n64))
  read(u) mask
  read(u) nt
  allocate(series(max(nt, 1)))
  if (nt > 0) read(u) series(1:nt)
  read(u) active_until
  close(u)
  call check(sparc_set_source(ctx, kind, mask, int(size(mask), c_size_t), series, nt, active_until), &
             'sparc_set_source')

  allocate(plane(int(cfg%nx) * int(cfg%ny)))
  inquire(iolength=reclen) plane
  open(newunit=v, file=trim(outfile), access='direct', form='unformatted', recl=reclen, &
       status='replace', action='write')
  rec = 0
  do i = 0, maxtime
    call check(sparc_step(ctx, i, 1), 'sparc_step')
    if (mod(i, steps) == 0) then
      call check(sparc_get_plane(ctx, SPARC_V_Z, cfg%o_rad, plane), 'sparc_get_plane')
      rec = rec + 1
      write(v, rec=rec) plane
      call check(sparc_norm(ctx, SPARC_V_Z, nrm), 'sparc_norm')
      print '(a,i8,a,es24.16)', ' time ', i, '  norm(v_z) ', nrm
    end if
  end do
  close(v)
  call sparc_destroy(ctx)

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= SPARC_OK) then
      print '(a,a,i4,a,a)', what, ' failed: ', rc, ' ', sparc_last_error(ctx)
      stop 1
    end if
  end subroutine check

end program sparc_timestepping
