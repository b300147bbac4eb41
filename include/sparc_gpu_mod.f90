! sparc_gpu_mod.f90 -- ISO_C_BINDING interface of libsparc_gpu for a Fortran
! host (the same entry points and constants as sparc_gpu.h).
!
! A Fortran driver replaces, in TIMESTEPPING (driver.f90:398-547),
!     call step()
! by
!     ierr = sparc_step(ctx, time, 1)
! after handing the arrays mp_Initialize_RHS leaves in module globals
! (physics.f90:34-404) to sparc_set_background / sparc_set_profile once.
! Arrays are passed as Fortran arrays in their natural (nx,ny,nz) order;
! plane numbers (e_rad, o_rad, k) are 1-based as in the reference.
module sparc_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: SPARC_ABI_VERSION = 1
  integer(c_int), parameter :: SPARC_OK = 0, SPARC_EINVAL = -1, SPARC_EHIP = -2, &
                               SPARC_ENOMEM = -3, SPARC_ESTATE = -4
  integer(c_int), parameter :: SPARC_FP64 = 0, SPARC_FP32 = 1
  integer(c_int), parameter :: SPARC_OPT_QUIET_PML_3D = 2, SPARC_OPT_QUIET_PML_3D_DISPL = 4, &
                               SPARC_OPT_MHD_PML_3D_DISPL = 10, SPARC_OPT_QUIET_PML_2D_DISPL = 17, &
                               SPARC_OPT_MHD_PML_2D_DISPL = 18, SPARC_OPT_MHD_PML_3D = 8
  integer(c_int), parameter :: SPARC_BX = 5, SPARC_MAXFIELDS = 8
  integer(c_int), parameter :: SPARC_RHO = 0, SPARC_P = 4
  integer(c_int), parameter :: SPARC_FORWARD = 0, SPARC_ADJOINT = 1
  integer(c_int), parameter :: SPARC_XI_X = 0, SPARC_XI_Y = 1, SPARC_XI_Z = 2, &
                               SPARC_V_X = 3, SPARC_V_Y = 4, SPARC_V_Z = 5, SPARC_NFIELDS = 6
  integer(c_int), parameter :: SPARC_ARR_A = 0, SPARC_ARR_SCR = 1, SPARC_ARR_PMLVARS = 2, &
                               SPARC_ARR_SCRPML = 3, SPARC_ARR_ALO = 4
  integer(c_int), parameter :: SPARC_BG_RHO0 = 0, SPARC_BG_RHOINV = 1, SPARC_BG_C2RHO0 = 2, &
                               SPARC_BG_GRADP0_Z = 3, SPARC_BG_GRADRHO0_Z = 4, SPARC_BG_AZ = 5, &
                               SPARC_BG_BZPML = 6, SPARC_BG_C2 = 7, SPARC_BG_BX = 8, &
                               SPARC_BG_BY = 9, SPARC_BG_BZ = 10, SPARC_BG_CURLBX = 11, &
                               SPARC_BG_CURLBY = 12, SPARC_BG_CURLBZ = 13, SPARC_BG_GRADP0_X = 14, &
                               SPARC_BG_GRADP0_Y = 15, SPARC_BG_GRADRHO0_X = 16, &
                               SPARC_BG_GRADRHO0_Y = 17, SPARC_BG_COUNT = 18
  integer(c_int), parameter :: SPARC_PROF_STRETCH = 0, SPARC_PROF_UNSTRETCH = 1, &
                               SPARC_PROF_G = 2, SPARC_PROF_Z = 3, SPARC_PROF_SOURCE_DEP = 4, &
                               SPARC_PROF_SPONGE = 5, SPARC_PROF_COUNT = 6
  integer(c_int), parameter :: SPARC_SRC_NONE = 0, SPARC_SRC_SERIES = 1, SPARC_SRC_LAGRANGE = 2
  integer(c_int), parameter :: SPARC_KERN_C2 = 0, SPARC_KERN_RHO = 1, SPARC_KERN_HESSIAN = 2, &
                               SPARC_KERN_VX = 3, SPARC_KERN_VY = 4, SPARC_KERN_VZ = 5, &
                               SPARC_KERN_BX = 6, SPARC_KERN_BY = 7, SPARC_KERN_BZ = 8, &
                               SPARC_KERN_PSI = 9

  ! sparc_config of sparc_gpu.h: 24 int32 then 16 double, no padding
  type, bind(C) :: sparc_config
    integer(c_int32_t) :: abi_version
    integer(c_int32_t) :: nx, ny, nz
    integer(c_int32_t) :: option
    integer(c_int32_t) :: precision
    integer(c_int32_t) :: npmlbot, npmltop
    integer(c_int32_t) :: damp_waves
    integer(c_int32_t) :: freq_filtering
    integer(c_int32_t) :: e_rad, o_rad
    integer(c_int32_t) :: mode
    integer(c_int32_t) :: cadforcing_step
    integer(c_int32_t) :: forcing_length
    integer(c_int32_t) :: num_steps
    integer(c_int32_t) :: compensated
    integer(c_int32_t) :: flow_kernels
    integer(c_int32_t) :: use_graphs
    integer(c_int32_t) :: plane_damping
    integer(c_int32_t) :: reserved_i(4)
    real(c_double) :: xlength, ylength
    real(c_double) :: timestep
    real(c_double) :: deltat
    real(c_double) :: stretchx, stretchy
    real(c_double) :: delta_width
    real(c_double) :: cadforcing
    real(c_double) :: field_scale
    real(c_double) :: reserved_d(7)
  end type sparc_config

  interface
    ! life cycle (Initialize_all allocations + INITIALIZE_STEP)
    integer(c_int) function sparc_create(cfg, ctx) bind(C, name="sparc_create")
      import :: c_int, c_ptr, sparc_config
      type(sparc_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    subroutine sparc_destroy(ctx) bind(C, name="sparc_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function sparc_last_error_c(ctx) bind(C, name="sparc_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function sparc_device_bytes(ctx, bytes) bind(C, name="sparc_device_bytes")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_size_t), intent(out) :: bytes
    end function
    integer(c_int) function sparc_footprint(cfg, bytes) bind(C, name="sparc_footprint")
      import :: c_int, c_size_t, sparc_config
      type(sparc_config), intent(in) :: cfg
      integer(c_size_t), intent(out) :: bytes
    end function

    ! set-up (mp_Initialize_RHS outputs, INIT_DAMPING, decay_coeff, sources)
    integer(c_int) function sparc_set_background(ctx, id, host, n) &
        bind(C, name="sparc_set_background")
      import :: c_int, c_ptr, c_double, c_size_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      real(c_double), intent(in) :: host(*)
      integer(c_size_t), value :: n
    end function
    integer(c_int) function sparc_set_profile(ctx, id, host, n) bind(C, name="sparc_set_profile")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      real(c_double), intent(in) :: host(*)
      integer(c_int), value :: n
    end function
    integer(c_int) function sparc_set_damping(ctx, rates, n) bind(C, name="sparc_set_damping")
      import :: c_int, c_ptr, c_double, c_size_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rates(*)
      integer(c_size_t), value :: n
    end function
    integer(c_int) function sparc_set_filter(ctx, decay_x, decay_y) bind(C, name="sparc_set_filter")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: decay_x(*), decay_y(*)
    end function
    integer(c_int) function sparc_set_source(ctx, kind, mask, nmask, series, nt, active_until) &
        bind(C, name="sparc_set_source")
      import :: c_int, c_ptr, c_double, c_size_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind
      real(c_double), intent(in) :: mask(*)
      integer(c_size_t), value :: nmask
      real(c_double), intent(in) :: series(*)
      integer(c_int), value :: nt, active_until
    end function

    integer(c_int) function sparc_set_source_planes(ctx, k_xy, n) &
        bind(C, name="sparc_set_source_planes")
      import :: c_int, c_ptr, c_int32_t, c_size_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: k_xy(*)
      integer(c_size_t), value :: n
    end function

    ! the hot path: nsteps calls of step() from time index t0
    integer(c_int) function sparc_step(ctx, t0, nsteps) bind(C, name="sparc_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: t0, nsteps
    end function
    integer(c_int) function sparc_sync(ctx) bind(C, name="sparc_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function sparc_reset(ctx) bind(C, name="sparc_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function sparc_apply_op(ctx, op, ibc, in, out) bind(C, name="sparc_apply_op")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: op, ibc
      real(c_double), intent(in) :: in(*)
      real(c_double), intent(inout) :: out(*)
    end function

    ! state access (module globals, the v_z slice, norm)
    integer(c_int) function sparc_get_state(ctx, arr, field, host) bind(C, name="sparc_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: arr, field
      real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function sparc_put_state(ctx, arr, field, host) bind(C, name="sparc_put_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: arr, field
      real(c_double), intent(in) :: host(*)
    end function
    integer(c_int) function sparc_get_plane(ctx, field, k, host) bind(C, name="sparc_get_plane")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, k
      real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function sparc_norm(ctx, field, out) bind(C, name="sparc_norm")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_double), intent(out) :: out
    end function
    integer(c_int) function sparc_sample(ctx, field, npts, ijk, host) bind(C, name="sparc_sample")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, npts
      integer(c_int32_t), intent(in) :: ijk(*)
      real(c_double), intent(out) :: host(*)
    end function
    ! the same without device allocation (at most 65536 points per call)
    integer(c_int) function sparc_sample_noalloc(ctx, field, npts, ijk, host) &
        bind(C, name="sparc_sample_noalloc")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, npts
      integer(c_int32_t), intent(in) :: ijk(*)
      real(c_double), intent(out) :: host(*)
    end function

    ! snapshots, checkpoints and kernels (WRITE_OUT_PARTIAL_STATE, PRODUCE_KERNELS)
    integer(c_int) function sparc_snapshots_alloc(ctx, nslots) bind(C, name="sparc_snapshots_alloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nslots
    end function
    integer(c_int) function sparc_snapshot_store(ctx, slot) bind(C, name="sparc_snapshot_store")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
    end function
    integer(c_int) function sparc_snapshots_copy(dst, src) bind(C, name="sparc_snapshots_copy")
      import :: c_int, c_ptr
      type(c_ptr), value :: dst, src
    end function
    integer(c_int) function sparc_kernel_accumulate_from(ctx, src, slot) &
        bind(C, name="sparc_kernel_accumulate_from")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, src
      integer(c_int), value :: slot
    end function
    integer(c_int) function sparc_checkpoints_alloc(ctx, nslots) bind(C, name="sparc_checkpoints_alloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nslots
    end function
    integer(c_int) function sparc_checkpoint_save(ctx, slot) bind(C, name="sparc_checkpoint_save")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
    end function
    integer(c_int) function sparc_checkpoint_load(ctx, slot) bind(C, name="sparc_checkpoint_load")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
    end function
    integer(c_int) function sparc_kernel_begin(ctx, grav, nz) bind(C, name="sparc_kernel_begin")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: grav(*)
      integer(c_int), value :: nz
    end function
    integer(c_int) function sparc_kernel_accumulate(ctx, slot) bind(C, name="sparc_kernel_accumulate")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
    end function
    integer(c_int) function sparc_kernel_get(ctx, which, host) bind(C, name="sparc_kernel_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      real(c_double), intent(out) :: host(*)
    end function

    ! instrumentation
    integer(c_int) function sparc_profile_enable(ctx, on) bind(C, name="sparc_profile_enable")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function sparc_profile_read(ctx, cls, ms, bytes, launches) &
        bind(C, name="sparc_profile_read")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: cls
      real(c_double), intent(out) :: ms, bytes
      integer(c_int64_t), intent(out) :: launches
    end function

    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  ! sparc_last_error as a Fortran string
  function sparc_last_error(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = sparc_last_error_c(ctx)
    if (.not. c_associated(p)) then
      msg = ""
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function sparc_last_error

end module sparc_gpu
