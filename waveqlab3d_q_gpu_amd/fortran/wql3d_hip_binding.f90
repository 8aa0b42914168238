! wql3d_hip_binding.f90 -- iso_c_binding interface of libwql3d_hip.so
! (include/wql3d.h) for the reference's Fortran host.
!
! The derived types mirror the C structs field for field (bind(C)); every
! entry point returns 0 or a WQL_E* code, which wql_check turns into an
! error stop carrying wql_last_error() (the reference stops the same way on
! a missing device array, rk_vector_openacc.f90:14-17).
module wql3d_hip_binding
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: WQL_OK = 0, WQL_EINVAL = 1, WQL_EHIP = 2, &
       WQL_ENOMEM = 3, WQL_ESTATE = 4, WQL_ENODEV = 5
  integer(c_int32_t), parameter :: WQL_CARTESIAN = 0, WQL_CURVILINEAR = 1
  integer(c_int32_t), parameter :: WQL_FP64 = 0, WQL_FP32_MIXED = 1, WQL_FP32_KAHAN = 2
  integer(c_int32_t), parameter :: WQL_BC_NONE = 0, WQL_BC_CHARACTERISTIC = 1, &
       WQL_BC_FREE_SURFACE = 2, WQL_BC_CLAMPED = 3
  integer(c_int32_t), parameter :: WQL_SRC_GAUSSIAN = 0, WQL_SRC_COSBELL = 1, WQL_SRC_RICKER = 2
  integer(c_int32_t), parameter :: WQL_SAT_BC = 0, WQL_SAT_ALPHA_THETA = 1
  integer(c_int32_t), parameter :: WQL_F = 1, WQL_DF = 2, WQL_GEOMETRY = 4, WQL_MATERIAL = 8, &
       WQL_MEMORY = 16, WQL_QINV = 32, WQL_PML = 64, WQL_FACES = 128, WQL_ALL = 255
  integer(c_int32_t), parameter :: WQL_IFACE_LOCKED = 0, WQL_IFACE_SLIP_WEAKENING_TPV5 = 1, &
       WQL_IFACE_SLIP_WEAKENING_TPV28 = 2, WQL_IFACE_SLIP_WEAKENING = 3, &
       WQL_IFACE_RATE_STATE_TPV101 = 4, WQL_IFACE_RATE_STATE_TPV102 = 5, &
       WQL_IFACE_LINEAR_FRICTION = 6, WQL_IFACE_RATE_STATE_RW = 7

  type, bind(C) :: wql_box
     integer(c_int32_t) :: lo(3), hi(3)
  end type wql_box

  type, bind(C) :: wql_block_desc
     integer(c_int32_t) :: nq, nr, ns, mesh, precision, nmech
     integer(c_int32_t) :: bc(6), pml(6), npml(6), arrays_on_device
     real(c_double)     :: hq, hr, hs, tau0, tau(8), weight(8)
     type(wql_box)      :: body
     type(c_ptr)        :: metricx, metricy, metricz, J, M, F, DF
     type(c_ptr)        :: eta(6), Deta(6), Qs_inv, Qp_inv
     type(wql_box)      :: pml_box(6)
     type(c_ptr)        :: pml_Q(6), pml_DQ(6)
     type(wql_box)      :: face_box(6)
     type(c_ptr)        :: n_l(6), n_m(6), n_n(6)
     integer(c_int32_t) :: sat_form
  end type wql_block_desc

  type, bind(C) :: wql_source
     integer(c_int32_t) :: type, order, alpha, near(3)
     real(c_double)     :: duration, t_init
     real(c_double)     :: mXX, mYY, mZZ, mXY, mXZ, mYZ
     real(c_double)     :: location(3), near_phys(3)
  end type wql_source

  type, bind(C) :: wql_iface_desc
     integer(c_int32_t) :: coupling, arrays_on_device
     type(c_ptr)        :: y, z, s0, T0L, T0R, fs, fd, Dc, W0
     real(c_double)     :: alpha
  end type wql_iface_desc

  interface
     ! accelerator_runtime seam
     integer(c_int) function wql_runtime_init(local_rank, device) bind(C)
       import; integer(c_int32_t), value :: local_rank; integer(c_int32_t) :: device
     end function
     integer(c_int) function wql_runtime_finalize() bind(C)
       import
     end function
     integer(c_int) function wql_capacity_check(predicted_bytes) bind(C)
       import; integer(c_int64_t), value :: predicted_bytes
     end function
     ! device_data seam
     integer(c_int) function wql_block_create(desc, blk) bind(C)
       import; type(wql_block_desc), intent(in) :: desc; type(c_ptr) :: blk
     end function
     integer(c_int64_t) function wql_block_device_bytes(blk) bind(C)
       import; type(c_ptr), value :: blk
     end function
     integer(c_int) function wql_block_upload(blk, mask) bind(C)
       import; type(c_ptr), value :: blk; integer(c_int32_t), value :: mask
     end function
     integer(c_int) function wql_block_download(blk, mask) bind(C)
       import; type(c_ptr), value :: blk; integer(c_int32_t), value :: mask
     end function
     integer(c_int) function wql_block_init_state(blk) bind(C)
       import; type(c_ptr), value :: blk
     end function
     subroutine wql_block_destroy(blk) bind(C)
       import; type(c_ptr), value :: blk
     end subroutine
     integer(c_int) function wql_block_find(host_F, blk) bind(C)
       import; type(c_ptr), value :: host_F; type(c_ptr) :: blk
     end function
     ! sources and receivers
     integer(c_int) function wql_block_set_sources(blk, n, src) bind(C)
       import; type(c_ptr), value :: blk; integer(c_int32_t), value :: n
       type(wql_source), intent(in) :: src(*)
     end function
     ! node coordinates (B%G%X without ghosts, x fastest), for mollified
     ! sources (wql_source%order = 0, the reference's mollify_source)
     integer(c_int) function wql_block_set_coordinates(blk, X, on_device) bind(C)
       import; type(c_ptr), value :: blk; real(c_double), intent(in) :: X(*)
       integer(c_int32_t), value :: on_device
     end function
     integer(c_int) function wql_block_set_receivers(blk, n, ijk, nslots) bind(C)
       import; type(c_ptr), value :: blk; integer(c_int32_t), value :: n, nslots
       integer(c_int32_t), intent(in) :: ijk(*)
     end function
     integer(c_int) function wql_block_get_seismograms(blk, nslots, out) bind(C)
       import; type(c_ptr), value :: blk; integer(c_int32_t), value :: nslots
       real(c_double) :: out(*)
     end function
     ! rk_vector seam
     integer(c_int) function wql_rk_scale(blk, A) bind(C)
       import; type(c_ptr), value :: blk; real(c_double), value :: A
     end function
     integer(c_int) function wql_rk_update(blk, Bdt) bind(C)
       import; type(c_ptr), value :: blk; real(c_double), value :: Bdt
     end function
     integer(c_int) function wql_array_enter(host, bytes) bind(C)
       import; type(c_ptr), value :: host; integer(c_int64_t), value :: bytes
     end function
     integer(c_int) function wql_array_exit(host) bind(C)
       import; type(c_ptr), value :: host
     end function
     integer(c_int) function wql_array_present(host, bytes) bind(C)
       import; type(c_ptr), value :: host; integer(c_int64_t), value :: bytes
     end function
     integer(c_int) function wql_array_update_device(host, bytes) bind(C)
       import; type(c_ptr), value :: host; integer(c_int64_t), value :: bytes
     end function
     integer(c_int) function wql_array_update_self(host, bytes) bind(C)
       import; type(c_ptr), value :: host; integer(c_int64_t), value :: bytes
     end function
     integer(c_int) function wql_scale_rate_present(rate, n, coefficient) bind(C)
       import; type(c_ptr), value :: rate; integer(c_int64_t), value :: n
       real(c_double), value :: coefficient
     end function
     integer(c_int) function wql_update_state_present(state, rate, n, increment) bind(C)
       import; type(c_ptr), value :: state, rate; integer(c_int64_t), value :: n
       real(c_double), value :: increment
     end function
     ! set_rates_elastic / traditional_cartesian_rhs / moment tensor / SAT seams
     integer(c_int) function wql_rhs(blk) bind(C)
       import; type(c_ptr), value :: blk
     end function
     integer(c_int) function wql_rhs_interior(blk, handled) bind(C)
       import; type(c_ptr), value :: blk; integer(c_int32_t) :: handled
     end function
     integer(c_int) function wql_set_moment_tensor(blk, t) bind(C)
       import; type(c_ptr), value :: blk; real(c_double), value :: t
     end function
     integer(c_int) function wql_sat(blk) bind(C)
       import; type(c_ptr), value :: blk
     end function
     ! fused paths
     integer(c_int) function wql_stage_fused(blk, A, Bdt, t) bind(C)
       import; type(c_ptr), value :: blk; real(c_double), value :: A, Bdt, t
     end function
     integer(c_int) function wql_time_step(blk, dt, nsteps, t) bind(C)
       import; type(c_ptr), value :: blk; real(c_double), value :: dt
       integer(c_int32_t), value :: nsteps; real(c_double) :: t
     end function
     ! diagnostics
     type(c_ptr) function wql_last_error() bind(C)
       import
     end function
     integer(c_int) function wql_block_max_state(blk, out) bind(C)
       import; type(c_ptr), value :: blk; real(c_double) :: out(3)
     end function
     integer(c_int) function wql_block_sync(blk) bind(C)
       import; type(c_ptr), value :: blk
     end function
     ! two-block interface
     integer(c_int) function wql_iface_create(L, R, desc, itf) bind(C)
       import; type(c_ptr), value :: L, R; type(wql_iface_desc), intent(in) :: desc
       type(c_ptr) :: itf
     end function
     subroutine wql_iface_destroy(itf) bind(C)
       import; type(c_ptr), value :: itf
     end subroutine
     integer(c_int) function wql_iface_time_step(itf, dt, nsteps, t) bind(C)
       import; type(c_ptr), value :: itf; real(c_double), value :: dt
       integer(c_int32_t), value :: nsteps; real(c_double) :: t
     end function
     integer(c_int) function wql_iface_set_frames(itf, nslots) bind(C)
       import; type(c_ptr), value :: itf; integer(c_int32_t), value :: nslots
     end function
     integer(c_int) function wql_iface_get_friction_state(itf, W) bind(C)
       import; type(c_ptr), value :: itf; real(c_double) :: W(*)
     end function
     integer(c_int) function wql_iface_get_frames(itf, nslots, out) bind(C)
       import; type(c_ptr), value :: itf; integer(c_int32_t), value :: nslots
       real(c_double) :: out(*)
     end function
     integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
       import; type(c_ptr), value :: s
     end function
  end interface

contains

  ! the message of the last failed call
  function wql_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, k
    p = wql_last_error()
    if (.not. c_associated(p)) then
       msg = ''
       return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do k = 1, n
       msg(k:k) = chars(k)
    end do
  end function wql_error_message

  ! stop with the library's message on a non-zero status
  subroutine wql_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc == WQL_OK) return
    write(*, '(A,I0,A,A,A,A)') 'RUN-HIP-', rc, ' in ', where, ': ', wql_error_message()
    error stop 'libwql3d_hip call failed'
  end subroutine wql_check

end module wql3d_hip_binding
