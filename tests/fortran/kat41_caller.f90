! kat41_caller.f90 -- a Fortran host driving libwql3d_hip.so through the
! iso_c_binding module, the way the reference's swap-in backend would: arrays
! with the reference's ghost bounds (nb = 3: -2 .. n+3 per axis,
! datatypes.f90 / mpi3dcomm.f90 cartesian3d_t mbq..pbq), passed by c_loc of
! their first element with their bounds in the descriptor.
!
! Usage: kat41_caller <input file>.  The input (stream, native endianness) is
! written by tests/test_gpu_fortran.py: n(3), nsteps, bc(6) (int32); dt, tau0
! (float64); the source (wql_source fields in order); then metricx, metricy,
! metricz (n x 3 each), J, M (n x 3) and, per face, n_l, n_m, n_n (face x 3),
! x fastest.  Prints the final max|F| (close_domain's shutdown figure,
! domain.f90:583-584), the library's max_state figure, and whether every
! ghost node kept its sentinel.
program kat41_caller
  use iso_c_binding
  use wql3d_hip_binding
  implicit none
  integer, parameter :: nb = 3
  real(c_double), parameter :: sentinel = 777.0_c_double
  integer(c_int32_t) :: n(3), nsteps, bc(6), device
  integer :: nq, nr, ns, fc, k, u
  real(c_double) :: dt, tau0, t, fmax, state(3)
  type(wql_source) :: src(1)
  type(wql_block_desc) :: d
  type(c_ptr) :: blk
  real(c_double), allocatable, target :: metricx(:,:,:,:), metricy(:,:,:,:), metricz(:,:,:,:)
  real(c_double), allocatable, target :: J(:,:,:), M(:,:,:,:), F(:,:,:,:), DF(:,:,:,:)
  type face_t
     real(c_double), allocatable :: nl(:,:,:), nm(:,:,:), nn(:,:,:)
  end type face_t
  type(face_t), target :: face(6)
  integer :: fn(2, 6)
  character(len=512) :: path
  logical :: ghosts_ok

  call get_command_argument(1, path)
  open(newunit=u, file=trim(path), access='stream', form='unformatted', status='old')
  read(u) n, nsteps, bc
  read(u) dt, tau0
  read(u) src(1)%type, src(1)%order, src(1)%alpha, src(1)%near
  read(u) src(1)%duration, src(1)%t_init, src(1)%mXX, src(1)%mYY, src(1)%mZZ, &
       src(1)%mXY, src(1)%mXZ, src(1)%mYZ, src(1)%location, src(1)%near_phys
  nq = n(1); nr = n(2); ns = n(3)
  allocate(metricx(1-nb:nq+nb, 1-nb:nr+nb, 1-nb:ns+nb, 3), source=sentinel)
  allocate(metricy, metricz, mold=metricx)
  metricy = sentinel; metricz = sentinel
  allocate(J(1-nb:nq+nb, 1-nb:nr+nb, 1-nb:ns+nb), source=sentinel)
  allocate(M(1-nb:nq+nb, 1-nb:nr+nb, 1-nb:ns+nb, 3), source=sentinel)
  allocate(F(1-nb:nq+nb, 1-nb:nr+nb, 1-nb:ns+nb, 9), source=sentinel)
  allocate(DF(1-nb:nq+nb, 1-nb:nr+nb, 1-nb:ns+nb, 9), source=sentinel)
  read(u) metricx(1:nq, 1:nr, 1:ns, :)
  read(u) metricy(1:nq, 1:nr, 1:ns, :)
  read(u) metricz(1:nq, 1:nr, 1:ns, :)
  read(u) J(1:nq, 1:nr, 1:ns)
  read(u) M(1:nq, 1:nr, 1:ns, :)
  ! face extents: q faces (nr, ns), r faces (nq, ns), s faces (nq, nr)
  fn(:, 1) = [nr, ns]; fn(:, 2) = [nr, ns]
  fn(:, 3) = [nq, ns]; fn(:, 4) = [nq, ns]
  fn(:, 5) = [nq, nr]; fn(:, 6) = [nq, nr]
  do fc = 1, 6
     allocate(face(fc)%nl(1-nb:fn(1,fc)+nb, 1-nb:fn(2,fc)+nb, 3), source=sentinel)
     allocate(face(fc)%nm, face(fc)%nn, mold=face(fc)%nl)
     read(u) face(fc)%nl(1:fn(1,fc), 1:fn(2,fc), :)
     read(u) face(fc)%nm(1:fn(1,fc), 1:fn(2,fc), :)
     read(u) face(fc)%nn(1:fn(1,fc), 1:fn(2,fc), :)
  end do
  close(u)
  F(1:nq, 1:nr, 1:ns, :) = 0.0_c_double        ! init_fields (fields.f90:38)
  DF(1:nq, 1:nr, 1:ns, :) = 1.0e40_c_double    ! fields.f90:39

  d%nq = nq; d%nr = nr; d%ns = ns
  d%mesh = WQL_CARTESIAN; d%precision = WQL_FP64; d%nmech = 0
  d%bc = bc; d%pml = 0; d%npml = 0; d%arrays_on_device = 0
  d%hq = 1.0_c_double / real(nq - 1, c_double)
  d%hr = 1.0_c_double / real(nr - 1, c_double)
  d%hs = 1.0_c_double / real(ns - 1, c_double)
  d%tau0 = tau0; d%tau = 0; d%weight = 0
  d%body%lo = [lbound(F, 1), lbound(F, 2), lbound(F, 3)]
  d%body%hi = [ubound(F, 1), ubound(F, 2), ubound(F, 3)]
  d%metricx = c_loc(metricx(1-nb, 1-nb, 1-nb, 1))
  d%metricy = c_loc(metricy(1-nb, 1-nb, 1-nb, 1))
  d%metricz = c_loc(metricz(1-nb, 1-nb, 1-nb, 1))
  d%J = c_loc(J(1-nb, 1-nb, 1-nb))
  d%M = c_loc(M(1-nb, 1-nb, 1-nb, 1))
  d%F = c_loc(F(1-nb, 1-nb, 1-nb, 1))
  d%DF = c_loc(DF(1-nb, 1-nb, 1-nb, 1))
  do k = 1, 6
     d%eta(k) = c_null_ptr; d%Deta(k) = c_null_ptr
     d%pml_Q(k) = c_null_ptr; d%pml_DQ(k) = c_null_ptr
     d%pml_box(k)%lo = 1; d%pml_box(k)%hi = 1
  end do
  d%Qs_inv = c_null_ptr; d%Qp_inv = c_null_ptr
  do fc = 1, 6
     d%face_box(fc)%lo = [1 - nb, 1 - nb, 1]
     d%face_box(fc)%hi = [fn(1, fc) + nb, fn(2, fc) + nb, 1]
     d%n_l(fc) = c_loc(face(fc)%nl(1-nb, 1-nb, 1))
     d%n_m(fc) = c_loc(face(fc)%nm(1-nb, 1-nb, 1))
     d%n_n(fc) = c_loc(face(fc)%nn(1-nb, 1-nb, 1))
  end do
  d%sat_form = WQL_SAT_BC

  call wql_check(wql_runtime_init(0_c_int32_t, device), 'initialize_accelerator_runtime')
  call wql_check(wql_block_create(d, blk), 'enter_domain_device_data')
  call wql_check(wql_block_upload(blk, WQL_ALL), 'enter_domain_device_data')
  call wql_check(wql_block_set_sources(blk, 1_c_int32_t, src), 'init_moment_tensor')
  t = 0.0_c_double
  call wql_check(wql_time_step(blk, dt, nsteps, t), 'time_step_RK')
  call wql_check(wql_block_download(blk, WQL_F), 'update_domain_host_data')
  call wql_check(wql_block_max_state(blk, state), 'close_domain')
  fmax = maxval(abs(F(1:nq, 1:nr, 1:ns, :)))
  ghosts_ok = count(F /= sentinel) == nq * nr * ns * 9
  write(*, '(A,ES24.16)') 'Elastic final state: max|field|=', fmax
  write(*, '(A,ES24.16)') 'library max_state: max|field|=', state(1)
  write(*, '(A,ES24.16)') 'final time t=', t
  write(*, '(A,L1)') 'ghost nodes untouched: ', ghosts_ok
  call wql_block_destroy(blk)
end program kat41_caller
