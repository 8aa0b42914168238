! device_data_hip.f90 -- HIP backend of the device_data seam, the swap-in for
! device_data_cpu.f90 / device_data_openacc.f90 (src/CMakeLists.txt:22-36),
! same public interface (device_data_cpu.f90:8-26).
!
! enter_domain_device_data creates one device block per block of the domain
! (wql_block_create from the block's allocatables, passed by c_loc of their
! first element with their ghost bounds, then wql_block_upload), and enters
! every RK state and rate array into the library's present table for the
! per-array rk_vector seam (rk_vector_hip.f90).  update_domain_host_data
! copies the device state back; exit_domain_device_data releases both.
! Uses the reference's datatypes module (and through it MPI), so it is
! compiled in the reference's own build.
module device_data
  use iso_c_binding
  use common, only : wp
  use datatypes, only : domain_type, block_type
  use wql3d_hip_binding
  implicit none
  private
  public :: enter_domain_device_data, assert_domain_device_data_present, &
       update_domain_host_data, exit_domain_device_data

  type(c_ptr), save :: blocks(2) = c_null_ptr

contains

  subroutine enter_domain_device_data(D)
    type(domain_type), intent(inout), target :: D
    integer :: i
    if (D%nblocks < 0) error stop 'invalid domain block count'
    do i = 1, D%nblocks
      if (.not. allocated(D%B(i)%F%F)) cycle
      call enter_block(D%B(i), trim(D%type_of_mesh), blocks(i))
      call rate_state_arrays(D%B(i), 'enter')
    end do
  end subroutine enter_domain_device_data

  subroutine assert_domain_device_data_present(D)
    type(domain_type), intent(inout), target :: D
    integer :: i
    do i = 1, D%nblocks
      if (.not. allocated(D%B(i)%F%F)) cycle
      if (.not. c_associated(blocks(i))) error stop 'block data is not present on device'
      call rate_state_arrays(D%B(i), 'assert')
    end do
  end subroutine assert_domain_device_data_present

  subroutine update_domain_host_data(D)
    type(domain_type), intent(inout), target :: D
    integer :: i
    do i = 1, D%nblocks
      if (.not. c_associated(blocks(i))) cycle
      call wql_check(wql_block_download(blocks(i), ior(ior(WQL_F, WQL_DF), &
           ior(WQL_MEMORY, WQL_PML))), 'update_domain_host_data')
    end do
  end subroutine update_domain_host_data

  subroutine exit_domain_device_data(D)
    type(domain_type), intent(inout), target :: D
    integer :: i
    do i = 1, D%nblocks
      if (.not. c_associated(blocks(i))) cycle
      call rate_state_arrays(D%B(i), 'exit')
      call wql_block_destroy(blocks(i))
      blocks(i) = c_null_ptr
    end do
  end subroutine exit_domain_device_data

  ! the descriptor of one block (datatypes.f90:24-247) and its device block
  subroutine enter_block(B, type_of_mesh, h)
    type(block_type), intent(inout), target :: B
    character(*), intent(in) :: type_of_mesh
    type(c_ptr), intent(out) :: h
    type(wql_block_desc) :: d
    integer :: f, lo(4)

    d%nq = B%G%C%nq; d%nr = B%G%C%nr; d%ns = B%G%C%ns
    d%mesh = merge(WQL_CURVILINEAR, WQL_CARTESIAN, type_of_mesh == 'curvilinear')
    d%precision = WQL_FP64
    d%arrays_on_device = 0
    d%bc = [B%boundary_vars%Lx, B%boundary_vars%Rx, B%boundary_vars%Ly, &
            B%boundary_vars%Ry, B%boundary_vars%Lz, B%boundary_vars%Rz]
    d%hq = B%G%hq; d%hr = B%G%hr; d%hs = B%G%hs
    d%tau0 = B%tau0
    d%tau = 0.0_c_double; d%weight = 0.0_c_double
    lo = lbound(B%F%F)
    d%body%lo = lo(1:3)
    d%body%hi = [ubound(B%F%F, 1), ubound(B%F%F, 2), ubound(B%F%F, 3)]
    d%metricx = c_loc(B%G%metricx(lo(1), lo(2), lo(3), 1))
    d%metricy = c_loc(B%G%metricy(lo(1), lo(2), lo(3), 1))
    d%metricz = c_loc(B%G%metricz(lo(1), lo(2), lo(3), 1))
    d%J = c_loc(B%G%J(lo(1), lo(2), lo(3)))
    d%M = c_loc(B%M%M(lo(1), lo(2), lo(3), 1))
    d%F = c_loc(B%F%F(lo(1), lo(2), lo(3), 1))
    d%DF = c_loc(B%F%DF(lo(1), lo(2), lo(3), 1))
    d%eta = c_null_ptr; d%Deta = c_null_ptr
    d%Qs_inv = c_null_ptr; d%Qp_inv = c_null_ptr
    d%nmech = 0
    if (B%M%anelastic_Q8) then
      d%nmech = 8
      d%tau(1:8) = B%M%tau_Q8; d%weight(1:8) = B%M%weight_Q8
      d%eta = [c_loc(B%M%eta4Q8), c_loc(B%M%eta5Q8), c_loc(B%M%eta6Q8), &
               c_loc(B%M%eta7Q8), c_loc(B%M%eta8Q8), c_loc(B%M%eta9Q8)]
      d%Deta = [c_loc(B%M%Deta4Q8), c_loc(B%M%Deta5Q8), c_loc(B%M%Deta6Q8), &
                c_loc(B%M%Deta7Q8), c_loc(B%M%Deta8Q8), c_loc(B%M%Deta9Q8)]
      d%Qs_inv = c_loc(B%M%Qs_inv_Q8); d%Qp_inv = c_loc(B%M%Qp_inv_Q8)
    else if (B%M%anelastic_Q) then
      d%nmech = 4
      d%tau(1:4) = B%M%tau_Q; d%weight(1:4) = B%M%weight_Q
      d%eta = [c_loc(B%M%eta4Q), c_loc(B%M%eta5Q), c_loc(B%M%eta6Q), &
               c_loc(B%M%eta7Q), c_loc(B%M%eta8Q), c_loc(B%M%eta9Q)]
      d%Deta = [c_loc(B%M%Deta4Q), c_loc(B%M%Deta5Q), c_loc(B%M%Deta6Q), &
                c_loc(B%M%Deta7Q), c_loc(B%M%Deta8Q), c_loc(B%M%Deta9Q)]
      d%Qs_inv = c_loc(B%M%Qs_inv_Q); d%Qp_inv = c_loc(B%M%Qp_inv_Q)
    end if
    do f = 1, 6
      d%pml(f) = 0; d%npml(f) = 0
      d%pml_Q(f) = c_null_ptr; d%pml_DQ(f) = c_null_ptr
      d%pml_box(f)%lo = 1; d%pml_box(f)%hi = 1
      if (B%PMLB(f)%pml .and. allocated(B%PMLB(f)%Q)) then
        d%pml(f) = 1; d%npml(f) = B%PMLB(f)%N_pml
        lo = lbound(B%PMLB(f)%Q)
        d%pml_box(f)%lo = lo(1:3)
        d%pml_box(f)%hi = [ubound(B%PMLB(f)%Q, 1), ubound(B%PMLB(f)%Q, 2), &
                           ubound(B%PMLB(f)%Q, 3)]
        d%pml_Q(f) = c_loc(B%PMLB(f)%Q(lo(1), lo(2), lo(3), 1))
        d%pml_DQ(f) = c_loc(B%PMLB(f)%DQ(lo(1), lo(2), lo(3), 1))
      end if
      d%n_l(f) = c_null_ptr; d%n_m(f) = c_null_ptr; d%n_n(f) = c_null_ptr
      d%face_box(f)%lo = 1; d%face_box(f)%hi = 1
      if (allocated(B%B(f)%n_l)) then
        d%face_box(f)%lo = [lbound(B%B(f)%n_l, 1), lbound(B%B(f)%n_l, 2), 1]
        d%face_box(f)%hi = [ubound(B%B(f)%n_l, 1), ubound(B%B(f)%n_l, 2), 1]
        d%n_l(f) = c_loc(B%B(f)%n_l(lbound(B%B(f)%n_l, 1), lbound(B%B(f)%n_l, 2), 1))
        d%n_m(f) = c_loc(B%B(f)%n_m(lbound(B%B(f)%n_m, 1), lbound(B%B(f)%n_m, 2), 1))
        d%n_n(f) = c_loc(B%B(f)%n_n(lbound(B%B(f)%n_n, 1), lbound(B%B(f)%n_n, 2), 1))
      end if
    end do
    d%sat_form = WQL_SAT_BC
    call wql_check(wql_block_create(d, h), 'enter_domain_device_data')
    call wql_check(wql_block_upload(h, WQL_ALL), 'enter_domain_device_data')
  end subroutine enter_block

  ! the arrays the rk_vector seam scales and updates (fields.f90:51-263):
  ! enter them into, check them in, or drop them from the present table
  subroutine rate_state_arrays(B, action)
    type(block_type), intent(inout), target :: B
    character(*), intent(in) :: action
    integer :: f
    call one(B%F%F); call one(B%F%DF)
    if (B%M%anelastic_Q8) then
      call one(B%M%eta4Q8); call one(B%M%eta5Q8); call one(B%M%eta6Q8)
      call one(B%M%eta7Q8); call one(B%M%eta8Q8); call one(B%M%eta9Q8)
      call one(B%M%Deta4Q8); call one(B%M%Deta5Q8); call one(B%M%Deta6Q8)
      call one(B%M%Deta7Q8); call one(B%M%Deta8Q8); call one(B%M%Deta9Q8)
    else if (B%M%anelastic_Q) then
      call one(B%M%eta4Q); call one(B%M%eta5Q); call one(B%M%eta6Q)
      call one(B%M%eta7Q); call one(B%M%eta8Q); call one(B%M%eta9Q)
      call one(B%M%Deta4Q); call one(B%M%Deta5Q); call one(B%M%Deta6Q)
      call one(B%M%Deta7Q); call one(B%M%Deta8Q); call one(B%M%Deta9Q)
    end if
    do f = 1, 6
      if (B%PMLB(f)%pml .and. allocated(B%PMLB(f)%Q)) then
        call one(B%PMLB(f)%Q); call one(B%PMLB(f)%DQ)
      end if
    end do
  contains
    subroutine one(a)
      real(kind=wp), allocatable, target, intent(inout) :: a(:,:,:,:)
      integer(c_int64_t) :: bytes
      bytes = int(size(a), c_int64_t) * int(storage_size(0.0_wp) / 8, c_int64_t)
      select case (action)
      case ('enter')
        call wql_check(wql_array_enter(c_loc(a), bytes), 'enter_domain_device_data')
      case ('assert')
        if (wql_array_present(c_loc(a), bytes) /= 1) &
             error stop 'RK state/rate array is not present on device'
      case ('exit')
        call wql_check(wql_array_exit(c_loc(a)), 'exit_domain_device_data')
      end select
    end subroutine one
  end subroutine rate_state_arrays

end module device_data
