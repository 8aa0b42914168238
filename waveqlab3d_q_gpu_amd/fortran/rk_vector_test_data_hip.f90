! rk_vector_test_data_hip.f90 -- HIP variant of the unit test's data module
! (tests/rk_vector_test_data_openacc.f90: acc enter data copyin / exit data
! delete), the swap-in that tests/rk_vector_kernel_test.f90 links with the
! HIP rk_vector backend: the two arrays get device mirrors through the
! library's present table.
module rk_vector_test_data
  use iso_c_binding
  use common, only : wp
  use wql3d_hip_binding, only : wql_runtime_init, wql_array_enter, wql_array_exit, wql_check
  implicit none
contains
  subroutine enter_test_data(state, rate)
    real(kind=wp), allocatable, target, intent(inout) :: state(:,:,:,:), rate(:,:,:,:)
    integer(c_int32_t) :: device
    call wql_check(wql_runtime_init(0_c_int32_t, device), 'enter_test_data')
    call wql_check(wql_array_enter(c_loc(state), int(size(state), c_int64_t) * &
         int(storage_size(0.0_wp) / 8, c_int64_t)), 'enter_test_data(state)')
    call wql_check(wql_array_enter(c_loc(rate), int(size(rate), c_int64_t) * &
         int(storage_size(0.0_wp) / 8, c_int64_t)), 'enter_test_data(rate)')
  end subroutine enter_test_data

  subroutine exit_test_data(state, rate)
    real(kind=wp), allocatable, target, intent(inout) :: state(:,:,:,:), rate(:,:,:,:)
    call wql_check(wql_array_exit(c_loc(rate)), 'exit_test_data(rate)')
    call wql_check(wql_array_exit(c_loc(state)), 'exit_test_data(state)')
  end subroutine exit_test_data
end module rk_vector_test_data
