! rk_vector_hip.f90 -- HIP backend of the rk_vector_kernels seam, the swap-in
! for rk_vector_cpu.f90 / rk_vector_openacc.f90 (selected by
! src/CMakeLists.txt:22-36).  Same interface (rk_vector_cpu.f90:7-19), same
! semantics as the OpenACC backend (rk_vector_openacc.f90:8-55): the array
! must be present on the device (entered by device_data or the unit test's
! rk_vector_test_data), the device copy is updated from the host, the kernel
! of libwql3d_hip.so runs on it (bit-exact with the CPU expression), and the
! host array receives the result.  Called once per array, 49 + 49 sites of
! fields.f90:51-263 per stage.
module rk_vector_kernels
  use iso_c_binding
  use common, only : wp
  use wql3d_hip_binding, only : wql_scale_rate_present, wql_update_state_present, wql_check
  implicit none
  private
  public :: scale_rate_array, update_state_array
contains
  subroutine scale_rate_array(rate, coefficient)
    real(kind=wp), allocatable, target, intent(inout) :: rate(:,:,:,:)
    real(kind=wp), intent(in) :: coefficient
    call wql_check(wql_scale_rate_present(c_loc(rate), int(size(rate), c_int64_t), &
         real(coefficient, c_double)), 'scale_rate_array')
  end subroutine scale_rate_array

  subroutine update_state_array(state, rate, increment)
    real(kind=wp), allocatable, target, intent(inout) :: state(:,:,:,:)
    real(kind=wp), allocatable, target, intent(inout) :: rate(:,:,:,:)
    real(kind=wp), intent(in) :: increment
    if (any(shape(state) /= shape(rate))) error stop 'RK state/rate shape mismatch'
    call wql_check(wql_update_state_present(c_loc(state), c_loc(rate), &
         int(size(state), c_int64_t), real(increment, c_double)), 'update_state_array')
  end subroutine update_state_array
end module rk_vector_kernels
