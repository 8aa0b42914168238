! traditional_cartesian_rhs_hip.f90 -- HIP backend of the
! traditional_cartesian_rhs_backend seam, the swap-in for
! traditional_cartesian_rhs_cpu.f90 / _openacc.f90 (src/CMakeLists.txt:22-36),
! same interface (traditional_cartesian_rhs_cpu.f90:8-15): when the block has
! a device mirror (created by device_data_hip.f90), the strict-interior RHS
! (RHS_Center -> JJU_x6_interior, Q4/Q8 included) runs on the GPU, DF += L(F)
! on the interior box, and handled = .true.; otherwise handled = .false. and
! the caller keeps its host path (RHS_Interior.f90:37-38).  Like the OpenACC
! backend (traditional_cartesian_rhs_openacc.f90:121) the rates are brought
! back to the host after the kernel.
! Uses the reference's datatypes module, so it is compiled in the
! reference's own build.
module traditional_cartesian_rhs_backend
  use iso_c_binding
  use common, only : wp
  use datatypes, only : block_type, block_grid_t, block_material
  use wql3d_hip_binding
  implicit none
  private
  public :: try_traditional_cartesian_rhs
contains
  subroutine try_traditional_cartesian_rhs(F, G, M, type_of_mesh, handled)
    type(block_type), intent(inout), target :: F
    type(block_grid_t), intent(in) :: G
    type(block_material), intent(inout) :: M
    character(len=*), intent(in) :: type_of_mesh
    logical, intent(out) :: handled
    type(c_ptr) :: blk
    integer(c_int32_t) :: done
    integer(c_int32_t) :: mask

    handled = .false.
    if (.not. allocated(F%F%F)) return
    if (wql_block_find(c_loc(F%F%F), blk) /= WQL_OK) return
    mask = ior(WQL_F, WQL_DF)
    if (F%M%anelastic_Q .or. F%M%anelastic_Q8) mask = ior(mask, WQL_MEMORY)
    call wql_check(wql_block_upload(blk, mask), 'try_traditional_cartesian_rhs')
    call wql_check(wql_rhs_interior(blk, done), 'try_traditional_cartesian_rhs')
    if (done == 0) return
    call wql_check(wql_block_download(blk, ior(WQL_DF, iand(mask, WQL_MEMORY))), &
         'try_traditional_cartesian_rhs')
    handled = .true.
  end subroutine try_traditional_cartesian_rhs
end module traditional_cartesian_rhs_backend
