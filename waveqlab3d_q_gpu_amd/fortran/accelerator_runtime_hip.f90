! accelerator_runtime_hip.f90 -- HIP backend of the accelerator_runtime
! seam, the swap-in for accelerator_runtime_cpu.f90 /
! accelerator_runtime_openacc.f90 (src/CMakeLists.txt:22-36), same public
! interface (accelerator_runtime_cpu.f90:19-64): the node-local rank (from
! MPI's shared-memory split, as the reference computes it) picks the GPU
! (wql_runtime_init binds local_rank mod device count), and the capacity
! check asks the device (wql_capacity_check, RUN-GPU-MEM-002).
! Uses the reference's MPI module, so it is compiled in the reference's own
! build (this image's flang cannot read the installed mpi.mod).
module accelerator_runtime
  use mpi
  use iso_c_binding
  use common, only : wp
  use wql3d_hip_binding, only : wql_runtime_init, wql_runtime_finalize, &
       wql_capacity_check, wql_check
  implicit none
  private

  integer, save :: node_comm = MPI_COMM_NULL
  integer, save :: local_rank = -1
  integer, save :: local_size = 0
  logical, save :: initialized = .false.

  public :: initialize_accelerator_runtime, finalize_accelerator_runtime, &
       enforce_device_memory_capacity

contains

  subroutine initialize_accelerator_runtime()
    integer :: ierr, world_rank
    integer(c_int32_t) :: device

    if (initialized) return
    call MPI_Comm_rank(MPI_COMM_WORLD, world_rank, ierr)
    call MPI_Comm_split_type(MPI_COMM_WORLD, MPI_COMM_TYPE_SHARED, world_rank, &
         MPI_INFO_NULL, node_comm, ierr)
    call MPI_Comm_rank(node_comm, local_rank, ierr)
    call MPI_Comm_size(node_comm, local_size, ierr)
    call wql_check(wql_runtime_init(int(local_rank, c_int32_t), device), &
         'initialize_accelerator_runtime')
    if (world_rank == 0) then
      write(*,'(/,A)') 'WaveQLab3D accelerator runtime'
      write(*,'(A)') '  backend: hip (libwql3d_hip.so, gfx950)'
      write(*,'(A,I0)') '  working precision bits: ', storage_size(0.0_wp)
      write(*,'(A,I0)') '  MPI ranks on rank-0 node: ', local_size
      write(*,'(A,I0)') '  rank-0 device: ', device
    end if
    initialized = .true.
  end subroutine initialize_accelerator_runtime

  subroutine finalize_accelerator_runtime()
    integer :: ierr

    if (.not.initialized) return
    call wql_check(wql_runtime_finalize(), 'finalize_accelerator_runtime')
    if (node_comm /= MPI_COMM_NULL) then
      call MPI_Comm_free(node_comm, ierr)
      node_comm = MPI_COMM_NULL
    end if
    initialized = .false.
  end subroutine finalize_accelerator_runtime

  subroutine enforce_device_memory_capacity(predicted_bytes)
    integer(kind=8), intent(in) :: predicted_bytes

    if (predicted_bytes < 0_8) error stop 'negative predicted device payload'
    call wql_check(wql_capacity_check(int(predicted_bytes, c_int64_t)), &
         'enforce_device_memory_capacity')
  end subroutine enforce_device_memory_capacity

end module accelerator_runtime
