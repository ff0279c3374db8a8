! cart_mpi (D3Q27) -- the reference's 3D topology module kept as API at one
! rank per GPU (cart_mpi_3d.f90:1-337): a 1x1x1 periodic communicator, every
! neighbour is this rank, the ghost fill of the populations happens on the
! device.  PassF / PassInt remain as host periodic fills of host arrays.
module cart_mpi
  use mpi
  implicit none
  integer, parameter :: dim = 3, nq = 27, ghost = 3, num_neighbors = 26

  integer nx, ny, nz
  integer ierr, num_procs, rank
  integer CART_COMM
  integer, dimension(dim) :: cart_coords, cart_num_procs
  logical, dimension(dim) :: cart_periods
  integer, dimension(dim) :: global_length, local_length, local_start, local_end
  integer array_size
  integer, dimension(0:num_neighbors) :: neighbor_procs

contains

  subroutine MPISetup
    call SetCartesian
    call SetLocalLength
    call SetNeighbors
  end subroutine MPISetup

  subroutine SetCartesian
    global_length(1) = nx + 1
    global_length(2) = ny + 1
    global_length(3) = nz + 1
    if (num_procs /= 1) then
       if (rank == 0) print *, 'imexlb: one MPI rank per GPU holds the whole lattice; run with -np 1'
       call MPI_ABORT(MPI_COMM_WORLD, 1, ierr)
    endif
    cart_num_procs = 1
    cart_periods = .true.
    call MPI_CART_CREATE(MPI_COMM_WORLD, dim, cart_num_procs, cart_periods, .true., CART_COMM, ierr)
    call MPI_COMM_RANK(CART_COMM, rank, ierr)
    call MPI_CART_COORDS(CART_COMM, rank, dim, cart_coords, ierr)
  end subroutine SetCartesian

  subroutine SetLocalLength
    integer i
    do i = 1, dim
       local_length(i) = global_length(i)
       local_start(i) = 0
       local_end(i) = global_length(i) - 1
       if (local_length(i) < ghost) then
          print *, 'Dimension = ', i, ' in Cartesian proc id =', rank, ' smaller than GhostCells'
          call MPI_ABORT(MPI_COMM_WORLD, 0, ierr)
       endif
    enddo
    array_size = 1
    do i = 1, dim
       array_size = array_size * (local_length(i) + 2 * ghost)
    enddo
  end subroutine SetLocalLength

  subroutine SetNeighbors
    neighbor_procs = rank
  end subroutine SetNeighbors

  pure integer function image_id(i, j, k) result(id)
    integer, intent(in) :: i, j, k
    integer x, y, z
    x = modulo(i - ghost, local_length(1))
    y = modulo(j - ghost, local_length(2))
    z = modulo(k - ghost, local_length(3))
    id = (x + ghost) + (local_length(1) + 2 * ghost) * ((y + ghost) + (local_length(2) + 2 * ghost) * (z + ghost))
  end function image_id

  pure logical function is_ghost(i, j, k)
    integer, intent(in) :: i, j, k
    is_ghost = i < ghost .or. i >= ghost + local_length(1) .or. j < ghost .or. j >= ghost + local_length(2) &
         .or. k < ghost .or. k >= ghost + local_length(3)
  end function is_ghost

  subroutine PassF(array)
    double precision, dimension(0:nq * array_size - 1) :: array
    integer i, j, k, id, src
    do k = 0, local_length(3) + 2 * ghost - 1
       do j = 0, local_length(2) + 2 * ghost - 1
          do i = 0, local_length(1) + 2 * ghost - 1
             if (.not. is_ghost(i, j, k)) cycle
             id = i + (local_length(1) + 2 * ghost) * (j + (local_length(2) + 2 * ghost) * k)
             src = image_id(i, j, k)
             array(id * nq:id * nq + nq - 1) = array(src * nq:src * nq + nq - 1)
          enddo
       enddo
    enddo
  end subroutine PassF

  subroutine PassInt(array)
    integer, dimension(0:array_size - 1) :: array
    integer i, j, k, id
    do k = 0, local_length(3) + 2 * ghost - 1
       do j = 0, local_length(2) + 2 * ghost - 1
          do i = 0, local_length(1) + 2 * ghost - 1
             if (.not. is_ghost(i, j, k)) cycle
             id = i + (local_length(1) + 2 * ghost) * (j + (local_length(2) + 2 * ghost) * k)
             array(id) = array(image_id(i, j, k))
          enddo
       enddo
    enddo
  end subroutine PassInt

end module cart_mpi
