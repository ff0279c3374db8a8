! imexlb_main -- driver with the reference's call sequence (main.f90:14-81):
! set-up, then iter = 0..max_step with Monitor every interv steps and the
! four phases per step, then the output file.  Elapsed time is wall time.
program imexlb_main
  use lbm
  implicit none
  integer gpuid
  double precision start, finish

  call MPI_INIT(ierr)
  call MPI_COMM_RANK(MPI_COMM_WORLD, rank, ierr)
  call MPI_COMM_SIZE(MPI_COMM_WORLD, num_procs, ierr)

  call DataInput
  call MPISetup
  call InitLattice
  call AllocateArrays
  call SetGeometry

  gpuid = mod(rank, omp_get_num_devices())
  call omp_set_default_device(gpuid)

  call InitUP
  call InitPDF

  count = 0
  start = MPI_WTIME()
  do iter = 0, max_step
     if (mod(iter, interv) == 0) then
        call Monitor
        count = count + 1
     endif
     call Collision
     call BoundaryCondition
     call Propagation
     call PostProcessing
  enddo
  finish = MPI_WTIME()
  if (rank == 0) write(*, *) 'Elapsed time: ', finish - start

  call WriteBinary
  call DeAllocateArrays
  call MPI_FINALIZE(ierr)
end program imexlb_main
