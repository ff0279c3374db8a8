! module mpi over the MPI library's Fortran include file, so that `use mpi`
! works with this compiler (the installed mpi.mod was written by another
! Fortran compiler).
module mpi
  implicit none
  include 'mpif.h'
end module mpi
