! Prints the 100 values InitUP draws (initialization.f90:262-264): the first
! random_number call of a process, unseeded, with this Fortran runtime.
program r_vec_flang
  implicit none
  double precision r_vec(0:99)
  integer i
  call random_number(r_vec)
  do i = 0, 99
     write(*, '(ES25.17)') r_vec(i)
  enddo
end program r_vec_flang
