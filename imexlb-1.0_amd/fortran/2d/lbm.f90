! lbm (D2Q9) -- the reference's algorithm module (lbm.f90:1-415) with the
! same procedures.  Each one enqueues the corresponding gfx950 kernels
! through the C ABI; Monitor and WriteBinary read results back.
module lbm
  use initialization

  ! maximum velocity on this rank
  double precision um
  ! writing offset of the output file
  integer(kind=MPI_OFFSET_KIND) offset

contains

  ! InitPDF (lbm.f90:24-48): f = feq(p, u) on fluid nodes.
  subroutine InitPDF
    call imx_check(imexlb_init_pdf(imx_ctx), 'InitPDF')
  end subroutine InitPDF

  ! Collision (lbm.f90:52-76)
  subroutine Collision
    call imx_check(imexlb_collision(imx_ctx), 'Collision')
  end subroutine Collision

  ! Propagation (lbm.f90:80-92)
  subroutine Propagation
    call imx_check(imexlb_propagation(imx_ctx), 'Propagation')
  end subroutine Propagation

  ! BoundaryCondition (lbm.f90:96-155) incl. the ghost fill PassF does
  subroutine BoundaryCondition
    call imx_check(imexlb_boundary_condition(imx_ctx), 'BoundaryCondition')
  end subroutine BoundaryCondition

  ! PostProcessing (lbm.f90:159-201)
  subroutine PostProcessing
    call imx_check(imexlb_post_processing(imx_ctx), 'PostProcessing')
  end subroutine PostProcessing

  ! Monitor (lbm.f90:205-227): maximum |(ux, uy)| over the fluid, printed.
  ! With IMEXLB_FORCES=1 it also prints the momentum-exchange drag and lift
  ! (lattice units) and their coefficients F / (rho0 uu^2 radius), which
  ! the reference does not compute; they need a completed step.
  subroutine Monitor
    double precision um_global, F(3)
    character(len=8) :: v
    integer st
    call imx_check(imexlb_monitor(imx_ctx, um), 'Monitor')
    call MPI_BARRIER(CART_COMM, ierr)
    call MPI_REDUCE(um, um_global, 1, MPI_DOUBLE_PRECISION, MPI_MAX, 0, CART_COMM, ierr)
    if (rank == 0) then
       write(*, *) "T=", iter / t_intv
       write(*, *) "Umax=", um_global / (charlength * t_intv)
       call get_environment_variable('IMEXLB_FORCES', v, status=st)
       if (st == 0 .and. trim(v) == '1' .and. iter > 0) then
          call imx_check(imexlb_forces(imx_ctx, F), 'Monitor forces')
          write(*, *) "Drag=", F(1), " Lift=", F(2)
          write(*, *) "Cd=", F(1) / (rho0 * uu**2 * radius), " Cl=", F(2) / (rho0 * uu**2 * radius)
       endif
       write(*, *) "-------------------"
    endif
  end subroutine Monitor

  ! WriteBinary (lbm.f90:231-413): Tecplot binary output_<count>.plt of
  ! x, y, p, ux, uy in physical units, header byte for byte as the
  ! reference writes it.  The variable limits are the true minima/maxima
  ! over fluid entries 1..F-1 (including 0).
  subroutine WriteBinary
    character filename*20, num*3
    integer, parameter :: numvars = 5
    integer i, j, idn, id, unit_no
    double precision uxmax, uymax, pmax, uxmin, uymin, pmin
    double precision, dimension(:), allocatable :: buffer

    call imx_check(imexlb_download_fields(imx_ctx, p, u, IMEXLB_LAYOUT_REF), 'WriteBinary')

    write(num, '(I3.3)') count
    write(filename, '(A)') "output_"//num//".plt"

    uxmax = 0.d0; uxmin = 0.d0; uymax = 0.d0; uymin = 0.d0; pmax = 0.d0; pmin = 0.d0
    do idn = 1, size_fluid - 1
       id = fluid_id(idn)
       uxmax = max(uxmax, u(id * dim)); uxmin = min(uxmin, u(id * dim))
       uymax = max(uymax, u(id * dim + 1)); uymin = min(uymin, u(id * dim + 1))
       pmax = max(pmax, p(id)); pmin = min(pmin, p(id))
    enddo
    if (rank /= 0) return

    open(newunit=unit_no, file=trim(filename), access='stream', form='unformatted', status='replace')
    write(unit_no) '#!TDV112'
    write(unit_no) 1, 0, ascii('Cylinder'), 0, numvars, &
         ichar('x'), 0, ichar('y'), 0, ichar('p'), 0, ascii('ux'), 0, ascii('uy'), 0
    write(unit_no) 299.0
    write(unit_no) ascii('ZONE '//num), 0, -1, 0
    write(unit_no) iter / t_intv
    write(unit_no) -1, 0, 0, 0, 0, nx + 1, ny + 1, 1, 0
    write(unit_no) 357.0, 299.0
    write(unit_no) 2, 2, 2, 2, 2, 0, 0, -1
    write(unit_no) 0.d0, nx / charlength, 0.d0, ny / charlength, &
         pmin * charlength / t_intv**2, pmax * charlength / t_intv**2, &
         uxmin / (charlength * t_intv), uxmax / (charlength * t_intv), &
         uymin / (charlength * t_intv), uymax / (charlength * t_intv)

    allocate(buffer(local_length(1) * local_length(2)))
    do j = 0, local_length(2) - 1
       do i = 0, local_length(1) - 1
          buffer(i + j * local_length(1) + 1) = (local_start(1) + i) / charlength
       enddo
    enddo
    write(unit_no) buffer
    do j = 0, local_length(2) - 1
       do i = 0, local_length(1) - 1
          buffer(i + j * local_length(1) + 1) = (local_start(2) + j) / charlength
       enddo
    enddo
    write(unit_no) buffer
    do j = 0, local_length(2) - 1
       do i = 0, local_length(1) - 1
          buffer(i + j * local_length(1) + 1) = p((i + ghost) + (local_length(1) + 2 * ghost) * (j + ghost)) &
               * charlength / t_intv**2
       enddo
    enddo
    write(unit_no) buffer
    do id = 0, dim - 1
       do j = 0, local_length(2) - 1
          do i = 0, local_length(1) - 1
             buffer(i + j * local_length(1) + 1) = &
                  u(((i + ghost) + (local_length(1) + 2 * ghost) * (j + ghost)) * dim + id) / (charlength * t_intv)
          enddo
       enddo
       write(unit_no) buffer
    enddo
    deallocate(buffer)
    close(unit_no)

  contains
    ! character codes of a string, one default integer each
    function ascii(s) result(o)
      character(len=*), intent(in) :: s
      integer, dimension(len(s)) :: o
      integer k
      do k = 1, len(s)
         o(k) = ichar(s(k:k))
      enddo
    end function ascii
  end subroutine WriteBinary

end module lbm
