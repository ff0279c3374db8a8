! lbm (D3Q27) -- the reference's 3D algorithm module (lbm_3d.f90:1-451) with
! the same procedures, each enqueueing gfx950 kernels through the C ABI.
module lbm
  use initialization

  double precision um
  integer(kind=MPI_OFFSET_KIND) offset

contains

  ! InitPDF (lbm_3d.f90:24-48)
  subroutine InitPDF
    call imx_check(imexlb_init_pdf(imx_ctx), 'InitPDF')
  end subroutine InitPDF

  ! Collision (lbm_3d.f90:52-76)
  subroutine Collision
    call imx_check(imexlb_collision(imx_ctx), 'Collision')
  end subroutine Collision

  ! Propagation (lbm_3d.f90:80-92)
  subroutine Propagation
    call imx_check(imexlb_propagation(imx_ctx), 'Propagation')
  end subroutine Propagation

  ! BoundaryCondition (lbm_3d.f90:96-157) incl. the ghost fill PassF does
  subroutine BoundaryCondition
    call imx_check(imexlb_boundary_condition(imx_ctx), 'BoundaryCondition')
  end subroutine BoundaryCondition

  ! PostProcessing (lbm_3d.f90:161-203)
  subroutine PostProcessing
    call imx_check(imexlb_post_processing(imx_ctx), 'PostProcessing')
  end subroutine PostProcessing

  ! Monitor (lbm_3d.f90:207-229): maximum |(ux, uy)| over the fluid.  With
  ! IMEXLB_FORCES=1 also the momentum-exchange force (lattice units) and the
  ! drag coefficient F_x / (0.5 rho0 uu^2 pi radius^2), which the reference
  ! does not compute; they need a completed step.
  subroutine Monitor
    double precision um_global, F(3)
    double precision, parameter :: pi = 3.14159265358979323846d0
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
          write(*, *) "Fx=", F(1), " Fy=", F(2), " Fz=", F(3)
          write(*, *) "Cd=", F(1) / (0.5d0 * rho0 * uu**2 * pi * radius**2)
       endif
       write(*, *) "-------------------"
    endif
  end subroutine Monitor

  ! WriteBinary (lbm_3d.f90:233-449): output_<count>.plt with x, y, z, p,
  ! ux, uy, uz in physical units, header byte for byte as the reference's.
  subroutine WriteBinary
    character filename*20, num*3
    integer, parameter :: numvars = 7
    integer i, j, k, idn, id, ind, unit_no, lx, ly, lz
    double precision vmax(0:3), vmin(0:3), conv
    double precision, dimension(:), allocatable :: buffer

    call imx_check(imexlb_download_fields(imx_ctx, p, u, IMEXLB_LAYOUT_REF), 'WriteBinary')

    write(num, '(I3.3)') count
    write(filename, '(A)') "output_"//num//".plt"

    ! limits over fluid entries 1..F-1: p, ux, uy, uz (starting from 0)
    vmax = 0.d0
    vmin = 0.d0
    do idn = 1, size_fluid - 1
       id = fluid_id(idn)
       vmax(0) = max(vmax(0), p(id)); vmin(0) = min(vmin(0), p(id))
       do ind = 0, dim - 1
          vmax(ind + 1) = max(vmax(ind + 1), u(id * dim + ind))
          vmin(ind + 1) = min(vmin(ind + 1), u(id * dim + ind))
       enddo
    enddo
    if (rank /= 0) return

    open(newunit=unit_no, file=trim(filename), access='stream', form='unformatted', status='replace')
    write(unit_no) '#!TDV112'
    write(unit_no) 1, 0, ascii('Sphere'), 0, numvars, &
         ichar('x'), 0, ichar('y'), 0, ichar('z'), 0, ichar('p'), 0, ascii('ux'), 0, ascii('uy'), 0, ascii('uz'), 0
    write(unit_no) 299.0
    write(unit_no) ascii('ZONE '//num), 0, -1, 0
    write(unit_no) iter / t_intv
    write(unit_no) -1, 0, 0, 0, 0, nx + 1, ny + 1, nz + 1, 0
    write(unit_no) 357.0, 299.0
    write(unit_no) 2, 2, 2, 2, 2, 2, 2, 0, 0, -1
    write(unit_no) 0.d0, nx / charlength, 0.d0, ny / charlength, 0.d0, nz / charlength, &
         vmin(0) * charlength / t_intv**2, vmax(0) * charlength / t_intv**2, &
         vmin(1) / (charlength * t_intv), vmax(1) / (charlength * t_intv), &
         vmin(2) / (charlength * t_intv), vmax(2) / (charlength * t_intv), &
         vmin(3) / (charlength * t_intv), vmax(3) / (charlength * t_intv)

    lx = local_length(1); ly = local_length(2); lz = local_length(3)
    allocate(buffer(lx * ly * lz))
    do ind = 1, 3  ! coordinates
       do k = 0, lz - 1
          do j = 0, ly - 1
             do i = 0, lx - 1
                select case (ind)
                case (1)
                   buffer(i + j * lx + k * lx * ly + 1) = (local_start(1) + i) / charlength
                case (2)
                   buffer(i + j * lx + k * lx * ly + 1) = (local_start(2) + j) / charlength
                case default
                   buffer(i + j * lx + k * lx * ly + 1) = (local_start(3) + k) / charlength
                end select
             enddo
          enddo
       enddo
       write(unit_no) buffer
    enddo
    do k = 0, lz - 1
       do j = 0, ly - 1
          do i = 0, lx - 1
             buffer(i + j * lx + k * lx * ly + 1) = &
                  p((i + ghost) + (lx + 2 * ghost) * (j + ghost) + (lx + 2 * ghost) * (ly + 2 * ghost) * (k + ghost)) &
                  * charlength / t_intv**2
          enddo
       enddo
    enddo
    write(unit_no) buffer
    conv = charlength * t_intv
    do ind = 0, dim - 1
       do k = 0, lz - 1
          do j = 0, ly - 1
             do i = 0, lx - 1
                buffer(i + j * lx + k * lx * ly + 1) = u(((i + ghost) + (lx + 2 * ghost) * (j + ghost) &
                     + (lx + 2 * ghost) * (ly + 2 * ghost) * (k + ghost)) * dim + ind) / conv
             enddo
          enddo
       enddo
       write(unit_no) buffer
    enddo
    deallocate(buffer)
    close(unit_no)

  contains
    function ascii(s) result(o)
      character(len=*), intent(in) :: s
      integer, dimension(len(s)) :: o
      integer kk
      do kk = 1, len(s)
         o(kk) = ichar(s(kk:kk))
      enddo
    end function ascii
  end subroutine WriteBinary

end module lbm
