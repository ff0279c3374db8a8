! initialization (D3Q27) -- the reference's 3D state and set-up module
! (initialization_3d.f90:1-310) with the same public variables and procedures.
! The lattice state lives on the GPU in a libimexlb.so context; the host
! arrays keep the reference's names and layout and are filled when the
! caller asks for them (SetGeometry, WriteBinary).
module initialization
  use cart_mpi
  use imexlb_capi
  implicit none

  double precision charlength, pos, re, uu, ma, radius
  double precision nu, rho0, dt, ome, tau, t_intv, inter, max_t
  integer max_step, iter, interv
  integer count

  double precision, dimension(0:nq - 1) :: t
  integer, dimension(0:nq * dim - 1) :: e

  double precision, dimension(:), allocatable :: f, fb
  double precision, dimension(:), allocatable :: p
  integer, dimension(:), allocatable :: geo
  integer, dimension(:), allocatable :: bl, br, bu, bd, bf, bb, b_user
  integer, dimension(:), allocatable :: fluid_id
  integer dy, dz
  integer size_fluid, l_size, r_size, u_size, d_size, f_size, b_size, user_size
  double precision, dimension(:), allocatable :: u

  ! the device lattice and the GPU it lives on
  type(c_ptr) :: imx_ctx = c_null_ptr
  integer :: imx_device = -1

contains

  ! Stop with the library's message when a C-ABI call fails.
  subroutine imx_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= IMEXLB_OK) then
       print *, 'imexlb: ', where, ' failed (', rc, '): ', trim(imexlb_error_text())
       call MPI_ABORT(MPI_COMM_WORLD, 1, ierr)
    endif
  end subroutine imx_check

  ! main.f90:31-32 ask OpenMP for the devices; here the devices are HIP GPUs.
  integer function omp_get_num_devices()
    integer(c_int) n
    call imx_check(imexlb_device_count(n), 'imexlb_device_count')
    omp_get_num_devices = n
  end function omp_get_num_devices

  subroutine omp_set_default_device(id)
    integer, intent(in) :: id
    if (imx_device >= 0 .and. id /= imx_device) then
       print *, 'imexlb: the lattice was placed on GPU ', imx_device, ', not ', id
       call MPI_ABORT(MPI_COMM_WORLD, 1, ierr)
    endif
    imx_device = id
  end subroutine omp_set_default_device

  ! Storage/arithmetic form and phase execution from the environment
  ! (IMEXLB_PRECISION = f64 | f32 | f16 | bf16 | bf16sr, IMEXLB_FUSED = 1 | 0);
  ! input.in keeps the reference's format.
  integer function imx_precision()
    character(len=16) :: v
    integer st
    imx_precision = IMEXLB_F64
    call get_environment_variable('IMEXLB_PRECISION', v, status=st)
    if (st /= 0) return
    select case (trim(v))
    case ('f32')
       imx_precision = IMEXLB_F32
    case ('f16')
       imx_precision = IMEXLB_F16_F32
    case ('bf16')
       imx_precision = IMEXLB_BF16_F32
    case ('bf16sr')
       imx_precision = IMEXLB_BF16SR_F32
    end select
  end function imx_precision

  ! DataInput (initialization_3d.f90:60-103): read input_3d.in, derive the scalars.
  subroutine DataInput
    open(100, file='input_3d.in', status='old')
    read(100, *) nx, ny, nz
    read(100, *) charlength, pos, re
    charlength = charlength * nx
    pos = pos * nx
    read(100, *) ome, dt, rho0
    nu = (2.d0 - ome) * dt / (6.d0 * ome)
    read(100, *) max_t, inter
    t_intv = charlength**2 / nu
    interv = inter * t_intv
    max_step = max_t * t_intv
    close(100)

    uu = re * nu / charlength
    tau = 3 * nu
    ma = uu * sqrt(3.d0)
    radius = charlength * 0.5d0

    if (rank == 0) then
       write(*, '(A24,F6.2)') "Reynolds Number      :", re
       write(*, '(A24,F6.2)') "Characteristic Length:", charlength
       write(*, '(A24,F6.2)') "Ma                   :", ma
       write(*, '(A24,F6.2)') "Kinematic Viscosity  :", nu
       write(*, '(A24,F6.2)') "Relaxation Parameter :", ome
       write(*, '(A24,F6.2)') "Relaxation Time      :", tau
       write(*, '(A24,F4.1)') "Density              :", rho0
       write(*, '(A24,I8)') "Maximum # of Steps   :", max_step
       write(*, '(A24,I8)') "Interval Size        :", interv
       write(*, *) "============================="
    endif
  end subroutine DataInput

  ! InitLattice (initialization_3d.f90:108-122): q = (ez+1)*9 + (ex+1)*3 + (ey+1)
  subroutine InitLattice
    integer i, j, k, iq
    do j = -1, 1
       do i = -1, 1
          do k = -1, 1
             iq = (k + 1) * 9 + (i + 1) * 3 + (j + 1)
             e(iq * dim) = i
             e(iq * dim + 1) = j
             e(iq * dim + 2) = k
             t(iq) = 8.d0 / 4.d0**(abs(i) + abs(j) + abs(k)) / 27.d0
          enddo
       enddo
    enddo
  end subroutine InitLattice

  ! AllocateArrays: host arrays of the reference's shapes, and the device
  ! lattice on this rank's GPU.
  subroutine AllocateArrays
    integer(c_int) ndev
    allocate(f(0:array_size * nq - 1), fb(0:array_size * nq - 1))
    allocate(p(0:array_size - 1), geo(0:array_size - 1))
    allocate(fluid_id(0:array_size - 1), b_user(0:array_size - 1))
    allocate(u(0:array_size * dim - 1))
    allocate(br(0:(local_length(2) + 2 * ghost) * (local_length(3) + 2 * ghost)))
    allocate(bl(0:(local_length(2) + 2 * ghost) * (local_length(3) + 2 * ghost)))
    allocate(bu(0:(local_length(2) + 2 * ghost) * (local_length(1) + 2 * ghost)))
    allocate(bd(0:(local_length(2) + 2 * ghost) * (local_length(1) + 2 * ghost)))
    allocate(bf(0:(local_length(1) + 2 * ghost) * (local_length(3) + 2 * ghost)))
    allocate(bb(0:(local_length(1) + 2 * ghost) * (local_length(3) + 2 * ghost)))
    p = 0.d0
    u = 0.d0
    call imx_check(imexlb_device_count(ndev), 'imexlb_device_count')
    if (ndev < 1) then
       print *, 'imexlb: no GPU visible'
       call MPI_ABORT(MPI_COMM_WORLD, 1, ierr)
    endif
    imx_device = mod(rank, int(ndev))
    call imx_check(imexlb_create(imx_ctx, dim, nx, ny, nz, imx_device, imx_precision()), 'imexlb_create')
  end subroutine AllocateArrays

  subroutine DeAllocateArrays
    integer(c_int) rc
    deallocate(f, fb, p, geo, fluid_id, b_user, u, br, bl, bu, bd, bf, bb)
    rc = imexlb_destroy(imx_ctx)
    imx_ctx = c_null_ptr
  end subroutine DeAllocateArrays

  ! SetGeometry (initialization_3d.f90:170-279): the library builds the mask,
  ! the boundary lists, PassInt(geo) and the body list b_user, and returns
  ! them in the reference's layout.
  subroutine SetGeometry
    integer(c_int64_t) n
    character(len=8) :: v
    integer st
    dy = local_length(1) + 2 * ghost
    dz = (local_length(1) + 2 * ghost) * (local_length(2) + 2 * ghost)
    call imx_check(imexlb_set_params(imx_ctx, ome, rho0, uu), 'imexlb_set_params')
    call imx_check(imexlb_set_geometry(imx_ctx, pos, radius), 'imexlb_set_geometry')
    call imx_check(imexlb_geometry_geo(imx_ctx, geo, int(array_size, c_int64_t)), 'imexlb_geometry_geo')
    call fetch(IMEXLB_LIST_FLUID, fluid_id, size_fluid)
    call fetch(IMEXLB_LIST_BL, bl, l_size)
    call fetch(IMEXLB_LIST_BR, br, r_size)
    call fetch(IMEXLB_LIST_BU, bu, u_size)
    call fetch(IMEXLB_LIST_BD, bd, d_size)
    call fetch(IMEXLB_LIST_BF, bf, f_size)
    call fetch(IMEXLB_LIST_BB, bb, b_size)
    call fetch(IMEXLB_LIST_USER, b_user, user_size)
    call get_environment_variable('IMEXLB_FUSED', v, status=st)
    if (st == 0 .and. trim(v) == '0') call imx_check(imexlb_set_fused(imx_ctx, 0), 'imexlb_set_fused')
  contains
    subroutine fetch(list, arr, nsize)
      integer(c_int), intent(in) :: list
      integer, dimension(0:) :: arr
      integer, intent(out) :: nsize
      call imx_check(imexlb_geometry_size(imx_ctx, list, n), 'imexlb_geometry_size')
      nsize = int(n)
      if (n > 0) call imx_check(imexlb_geometry_list(imx_ctx, list, arr, int(size(arr), c_int64_t)), &
           'imexlb_geometry_list')
    end subroutine fetch
  end subroutine SetGeometry

  ! InitUP (initialization_3d.f90:282-308): the 100 perturbation values come
  ! from the Fortran runtime's random_number, the field is set on the GPU.
  subroutine InitUP
    double precision r_vec(0:99)
    call random_number(r_vec)
    call imx_check(imexlb_init_up(imx_ctx, r_vec), 'imexlb_init_up')
  end subroutine InitUP

end module initialization
