! imexlb_capi.f90 -- ISO_C_BINDING interfaces to libimexlb.so (the C ABI
! declared in include/imexlb.h) for the Fortran drop-in modules cart_mpi,
! initialization and lbm.
module imexlb_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: IMEXLB_OK = 0
  integer(c_int), parameter :: IMEXLB_F64 = 0, IMEXLB_F32 = 1, IMEXLB_F16_F32 = 2, IMEXLB_BF16_F32 = 3, &
                                IMEXLB_BF16SR_F32 = 4
  integer(c_int), parameter :: IMEXLB_LAYOUT_REF = 0, IMEXLB_LAYOUT_INTERIOR = 1
  integer(c_int), parameter :: IMEXLB_PDF_F = 0, IMEXLB_PDF_FB = 1
  integer(c_int), parameter :: IMEXLB_LIST_FLUID = 0, IMEXLB_LIST_BL = 1, IMEXLB_LIST_BR = 2, &
       IMEXLB_LIST_BU = 3, IMEXLB_LIST_BD = 4, IMEXLB_LIST_BF = 5, IMEXLB_LIST_BB = 6, IMEXLB_LIST_USER = 7

  interface
     function imexlb_last_error() bind(C, name="imexlb_last_error") result(s)
       import :: c_ptr
       type(c_ptr) :: s
     end function
     function imexlb_device_count(n) bind(C, name="imexlb_device_count") result(rc)
       import :: c_int
       integer(c_int) :: n
       integer(c_int) :: rc
     end function
     function imexlb_create(ctx, dim, nx, ny, nz, device, precision) bind(C, name="imexlb_create") result(rc)
       import :: c_ptr, c_int
       type(c_ptr) :: ctx
       integer(c_int), value :: dim, nx, ny, nz, device, precision
       integer(c_int) :: rc
     end function
     function imexlb_destroy(ctx) bind(C, name="imexlb_destroy") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function
     function imexlb_set_params(ctx, ome, rho0, uu) bind(C, name="imexlb_set_params") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: ome, rho0, uu
       integer(c_int) :: rc
     end function
     function imexlb_set_geometry(ctx, pos, radius) bind(C, name="imexlb_set_geometry") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: pos, radius
       integer(c_int) :: rc
     end function
     function imexlb_geometry_size(ctx, list, n) bind(C, name="imexlb_geometry_size") result(rc)
       import :: c_ptr, c_int, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: list
       integer(c_int64_t) :: n
       integer(c_int) :: rc
     end function
     function imexlb_geometry_list(ctx, list, out, cap) bind(C, name="imexlb_geometry_list") result(rc)
       import :: c_ptr, c_int, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: list
       integer(c_int32_t) :: out(*)
       integer(c_int64_t), value :: cap
       integer(c_int) :: rc
     end function
     function imexlb_geometry_geo(ctx, geo, cap) bind(C, name="imexlb_geometry_geo") result(rc)
       import :: c_ptr, c_int, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t) :: geo(*)
       integer(c_int64_t), value :: cap
       integer(c_int) :: rc
     end function
     function imexlb_init_up(ctx, r_vec) bind(C, name="imexlb_init_up") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: r_vec(*)
       integer(c_int) :: rc
     end function
     function imexlb_upload_fields(ctx, p, u, layout) bind(C, name="imexlb_upload_fields") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: p(*), u(*)
       integer(c_int), value :: layout
       integer(c_int) :: rc
     end function
     function imexlb_set_fused(ctx, fused) bind(C, name="imexlb_set_fused") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: fused
       integer(c_int) :: rc
     end function
     function imexlb_monitor(ctx, um) bind(C, name="imexlb_monitor") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: um
       integer(c_int) :: rc
     end function
     function imexlb_forces(ctx, F) bind(C, name="imexlb_forces") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: F(3)
       integer(c_int) :: rc
     end function
     function imexlb_download_fields(ctx, p, u, layout) bind(C, name="imexlb_download_fields") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: p(*), u(*)
       integer(c_int), value :: layout
       integer(c_int) :: rc
     end function
     function imexlb_download_pdf(ctx, which, out, layout) bind(C, name="imexlb_download_pdf") result(rc)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: which
       real(c_double) :: out(*)
       integer(c_int), value :: layout
       integer(c_int) :: rc
     end function
  end interface

  ! the four per-step procedures and InitPDF share one signature
  interface
     function imexlb_init_pdf(ctx) bind(C, name="imexlb_init_pdf") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function
     function imexlb_collision(ctx) bind(C, name="imexlb_collision") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function
     function imexlb_boundary_condition(ctx) bind(C, name="imexlb_boundary_condition") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function
     function imexlb_propagation(ctx) bind(C, name="imexlb_propagation") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function
     function imexlb_post_processing(ctx) bind(C, name="imexlb_post_processing") result(rc)
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function
  end interface

contains

  ! Text of the last error raised by the library.
  function imexlb_error_text() result(msg)
    character(len=512) :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    p = imexlb_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [512])
    do i = 1, 512
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    enddo
  end function imexlb_error_text

end module imexlb_capi
