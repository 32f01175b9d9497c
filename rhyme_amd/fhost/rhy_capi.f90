! rhy_capi.f90 -- ISO_C_BINDING view of include/rhyme_hip.h.
!
! The derived type rhy_params_t and every interface below mirror the C
! declarations field for field; the Fortran host modules (rhyme_host.f90)
! are the only users.
module rhy_capi
   use, intrinsic :: iso_c_binding
   implicit none

   integer(c_int), parameter :: RHY_OK = 0
   integer(c_int), parameter :: RHY_FP64 = 0, RHY_MIXED = 1
   integer, parameter :: RHY_MAX_CMP = 16

   type, bind(C) :: rhy_params_t
      integer(c_int32_t) :: ndim = 0
      integer(c_int32_t) :: limiter = 2
      integer(c_int32_t) :: solver_type = 2
      integer(c_int32_t) :: n_iteration = 10000
      real(c_double) :: gamma = 0
      real(c_double) :: kb_amu = 0
      real(c_double) :: omega = 0
      real(c_double) :: tolerance = 1d-8
      real(c_double) :: w_vacuum_rho = 0
      real(c_double) :: w_vacuum_p = 0
      real(c_double) :: courant = .8d0
      integer(c_int32_t) :: bc_types(6) = 1
      integer(c_int32_t) :: ncmp = 0
      integer(c_int32_t) :: reserved = 0
      real(c_double) :: cons_inflows(RHY_MAX_CMP, 6) = 0
   end type rhy_params_t

   interface
      integer(c_int) function rhy_init(device) bind(C, name="rhy_init")
         import :: c_int, c_int32_t
         integer(c_int32_t), value :: device
      end function rhy_init

      type(c_ptr) function rhy_last_error() bind(C, name="rhy_last_error")
         import :: c_ptr
      end function rhy_last_error

      integer(c_int) function rhy_box_create(ndim, dims, ncmp, precision, out) &
         bind(C, name="rhy_box_create")
         import :: c_int, c_int32_t, c_ptr
         integer(c_int32_t), value :: ndim, ncmp, precision
         integer(c_int32_t), intent(in) :: dims(3)
         type(c_ptr), intent(out) :: out
      end function rhy_box_create

      integer(c_int) function rhy_box_resize(box, dims) bind(C, name="rhy_box_resize")
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: box
         integer(c_int32_t), intent(in) :: dims(3)
      end function rhy_box_resize

      integer(c_int) function rhy_box_destroy(box) bind(C, name="rhy_box_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: box
      end function rhy_box_destroy

      integer(c_int) function rhy_box_upload(box, cells) bind(C, name="rhy_box_upload")
         import :: c_int, c_ptr
         type(c_ptr), value :: box
         type(c_ptr), value :: cells
      end function rhy_box_upload

      integer(c_int) function rhy_box_download(box, cells) bind(C, name="rhy_box_download")
         import :: c_int, c_ptr
         type(c_ptr), value :: box
         type(c_ptr), value :: cells
      end function rhy_box_download

      integer(c_int) function rhy_bc_set_boundaries(box, p) &
         bind(C, name="rhy_bc_set_boundaries")
         import :: c_int, c_ptr, rhy_params_t
         type(c_ptr), value :: box
         type(rhy_params_t), intent(in) :: p
      end function rhy_bc_set_boundaries

      integer(c_int) function rhy_bc_set_face(box, p, face) bind(C, name="rhy_bc_set_face")
         import :: c_int, c_int32_t, c_ptr, rhy_params_t
         type(c_ptr), value :: box
         type(rhy_params_t), intent(in) :: p
         integer(c_int32_t), value :: face
      end function rhy_bc_set_face

      integer(c_int) function rhy_cfl_time_step(box, p, dx, dt) &
         bind(C, name="rhy_cfl_time_step")
         import :: c_int, c_ptr, c_double, rhy_params_t
         type(c_ptr), value :: box
         type(rhy_params_t), intent(in) :: p
         real(c_double), intent(in) :: dx(3)
         real(c_double), intent(out) :: dt
      end function rhy_cfl_time_step

      integer(c_int) function rhy_mh_solve(box, p, dx, dt) bind(C, name="rhy_mh_solve")
         import :: c_int, c_ptr, c_double, rhy_params_t
         type(c_ptr), value :: box
         type(rhy_params_t), intent(in) :: p
         real(c_double), intent(in) :: dx(3)
         real(c_double), value :: dt
      end function rhy_mh_solve

      integer(c_int) function rhy_riemann_batch(p, n, l, r, axis, u) &
         bind(C, name="rhy_riemann_batch")
         import :: c_int, c_int64_t, c_int32_t, c_double, rhy_params_t
         type(rhy_params_t), intent(in) :: p
         integer(c_int64_t), value :: n
         real(c_double), intent(in) :: l(*), r(*)
         integer(c_int32_t), intent(in) :: axis(*)
         real(c_double), intent(out) :: u(*)
      end function rhy_riemann_batch

      integer(c_int) function rhy_limiter_batch(p, n, ul, u, ur, delta) &
         bind(C, name="rhy_limiter_batch")
         import :: c_int, c_int64_t, c_double, rhy_params_t
         type(rhy_params_t), intent(in) :: p
         integer(c_int64_t), value :: n
         real(c_double), intent(in) :: ul(*), u(*), ur(*)
         real(c_double), intent(out) :: delta(*)
      end function rhy_limiter_batch

      integer(c_int) function rhy_half_step_batch(p, n, u, delta, axis, dx, dt, l, r) &
         bind(C, name="rhy_half_step_batch")
         import :: c_int, c_int64_t, c_int32_t, c_double, rhy_params_t
         type(rhy_params_t), intent(in) :: p
         integer(c_int64_t), value :: n
         real(c_double), intent(in) :: u(*), delta(*)
         integer(c_int32_t), intent(in) :: axis(*)
         real(c_double), value :: dx, dt
         real(c_double), intent(out) :: l(*), r(*)
      end function rhy_half_step_batch
   end interface

contains

   ! Text of the library's last error as a Fortran string.
   function rhy_error_text() result(msg)
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: n
      p = rhy_last_error()
      if (.not. c_associated(p)) then
         msg = ''
         return
      end if
      call c_f_pointer(p, chars, [4096])
      n = 0
      do while (n < 4096)
         if (chars(n + 1) == c_null_char) exit
         n = n + 1
      end do
      allocate (character(len=n) :: msg)
      msg = transfer(chars(1:n), msg)
   end function rhy_error_text
end module rhy_capi
