! rhy_capi.f90 -- ISO_C_BINDING view of include/rhyme_hip.h.
!
! The derived types rhy_params_t and rhy_ic_t and every interface below
! mirror the C declarations field for field.  The Fortran host modules
! (rhyme_host.f90) use the box and unit-level entry points; the run loop,
! batch, pool and painting interfaces are here for Fortran drivers that
! advance many boxes or serve jobs.
module rhy_capi
   use, intrinsic :: iso_c_binding
   implicit none

   integer(c_int), parameter :: RHY_OK = 0, RHY_ERR_OOM = -3
   integer(c_int), parameter :: RHY_FP64 = 0, RHY_MIXED = 1
   integer(c_int32_t), parameter :: RHY_IC_TWO_STATE = 1, RHY_IC_BLAST = 2
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

   ! riemann_problem_solution_t flattened (include/rhyme_hip.h rhy_rp_solution_t)
   type, bind(C) :: rhy_rp_solution_t
      real(c_double) :: l_rho = 0, l_v(3) = 0, l_p = 0, l_cs = 0
      real(c_double) :: r_rho = 0, r_v(3) = 0, r_p = 0, r_cs = 0
      real(c_double) :: star_p = 0, star_u = 0
      integer(c_int32_t) :: l_is_shock = 0, r_is_shock = 0
      real(c_double) :: l_shock_rho = 0, l_shock_speed = 0
      real(c_double) :: l_fan_rho = 0, l_fan_cs = 0, l_fan_speedH = 0, l_fan_speedT = 0
      real(c_double) :: r_shock_rho = 0, r_shock_speed = 0
      real(c_double) :: r_fan_rho = 0, r_fan_cs = 0, r_fan_speedH = 0, r_fan_speedT = 0
      real(c_double) :: l_f = 0, r_f = 0
   end type rhy_rp_solution_t

   ! region ids of rhy_riemann_region_batch
   integer(c_int32_t), parameter :: RHY_W_L = 1, RHY_W_R = 2, RHY_W_STARL_SHO = 3, &
                                    RHY_W_STARR_SHO = 4, RHY_W_STARL_FAN = 5, &
                                    RHY_W_STARR_FAN = 6, RHY_W_LFAN = 7, RHY_W_RFAN = 8

   type, bind(C) :: rhy_ic_t
      integer(c_int32_t) :: kind = 0
      integer(c_int32_t) :: axis = 0
      real(c_double) :: gamma = 0
      real(c_double) :: rho_l = 0, p_l = 0, rho_r = 0, p_r = 0
      real(c_double) :: radius = 0, jitter = 0, p0 = 0, energy = 0
      integer(c_int64_t) :: seed = 0
   end type rhy_ic_t

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

      integer(c_int) function rhy_riemann_solve_batch(p, n, l, r, axis, dx, dt, u) &
         bind(C, name="rhy_riemann_solve_batch")
         import :: c_int, c_int64_t, c_int32_t, c_double, rhy_params_t
         type(rhy_params_t), intent(in) :: p
         integer(c_int64_t), value :: n
         real(c_double), intent(in) :: l(*), r(*)
         integer(c_int32_t), intent(in) :: axis(*)
         real(c_double), intent(in) :: dx(*), dt(*)
         real(c_double), intent(out) :: u(*)
      end function rhy_riemann_solve_batch

      integer(c_int) function rhy_riemann_star_batch(p, n, sol, axis) &
         bind(C, name="rhy_riemann_star_batch")
         import :: c_int, c_int64_t, c_int32_t, rhy_params_t, rhy_rp_solution_t
         type(rhy_params_t), intent(in) :: p
         integer(c_int64_t), value :: n
         type(rhy_rp_solution_t), intent(inout) :: sol(*)
         integer(c_int32_t), intent(in) :: axis(*)
      end function rhy_riemann_star_batch

      integer(c_int) function rhy_riemann_sampling_batch(p, n, sol, axis, dx, dt, u) &
         bind(C, name="rhy_riemann_sampling_batch")
         import :: c_int, c_int64_t, c_int32_t, c_double, rhy_params_t, rhy_rp_solution_t
         type(rhy_params_t), intent(in) :: p
         integer(c_int64_t), value :: n
         type(rhy_rp_solution_t), intent(in) :: sol(*)
         integer(c_int32_t), intent(in) :: axis(*)
         real(c_double), intent(in) :: dx(*), dt(*)
         real(c_double), intent(out) :: u(*)
      end function rhy_riemann_sampling_batch

      integer(c_int) function rhy_riemann_region_batch(p, n, sol, region, axis, dxdt, u) &
         bind(C, name="rhy_riemann_region_batch")
         import :: c_int, c_int64_t, c_int32_t, c_double, rhy_params_t, rhy_rp_solution_t
         type(rhy_params_t), intent(in) :: p
         integer(c_int64_t), value :: n
         type(rhy_rp_solution_t), intent(in) :: sol(*)
         integer(c_int32_t), intent(in) :: region(*), axis(*)
         real(c_double), intent(in) :: dxdt(*)
         real(c_double), intent(out) :: u(*)
      end function rhy_riemann_region_batch

      integer(c_int) function rhy_irs_exec_batch(p, n, l, r, p_star, l_f, r_f) &
         bind(C, name="rhy_irs_exec_batch")
         import :: c_int, c_int64_t, c_double, rhy_params_t
         type(rhy_params_t), intent(in) :: p
         integer(c_int64_t), value :: n
         real(c_double), intent(in) :: l(*), r(*)
         real(c_double), intent(out) :: p_star(*), l_f(*), r_f(*)
      end function rhy_irs_exec_batch

      integer(c_int) function rhy_irs_guess_batch(p, n, l, r, guesses) &
         bind(C, name="rhy_irs_guess_batch")
         import :: c_int, c_int64_t, c_double, rhy_params_t
         type(rhy_params_t), intent(in) :: p
         integer(c_int64_t), value :: n
         real(c_double), intent(in) :: l(*), r(*)
         real(c_double), intent(out) :: guesses(*)
      end function rhy_irs_guess_batch

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

      integer(c_int) function rhy_box_sync(box) bind(C, name="rhy_box_sync")
         import :: c_int, c_ptr
         type(c_ptr), value :: box
      end function rhy_box_sync

      integer(c_int) function rhy_run(box, p, dx, nsteps, dts) bind(C, name="rhy_run")
         import :: c_int, c_int32_t, c_ptr, c_double, rhy_params_t
         type(c_ptr), value :: box
         type(rhy_params_t), intent(in) :: p
         real(c_double), intent(in) :: dx(3)
         integer(c_int32_t), value :: nsteps
         real(c_double), intent(out) :: dts(*)
      end function rhy_run

      integer(c_int) function rhy_box_paint(box, ic) bind(C, name="rhy_box_paint")
         import :: c_int, c_ptr, rhy_ic_t
         type(c_ptr), value :: box
         type(rhy_ic_t), intent(in) :: ic
      end function rhy_box_paint

      integer(c_int) function rhy_batch_create(ndim, nboxes, dims, dx, ncmp, precision, out) &
         bind(C, name="rhy_batch_create")
         import :: c_int, c_int32_t, c_double, c_ptr
         integer(c_int32_t), value :: ndim, nboxes, ncmp, precision
         integer(c_int32_t), intent(in) :: dims(3, *)
         real(c_double), intent(in) :: dx(3, *)
         type(c_ptr), intent(out) :: out
      end function rhy_batch_create

      integer(c_int) function rhy_batch_destroy(batch) bind(C, name="rhy_batch_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
      end function rhy_batch_destroy

      integer(c_int) function rhy_batch_upload(batch, box, cells) &
         bind(C, name="rhy_batch_upload")
         import :: c_int, c_int32_t, c_double, c_ptr
         type(c_ptr), value :: batch
         integer(c_int32_t), value :: box
         real(c_double), intent(in) :: cells(*)
      end function rhy_batch_upload

      integer(c_int) function rhy_batch_download(batch, box, cells) &
         bind(C, name="rhy_batch_download")
         import :: c_int, c_int32_t, c_double, c_ptr
         type(c_ptr), value :: batch
         integer(c_int32_t), value :: box
         real(c_double), intent(out) :: cells(*)
      end function rhy_batch_download

      integer(c_int) function rhy_batch_run_async(batch, p, nsteps, use_graph) &
         bind(C, name="rhy_batch_run_async")
         import :: c_int, c_int32_t, c_ptr, rhy_params_t
         type(c_ptr), value :: batch
         type(rhy_params_t), intent(in) :: p
         integer(c_int32_t), value :: nsteps, use_graph
      end function rhy_batch_run_async

      integer(c_int) function rhy_batch_sync(batch) bind(C, name="rhy_batch_sync")
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
      end function rhy_batch_sync

      integer(c_int) function rhy_pool_create(ndim, ncmp, precision, budget_bytes, &
                                              max_boxes, out) bind(C, name="rhy_pool_create")
         import :: c_int, c_int32_t, c_int64_t, c_ptr
         integer(c_int32_t), value :: ndim, ncmp, precision, max_boxes
         integer(c_int64_t), value :: budget_bytes
         type(c_ptr), intent(out) :: out
      end function rhy_pool_create

      integer(c_int) function rhy_pool_add_ic(pool, dims, dx, ic, limiter, bc_types, slot) &
         bind(C, name="rhy_pool_add_ic")
         import :: c_int, c_int32_t, c_double, c_ptr, rhy_ic_t
         type(c_ptr), value :: pool
         integer(c_int32_t), intent(in) :: dims(3)
         real(c_double), intent(in) :: dx(3)
         type(rhy_ic_t), intent(in) :: ic
         integer(c_int32_t), value :: limiter
         integer(c_int32_t), intent(in) :: bc_types(6)
         integer(c_int32_t), intent(out) :: slot
      end function rhy_pool_add_ic

      integer(c_int) function rhy_pool_remove(pool, slot) bind(C, name="rhy_pool_remove")
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: pool
         integer(c_int32_t), value :: slot
      end function rhy_pool_remove

      ! snapshot payloads: the FP32 data of rhyme_chombo_write_level_data and
      ! the restart reader (ncmp * product(dims) values per box)
      integer(c_int) function rhy_box_snapshot_pack(box, data) &
         bind(C, name="rhy_box_snapshot_pack")
         import :: c_int, c_ptr, c_float
         type(c_ptr), value :: box
         real(c_float), intent(out) :: data(*)
      end function rhy_box_snapshot_pack

      integer(c_int) function rhy_box_snapshot_unpack(box, data) &
         bind(C, name="rhy_box_snapshot_unpack")
         import :: c_int, c_ptr, c_float
         type(c_ptr), value :: box
         real(c_float), intent(in) :: data(*)
      end function rhy_box_snapshot_unpack

      integer(c_int) function rhy_batch_snapshot_length(batch, n) &
         bind(C, name="rhy_batch_snapshot_length")
         import :: c_int, c_int64_t, c_ptr
         type(c_ptr), value :: batch
         integer(c_int64_t), intent(out) :: n
      end function rhy_batch_snapshot_length

      integer(c_int) function rhy_batch_snapshot_pack(batch, data) &
         bind(C, name="rhy_batch_snapshot_pack")
         import :: c_int, c_ptr, c_float
         type(c_ptr), value :: batch
         real(c_float), intent(out) :: data(*)
      end function rhy_batch_snapshot_pack

      integer(c_int) function rhy_batch_snapshot_unpack(batch, data) &
         bind(C, name="rhy_batch_snapshot_unpack")
         import :: c_int, c_ptr, c_float
         type(c_ptr), value :: batch
         real(c_float), intent(in) :: data(*)
      end function rhy_batch_snapshot_unpack

      ! two-level hierarchy (ratio 2): lo(3, nboxes) places the fine boxes
      integer(c_int) function rhy_batch_cfl_dts(batch, p, dts) bind(C, name="rhy_batch_cfl_dts")
         import :: c_int, c_double, c_ptr, rhy_params_t
         type(c_ptr), value :: batch
         type(rhy_params_t), intent(in) :: p
         real(c_double), intent(out) :: dts(*)
      end function rhy_batch_cfl_dts

      integer(c_int) function rhy_batch_run_dt_async(batch, p, dt, nsteps) &
         bind(C, name="rhy_batch_run_dt_async")
         import :: c_int, c_int32_t, c_double, c_ptr, rhy_params_t
         type(c_ptr), value :: batch
         type(rhy_params_t), intent(in) :: p
         real(c_double), value :: dt
         integer(c_int32_t), value :: nsteps
      end function rhy_batch_run_dt_async

      integer(c_int) function rhy_amr_prolong(coarse, fine, lo, half) bind(C, name="rhy_amr_prolong")
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: coarse, fine
         integer(c_int32_t), intent(in) :: lo(3, *)
         integer(c_int32_t), value :: half
      end function rhy_amr_prolong

      integer(c_int) function rhy_amr_restrict(coarse, fine, lo) bind(C, name="rhy_amr_restrict")
         import :: c_int, c_int32_t, c_ptr
         type(c_ptr), value :: coarse, fine
         integer(c_int32_t), intent(in) :: lo(3, *)
      end function rhy_amr_restrict
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
