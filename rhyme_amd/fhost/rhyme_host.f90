! rhyme_host.f90 -- Fortran host modules of the hydro hot path.
!
! These modules keep the names, derived types, defaults and procedure
! argument lists of the reference's hot-path interfaces so that a Rhyme
! driver (src/main.f90:106-177) links against them unchanged; every
! procedure body calls the gfx950 library through rhy_capi (ISO_C_BINDING).
! File:line citations are relative to the reference's src/.
!
! Residency: the first call that touches a box creates its device copy from
! box%cells; from then on the device copy is authoritative.  After writing
! box%cells on the host call rhyme_samr_upload(samr); before reading them
! call rhyme_samr_download(samr) (the reference's host readers -- chombo
! output, report, sanity check, stabilizer, src/main.f90:122-155 -- are the
! places that need it).
!
! Compile with -cpp -DNDIM=<1|2|3> [-DNSPE=3], as the reference is
! (config.cmake:8-18).

#ifndef NSPE
#define NSPE 3
#endif
#define NCMP (1 + NDIM + 1 + 1 + NSPE)

module rhyme_units
   ! component indices of rhyme_units/src/rhyme_units.f90:32-57
   implicit none

   type, private :: component_indices_t
      integer :: rho = 1
      integer :: rho_u = 2, u = 2
#if NDIM > 1
      integer :: rho_v = 3, v = 3
#endif
#if NDIM > 2
      integer :: rho_w = 4, w = 4
#endif
      integer :: e_tot = 1 + NDIM + 1, p = 1 + NDIM + 1
      integer :: temp = 1 + NDIM + 1 + 1
      integer :: ntr_frac_0 = 1 + NDIM + 1 + 1 + 1
   end type component_indices_t

   type(component_indices_t), parameter :: cid = component_indices_t()

   ! only the values of kB and amu in code units cross into the hot path
   ! (rhyme_thermo_base.f90:130)
   type nombre_t
      real(kind=8) :: v = 0d0
   end type nombre_t

   type units_t
      type(nombre_t) :: kb, amu
   end type units_t
end module rhyme_units

module rhyme_logger
   ! Minimal logger with the calls the hot path makes (logger_t,
   ! rhyme_logger/src/rhyme_logger.f90:24-68): begin/end_section, log, warn, err.
   implicit none

   integer, save :: rhy_host_nerr = 0
   character(len=1024), save :: rhy_host_last_err = ''

   type logger_t
      integer :: secid = 0
      logical :: quiet = .true.
      integer :: nerr = 0
      character(len=1024) :: last_err = ''
   contains
      procedure :: begin_section => logger_begin_section
      procedure :: end_section => logger_end_section
      procedure :: log => logger_log
      procedure :: warn => logger_warn
      procedure :: err => logger_err
   end type logger_t

contains

   subroutine logger_begin_section(logger, name)
      class(logger_t), intent(inout) :: logger
      class(*), intent(in) :: name
      logger%secid = logger%secid + 1
   end subroutine logger_begin_section

   subroutine logger_end_section(logger, print_duration)
      class(logger_t), intent(inout) :: logger
      logical, intent(in), optional :: print_duration
      logger%secid = max(0, logger%secid - 1)
   end subroutine logger_end_section

   subroutine logger_log(logger, message)
      class(logger_t), intent(inout) :: logger
      character(len=*), intent(in) :: message
      if (.not. logger%quiet) print '(A)', trim(message)
   end subroutine logger_log

   subroutine logger_warn(logger, message)
      class(logger_t), intent(inout) :: logger
      character(len=*), intent(in) :: message
      print '(A,A)', 'warning: ', trim(message)
   end subroutine logger_warn

   subroutine logger_err(logger, message)
      class(logger_t), intent(inout) :: logger
      character(len=*), intent(in) :: message
      logger%nerr = logger%nerr + 1
      logger%last_err = message
      write (0, '(A,A)') 'error: ', trim(message)
   end subroutine logger_err

   ! Errors of the procedures whose reference interfaces take no logger
   ! (rhyme_cfl_time_step, rhyme_samr_bc_set_*, the two solver variants, the
   ! unit procedures of the Riemann solver, limiter and half step): printed
   ! like logger%err and counted here, readable through rhy_host_nerr and
   ! rhy_host_last_err.  The host arrays such a procedure would have written
   ! are left as they were.
   subroutine rhy_host_error(message)
      character(len=*), intent(in) :: message
      rhy_host_nerr = rhy_host_nerr + 1
      rhy_host_last_err = message
      write (0, '(A,A)') 'error: ', trim(message)
   end subroutine rhy_host_error
end module rhyme_logger

module rhyme_thermo_base
   ! rhyme_thermo_base/src/rhyme_thermo_base.f90:8-26,118-134
   use rhyme_units
   use rhyme_logger
   implicit none

   type, private :: thermo_base_indices_t
      integer :: unset = -1
      integer :: monatomic = 1
      integer :: diatomic = 2
      integer :: polyatomic = 3
   end type thermo_base_indices_t

   type(thermo_base_indices_t), parameter :: thid = thermo_base_indices_t()

   real(kind=8), parameter :: ig_gamma(thid%monatomic:thid%polyatomic) = &
                              [5.d0/3.d0, 7.d0/5.d0, 4.d0/3.d0]

   integer, private :: state_of_matter = -1
   real(kind=8), private :: kb_amu = 0.d0

   type thermo_base_t
      integer :: state_of_matter = -1
   end type thermo_base_t

contains

   subroutine rhyme_thermo_base_init(thermo, units, logger)
      type(thermo_base_t), intent(in) :: thermo
      type(units_t), intent(in) :: units
      type(logger_t), intent(inout) :: logger
      state_of_matter = thermo%state_of_matter
      kb_amu = units%kb%v/units%amu%v
      call logger%log('thermo_base initialised')
   end subroutine rhyme_thermo_base_init

   pure function get_gamma() result(g)
      real(kind=8) :: g
      if (state_of_matter >= thid%monatomic .and. state_of_matter <= thid%polyatomic) then
         g = ig_gamma(state_of_matter)
      else
         g = -huge(0d0)
      end if
   end function get_gamma

   pure function get_kb_amu() result(v)
      real(kind=8) :: v
      v = kb_amu
   end function get_kb_amu
end module rhyme_thermo_base

module rhyme_samr
   ! rhyme_samr/src/rhyme_samr.f90:23-65 and rhyme_samr_init_box.f90:29-33
   use rhyme_units
   implicit none

   type, private :: samr_indices_t
      integer :: ghost = -1
      integer :: unset = -10
      integer :: max_nlevels = 23
      integer :: x = 1, left = 1, right = 2
#if NDIM > 1
      integer :: y = 2, bottom = 3, top = 4
#endif
#if NDIM > 2
      integer :: z = 3, back = 5, front = 6
#endif
   end type samr_indices_t

   type(samr_indices_t), parameter :: samrid = samr_indices_t()

   type samr_box_t
      integer :: dims(NDIM) = samrid%unset
      integer :: level = samrid%unset
      integer :: number = samrid%unset
      integer :: left_edge(NDIM) = samrid%unset
      integer :: right_edge(NDIM) = samrid%unset
#if NDIM == 1
      integer, allocatable :: flags(:)
      real(kind=8), allocatable :: cells(:, :)
#elif NDIM == 2
      integer, allocatable :: flags(:, :)
      real(kind=8), allocatable :: cells(:, :, :)
#else
      integer, allocatable :: flags(:, :, :)
      real(kind=8), allocatable :: cells(:, :, :, :)
#endif
   end type samr_box_t

   type samr_level_t
      integer :: level = samrid%unset
      integer :: nboxes = 0
      integer :: max_nboxes = 0
      integer :: iteration = 0
      real(kind=8) :: refine_factor = 2d0
      real(kind=8) :: t = 0.d0
      real(kind=8) :: dt = 0.d0, dx(NDIM) = 0.d0
      type(samr_box_t), allocatable :: boxes(:)
   end type samr_level_t

   type samr_t
      logical :: initialized = .false.
      integer :: nlevels = 0
      integer :: base_grid(NDIM) = 0
      integer :: ghost_cells(NDIM) = 2
      integer :: max_nboxes(0:23) = 0
      real(kind=8) :: box_lengths(NDIM) = 1d0
      type(samr_level_t) :: levels(0:23)
   end type samr_t

contains

   subroutine rhyme_samr_init_box(samr, l, b, dims, ledges, redges)
      type(samr_t), intent(inout) :: samr
      integer, intent(in) :: l, b, dims(NDIM), ledges(NDIM), redges(NDIM)
      integer :: lb(NDIM), ub(NDIM)

      samr%levels(l)%boxes(b)%level = l
      samr%levels(l)%boxes(b)%number = b
      lb = -samr%ghost_cells + 1
      ub = dims + samr%ghost_cells
#if NDIM == 1
      allocate (samr%levels(l)%boxes(b)%flags(lb(1):ub(1)))
      allocate (samr%levels(l)%boxes(b)%cells(lb(1):ub(1), NCMP))
#elif NDIM == 2
      allocate (samr%levels(l)%boxes(b)%flags(lb(1):ub(1), lb(2):ub(2)))
      allocate (samr%levels(l)%boxes(b)%cells(lb(1):ub(1), lb(2):ub(2), NCMP))
#else
      allocate (samr%levels(l)%boxes(b)%flags(lb(1):ub(1), lb(2):ub(2), lb(3):ub(3)))
      allocate (samr%levels(l)%boxes(b)%cells(lb(1):ub(1), lb(2):ub(2), lb(3):ub(3), NCMP))
#endif
      samr%levels(l)%boxes(b)%cells = 0d0
      samr%levels(l)%nboxes = samr%levels(l)%nboxes + 1
      samr%levels(l)%boxes(b)%dims = dims
      samr%levels(l)%boxes(b)%left_edge = ledges
      samr%levels(l)%boxes(b)%right_edge = redges
   end subroutine rhyme_samr_init_box
end module rhyme_samr

module rhy_device
   ! Registry of device copies, one per (level, box).  Host mirror sync is
   ! explicit: rhyme_samr_upload / rhyme_samr_download.
   use, intrinsic :: iso_c_binding
   use rhy_capi
   use rhyme_samr
   use rhyme_logger
   implicit none

   type rhy_slot_t
      type(c_ptr) :: h = c_null_ptr
      integer :: dims(NDIM) = -1
   end type rhy_slot_t

   type(rhy_slot_t), save, private :: slots(0:23, 1:1024)
   integer(c_int32_t), save :: rhy_precision = RHY_FP64

contains

   ! Device handle of a box, creating (and uploading) it on first use or
   ! after a shape change.  Returns c_null_ptr and a message on failure.
   function rhy_box_handle(box, msg) result(h)
      type(samr_box_t), intent(inout), target :: box
      character(len=:), allocatable, intent(out) :: msg
      type(c_ptr) :: h
      integer(c_int32_t) :: d3(3)
      integer(c_int) :: rc
      integer :: l, b

      msg = ''
      h = c_null_ptr
      l = box%level
      b = box%number
      if (l < 0 .or. l > 23 .or. b < 1 .or. b > 1024) then
         msg = 'box level/number out of range'
         return
      end if
      d3 = 1
      d3(1:NDIM) = box%dims
      if (c_associated(slots(l, b)%h)) then
         if (all(slots(l, b)%dims == box%dims)) then
            h = slots(l, b)%h
            return
         end if
         rc = rhy_box_resize(slots(l, b)%h, d3)
         if (rc /= RHY_OK) then
            msg = rhy_error_text()
            return
         end if
      else
         rc = rhy_box_create(int(NDIM, c_int32_t), d3, int(NCMP, c_int32_t), rhy_precision, &
                             slots(l, b)%h)
         if (rc /= RHY_OK) then
            msg = rhy_error_text()
            slots(l, b)%h = c_null_ptr
            return
         end if
      end if
      slots(l, b)%dims = box%dims
      rc = rhy_box_upload(slots(l, b)%h, c_loc(box%cells))
      if (rc /= RHY_OK) then
         msg = rhy_error_text()
         return
      end if
      h = slots(l, b)%h
   end function rhy_box_handle

   subroutine rhyme_samr_box_upload(box)
      type(samr_box_t), intent(inout), target :: box
      type(c_ptr) :: h
      character(len=:), allocatable :: msg
      integer(c_int) :: rc
      h = rhy_box_handle(box, msg)
      if (.not. c_associated(h)) then
         call rhy_host_error('rhyme_samr_upload: '//msg)
         return
      end if
      rc = rhy_box_upload(h, c_loc(box%cells))
      if (rc /= RHY_OK) call rhy_host_error('rhyme_samr_upload: '//rhy_error_text())
   end subroutine rhyme_samr_box_upload

   subroutine rhyme_samr_box_download(box)
      type(samr_box_t), intent(inout), target :: box
      type(c_ptr) :: h
      character(len=:), allocatable :: msg
      integer(c_int) :: rc
      h = rhy_box_handle(box, msg)
      if (.not. c_associated(h)) then
         call rhy_host_error('rhyme_samr_download: '//msg)
         return
      end if
      rc = rhy_box_download(h, c_loc(box%cells))
      if (rc /= RHY_OK) call rhy_host_error('rhyme_samr_download: '//rhy_error_text())
   end subroutine rhyme_samr_box_download

   subroutine rhyme_samr_upload(samr)
      type(samr_t), intent(inout) :: samr
      integer :: l, b
      do l = 0, samr%nlevels - 1
         do b = 1, samr%levels(l)%nboxes
            call rhyme_samr_box_upload(samr%levels(l)%boxes(b))
         end do
      end do
   end subroutine rhyme_samr_upload

   subroutine rhyme_samr_download(samr)
      type(samr_t), intent(inout) :: samr
      integer :: l, b
      do l = 0, samr%nlevels - 1
         do b = 1, samr%levels(l)%nboxes
            call rhyme_samr_box_download(samr%levels(l)%boxes(b))
         end do
      end do
   end subroutine rhyme_samr_download

   ! FP32 payload of a level, box after box, as rhyme_chombo_write_level_data
   ! fills chws%data (rhyme_chombo_write_level_data.f90:41-71), packed on the
   ! device.  data must hold NCMP * sum(product(dims)) values; ok is false
   ! (and msg set) when a box has no device copy or the call fails.
   subroutine rhyme_samr_level_snapshot(level, data, ok, msg)
      type(samr_level_t), intent(inout), target :: level
      real(kind=4), intent(out), target :: data(:)
      logical, intent(out) :: ok
      character(len=:), allocatable, intent(out) :: msg
      type(c_ptr) :: h
      integer(c_int) :: rc
      integer :: b, offset, n
      ok = .false.
      msg = ''
      offset = 1
      do b = 1, level%nboxes
         n = NCMP*product(level%boxes(b)%dims)
         if (offset + n - 1 > size(data)) then
            msg = 'snapshot buffer too small'
            return
         end if
         h = rhy_box_handle(level%boxes(b), msg)
         if (.not. c_associated(h)) return
         rc = rhy_box_snapshot_pack(h, data(offset:offset + n - 1))
         if (rc /= RHY_OK) then
            msg = rhy_error_text()
            return
         end if
         offset = offset + n
      end do
      ok = .true.
   end subroutine rhyme_samr_level_snapshot

   subroutine rhy_device_release_all()
      integer :: l, b
      integer(c_int) :: rc
      do l = 0, 23
         do b = 1, 1024
            if (c_associated(slots(l, b)%h)) then
               rc = rhy_box_destroy(slots(l, b)%h)
               slots(l, b)%h = c_null_ptr
               slots(l, b)%dims = -1
            end if
         end do
      end do
   end subroutine rhy_device_release_all
end module rhy_device

module rhyme_slope_limiter
   ! rhyme_slope_limiter/src/rhyme_slope_limiter.f90:6-23
   use, intrinsic :: iso_c_binding
   use rhy_capi
   use rhyme_units
   use rhyme_thermo_base
   use rhyme_logger
   implicit none

   type, private :: slope_limiter_indices_t
      integer :: van_Leer = 1, minmod = 2, van_albada = 3, superbee = 4
   end type slope_limiter_indices_t

   type(slope_limiter_indices_t), parameter :: slid = slope_limiter_indices_t()

   type slope_limiter_t
      real(kind=8) :: w = 0.d0
      integer :: type = slid%minmod
   end type slope_limiter_t

contains

   subroutine rhyme_slope_limiter_run(sl, ul, u, ur, delta)
      class(slope_limiter_t), intent(in) :: sl
      real(kind=8), dimension(cid%rho:cid%e_tot), intent(in) :: ul, u, ur
      real(kind=8), dimension(cid%rho:cid%e_tot), intent(out) :: delta
      type(rhy_params_t) :: p
      integer(c_int) :: rc
      p%ndim = NDIM
      p%limiter = sl%type
      p%omega = sl%w
      p%gamma = get_gamma()
      p%kb_amu = get_kb_amu()
      p%ncmp = NCMP
      rc = rhy_limiter_batch(p, 1_c_int64_t, ul, u, ur, delta)
      if (rc /= RHY_OK) then
         delta = 0d0
         call rhy_host_error('rhyme_slope_limiter_run: '//rhy_error_text())
      end if
   end subroutine rhyme_slope_limiter_run
end module rhyme_slope_limiter

module rhyme_irs
   ! rhyme_riemann_problem/src/rhyme_irs/src/rhyme_irs.f90:16-52: the exact
   ! iterative Riemann solver's p* and its starting guesses, on the device.
   ! The gamma constants the reference keeps module-private (:9-14) come from
   ! rhyme_thermo_base at each call.  rhyme_irs_guess_p_star is not pure here
   ! (it calls the device); its only reference caller is rhyme_irs_exec.
   use, intrinsic :: iso_c_binding
   use rhy_capi
   use rhyme_units
   use rhyme_thermo_base
   use rhyme_logger
   implicit none

   type irs_t
      integer :: n_iteration = 10000
      real(kind=8) :: tolerance = 1d-8
   end type irs_t

contains

   subroutine rhyme_irs_init(irs, thermo, logger)
      type(irs_t), intent(inout) :: irs
      type(thermo_base_t), intent(in) :: thermo
      type(logger_t), intent(inout) :: logger
      integer(c_int) :: rc
      rc = rhy_init(0_c_int32_t)
      if (rc /= RHY_OK) call logger%err('rhy_init: '//rhy_error_text())
   end subroutine rhyme_irs_init

   subroutine rhy_irs_params(irs, p)
      type(irs_t), intent(in) :: irs
      type(rhy_params_t), intent(inout) :: p
      p%ndim = NDIM
      p%gamma = get_gamma()
      p%kb_amu = get_kb_amu()
      p%n_iteration = irs%n_iteration
      p%tolerance = irs%tolerance
      p%ncmp = NCMP
   end subroutine rhy_irs_params

   ! rhyme_irs_exec.f90:3-45.  On a device error p_star is 0, l_f and r_f
   ! keep their values, and the error is reported (rhy_host_error).
   function rhyme_irs_exec(irs, l_rho, l_v, l_p, l_cs, l_f, r_rho, r_v, r_p, r_cs, r_f, &
                           axis) result(p_star)
      type(irs_t), intent(in) :: irs
      real(kind=8), intent(in) :: l_rho, l_v(NDIM), l_p, l_cs, r_rho, r_v(NDIM), r_p, r_cs
      real(kind=8), intent(inout) :: l_f, r_f
      integer, intent(in) :: axis
      real(kind=8) :: p_star
      type(rhy_params_t) :: p
      real(c_double) :: l(4), r(4), ps(1), lf(1), rf(1)
      integer(c_int) :: rc
      p_star = 0d0
      if (axis < 1 .or. axis > NDIM) then
         call rhy_host_error('rhyme_irs_exec: axis out of range')
         return
      end if
      call rhy_irs_params(irs, p)
      l = [l_rho, l_v(axis), l_p, l_cs]
      r = [r_rho, r_v(axis), r_p, r_cs]
      rc = rhy_irs_exec_batch(p, 1_c_int64_t, l, r, ps, lf, rf)
      if (rc /= RHY_OK) then
         call rhy_host_error('rhyme_irs_exec: '//rhy_error_text())
         return
      end if
      p_star = ps(1)
      l_f = lf(1)
      r_f = rf(1)
   end function rhyme_irs_exec

   ! rhyme_irs_guess_p_star.f90:3-72 (irs supplies nothing to the guesses in
   ! the reference; it is not an argument there either).
   function rhyme_irs_guess_p_star(l_rho, l_v, l_p, l_cs, r_rho, r_v, r_p, r_cs, axis) &
      result(p_star)
      real(kind=8), intent(in) :: l_rho, l_v(NDIM), l_p, l_cs, r_rho, r_v(NDIM), r_p, r_cs
      integer, intent(in) :: axis
      real(kind=8) :: p_star(6)
      type(rhy_params_t) :: p
      real(c_double) :: l(4), r(4), g(6)
      integer(c_int) :: rc
      p_star = 0d0
      if (axis < 1 .or. axis > NDIM) then
         call rhy_host_error('rhyme_irs_guess_p_star: axis out of range')
         return
      end if
      call rhy_irs_params(irs_t(), p)
      l = [l_rho, l_v(axis), l_p, l_cs]
      r = [r_rho, r_v(axis), r_p, r_cs]
      rc = rhy_irs_guess_batch(p, 1_c_int64_t, l, r, g)
      if (rc /= RHY_OK) then
         call rhy_host_error('rhyme_irs_guess_p_star: '//rhy_error_text())
         return
      end if
      p_star = g
   end function rhyme_irs_guess_p_star
end module rhyme_irs

module rhyme_riemann_problem
   ! rhyme_riemann_problem/src/rhyme_riemann_problem.f90:18-133: the types of
   ! the Riemann problem and its solution, and the module procedures, each
   ! one call into the device library.  The sampling, star-state and region
   ! functions the reference declares pure are not pure here (they call the
   ! device); their reference callers (rhyme_riemann_problem_solve and the
   ! tests) do not rely on purity.
   use, intrinsic :: iso_c_binding
   use rhy_capi
   use rhyme_units
   use rhyme_thermo_base
   use rhyme_irs
   use rhyme_logger
   implicit none

   type, private :: riemann_problem_indices_t
      integer :: exact_rs = 1, deep_rs = 2
   end type riemann_problem_indices_t

   type(riemann_problem_indices_t), parameter :: rpid = riemann_problem_indices_t()

   type riemann_problem_t
      real(kind=8), dimension(cid%rho:cid%p) :: w_vacuum = 0d0
      real(kind=8) :: tolerance = 1d-8
      integer :: n_iteration = 1000, solver = rpid%exact_rs
      character(len=1024) :: path = ""
      type(irs_t) :: irs = irs_t()
   end type riemann_problem_t

   type rp_shock_t
      real(kind=8) :: rho = 0.d0, speed = 0.d0
   end type rp_shock_t

   type rp_fan_t
      real(kind=8) :: rho = 0.d0, cs = 0.d0, speedH = 0.d0, speedT = 0.d0
   end type rp_fan_t

   type rp_star_side_t
      logical :: is_shock = .false.
      type(rp_shock_t) :: shock
      type(rp_fan_t) :: fan
      real(kind=8) :: f, fprime
   end type rp_star_side_t

   type rp_star_t
      real(kind=8) :: u = 0.d0, p = 0.d0
      type(rp_star_side_t) :: left, right
   end type rp_star_t

   type rp_side_t
      real(kind=8) :: rho = 0.d0, v(NDIM) = 0.d0, p = 0.d0
      real(kind=8) :: cs = 0.d0
   end type rp_side_t

   type riemann_problem_solution_t
      type(rp_star_t) :: star
      type(rp_side_t) :: left, right
   end type riemann_problem_solution_t

contains

   subroutine rhyme_riemann_problem_init(rp, units, thermo, logger)
      type(riemann_problem_t), intent(inout) :: rp
      type(units_t), intent(in) :: units
      type(thermo_base_t), intent(in) :: thermo
      type(logger_t), intent(inout) :: logger
      integer(c_int) :: rc
      if (rp%solver /= rpid%exact_rs) call logger%err('Unknown solver: only ExactRS is supported')
      rc = rhy_init(0_c_int32_t)
      if (rc /= RHY_OK) call logger%err('rhy_init: '//rhy_error_text())
   end subroutine rhyme_riemann_problem_init

   subroutine rhy_rp_params(rp, p)
      type(riemann_problem_t), intent(in) :: rp
      type(rhy_params_t), intent(inout) :: p
      p%ndim = NDIM
      p%gamma = get_gamma()
      p%kb_amu = get_kb_amu()
      p%n_iteration = rp%irs%n_iteration
      p%tolerance = rp%irs%tolerance
      p%w_vacuum_rho = rp%w_vacuum(cid%rho)
      p%w_vacuum_p = rp%w_vacuum(cid%p)
      p%ncmp = NCMP
   end subroutine rhy_rp_params

   ! rhyme_riemann_problem_solve.f90:3-104: the conserved state at
   ! x/t = dx/dt (dx = 0: the Godunov state of the MUSCL-Hancock loops).  On
   ! a device error u is left as it was and the error reported.
   subroutine rhyme_riemann_problem_solve(rp, l, r, dx, dt, axis, u)
      type(riemann_problem_t), intent(in) :: rp
      real(kind=8), dimension(cid%rho:cid%e_tot), intent(in) :: l, r
      real(kind=8), intent(in) :: dx, dt
      integer, intent(in) :: axis
      real(kind=8), intent(inout) :: u(cid%rho:cid%e_tot)
      type(rhy_params_t) :: p
      integer(c_int32_t) :: ax(1)
      real(c_double) :: dx1(1), dt1(1), uo(cid%rho:cid%e_tot)
      integer(c_int) :: rc
      call rhy_rp_params(rp, p)
      ax(1) = axis
      dx1(1) = dx
      dt1(1) = dt
      rc = rhy_riemann_solve_batch(p, 1_c_int64_t, l, r, ax, dx1, dt1, uo)
      if (rc /= RHY_OK) then
         call rhy_host_error('rhyme_riemann_problem_solve: '//rhy_error_text())
         return
      end if
      u = uo
   end subroutine rhyme_riemann_problem_solve

   subroutine rhy_solution_to_c(s, c)
      type(riemann_problem_solution_t), intent(in) :: s
      type(rhy_rp_solution_t), intent(out) :: c
      c%l_rho = s%left%rho
      c%l_v(1:NDIM) = s%left%v
      c%l_p = s%left%p
      c%l_cs = s%left%cs
      c%r_rho = s%right%rho
      c%r_v(1:NDIM) = s%right%v
      c%r_p = s%right%p
      c%r_cs = s%right%cs
      c%star_p = s%star%p
      c%star_u = s%star%u
      c%l_is_shock = merge(1, 0, s%star%left%is_shock)
      c%r_is_shock = merge(1, 0, s%star%right%is_shock)
      c%l_shock_rho = s%star%left%shock%rho
      c%l_shock_speed = s%star%left%shock%speed
      c%l_fan_rho = s%star%left%fan%rho
      c%l_fan_cs = s%star%left%fan%cs
      c%l_fan_speedH = s%star%left%fan%speedH
      c%l_fan_speedT = s%star%left%fan%speedT
      c%r_shock_rho = s%star%right%shock%rho
      c%r_shock_speed = s%star%right%shock%speed
      c%r_fan_rho = s%star%right%fan%rho
      c%r_fan_cs = s%star%right%fan%cs
      c%r_fan_speedH = s%star%right%fan%speedH
      c%r_fan_speedT = s%star%right%fan%speedT
      c%l_f = s%star%left%f
      c%r_f = s%star%right%f
   end subroutine rhy_solution_to_c

   ! the star members of a device record into sol (sides untouched)
   subroutine rhy_star_from_c(c, s)
      type(rhy_rp_solution_t), intent(in) :: c
      type(riemann_problem_solution_t), intent(inout) :: s
      s%star%p = c%star_p
      s%star%u = c%star_u
      s%star%left%is_shock = c%l_is_shock /= 0
      s%star%right%is_shock = c%r_is_shock /= 0
      s%star%left%shock%rho = c%l_shock_rho
      s%star%left%shock%speed = c%l_shock_speed
      s%star%left%fan%rho = c%l_fan_rho
      s%star%left%fan%cs = c%l_fan_cs
      s%star%left%fan%speedH = c%l_fan_speedH
      s%star%left%fan%speedT = c%l_fan_speedT
      s%star%right%shock%rho = c%r_shock_rho
      s%star%right%shock%speed = c%r_shock_speed
      s%star%right%fan%rho = c%r_fan_rho
      s%star%right%fan%cs = c%r_fan_cs
      s%star%right%fan%speedH = c%r_fan_speedH
      s%star%right%fan%speedT = c%r_fan_speedT
      s%star%left%f = c%l_f
      s%star%right%f = c%r_f
   end subroutine rhy_star_from_c

   ! rhyme_riemann_problem_star.f90:3-66: sol's sides in, its star out.  On
   ! a device error sol is left as it was and the error reported.
   subroutine rhyme_riemann_problem_star(rp, sol, axis)
      type(riemann_problem_t), intent(in) :: rp
      type(riemann_problem_solution_t), intent(inout) :: sol
      integer, intent(in) :: axis
      type(rhy_params_t) :: p
      type(rhy_rp_solution_t) :: c(1)
      integer(c_int32_t) :: ax(1)
      integer(c_int) :: rc
      if (rp%solver /= rpid%exact_rs) return
      call rhy_rp_params(rp, p)
      call rhy_solution_to_c(sol, c(1))
      ax(1) = axis
      rc = rhy_riemann_star_batch(p, 1_c_int64_t, c, ax)
      if (rc /= RHY_OK) then
         call rhy_host_error('rhyme_riemann_problem_star: '//rhy_error_text())
         return
      end if
      call rhy_star_from_c(c(1), sol)
   end subroutine rhyme_riemann_problem_star

   ! gamma and kB/amu for the procedures without a riemann_problem_t
   subroutine rhy_gas_params(p)
      type(rhy_params_t), intent(inout) :: p
      p%ndim = NDIM
      p%gamma = get_gamma()
      p%kb_amu = get_kb_amu()
      p%ncmp = NCMP
   end subroutine rhy_gas_params

   ! rhyme_riemann_problem_sampling.f90:3-48.  On a device error u is 0 and
   ! the error reported.
   subroutine rhyme_riemann_problem_sampling(solution, axis, dx, dt, u)
      type(riemann_problem_solution_t), intent(in) :: solution
      integer, intent(in) :: axis
      real(kind=8), intent(in) :: dx, dt
      real(kind=8), intent(out) :: u(cid%rho:cid%e_tot)
      type(rhy_params_t) :: p
      type(rhy_rp_solution_t) :: c(1)
      integer(c_int32_t) :: ax(1)
      real(c_double) :: dx1(1), dt1(1)
      integer(c_int) :: rc
      call rhy_gas_params(p)
      call rhy_solution_to_c(solution, c(1))
      ax(1) = axis
      dx1(1) = dx
      dt1(1) = dt
      rc = rhy_riemann_sampling_batch(p, 1_c_int64_t, c, ax, dx1, dt1, u)
      if (rc /= RHY_OK) then
         u = 0d0
         call rhy_host_error('rhyme_riemann_problem_sampling: '//rhy_error_text())
      end if
   end subroutine rhyme_riemann_problem_sampling

   ! one region state of a solution (RHY_W_*), x/t = dxdt for the fans
   function rhy_region(s, region, axis, dxdt, who) result(u)
      type(riemann_problem_solution_t), intent(in) :: s
      integer(c_int32_t), intent(in) :: region
      integer, intent(in) :: axis
      real(kind=8), intent(in) :: dxdt
      character(len=*), intent(in) :: who
      real(kind=8) :: u(cid%rho:cid%e_tot)
      type(rhy_params_t) :: p
      type(rhy_rp_solution_t) :: c(1)
      integer(c_int32_t) :: ax(1), reg(1)
      real(c_double) :: d(1)
      integer(c_int) :: rc
      call rhy_gas_params(p)
      call rhy_solution_to_c(s, c(1))
      ax(1) = axis
      reg(1) = region
      d(1) = dxdt
      rc = rhy_riemann_region_batch(p, 1_c_int64_t, c, reg, ax, d, u)
      if (rc /= RHY_OK) then
         u = 0d0
         call rhy_host_error(who//': '//rhy_error_text())
      end if
   end function rhy_region

   function rhy_side_solution(side) result(s)
      type(rp_side_t), intent(in) :: side
      type(riemann_problem_solution_t) :: s
      s%left = side
   end function rhy_side_solution

   ! rhyme_riemann_problem_w_k.f90:3-11
   function riemann_problem_w_k(s) result(u)
      type(rp_side_t), intent(in) :: s
      real(kind=8) :: u(cid%rho:cid%e_tot)
      u = rhy_region(rhy_side_solution(s), RHY_W_L, 1, 0d0, 'riemann_problem_w_k')
   end function riemann_problem_w_k

   ! rhyme_riemann_problem_rp_side_to_cons.f90:3-11
   function riemann_problem_rp_side_to_cons(s) result(u)
      type(rp_side_t), intent(in) :: s
      real(kind=8) :: u(cid%rho:cid%e_tot)
      u = rhy_region(rhy_side_solution(s), RHY_W_L, 1, 0d0, 'riemann_problem_rp_side_to_cons')
   end function riemann_problem_rp_side_to_cons

   ! rhyme_riemann_problem_w_kfan.f90:3-32
   function riemann_problem_w_kfan(s, dxdt, axis, is_right) result(u)
      type(rp_side_t), intent(in) :: s
      real(kind=8), intent(in) :: dxdt
      integer, intent(in) :: axis
      logical, intent(in) :: is_right
      real(kind=8) :: u(cid%rho:cid%e_tot)
      type(riemann_problem_solution_t) :: sol
      if (is_right) then
         sol%right = s
         u = rhy_region(sol, RHY_W_RFAN, axis, dxdt, 'riemann_problem_w_kfan')
      else
         sol%left = s
         u = rhy_region(sol, RHY_W_LFAN, axis, dxdt, 'riemann_problem_w_kfan')
      end if
   end function riemann_problem_w_kfan

   ! rhyme_riemann_problem_w_starL_sho.f90:3-15 and siblings
   function riemann_problem_w_starl_sho(s, axis) result(u)
      type(riemann_problem_solution_t), intent(in) :: s
      integer, intent(in) :: axis
      real(kind=8) :: u(cid%rho:cid%e_tot)
      u = rhy_region(s, RHY_W_STARL_SHO, axis, 0d0, 'riemann_problem_w_starl_sho')
   end function riemann_problem_w_starl_sho

   function riemann_problem_w_starr_sho(s, axis) result(u)
      type(riemann_problem_solution_t), intent(in) :: s
      integer, intent(in) :: axis
      real(kind=8) :: u(cid%rho:cid%e_tot)
      u = rhy_region(s, RHY_W_STARR_SHO, axis, 0d0, 'riemann_problem_w_starr_sho')
   end function riemann_problem_w_starr_sho

   function riemann_problem_w_starl_fan(s, axis) result(u)
      type(riemann_problem_solution_t), intent(in) :: s
      integer, intent(in) :: axis
      real(kind=8) :: u(cid%rho:cid%e_tot)
      u = rhy_region(s, RHY_W_STARL_FAN, axis, 0d0, 'riemann_problem_w_starl_fan')
   end function riemann_problem_w_starl_fan

   function riemann_problem_w_starr_fan(s, axis) result(u)
      type(riemann_problem_solution_t), intent(in) :: s
      integer, intent(in) :: axis
      real(kind=8) :: u(cid%rho:cid%e_tot)
      u = rhy_region(s, RHY_W_STARR_FAN, axis, 0d0, 'riemann_problem_w_starr_fan')
   end function riemann_problem_w_starr_fan
end module rhyme_riemann_problem

module rhyme_cfl
   ! rhyme_cfl/src/rhyme_cfl.f90:7-24.  rhyme_cfl_time_step is not pure
   ! here: it reads the device copy of level-0 box 1.
   use, intrinsic :: iso_c_binding
   use rhy_capi
   use rhy_device
   use rhyme_samr
   use rhyme_thermo_base
   use rhyme_logger
   implicit none

   type cfl_t
      real(kind=8) :: courant_number = .8d0
   end type cfl_t

contains

   subroutine rhyme_cfl_init(cfl, thermo, samr, logger)
      type(cfl_t), intent(inout) :: cfl
      type(thermo_base_t), intent(in) :: thermo
      type(samr_t), intent(in) :: samr
      type(logger_t), intent(inout) :: logger
      call logger%log('cfl initialised')
   end subroutine rhyme_cfl_init

   ! rhyme_cfl_time_step.f90:3-60.  On an error dt is -1 and the error is
   ! reported (rhy_host_error): a caller's step with it would be rejected.
   function rhyme_cfl_time_step(c, samr) result(dt)
      real(kind=8), intent(in) :: c
      type(samr_t), intent(inout) :: samr
      real(kind=8) :: dt
      type(rhy_params_t) :: p
      type(c_ptr) :: h
      character(len=:), allocatable :: msg
      real(c_double) :: dx3(3), dtc
      integer(c_int) :: rc
      dt = -1d0
      h = rhy_box_handle(samr%levels(0)%boxes(1), msg)
      if (.not. c_associated(h)) then
         call rhy_host_error('rhyme_cfl_time_step: '//msg)
         return
      end if
      p%ndim = NDIM
      p%gamma = get_gamma()
      p%kb_amu = get_kb_amu()
      p%courant = c
      p%ncmp = NCMP
      dx3 = 1d0
      dx3(1:NDIM) = samr%levels(0)%dx
      rc = rhy_cfl_time_step(h, p, dx3, dtc)
      if (rc /= RHY_OK) then
         call rhy_host_error('rhyme_cfl_time_step: '//rhy_error_text())
         return
      end if
      dt = dtc
   end function rhyme_cfl_time_step
end module rhyme_cfl

module rhyme_samr_bc
   ! rhyme_samr_bc/src/rhyme_samr_bc.f90:9-72; not pure here (device copy).
   use, intrinsic :: iso_c_binding
   use rhy_capi
   use rhy_device
   use rhyme_units
   use rhyme_samr
   use rhyme_thermo_base
   use rhyme_logger
   implicit none

   type, private :: samr_bc_indices_t
      integer :: unset = -1
      integer :: reflective = 1, outflow = 2, periodic = 3, inflow = 4
      integer :: left = 1, right = 2
#if NDIM > 1
      integer :: bottom = 3, top = 4
#endif
#if NDIM > 2
      integer :: back = 5, front = 6
#endif
   end type samr_bc_indices_t

   type(samr_bc_indices_t), parameter :: bcid = samr_bc_indices_t()

   type samr_bc_t
      integer :: types(2*NDIM) = bcid%reflective
      real(kind=8) :: prim_inflows(NCMP, 2*NDIM) = 0d0
      real(kind=8) :: cons_inflows(NCMP, 2*NDIM) = 0d0
   end type samr_bc_t

contains

   ! rhyme_samr_bc_init.f90:53-87: inflow primitive -> conserved (non-hydro
   ! components copied) and every ghost cell of vars 1..NDIM+2 zeroed (on
   ! the host mirror, before the device copy exists or is re-uploaded).
   subroutine rhyme_samr_bc_init(bc, thermo, samr, logger)
      type(samr_bc_t), intent(inout) :: bc
      type(thermo_base_t), intent(in) :: thermo
      type(samr_t), intent(inout) :: samr
      type(logger_t), intent(inout) :: logger
      integer :: i, b, c
      real(kind=8) :: g, s
      g = get_gamma()
      do i = 1, 2*NDIM
         if (bc%types(i) == bcid%inflow) then
            s = 0d0
            do c = 0, NDIM - 1
               s = s + bc%prim_inflows(cid%u + c, i)**2
            end do
            bc%cons_inflows(cid%rho, i) = bc%prim_inflows(cid%rho, i)
            bc%cons_inflows(cid%rho_u:cid%rho_u + NDIM - 1, i) = &
               bc%prim_inflows(cid%rho, i)*bc%prim_inflows(cid%u:cid%u + NDIM - 1, i)
            bc%cons_inflows(cid%e_tot, i) = .5d0*bc%prim_inflows(cid%rho, i)*s &
                                            + bc%prim_inflows(cid%p, i)/(g - 1.d0)
            bc%cons_inflows(cid%e_tot + 1:NCMP, i) = bc%prim_inflows(cid%e_tot + 1:NCMP, i)
         end if
      end do
      do b = 1, samr%levels(0)%nboxes
         call zero_ghosts(samr%levels(0)%boxes(b))
      end do
      call logger%log('samr_bc initialised')
   end subroutine rhyme_samr_bc_init

   subroutine zero_ghosts(box)
      type(samr_box_t), intent(inout) :: box
      integer :: c
      integer :: lo(NDIM), hi(NDIM)
      lo = 1
      hi = box%dims
      do c = cid%rho, cid%e_tot
#if NDIM == 1
         box%cells(:lo(1) - 1, c) = 0d0
         box%cells(hi(1) + 1:, c) = 0d0
#elif NDIM == 2
         box%cells(:lo(1) - 1, :, c) = 0d0
         box%cells(hi(1) + 1:, :, c) = 0d0
         box%cells(:, :lo(2) - 1, c) = 0d0
         box%cells(:, hi(2) + 1:, c) = 0d0
#else
         box%cells(:lo(1) - 1, :, :, c) = 0d0
         box%cells(hi(1) + 1:, :, :, c) = 0d0
         box%cells(:, :lo(2) - 1, :, c) = 0d0
         box%cells(:, hi(2) + 1:, :, c) = 0d0
         box%cells(:, :, :lo(3) - 1, c) = 0d0
         box%cells(:, :, hi(3) + 1:, c) = 0d0
#endif
      end do
   end subroutine zero_ghosts

   subroutine rhy_bc_params(bc, p)
      type(samr_bc_t), intent(in) :: bc
      type(rhy_params_t), intent(inout) :: p
      integer :: f
      p%ndim = NDIM
      p%gamma = get_gamma()
      p%kb_amu = get_kb_amu()
      p%ncmp = NCMP
      p%bc_types = -1
      p%bc_types(1:2*NDIM) = bc%types
      do f = 1, 2*NDIM
         p%cons_inflows(1:NCMP, f) = bc%cons_inflows(:, f)
      end do
   end subroutine rhy_bc_params

   subroutine rhyme_samr_bc_set_boundaries(bc, samr)
      type(samr_bc_t), intent(in) :: bc
      type(samr_t), intent(inout) :: samr
      call set_faces(bc, samr%levels(0)%boxes(1), -1)
   end subroutine rhyme_samr_bc_set_boundaries

   subroutine set_faces(bc, box, face)
      type(samr_bc_t), intent(in) :: bc
      type(samr_box_t), intent(inout) :: box
      integer, intent(in) :: face
      type(rhy_params_t) :: p
      type(c_ptr) :: h
      character(len=:), allocatable :: msg
      integer(c_int) :: rc
      h = rhy_box_handle(box, msg)
      if (.not. c_associated(h)) then
         call rhy_host_error('rhyme_samr_bc_set_boundaries: '//msg)
         return
      end if
      call rhy_bc_params(bc, p)
      if (face < 0) then
         rc = rhy_bc_set_boundaries(h, p)
      else
         rc = rhy_bc_set_face(h, p, int(face, c_int32_t))
      end if
      if (rc /= RHY_OK) call rhy_host_error('rhyme_samr_bc_set_boundaries: '//rhy_error_text())
   end subroutine set_faces

   subroutine rhyme_samr_bc_set_left_boundary(bc, box)
      type(samr_bc_t), intent(in) :: bc
      type(samr_box_t), intent(inout) :: box
      call set_faces(bc, box, 0)
   end subroutine rhyme_samr_bc_set_left_boundary

   subroutine rhyme_samr_bc_set_right_boundary(bc, box)
      type(samr_bc_t), intent(in) :: bc
      type(samr_box_t), intent(inout) :: box
      call set_faces(bc, box, 1)
   end subroutine rhyme_samr_bc_set_right_boundary

#if NDIM > 1
   subroutine rhyme_samr_bc_set_bottom_boundary(bc, box)
      type(samr_bc_t), intent(in) :: bc
      type(samr_box_t), intent(inout) :: box
      call set_faces(bc, box, 2)
   end subroutine rhyme_samr_bc_set_bottom_boundary

   subroutine rhyme_samr_bc_set_top_boundary(bc, box)
      type(samr_bc_t), intent(in) :: bc
      type(samr_box_t), intent(inout) :: box
      call set_faces(bc, box, 3)
   end subroutine rhyme_samr_bc_set_top_boundary
#endif

#if NDIM > 2
   subroutine rhyme_samr_bc_set_back_boundary(bc, box)
      type(samr_bc_t), intent(in) :: bc
      type(samr_box_t), intent(inout) :: box
      call set_faces(bc, box, 4)
   end subroutine rhyme_samr_bc_set_back_boundary

   subroutine rhyme_samr_bc_set_front_boundary(bc, box)
      type(samr_bc_t), intent(in) :: bc
      type(samr_box_t), intent(inout) :: box
      call set_faces(bc, box, 5)
   end subroutine rhyme_samr_bc_set_front_boundary
#endif
end module rhyme_samr_bc

module rhyme_mh_workspace
   ! rhyme_mh_workspace/src/rhyme_mh_workspace.f90:18-43.  The scratch of
   ! the reference (ul, ur, fr or u) lives inside the device box; this type
   ! keeps the reference's shape and its `type` selector.
   use rhyme_samr
   use rhyme_logger
   implicit none

   type mh_workspace_indices_t
      integer :: cpu_intensive = 1, memory_intensive = 2
   end type mh_workspace_indices_t

   type(mh_workspace_indices_t), parameter :: mhwsid = mh_workspace_indices_t()

   type mh_workspace_t
      integer :: nlevels = 0
      integer :: type = mhwsid%memory_intensive
      logical :: initialized = .false.
   end type mh_workspace_t

contains

   subroutine rhyme_mh_workspace_init(mhws, samr, logger)
      type(mh_workspace_t), intent(inout) :: mhws
      type(samr_t), intent(in) :: samr
      type(logger_t), intent(inout) :: logger
      if (mhws%initialized) call logger%warn('Trying to re-initialize mh_workspace object')
      mhws%nlevels = samr%nlevels
      mhws%initialized = .true.
   end subroutine rhyme_mh_workspace_init

   subroutine rhyme_mh_workspace_check(mhws, box)
      type(mh_workspace_t), intent(inout) :: mhws
      type(samr_box_t), intent(in) :: box
      ! device scratch is sized with the box (rhy_box_create / rhy_box_resize)
   end subroutine rhyme_mh_workspace_check
end module rhyme_mh_workspace

module rhyme_muscl_hancock
   ! rhyme_muscl_hancock/src/rhyme_muscl_hancock.f90:9-64
   use, intrinsic :: iso_c_binding
   use rhy_capi
   use rhy_device
   use rhyme_units
   use rhyme_samr
   use rhyme_thermo_base
   use rhyme_slope_limiter
   use rhyme_riemann_problem
   use rhyme_mh_workspace
   use rhyme_logger
   implicit none

   type, private :: rhyme_muscl_hancock_indices_t
      integer :: cpu_intensive = mhwsid%cpu_intensive
      integer :: memory_intensive = mhwsid%memory_intensive
   end type rhyme_muscl_hancock_indices_t

   type(rhyme_muscl_hancock_indices_t), parameter :: mhid = rhyme_muscl_hancock_indices_t()

   type muscl_hancock_t
      integer :: solver_type = mhid%memory_intensive
   end type muscl_hancock_t

contains

   subroutine rhyme_muscl_hancock_init(mh, samr, ws, logger)
      type(muscl_hancock_t), intent(inout) :: mh
      type(samr_t), intent(in) :: samr
      type(mh_workspace_t), intent(inout) :: ws
      type(logger_t), intent(inout) :: logger
      call rhyme_mh_workspace_init(ws, samr, logger)
   end subroutine rhyme_muscl_hancock_init

   ! One device step of the box; ok is false and msg set on any failure (no
   ! device copy, or the library rejected the call: the device state and
   ! box%cells are then as they were).
   subroutine solve_with(solver_type, box, dx, dt, rp, sl, ws, ok, msg)
      integer, intent(in) :: solver_type
      type(samr_box_t), intent(inout) :: box
      real(kind=8), intent(in) :: dx(NDIM), dt
      type(riemann_problem_t), intent(in) :: rp
      type(slope_limiter_t), intent(in) :: sl
      type(mh_workspace_t), intent(inout) :: ws
      logical, intent(out) :: ok
      character(len=:), allocatable, intent(out) :: msg
      type(rhy_params_t) :: p
      type(c_ptr) :: h
      real(c_double) :: dx3(3)
      integer(c_int) :: rc

      ok = .false.
      h = rhy_box_handle(box, msg)
      if (.not. c_associated(h)) return
      call rhy_rp_params(rp, p)
      p%limiter = sl%type
      p%omega = sl%w
      p%solver_type = solver_type
      dx3 = 1d0
      dx3(1:NDIM) = dx
      rc = rhy_mh_solve(h, p, dx3, dt)
      if (rc /= RHY_OK) then
         msg = rhy_error_text()
         return
      end if
      ok = .true.
   end subroutine solve_with

   subroutine rhyme_muscl_hancock_solve_memory_intensive(box, dx, dt, rp, sl, ws)
      type(samr_box_t), intent(inout) :: box
      real(kind=8), intent(in) :: dx(NDIM), dt
      type(riemann_problem_t), intent(inout) :: rp
      type(slope_limiter_t), intent(in) :: sl
      type(mh_workspace_t), intent(inout) :: ws
      logical :: ok
      character(len=:), allocatable :: msg
      call solve_with(mhid%memory_intensive, box, dx, dt, rp, sl, ws, ok, msg)
      if (.not. ok) call rhy_host_error('rhyme_muscl_hancock_solve_memory_intensive: '//msg)
   end subroutine rhyme_muscl_hancock_solve_memory_intensive

   subroutine rhyme_muscl_hancock_solve_cpu_intensive(box, dx, dt, rp, sl, ws)
      type(samr_box_t), intent(inout) :: box
      real(kind=8), intent(in) :: dx(NDIM), dt
      type(riemann_problem_t), intent(inout) :: rp
      type(slope_limiter_t), intent(in) :: sl
      type(mh_workspace_t), intent(inout) :: ws
      logical :: ok
      character(len=:), allocatable :: msg
      call solve_with(mhid%cpu_intensive, box, dx, dt, rp, sl, ws, ok, msg)
      if (.not. ok) call rhy_host_error('rhyme_muscl_hancock_solve_cpu_intensive: '//msg)
   end subroutine rhyme_muscl_hancock_solve_cpu_intensive

   ! rhyme_muscl_hancock_solve.f90:17-26
   subroutine rhyme_muscl_hancock_solve(mh, box, dx, dt, rp, sl, ws, logger)
      type(muscl_hancock_t), intent(inout) :: mh
      type(samr_box_t), intent(inout) :: box
      real(kind=8), intent(in) :: dx(NDIM), dt
      type(riemann_problem_t), intent(inout) :: rp
      type(slope_limiter_t), intent(in) :: sl
      type(mh_workspace_t), intent(inout) :: ws
      type(logger_t), intent(inout) :: logger
      logical :: ok
      character(len=:), allocatable :: msg

      call logger%begin_section('solver')
      select case (mh%solver_type)
      case (mhid%cpu_intensive, mhid%memory_intensive)
         call solve_with(mh%solver_type, box, dx, dt, rp, sl, ws, ok, msg)
         if (.not. ok) call logger%err('rhyme_muscl_hancock_solve: '//msg)
      case DEFAULT
         call logger%err('Unknown solver type')
         return
      end select
      call logger%end_section
   end subroutine rhyme_muscl_hancock_solve

   ! rhyme_muscl_hancock_half_step_extrapolation.f90:3-24 (one cell on the device)
   subroutine rhyme_muscl_hancock_half_step_extrapolation(u, delta, axis, dx, dt, l, r)
      real(kind=8), dimension(cid%rho:cid%e_tot), intent(in) :: u, delta
      integer, intent(in) :: axis
      real(kind=8), intent(in) :: dx, dt
      real(kind=8), dimension(cid%rho:cid%e_tot), intent(out) :: l, r
      type(rhy_params_t) :: p
      integer(c_int32_t) :: ax(1)
      integer(c_int) :: rc
      p%ndim = NDIM
      p%gamma = get_gamma()
      p%kb_amu = get_kb_amu()
      p%ncmp = NCMP
      ax(1) = axis
      rc = rhy_half_step_batch(p, 1_c_int64_t, u, delta, ax, dx, dt, l, r)
      if (rc /= RHY_OK) then
         l = u
         r = u
         call rhy_host_error('rhyme_muscl_hancock_half_step_extrapolation: '//rhy_error_text())
      end if
   end subroutine rhyme_muscl_hancock_half_step_extrapolation
end module rhyme_muscl_hancock
