! rhy_fhost_units.f90 -- calls the Fortran host's unit procedures the way
! the reference's own tests call them, and forces the errors the host must
! report.
!
! usage: units <mode> <in> <out>
!   irs       in: "n n_iteration tolerance", then n lines
!             "l_rho l_u l_p l_cs r_rho r_u r_p r_cs" (axis 1);
!             out: n x (p_star, l_f, r_f) from rhyme_irs_exec
!   star      in: "n n_iteration tolerance", then n lines of sides
!             "l_rho l_v(1:NDIM) l_p l_cs r_rho r_v(1:NDIM) r_p r_cs axis";
!             out: n x 18 star values of rhyme_riemann_problem_star
!   sampling  in: "m n_iteration tolerance", a sides line (as for star), the
!             conserved l(1:NDIM+2) and r(1:NDIM+2), then m lines "dx dt";
!             out: m x (sampling u, solve u), NDIM+2 values each
!   errors    in: "kb amu" (unused file contents beyond that);
!             out: 10 values, the error counts and checks listed below
! Gas: diatomic unless the in file's mode says otherwise (state_of_matter
! read from the environment variable RHY_SOM, default 2).
program rhy_fhost_units
   use rhyme_units
   use rhyme_logger
   use rhyme_thermo_base
   use rhyme_samr
   use rhy_device
   use rhyme_slope_limiter
   use rhyme_irs
   use rhyme_riemann_problem
   use rhyme_cfl
   use rhyme_samr_bc
   use rhyme_mh_workspace
   use rhyme_muscl_hancock
   implicit none

   character(len=1024) :: mode, in_file, out_file, env
   type(logger_t) :: logger
   type(units_t) :: units
   type(thermo_base_t) :: thermo
   type(riemann_problem_t) :: rp
   integer :: u_in, u_out, n, i, nit, som, stat
   real(kind=8) :: tol

   call get_command_argument(1, mode)
   call get_command_argument(2, in_file)
   call get_command_argument(3, out_file)
   call get_environment_variable('RHY_SOM', env, status=stat)
   som = 2
   if (stat == 0) read (env, *) som

   units%kb%v = 1d0
   units%amu%v = 1d0
   thermo%state_of_matter = som
   call rhyme_thermo_base_init(thermo, units, logger)

   open (newunit=u_in, file=trim(in_file), status='old', action='read')
   open (newunit=u_out, file=trim(out_file), access='stream', form='unformatted', &
         status='replace')

   select case (trim(mode))
   case ('irs')
      call do_irs()
   case ('star')
      call do_star()
   case ('sampling')
      call do_sampling()
   case ('errors')
      call do_errors()
   case default
      stop 2
   end select
   close (u_in)
   close (u_out)
   call rhy_device_release_all()
   if (logger%nerr > 0 .and. trim(mode) /= 'errors') stop 3

contains

   subroutine read_head()
      read (u_in, *) n, nit, tol
      rp%irs%n_iteration = nit
      rp%irs%tolerance = tol
      call rhyme_riemann_problem_init(rp, units, thermo, logger)
   end subroutine read_head

   subroutine read_sides(sol, axis)
      type(riemann_problem_solution_t), intent(out) :: sol
      integer, intent(out) :: axis
      read (u_in, *) sol%left%rho, sol%left%v, sol%left%p, sol%left%cs, &
         sol%right%rho, sol%right%v, sol%right%p, sol%right%cs, axis
   end subroutine read_sides

   subroutine do_irs()
      real(kind=8) :: lr, lu, lp, lc, rr, ru, rpr, rc, lv(NDIM), rv(NDIM), ps, lf, rf
      call read_head()
      call rhyme_irs_init(rp%irs, thermo, logger)
      do i = 1, n
         read (u_in, *) lr, lu, lp, lc, rr, ru, rpr, rc
         lv = 0d0
         rv = 0d0
         lv(1) = lu
         rv(1) = ru
         lf = 0d0
         rf = 0d0
         ps = rhyme_irs_exec(rp%irs, lr, lv, lp, lc, lf, rr, rv, rpr, rc, rf, 1)
         write (u_out) ps, lf, rf
      end do
   end subroutine do_irs

   subroutine write_star(s)
      type(riemann_problem_solution_t), intent(in) :: s
      write (u_out) s%star%p, s%star%u, merge(1d0, 0d0, s%star%left%is_shock), &
         s%star%left%shock%rho, s%star%left%shock%speed, s%star%left%fan%rho, &
         s%star%left%fan%cs, s%star%left%fan%speedH, s%star%left%fan%speedT, &
         merge(1d0, 0d0, s%star%right%is_shock), s%star%right%shock%rho, &
         s%star%right%shock%speed, s%star%right%fan%rho, s%star%right%fan%cs, &
         s%star%right%fan%speedH, s%star%right%fan%speedT, s%star%left%f, s%star%right%f
   end subroutine write_star

   subroutine do_star()
      type(riemann_problem_solution_t) :: sol
      integer :: axis
      call read_head()
      do i = 1, n
         call read_sides(sol, axis)
         call rhyme_riemann_problem_star(rp, sol, axis)
         call write_star(sol)
      end do
   end subroutine do_star

   subroutine do_sampling()
      type(riemann_problem_solution_t) :: sol
      real(kind=8) :: l(cid%rho:cid%e_tot), r(cid%rho:cid%e_tot), us(cid%rho:cid%e_tot)
      real(kind=8) :: uv(cid%rho:cid%e_tot), dx, dt
      integer :: axis
      call read_head()
      call read_sides(sol, axis)
      read (u_in, *) l
      read (u_in, *) r
      call rhyme_riemann_problem_star(rp, sol, axis)
      do i = 1, n
         read (u_in, *) dx, dt
         call rhyme_riemann_problem_sampling(sol, axis, dx, dt, us)
         uv = 0d0
         call rhyme_riemann_problem_solve(rp, l, r, dx, dt, axis, uv)
         write (u_out) us, uv
      end do
   end subroutine do_sampling

   ! Forced errors.  out: (1) logger%nerr after rhyme_muscl_hancock_solve with
   ! an unknown limiter, (2) 1 if that box's cells came back unchanged,
   ! (3) rhy_host_nerr after the same through the memory_intensive variant,
   ! (4) 1 if the cells are still unchanged, (5) dt of rhyme_cfl_time_step on
   ! a box without a valid level, (6) rhy_host_nerr after it and after
   ! rhyme_samr_bc_set_boundaries on that box, (7) 1 if
   ! rhyme_riemann_problem_solve with an axis past NDIM left u unchanged,
   ! (8) rhy_host_nerr after it, the limiter with an unknown type, the half
   ! step and rhyme_irs_exec with axis 0, (9) logger%nerr at the end,
   ! (10) 1 if a valid step after all of that still advanced the box.
   subroutine do_errors()
      type(samr_t) :: samr, bad
      type(slope_limiter_t) :: sl
      type(mh_workspace_t) :: ws
      type(muscl_hancock_t) :: mh
      type(samr_bc_t) :: bc
      real(kind=8), allocatable :: before(:)
      real(kind=8) :: dx(NDIM), dt, l(cid%rho:cid%e_tot), r(cid%rho:cid%e_tot)
      real(kind=8) :: u(cid%rho:cid%e_tot), d(cid%rho:cid%e_tot), ps, lf, rf, vz(NDIM)
      real(kind=8) :: hl(cid%rho:cid%e_tot), hr(cid%rho:cid%e_tot)
      integer :: dims(NDIM), k
      logical :: same, u_kept

      dims = 8
      samr%nlevels = 1
      samr%base_grid = dims
      samr%levels(0)%level = 0
      samr%levels(0)%dx = 1d0/8
      allocate (samr%levels(0)%boxes(1))
      call rhyme_samr_init_box(samr, 0, 1, dims, [(1, k=1, NDIM)], dims)
      ! a smooth state (rho 1 + x/10, at rest) in every interior cell
      call fill(samr%levels(0)%boxes(1))
      rp%irs%n_iteration = 100
      rp%irs%tolerance = 1d-6
      call rhyme_riemann_problem_init(rp, units, thermo, logger)
      call rhyme_samr_bc_init(bc, thermo, samr, logger)
      call rhyme_muscl_hancock_init(mh, samr, ws, logger)
      call rhyme_samr_upload(samr)
      before = reshape(samr%levels(0)%boxes(1)%cells, [size(samr%levels(0)%boxes(1)%cells)])
      dx = samr%levels(0)%dx

      sl%type = 99
      call rhyme_muscl_hancock_solve(mh, samr%levels(0)%boxes(1), dx, 1d-3, rp, sl, ws, logger)
      call rhyme_samr_download(samr)
      same = all(reshape(samr%levels(0)%boxes(1)%cells, [size(before)]) == before)
      write (u_out) real(logger%nerr, 8), merge(1d0, 0d0, same)

      call rhyme_muscl_hancock_solve_memory_intensive(samr%levels(0)%boxes(1), dx, 1d-3, rp, &
                                                      sl, ws)
      call rhyme_samr_download(samr)
      same = all(reshape(samr%levels(0)%boxes(1)%cells, [size(before)]) == before)
      write (u_out) real(rhy_host_nerr, 8), merge(1d0, 0d0, same)

      bad%nlevels = 1
      bad%levels(0)%dx = 1d0/8
      allocate (bad%levels(0)%boxes(1))
      bad%levels(0)%boxes(1)%dims = dims
      dt = rhyme_cfl_time_step(0.8d0, bad)
      call rhyme_samr_bc_set_boundaries(bc, bad)
      write (u_out) dt, real(rhy_host_nerr, 8)

      l = 0d0
      l(cid%rho) = 1d0
      l(cid%e_tot) = 2.5d0
      r = l
      u = -7d0
      call rhyme_riemann_problem_solve(rp, l, r, 0d0, 1d0, NDIM + 1, u)
      u_kept = all(u == -7d0)
      sl%type = 99
      call rhyme_slope_limiter_run(sl, l, l, r, d)
      d = 0d0
      call rhyme_muscl_hancock_half_step_extrapolation(l, d, 0, 1d0, 1d0, hl, hr)
      vz = 0d0
      lf = 0d0
      rf = 0d0
      ps = rhyme_irs_exec(rp%irs, 1d0, vz, 1d0, 1d0, lf, 1d0, vz, 1d0, 1d0, rf, 0)
      write (u_out) merge(1d0, 0d0, u_kept), real(rhy_host_nerr, 8)

      sl%type = 2
      call rhyme_samr_bc_set_boundaries(bc, samr)
      call rhyme_muscl_hancock_solve(mh, samr%levels(0)%boxes(1), dx, 1d-3, rp, sl, ws, logger)
      call rhyme_samr_download(samr)
      same = all(reshape(samr%levels(0)%boxes(1)%cells, [size(before)]) == before)
      write (u_out) real(logger%nerr, 8), merge(0d0, 1d0, same)
   end subroutine do_errors

   subroutine fill(box)
      type(samr_box_t), intent(inout) :: box
      integer :: i1
#if NDIM == 1
      do i1 = 1, box%dims(1)
         box%cells(i1, cid%rho) = 1d0 + 0.1d0*i1
         box%cells(i1, cid%e_tot) = 2.5d0
      end do
#elif NDIM == 2
      do i1 = 1, box%dims(1)
         box%cells(i1, 1:box%dims(2), cid%rho) = 1d0 + 0.1d0*i1
         box%cells(i1, 1:box%dims(2), cid%e_tot) = 2.5d0
      end do
#else
      do i1 = 1, box%dims(1)
         box%cells(i1, 1:box%dims(2), 1:box%dims(3), cid%rho) = 1d0 + 0.1d0*i1
         box%cells(i1, 1:box%dims(2), 1:box%dims(3), cid%e_tot) = 2.5d0
      end do
#endif
   end subroutine fill
end program rhy_fhost_units
