! rhy_fhost_driver.f90 -- drives the Fortran host modules the way the
! reference's main loop does (src/main.f90:106-136): dt from
! rhyme_cfl_time_step, rhyme_samr_bc_set_boundaries, then
! rhyme_muscl_hancock_solve on every box, for nsteps steps.
!
! usage: driver <namelist file> <cells in> <cells out>
!   cells in/out: the box's cells(:, ..., 1:NCMP) array as raw FP64 in
!   Fortran order, ghosts included; cells out is followed by the nsteps dts
!   and by the level's FP32 snapshot payload (rhyme_samr_level_snapshot).
program rhy_fhost_driver
   use rhyme_units
   use rhyme_logger
   use rhyme_thermo_base
   use rhyme_samr
   use rhy_device
   use rhyme_slope_limiter
   use rhyme_riemann_problem
   use rhyme_cfl
   use rhyme_samr_bc
   use rhyme_mh_workspace
   use rhyme_muscl_hancock
   implicit none

   integer :: dims(3) = 1, nsteps = 1, state_of_matter = 1, limiter = 2, solver_type = 2
   integer :: n_iteration = 100, bc_types(6) = 1, precision = 0
   real(kind=8) :: kb = 1d0, amu = 1d0, omega = 0d0, tolerance = 1d-6, courant = .8d0
   real(kind=8) :: w_vacuum_rho = 0d0, w_vacuum_p = 0d0, box_length = 1d0
   namelist /run/ dims, nsteps, state_of_matter, limiter, solver_type, n_iteration, bc_types, &
      kb, amu, omega, tolerance, courant, w_vacuum_rho, w_vacuum_p, box_length, precision

   type(logger_t) :: logger
   type(units_t) :: units
   type(thermo_base_t) :: thermo
   type(samr_t) :: samr
   type(slope_limiter_t) :: sl
   type(riemann_problem_t) :: rp
   type(cfl_t) :: cfl
   type(samr_bc_t) :: bc
   type(mh_workspace_t) :: ws
   type(muscl_hancock_t) :: mh
   character(len=1024) :: nml_file, in_file, out_file
   integer :: u, s, l, b
   real(kind=8), allocatable :: dts(:)
   real(kind=4), allocatable :: payload(:)
   logical :: ok
   character(len=:), allocatable :: msg

   call get_command_argument(1, nml_file)
   call get_command_argument(2, in_file)
   call get_command_argument(3, out_file)
   open (newunit=u, file=trim(nml_file), status='old', action='read')
   read (u, nml=run)
   close (u)

   units%kb%v = kb
   units%amu%v = amu
   thermo%state_of_matter = state_of_matter
   call rhyme_thermo_base_init(thermo, units, logger)

   samr%nlevels = 1
   samr%base_grid = dims(1:NDIM)
   samr%ghost_cells = 2
   samr%box_lengths = box_length
   samr%levels(0)%level = 0
   samr%levels(0)%max_nboxes = 1
   samr%levels(0)%dx = box_length/dims(1:NDIM)
   allocate (samr%levels(0)%boxes(1))
   call rhyme_samr_init_box(samr, 0, 1, dims(1:NDIM), [(1, l=1, NDIM)], dims(1:NDIM))

   open (newunit=u, file=trim(in_file), access='stream', form='unformatted', status='old')
   read (u) samr%levels(0)%boxes(1)%cells
   close (u)

   sl%type = limiter
   sl%w = omega
   rp%w_vacuum(cid%rho) = w_vacuum_rho
   rp%w_vacuum(cid%p) = w_vacuum_p
   rp%irs%tolerance = tolerance
   rp%irs%n_iteration = n_iteration
   cfl%courant_number = courant
   bc%types = bc_types(1:2*NDIM)
   mh%solver_type = solver_type
   ws%type = solver_type
   rhy_precision = precision

   call rhyme_riemann_problem_init(rp, units, thermo, logger)
   call rhyme_samr_bc_init(bc, thermo, samr, logger)
   call rhyme_muscl_hancock_init(mh, samr, ws, logger)
   call rhyme_cfl_init(cfl, thermo, samr, logger)
   call rhyme_samr_upload(samr)

   allocate (dts(nsteps))
   do s = 1, nsteps
      samr%levels(0)%dt = rhyme_cfl_time_step(cfl%courant_number, samr)
      dts(s) = samr%levels(0)%dt
      call rhyme_samr_bc_set_boundaries(bc, samr)
      do l = samr%nlevels - 1, 0, -1
         do b = 1, samr%levels(l)%nboxes
            call rhyme_muscl_hancock_solve(mh, samr%levels(l)%boxes(b), &
                                           samr%levels(l)%dx, samr%levels(l)%dt, rp, sl, ws, logger)
         end do
      end do
      samr%levels(0)%t = samr%levels(0)%t + samr%levels(0)%dt
      samr%levels(0)%iteration = samr%levels(0)%iteration + 1
   end do

   call rhyme_samr_download(samr)
   allocate (payload(size(samr%levels(0)%boxes(1)%cells, dim=NDIM + 1)*product(dims(1:NDIM))))
   call rhyme_samr_level_snapshot(samr%levels(0), payload, ok, msg)
   if (.not. ok) call logger%err('snapshot: '//msg)
   open (newunit=u, file=trim(out_file), access='stream', form='unformatted', status='replace')
   write (u) samr%levels(0)%boxes(1)%cells
   write (u) dts
   write (u) payload
   close (u)
   call rhy_device_release_all()
   if (logger%nerr > 0) stop 3
end program rhy_fhost_driver
