!*************************************************************************
! opf_main.f90 -- the reference's main loop (main.f90:18-46) over the
! procedures of opf_path.
!
! usage: opf_main N K M Re alpha dt nsteps state_in state_out [series]
!   dt < 0 selects the adaptive time step (d_timestep < 0, main.f90:24-29)
!   with |dt| as the initial tim_dt.
!   state_in / state_out: raw stream files holding vel_ur, vel_ut, vel_uz as
!   Re(1:N,0:H-1) then Im(1:N,0:H-1) per component, or state files of
!   io_save_state (netCDF classic) when the name ends in '.cdf.dat'; a state
!   file read takes tim_t from the file (d_time < 0).  series: optional text
!   file with one line per step: step, correctors used, last dterr, tim_dt.
!*************************************************************************
program opf_main
   use opf_path
   implicit none
   integer :: N, K, M, nsteps, u, ncorr, its, stepdev
   double precision :: Re, alpha, dt, dterr_dev
   character(len=512) :: arg, fin, fout, fser
   logical :: want_series

   if (command_argument_count() < 9) stop 'usage: opf_main N K M Re alpha dt nsteps state_in state_out [series]'
   call get_command_argument(1, arg); read(arg, *) N
   call get_command_argument(2, arg); read(arg, *) K
   call get_command_argument(3, arg); read(arg, *) M
   call get_command_argument(4, arg); read(arg, *) Re
   call get_command_argument(5, arg); read(arg, *) alpha
   call get_command_argument(6, arg); read(arg, *) dt
   call get_command_argument(7, arg); read(arg, *) nsteps
   call get_command_argument(8, fin)
   call get_command_argument(9, fout)
   want_series = command_argument_count() >= 10
   if (want_series) call get_command_argument(10, fser)

   d_timestep = dt
   call vel_precompute(N, K, M, 1, Re, alpha, dabs(dt))
   call vel_matrices()
   if (is_state_file(fin)) then
      call io_load_state(fin, -1d0)
   else
      open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
      read(u) vel_ur%Re, vel_ur%Im
      read(u) vel_ut%Re, vel_ut%Im
      read(u) vel_uz%Re, vel_uz%Im
      close(u)
      call vel_put_state()
   end if
   if (want_series) open(newunit=u, file=trim(fser), status='replace')

   i_maxtstep = nsteps
   do while (tim_step < i_maxtstep)
      call vel_imposesym()
      call vel_transform()
      call vel_nonlinear()
      if (d_timestep < 0d0) then
         call vel_maxtstep()
         call tim_new_tstep()
         if (tim_new_dt) call vel_matrices()
      end if
      call vel_predictor()
      tim_it = 1
      its = 0
      do while (tim_it /= 0)
         call vel_transform()
         call vel_nonlinear()
         call vel_corrector()
         its = its + 1
         if (want_series) dterr_dev = tim_dterr
         call tim_check_cgce()
      end do
      if (want_series) write(u, '(I8,I4,2ES26.17E3)') tim_step + 1, its, dterr_dev, tim_dt
      tim_t = tim_t + tim_dt
      tim_step = tim_step + 1
   end do
   if (want_series) close(u)

   call vel_status(stepdev, ncorr, dterr_dev)
   if (is_state_file(fout)) then
      call io_save_state(fout)
   else
      call vel_get_state()
      open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
      write(u) vel_ur%Re, vel_ur%Im
      write(u) vel_ut%Re, vel_ut%Im
      write(u) vel_uz%Re, vel_uz%Im
      close(u)
   end if
   call vel_cleanup()

contains

   logical function is_state_file(name)
      character(len=*), intent(in) :: name
      integer :: n
      n = len_trim(name)
      is_state_file = n >= 8
      if (is_state_file) is_state_file = name(n-7:n) == '.cdf.dat'
   end function is_state_file
end program opf_main
