!*************************************************************************
! Harness driver for the compiled reference (oracle/_ref).
!
! Test infrastructure only.  It is linked against the reference's own core
! modules (parameters, mpif, meshs, variables, transform, timestep,
! velocity), compiled from the sources in place by oracle/build_ref.sh, and
! replaces only main.f90 / io.f90: the time loop below follows main.f90:18-46
! call for call, and the diagnostics follow io.f90:501-524 (energy) and
! io.f90:627-639 (friction velocity).  netCDF-Fortran is absent, so the
! state is exchanged as raw little-endian float64 in the reference's own
! coll layout: Re(1:N,0:H-1) then Im(1:N,0:H-1) for ur, ut, uz.
!
! The var_null(2) points of main.f90:23,38 call plugin_hook: nothing
! (oracle/plugin_none.f90) or the FLAG = 2 branch of a plugin's util
! (oracle/plugin_nonnewtonian.f90, oracle/plugin_les.f90), which then reads
! params_model.in from the working directory.
!
! usage: opf_ref_<cfg> state_in out_prefix Re alpha dt nsteps [dump steps...]
!   dt < 0 selects the adaptive time step of main.f90:24-29 (d_timestep < 0)
!   with |dt| as the initial tim_dt: vel_maxtstep, tim_new_tstep and, on a
!   change, vel_matrices before every predictor.
!   writes out_prefix_series.txt   one line per step:
!          step t E Ek0 Em0 Pr0(=Ub) Uc Ufr ncorr dterr_last
!   out_prefix_dt.txt              adaptive runs, one line per step after
!                                  tim_new_tstep: step tim_dt tim_cfl_dt
!                                  tim_cfl_dir tim_corr_dt tim_new_dt
!                                  (io_write_timestep, io.f90:645-657)
!   out_prefix_stage_*.bin         step-1 intermediates (N after each
!                                  nonlinear evaluation, u after predictor
!                                  and after each corrector)
!   out_prefix_state_XXXXXX.bin    state after the listed steps
!   out_prefix_timing.txt          wall seconds for the time loop
!   nsteps = -1: no time loop; out_prefix_sym_I.bin holds var_imposesym(I)
!   (variables.f90:538-618) of the input state for I = 0..3.
!   nsteps = -2: no time loop; out_prefix_mesh.bin holds mes_D of
!   meshs.f90:48-58 as mes_precompute left it: r(:,1), intrdr, dr0, dr1,
!   dr(1)%M, dr(2)%M, radLap%M (column-major float64, in that order).
!   OPF_ORBIT="ndts T sth sz" in the environment: instead of the loop, the
!   call sequence of steporbit (newton.f90:721-775): for ndts /= 1,
!   tim_dt = T/ndts and vel_matrices; ndts steps of imposesym, transform,
!   nonlinear, predictor, transform, nonlinear, corrector, imposesym,
!   tim_check_cgce with tim_it = 1; for ndts /= 1 the final state shifted by
!   var_coll_shift(-sth, -sz).  Writes out_prefix_orbit.bin.
!*************************************************************************
 program opf_ref_driver
   use velocity
   implicit none
   character(len=512) :: infile, prefix, arg
   double precision   :: re_in, alpha_in, dt_in
   integer            :: nsteps, ndump, i, ncorr, nstage
   integer, allocatable :: dumps(:)
   integer(8)         :: c0, c1, crate
   double precision   :: dterr_last, E, Ek0, Em0, Uc, Ufr, Ub
   integer, parameter :: useries = 41, utim = 42, udt = 43
   logical            :: adaptive

   if(command_argument_count() < 6) stop 'usage: state_in prefix Re alpha dt nsteps [dumps]'
   call get_command_argument(1, infile)
   call get_command_argument(2, prefix)
   call get_command_argument(3, arg);  read(arg,*) re_in
   call get_command_argument(4, arg);  read(arg,*) alpha_in
   call get_command_argument(5, arg);  read(arg,*) dt_in
   call get_command_argument(6, arg);  read(arg,*) nsteps
   ndump = command_argument_count() - 6
   allocate(dumps(max(ndump,1)))
   dumps = -1
   do i = 1, ndump
      call get_command_argument(6+i, arg);  read(arg,*) dumps(i)
   end do

   call mpi_precompute()
   call par_precompute()
   d_Re    = re_in
   d_alpha = alpha_in
   call mes_precompute()
   call var_precompute()
   call tra_precompute()
   call tim_precompute()
   call vel_precompute()
   adaptive = dt_in < 0d0
   tim_dt = dabs(dt_in)
   call read_state(trim(infile))
   if(nsteps == -2) then
      call dump_mesh()
      stop
   end if
   if(nsteps < 0) then
      call dump_symmetries()
      stop
   end if
   call get_environment_variable('OPF_ORBIT', arg, status=i)
   if(i == 0) then
      call run_orbit(trim(arg))
      stop
   end if
   call vel_matrices()

   open(useries, file=trim(prefix)//'_series.txt', status='replace')
   if(adaptive) open(udt, file=trim(prefix)//'_dt.txt', status='replace')
   nstage = 0
   call system_clock(c0, crate)

   do while(tim_step < nsteps)
      call vel_imposesym()
      call vel_transform()
      call vel_nonlinear()
      call plugin_hook()                      ! var_null(2), main.f90:23
      if(tim_step==0) call dump_stage('N', vel_Nr, vel_Nt, vel_Nz)
      if(adaptive) then                       ! main.f90:24-29
         call vel_maxtstep()
         call tim_new_tstep()
         if(tim_new_dt) call vel_matrices()
         write(udt,'(I8,2ES26.17E3,I3,ES26.17E3,I3)') tim_step, tim_dt, tim_cfl_dt,  &
            tim_cfl_dir, tim_corr_dt, merge(1, 0, tim_new_dt)
      end if
      call diagnostics(E, Ek0, Em0, Uc, Ufr)
      Ub = vel_Pr0            ! io_write_friction reads it here (io.f90:631)
      call vel_predictor()
      if(tim_step==0) call dump_stage('u', vel_ur, vel_ut, vel_uz)
      tim_it = 1
      ncorr  = 0
      do while(tim_it/=0)
         call vel_transform()
         call vel_nonlinear()
         call plugin_hook()                   ! var_null(2), main.f90:38
         if(tim_step==0) call dump_stage('N', vel_Nr, vel_Nt, vel_Nz)
         call vel_corrector()
         if(tim_step==0) call dump_stage('u', vel_ur, vel_ut, vel_uz)
         ncorr = ncorr + 1
         dterr_last = tim_dterr
         call tim_check_cgce()
      end do
      if(tim_step < 0) stop 'corrector failure'
      write(useries,'(I8,7ES26.17E3,I4,ES26.17E3)') tim_step, tim_t, E, Ek0, Em0,  &
         Ub, Uc, Ufr, ncorr, dterr_last
      tim_t    = tim_t    + tim_dt
      tim_step = tim_step + 1
      do i = 1, ndump
         if(dumps(i)==tim_step) call write_state(tim_step)
      end do
   end do

   call system_clock(c1)
   close(useries)
   if(adaptive) close(udt)
   open(utim, file=trim(prefix)//'_timing.txt', status='replace')
   write(utim,'(ES26.17E3,I8)') dble(c1-c0)/dble(crate), nsteps
   close(utim)

 contains

   subroutine diagnostics(E, Ek0, Em0, Uc, Ufr)
      ! io_write_energy (io.f90:501-524) and io_write_friction (io.f90:627-639)
      double precision, intent(out) :: E, Ek0, Em0, Uc, Ufr
      double precision :: E_, Ek(0:i_K1), Em(0:i_M1)
      call var_coll_norm(vel_ur, E,Ek,Em)
      Ek0 = Ek(0)
      Em0 = Em(0)
      call var_coll_norm(vel_ut, E_,Ek,Em)
      E   = E + E_
      Ek0 = Ek0 + Ek(0)
      Em0 = Em0 + Em(0)
      call var_coll_norm(vel_uz, E_,Ek,Em)
      E   = E + E_
      Ek0 = Ek0 + Ek(0)
      Em0 = Em0 + Em(0)
      Uc  = 1d0 + dot_product(vel_uz%Re(1:1+i_KL,0),mes_D%dr0(:,0))
      Ufr = dot_product(vel_uz%Re(i_N-i_KL:i_N,0),mes_D%dr1(:,1))
      Ufr = dsqrt(dabs((Ufr-2d0)/d_Re))
   end subroutine diagnostics

   subroutine read_state(fn)
      character(*), intent(in) :: fn
      integer :: u
      open(newunit=u, file=fn, access='stream', form='unformatted', status='old')
      call var_coll_init(vel_ur)
      call var_coll_init(vel_ut)
      call var_coll_init(vel_uz)
      read(u) vel_ur%Re(:,0:i_H1), vel_ur%Im(:,0:i_H1)
      read(u) vel_ut%Re(:,0:i_H1), vel_ut%Im(:,0:i_H1)
      read(u) vel_uz%Re(:,0:i_H1), vel_uz%Im(:,0:i_H1)
      close(u)
   end subroutine read_state

   subroutine write_coll3(fn, a, b, c)
      character(*), intent(in) :: fn
      type (coll),  intent(in) :: a, b, c
      integer :: u
      open(newunit=u, file=fn, access='stream', form='unformatted', status='replace')
      write(u) a%Re(:,0:i_H1), a%Im(:,0:i_H1)
      write(u) b%Re(:,0:i_H1), b%Im(:,0:i_H1)
      write(u) c%Re(:,0:i_H1), c%Im(:,0:i_H1)
      close(u)
   end subroutine write_coll3

   subroutine write_state(step)
      integer, intent(in) :: step
      character(6) :: cnum
      write(cnum,'(I6.6)') step
      call write_coll3(trim(prefix)//'_state_'//cnum//'.bin', vel_ur, vel_ut, vel_uz)
   end subroutine write_state

   subroutine run_orbit(spec)
      character(*), intent(in) :: spec
      integer :: ndts, ndt
      double precision :: T, sth, sz
      read(spec,*) ndts, T, sth, sz
      if(ndts /= 1) tim_dt = T / dble(ndts)
      call vel_matrices()
      do ndt = 1, ndts
         call vel_imposesym()
         call vel_transform()
         call vel_nonlinear()
         call vel_predictor()
         call vel_transform()
         call vel_nonlinear()
         call vel_corrector()
         call vel_imposesym()
         tim_it = 1
         call tim_check_cgce()
         if(ndts /= 1) then
            tim_t    = tim_t    + tim_dt
            tim_step = tim_step + 1
            if(ndt == ndts) then
               call var_coll_shift(-sth, -sz, vel_ur, vel_ur)
               call var_coll_shift(-sth, -sz, vel_ut, vel_ut)
               call var_coll_shift(-sth, -sz, vel_uz, vel_uz)
            end if
         end if
      end do
      call write_coll3(trim(prefix)//'_orbit.bin', vel_ur, vel_ut, vel_uz)
   end subroutine run_orbit

   subroutine dump_symmetries()
      type (coll) :: a, b, c
      character(1) :: ci
      integer :: isym
      do isym = 0, 3
         call var_coll_copy(vel_ur, a)
         call var_coll_copy(vel_ut, b)
         call var_coll_copy(vel_uz, c)
         call var_imposesym(isym, a, b, c)
         write(ci,'(I1)') isym
         call write_coll3(trim(prefix)//'_sym_'//ci//'.bin', a, b, c)
      end do
   end subroutine dump_symmetries

   subroutine dump_mesh()
      integer :: u
      open(newunit=u, file=trim(prefix)//'_mesh.bin', access='stream', form='unformatted',  &
         status='replace')
      write(u) mes_D%r(:,1), mes_D%intrdr, mes_D%dr0, mes_D%dr1
      write(u) mes_D%dr(1)%M, mes_D%dr(2)%M, mes_D%radLap%M
      close(u)
   end subroutine dump_mesh

   subroutine dump_stage(tag, a, b, c)
      character(*), intent(in) :: tag
      type (coll),  intent(in) :: a, b, c
      character(2) :: cnum
      nstage = nstage + 1
      write(cnum,'(I2.2)') nstage
      call write_coll3(trim(prefix)//'_stage_'//cnum//tag//'.bin', a, b, c)
   end subroutine dump_stage

 end program opf_ref_driver
