!*************************************************************************
! io.f90 -- module io of the drop-in set: the run-time files and state
! files of io.f90 with the reference's names and formats, without a netCDF
! library (libopf reads and writes the netCDF classic state files itself).
!   state.cdf.in        read by io_load_state (io.f90:98-165)
!   stateNNNN.cdf.dat   written by io_save_state every i_save_rate1 steps
!                       and at the end (io.f90:295-340)
!   vel_energy.dat      t, E, Ek(0), Em(0) every i_save_rate2 steps
!   vel_friction.dat    t, Ub (vel_Pr0 at constant flux), Uc, u_tau
!   tim_step.dat        the time-step record of io_write_timestep
! The spectra, mean profiles, point velocities and total energy/input/
! dissipation series (io_save_spectrum, io_save_meanprof,
! io_write_pointvel, io_write_totEID) are not written.
!*************************************************************************
 module io
   use iso_c_binding
   use velocity
   implicit none
   save

   character(200)        :: io_statefile
   integer               :: io_save1, io_save2
   integer,     private  :: io_KE, io_dt, io_fr

 contains

   subroutine io_precompute()
      io_statefile = 'state.cdf.in'
      io_save1 = 0
      io_save2 = 0
      io_dt    = 10
      io_KE    = 20
      io_fr    = 50
   end subroutine io_precompute

   subroutine io_openfiles()
      character(10), save :: s = 'unknown', a = 'sequential'
      if(io_dt/=0) open(io_dt, status=s, access=a, file='tim_step.dat')
      if(io_KE/=0) open(io_KE, status=s, access=a, file='vel_energy.dat')
      if(io_fr/=0) open(io_fr, status=s, access=a, file='vel_friction.dat')
      a = 'append'
   end subroutine io_openfiles

   subroutine io_closefiles()
      if(io_dt/=0) close(io_dt)
      if(io_KE/=0) close(io_KE)
      if(io_fr/=0) close(io_fr)
   end subroutine io_closefiles

   ! io_write2files (io.f90:63-90)
   subroutine io_write2files()
      if(modulo(tim_step, i_save_rate1)==0) then
         call io_save_state()
         io_save1 = io_save1 + 1
      end if
      if(modulo(tim_step, i_save_rate2)==0) then
         if(io_KE/=0) call io_write_energy()
         if(io_fr/=0) call io_write_friction()
         if(io_dt/=0 .and. d_timestep>0d0) call io_write_timestep()
         io_save2 = io_save2 + 1
      end if
      if(io_dt/=0 .and. tim_new_dt) call io_write_timestep()
      if(modulo(tim_step, i_save_rate2*50)==0) then
         call io_closefiles()
         call io_openfiles()
      end if
   end subroutine io_write2files

   ! io_load_state (io.f90:98-165): state.cdf.in onto this grid (harmonics
   ! matched by (k, m), radial interpolation when the points differ); tim_t
   ! from the file when d_time < 0, tim_dt from it when d_timestep < 0
   subroutine io_load_state()
      real(c_double) :: t, Re, alpha, dt, dtcor, dtcfl
      real(c_double) :: r(i_N)
      integer(c_int) :: rc
      logical :: ex
      inquire(file=trim(io_statefile), exist=ex)
      if(.not. ex) then
         open(99, file='PRECOMPUTING')
         close(99, status='delete')
         print*, 'state file not found!: '//trim(io_statefile)
         stop 'io_load_state: file not found!'
      end if
      r = mes_D%r(:,1)
      rc = opf_state_file_read(trim(io_statefile)//c_null_char, int(i_N, c_int), int(i_K, c_int),  &
                               int(i_M, c_int), int(i_Mp, c_int), r, t, Re, alpha, dt, dtcor, dtcfl,  &
                               vel_ur%Re, vel_ur%Im, vel_ut%Re, vel_ut%Im, vel_uz%Re, vel_uz%Im)
      call opf_need(rc, 'io_load_state')
      if(d_time<0d0) tim_t = t
      if(dabs(tim_t-t)>1d-8)     print*, ' t    :', t, ' --> ', tim_t
      if(dabs(d_Re-Re)>1d-8)     print*, ' Re   :', Re, ' --> ', d_Re
      if(dabs(d_alpha-alpha)>1d-8) print*, ' alpha:', alpha, ' --> ', d_alpha
      if(d_timestep<0d0) tim_dt = dt
      if(d_timestep>0d0) tim_dt = d_timestep
      tim_corr_dt = dtcor
      tim_cfl_dt  = dtcfl
      call vel_put_state()
   end subroutine io_load_state

   ! io_save_state (io.f90:295-340): stateNNNN.cdf.dat of the device state
   subroutine io_save_state()
      character(4) :: cnum
      real(c_double) :: r(i_N)
      write(cnum,'(I4.4)') io_save1
      print*, ' saving state'//cnum//'  t=', tim_t
      call vel_get_state()
      r = mes_D%r(:,1)
      call opf_need(opf_state_file_write('state'//cnum//'.cdf.dat'//c_null_char, int(i_N, c_int),  &
                                         int(i_K, c_int), int(i_M, c_int), int(i_Mp, c_int), r, tim_t,  &
                                         d_Re, d_alpha, tim_dt, tim_corr_dt, tim_cfl_dt, vel_ur%Re, vel_ur%Im,  &
                                         vel_ut%Re, vel_ut%Im, vel_uz%Re, vel_uz%Im), 'io_save_state')
   end subroutine io_save_state

   ! io_write_energy (io.f90:501-524): the norms of the device state
   subroutine io_write_energy()
      double precision :: E, Ek0, Em0, E_, Ek(0:i_K1), Em(0:i_M1)
      integer(c_int) :: comp
      E = 0d0; Ek0 = 0d0; Em0 = 0d0
      do comp = 0, 2
         call opf_need(opf_var_norm(opf_h, opf_slot, comp, E_, Ek, Em), 'io_write_energy')
         E   = E + E_
         Ek0 = Ek0 + Ek(0)
         Em0 = Em0 + Em(0)
      end do
      write(io_KE,'(4e20.12)') tim_t, E, Ek0, Em0
      if(E-Ek0>d_minE3d .or. tim_t<20d0) return
      print*, 'io_write_energy: Relaminarised!'
      open(99, file='RUNNING')
      close(99, status='delete')
   end subroutine io_write_energy

   ! io_write_friction (io.f90:627-639)
   subroutine io_write_friction()
      double precision :: Ub, Uc, Ufr
      call opf_need(opf_friction(opf_h, opf_slot, Ub, Uc, Ufr), 'io_write_friction')
      if(b_const_flux) vel_Pr0 = Ub
      write(io_fr,'(4e20.12)') tim_t, Ub, Uc, Ufr
   end subroutine io_write_friction

   ! io_write_timestep (io.f90:645-657)
   subroutine io_write_timestep()
   10 format(3e18.10,I2)
   11 format(1e18.10,I8,3e17.9,I2)
      if(tim_new_dt) then
         write(io_dt,11) tim_t, tim_step, tim_dt, tim_corr_dt, tim_cfl_dt, tim_cfl_dir
      else
         call vel_maxtstep()
         write(io_dt,10) tim_t, tim_corr_dt, tim_cfl_dt, tim_cfl_dir
      end if
   end subroutine io_write_timestep

 end module io
