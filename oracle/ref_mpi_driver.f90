!*************************************************************************
! ref_mpi_driver.f90 -- timing driver for the reference core built with
! MPI ranks (_Nr x _Ns of parallel.h:5-6, mpi.f90:22-31).  Test and
! benchmark infrastructure only: it is the CPU baseline the GPU path is
! timed against.
!
! usage: mpiexec -n NP opf_ref_mpi_NAME state_in prefix Re alpha dt nsteps
!
! Every rank reads the whole initial state (the raw binary of
! ref_driver.f90: ur, ut, uz, each Re(1:N,0:H-1) then Im) and keeps its own
! harmonics var_H%pH0 .. var_H%pH0 + var_H%pH1 (variables.f90:64-77).  The
! loop is main.f90:18-46 with a fixed time step and no output; the
! transposes between the ranks are the reference's own
! (variables.f90:200-370).  Rank 0 writes
!   prefix_timing.txt   wall seconds of the loop (mpi_wtime), steps, ranks
!   prefix_series.txt   step and kinetic energy (var_coll_norm of the three
!                       components, io.f90:501-524) after the last step
!*************************************************************************
 program opf_ref_mpi_driver
   use velocity
   implicit none
   character(len=512) :: infile, prefix, arg
   double precision   :: re_in, alpha_in, dt_in, t0, t1, E, E_
   double precision   :: Ek(0:i_K1), Em(0:i_M1)
   integer            :: nsteps, u

   if(command_argument_count() < 6) stop 'usage: state_in prefix Re alpha dt nsteps'
   call get_command_argument(1, infile)
   call get_command_argument(2, prefix)
   call get_command_argument(3, arg);  read(arg,*) re_in
   call get_command_argument(4, arg);  read(arg,*) alpha_in
   call get_command_argument(5, arg);  read(arg,*) dt_in
   call get_command_argument(6, arg);  read(arg,*) nsteps

   call mpi_precompute()
   call par_precompute()
   d_Re    = re_in
   d_alpha = alpha_in
   call mes_precompute()
   call var_precompute()
   call tra_precompute()
   call tim_precompute()
   call vel_precompute()
   tim_dt = dt_in
   call read_own_harmonics(trim(infile))
   call vel_matrices()

   call mpi_barrier(mpi_comm_world, mpi_er)
   t0 = mpi_wtime()
   do while(tim_step < nsteps)
      call vel_imposesym()
      call vel_transform()
      call vel_nonlinear()
      call vel_predictor()
      tim_it = 1
      do while(tim_it/=0)
         call vel_transform()
         call vel_nonlinear()
         call vel_corrector()
         call tim_check_cgce()
      end do
      if(tim_step < 0) stop 'corrector failure'
      tim_t    = tim_t    + tim_dt
      tim_step = tim_step + 1
   end do
   call mpi_barrier(mpi_comm_world, mpi_er)
   t1 = mpi_wtime()

   call var_coll_norm(vel_ur, E, Ek, Em)
   call var_coll_norm(vel_ut, E_, Ek, Em)
   E = E + E_
   call var_coll_norm(vel_uz, E_, Ek, Em)
   E = E + E_
   if(mpi_rnk == 0) then
      open(newunit=u, file=trim(prefix)//'_timing.txt', status='replace')
      write(u,'(ES26.17E3,2I8)') t1 - t0, nsteps, mpi_sze
      close(u)
      open(newunit=u, file=trim(prefix)//'_series.txt', status='replace')
      write(u,'(I8,ES26.17E3)') tim_step, E
      close(u)
   end if
   call mpi_finalize(mpi_er)

 contains

   subroutine read_own_harmonics(fn)
      character(*), intent(in) :: fn
      double precision, allocatable :: buf(:,:)
      integer :: v, c, h0, h1
      h0 = var_H%pH0
      h1 = var_H%pH0 + var_H%pH1
      allocate(buf(i_N, 0:i_H1))
      call var_coll_init(vel_ur)
      call var_coll_init(vel_ut)
      call var_coll_init(vel_uz)
      open(newunit=v, file=fn, access='stream', form='unformatted', status='old')
      do c = 1, 6
         read(v) buf
         select case(c)
         case(1); vel_ur%Re(:,0:var_H%pH1) = buf(:,h0:h1)
         case(2); vel_ur%Im(:,0:var_H%pH1) = buf(:,h0:h1)
         case(3); vel_ut%Re(:,0:var_H%pH1) = buf(:,h0:h1)
         case(4); vel_ut%Im(:,0:var_H%pH1) = buf(:,h0:h1)
         case(5); vel_uz%Re(:,0:var_H%pH1) = buf(:,h0:h1)
         case(6); vel_uz%Im(:,0:var_H%pH1) = buf(:,h0:h1)
         end select
      end do
      close(v)
      deallocate(buf)
   end subroutine read_own_harmonics

 end program opf_ref_mpi_driver
