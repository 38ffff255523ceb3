!*************************************************************************
! opf_path.f90 -- the reference's time-step procedures over libopf.so.
!
! The procedures keep the names, the argument-free calling convention and
! the module state of the reference (velocity.f90, timestep.f90,
! transform.fftw3.f90): a host written against those modules runs the same
! sequence of calls (main.f90:18-46) with the work done on the GPU.  The
! state lives in device memory; vel_ur/vel_ut/vel_uz are host mirrors that
! vel_get_state / vel_put_state synchronise (the io_* points of the
! reference).  Sizes are run-time values here (compile-time parameters in
! parameters.f90:17-20).
!
! Error behaviour: a failing call stops the program with the reference's
! message (timestep.f90:93,117,120 stop on LAPACK failures).
!*************************************************************************
module opf_path
   use iso_c_binding
   use opf_c
   implicit none

   ! sizes (parameters.f90:17-20) and the number of harmonics (var_H)
   integer :: i_N, i_K, i_M, i_Mp, i_H

   ! parameters.f90:36-41
   double precision :: d_Re, d_alpha
   double precision :: d_implicit = 0.51d0
   double precision :: d_dterr = 1d-5
   double precision :: d_courant = 0.5d0
   integer :: i_maxtstep = -1
   logical :: b_const_flux = .true.
   double precision :: d_timestep = 0.01d0   ! < 0: adaptive (main.f90:24-29)
   double precision :: d_maxdt = 1d99

   ! the coll type of variables.f90:37-40 with run-time extents
   type coll
      real(c_double), allocatable :: Re(:,:), Im(:,:)   ! (1:i_N, 0:i_H-1)
   end type coll

   type(coll) :: vel_ur, vel_ut, vel_uz
   double precision :: vel_Pr0
   double precision, allocatable, target :: mes_r(:)   ! mes_D%r(:,1)

   ! timestep.f90:12-24
   double precision :: tim_t, tim_dt, tim_dterr, tim_corr_dt, tim_cfl_dt
   integer :: tim_it, tim_step, tim_pcn, tim_cfl_dir
   logical :: tim_new_dt

   type(c_ptr), private :: h = c_null_ptr
   integer(c_int), parameter, private :: slot = 0

contains

   subroutine opf_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc /= OPF_OK) then
         print*, trim(where)//': '//opf_error_message()
         if (rc == OPF_ERR_SINGULAR) stop 'tim_lumesh_init'
         stop 'opf'
      end if
   end subroutine opf_check

   subroutine coll_alloc(a)
      type(coll), intent(inout) :: a
      if (allocated(a%Re)) deallocate(a%Re, a%Im)
      allocate(a%Re(i_N, 0:i_H-1), a%Im(i_N, 0:i_H-1))
      a%Re = 0d0
      a%Im = 0d0
   end subroutine coll_alloc

   ! vel_precompute / var_precompute / tra_precompute / mes_precompute
   ! (velocity.f90:47-55, variables.f90:60-88, transform.fftw3.f90:35-59,
   ! meshs.f90:66-103) for the default mesh
   subroutine vel_precompute(N, K, M, Mp, Re, alpha, dt)
      integer, intent(in) :: N, K, M, Mp
      double precision, intent(in) :: Re, alpha, dt
      integer(c_int) :: n_, k_, m_, mp_, hh, cap
      integer(c_size_t) :: bytes
      call opf_check(opf_ctx_create(h, int(N, c_int), int(K, c_int), int(M, c_int), int(Mp, c_int), 1_c_int, &
                                 c_null_ptr), 'vel_precompute')
      call opf_check(opf_ctx_dims(h, n_, k_, m_, mp_, hh, cap, bytes), 'vel_precompute')
      i_N = N; i_K = K; i_M = M; i_Mp = Mp; i_H = hh
      d_Re = Re; d_alpha = alpha
      tim_dt = dt
      tim_t = 0d0; tim_step = 0; tim_it = 0; tim_dterr = 0d0; tim_corr_dt = 0d0; tim_pcn = -1
      tim_cfl_dt = 0d0; tim_cfl_dir = 0; tim_new_dt = .false.
      call coll_alloc(vel_ur); call coll_alloc(vel_ut); call coll_alloc(vel_uz)
      if (allocated(mes_r)) deallocate(mes_r)
      allocate(mes_r(i_N))
      call opf_check(opf_mesh_get(h, c_loc(mes_r), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), &
                     'vel_precompute')
   end subroutine vel_precompute

   ! vel_matrices (velocity.f90:120-146): factors for tim_dt
   subroutine vel_matrices()
      integer(c_int) :: cf
      cf = 0
      if (b_const_flux) cf = 1
      call opf_check(opf_slot_setup(h, slot, d_Re, d_alpha, tim_dt, d_implicit, cf), 'vel_matrices')
   end subroutine vel_matrices

   ! host mirrors -> device (after io_load_state)
   subroutine vel_put_state()
      call opf_check(opf_put_state(h, slot, vel_ur%Re, vel_ur%Im, vel_ut%Re, vel_ut%Im, vel_uz%Re, vel_uz%Im, &
                                tim_t), 'vel_put_state')
   end subroutine vel_put_state

   ! device -> host mirrors (before io_save_state / io_write2files)
   subroutine vel_get_state()
      real(c_double) :: t
      call opf_check(opf_get_state(h, slot, vel_ur%Re, vel_ur%Im, vel_ut%Re, vel_ut%Im, vel_uz%Re, vel_uz%Im, &
                                t), 'vel_get_state')
   end subroutine vel_get_state

   ! io_load_state (io.f90:98-165): the file onto the grid and into the host
   ! mirrors; tim_t from the file when d_time < 0, tim_dt when d_timestep < 0
   ! (then vel_matrices), tim_corr_dt and tim_cfl_dt; then vel_put_state.
   subroutine io_load_state(fname, d_time)
      character(len=*), intent(in) :: fname
      double precision, intent(in) :: d_time
      real(c_double) :: t, Re, alpha, dt, dtcor, dtcfl
      call opf_check(opf_state_file_read(trim(fname)//c_null_char, int(i_N, c_int), int(i_K, c_int), &
                                         int(i_M, c_int), int(i_Mp, c_int), mes_r, t, Re, alpha, dt, dtcor, &
                                         dtcfl, vel_ur%Re, vel_ur%Im, vel_ut%Re, vel_ut%Im, vel_uz%Re, &
                                         vel_uz%Im), 'io_load_state')
      if (d_time < 0d0) tim_t = t
      if (d_timestep < 0d0 .and. dt > 0d0 .and. dt /= tim_dt) then
         tim_dt = dt
         call vel_matrices()
      end if
      tim_corr_dt = dtcor
      tim_cfl_dt = dtcfl
      call vel_put_state()
   end subroutine io_load_state

   ! io_save_state (io.f90:295-340) to the file fname, from the device state
   subroutine io_save_state(fname)
      character(len=*), intent(in) :: fname
      call vel_get_state()
      call opf_check(opf_state_file_write(trim(fname)//c_null_char, int(i_N, c_int), int(i_K, c_int), &
                                          int(i_M, c_int), int(i_Mp, c_int), mes_r, tim_t, d_Re, d_alpha, &
                                          tim_dt, tim_corr_dt, tim_cfl_dt, vel_ur%Re, vel_ur%Im, vel_ut%Re, &
                                          vel_ut%Im, vel_uz%Re, vel_uz%Im), 'io_save_state')
   end subroutine io_save_state

   subroutine vel_imposesym()                       ! velocity.f90:61-74
      call opf_check(opf_vel_imposesym(h, slot), 'vel_imposesym')
   end subroutine vel_imposesym

   subroutine vel_transform()                       ! velocity.f90:304-317
      call opf_check(opf_vel_transform(h, slot), 'vel_transform')
   end subroutine vel_transform

   subroutine vel_nonlinear()                       ! velocity.f90:323-335
      call opf_check(opf_vel_nonlinear(h, slot), 'vel_nonlinear')
   end subroutine vel_nonlinear

   subroutine vel_predictor()                       ! velocity.f90:400-410
      call opf_check(opf_vel_predictor(h, slot), 'vel_predictor')
   end subroutine vel_predictor

   ! vel_corrector (velocity.f90:416-432) with tim_measurecorr
   ! (timestep.f90:190-207): tim_dterr = max(tim_dterr, sqrt(max|du|^2))
   subroutine vel_corrector()
      real(c_double) :: e
      call opf_check(opf_vel_corrector(h, slot, e), 'vel_corrector')
      tim_dterr = max(tim_dterr, e)
   end subroutine vel_corrector

   ! vel_maxtstep (velocity.f90:438-481): the CFL limit of the physical
   ! fields of the last vel_transform, computed on the device
   subroutine vel_maxtstep()
      real(c_double) :: cfl
      integer(c_int) :: dir
      call opf_check(opf_vel_maxtstep(h, slot, cfl, dir), 'vel_maxtstep')
      tim_cfl_dt = cfl
      tim_cfl_dir = dir
   end subroutine vel_maxtstep

   ! tim_new_tstep (timestep.f90:259-278): the new tim_dt and tim_new_dt
   subroutine tim_new_tstep()
      double precision :: dt
      integer, save :: i = 0
      dt = min(tim_dt * 1.11d0, d_maxdt)
      if (tim_step == 0 .and. tim_corr_dt == 0d0) dt = min(dt, tim_cfl_dt * 0.1d0)
      if (tim_cfl_dt > 0d0) dt = min(dt, tim_cfl_dt * d_courant)
      if (tim_corr_dt > 0d0) dt = min(dt, tim_corr_dt * 0.95d0)
      i = i - 1
      tim_new_dt = (dt < tim_dt * 0.95d0 .or. (dt > tim_dt * 1.10d0 .and. i < 0) &
                    .or. (dt == d_maxdt .and. tim_dt < dt .and. i < 0))
      if (tim_new_dt) then
         if (dt >= tim_dt) i = 99
         tim_dt = dt
      end if
   end subroutine tim_new_tstep

   ! tim_check_cgce (timestep.f90:214-253): the corrector iteration control
   subroutine tim_check_cgce()
      double precision, save :: lasterr
      if (tim_it == 1) then
         tim_corr_dt = tim_dt * dsqrt(d_dterr / tim_dterr)
         lasterr = 1d99
      end if
      if (tim_dt < 1d-9 .and. tim_step > 30) then
         print*, 'tim_check_cgce: dt --> 0 !!!?'
         tim_step = i_maxtstep - 1
         tim_it = 0
      else if (tim_it == tim_pcn) then
         tim_it = 0
      else if (tim_it > 10) then
         print*, 'tim_check_cgce: too many its!!!'
         tim_step = i_maxtstep - 1
         tim_it = 0
      else if (tim_dterr > lasterr) then
         print*, 'tim_check_cgce: increasing error!!!'
         if (tim_dterr > 2d0 * d_dterr) tim_step = i_maxtstep - 1
         if (tim_dterr < 2d0 * d_dterr) tim_corr_dt = tim_dt / (1d1 * d_courant)
         tim_it = 0
      else if (tim_dterr > d_dterr) then
         lasterr = tim_dterr
         tim_it = tim_it + 1
      else
         tim_it = 0
      end if
      tim_dterr = 0d0
   end subroutine tim_check_cgce

   ! tra_coll2phys / tra_phys2coll of one field (variables.f90:184-197,
   ! transform.fftw3.f90:65-147); p(0:3K-1, 0:3M-1, 1:N)
   subroutine tra_coll2phys(c, p)
      type(coll), intent(in) :: c
      real(c_double), intent(out) :: p(0:, 0:, :)
      call opf_check(opf_tra_coll2phys(h, c%Re, c%Im, p), 'tra_coll2phys')
   end subroutine tra_coll2phys

   subroutine tra_phys2coll(p, c)
      real(c_double), intent(in) :: p(0:, 0:, :)
      type(coll), intent(inout) :: c
      call opf_check(opf_tra_phys2coll(h, p, c%Re, c%Im), 'tra_phys2coll')
   end subroutine tra_phys2coll

   ! tim_lumesh_invert (timestep.f90:102-123) with factor `which`
   ! (0 LDp, 1 LDm, 2 LDz, 3 LNp) of the current matrices
   subroutine tim_lumesh_invert(which, b)
      integer, intent(in) :: which
      type(coll), intent(inout) :: b
      call opf_check(opf_tim_lumesh_invert(h, slot, int(which, c_int), b%Re, b%Im), 'tim_lumesh_invert')
   end subroutine tim_lumesh_invert

   ! var_coll_norm (variables.f90:718-751) of component comp (0..2)
   subroutine var_coll_norm(comp, E, Ek, Em)
      integer, intent(in) :: comp
      double precision, intent(out) :: E, Ek(0:i_K-1), Em(0:i_M-1)
      call opf_check(opf_var_norm(h, slot, int(comp, c_int), E, Ek, Em), 'var_coll_norm')
   end subroutine var_coll_norm

   ! io_write_friction quantities (io.f90:627-639)
   subroutine vel_friction(Ub, Uc, Ufr)
      double precision, intent(out) :: Ub, Uc, Ufr
      call opf_check(opf_friction(h, slot, Ub, Uc, Ufr), 'io_write_friction')
   end subroutine vel_friction

   ! vel_Pr0 and the step counters as the device holds them
   subroutine vel_status(step, ncorr, dterr)
      integer, intent(out) :: step, ncorr
      double precision, intent(out) :: dterr
      integer(c_int) :: s, st, nc
      real(c_double) :: t, de, pr0
      call opf_check(opf_slot_status(h, slot, s, t, st, nc, de, pr0), 'vel_status')
      step = s; ncorr = nc; dterr = de; vel_Pr0 = pr0
   end subroutine vel_status

   subroutine vel_cleanup()
      if (c_associated(h)) call opf_check(opf_ctx_destroy(h), 'cleanup')
      h = c_null_ptr
   end subroutine vel_cleanup

end module opf_path
