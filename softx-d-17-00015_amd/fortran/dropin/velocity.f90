!*************************************************************************
! velocity.f90 -- module velocity of the drop-in set: the time-step
! procedures of velocity.f90 with the reference's names, argument-free
! calling convention and module state, done on the device.  The state
! (u, N, the factor sets LDp, LDm, LDz, LNp and the influence matrices)
! lives in device memory; vel_ur, vel_ut, vel_uz are host copies that
! vel_get_state / vel_put_state synchronise, which the io module does at
! its load and save points.  vel_r .. vel_curlz and vel_Nr .. vel_Nz are
! filled on request (vel_get_phys, vel_get_nonlinear) for code that reads
! them on the host.
!*************************************************************************
 module velocity
   use iso_c_binding
   use variables
   use transform
   use timestep
   implicit none
   save

   type (phys) :: vel_r
   type (phys) :: vel_t
   type (phys) :: vel_z
   type (phys) :: vel_curlr
   type (phys) :: vel_curlt
   type (phys) :: vel_curlz
   type (coll) :: vel_ur
   type (coll) :: vel_ut
   type (coll) :: vel_uz
   type (coll) :: vel_Nr
   type (coll) :: vel_Nt
   type (coll) :: vel_Nz
   double precision :: vel_nu
   double precision :: vel_Pr0
   double precision :: vel_U(i_N)
   double precision :: vel_Up(i_N)

   type (lumesh), private :: LDp(0:i_pH1), LDm(0:i_pH1)
   type (lumesh), private :: LDz(0:i_pH1), LNp(0:i_pH1)

 contains

   ! vel_precompute (velocity.f90:47-55)
   subroutine vel_precompute()
      call opf_context()
      call var_coll_init(vel_ur)
      call var_coll_init(vel_ut)
      call var_coll_init(vel_uz)
      vel_nu  = 1d0
      vel_Pr0 = 0d0
      vel_U   = 1d0 - mes_D%r(:,2)
      vel_Up  = - 2d0*mes_D%r(:,1)
   end subroutine vel_precompute

   ! vel_matrices (velocity.f90:120-146): LNp, LDp, LDm, LDz for tim_dt, the
   ! influence matrices of vel_adjPPE(0) and vel_adjustFlux(0), on the device
   subroutine vel_matrices()
      integer(c_int) :: cf
      cf = 0
      if(b_const_flux) cf = 1
      call opf_context()
      call opf_need(opf_slot_symmetry(opf_h, opf_slot, merge(1_c_int, 0_c_int, b_mirrorsym),  &
                                      merge(1_c_int, 0_c_int, b_shiftrefl), merge(1_c_int, 0_c_int, b_shiftrott)),  &
                    'vel_matrices')
      call opf_need(opf_slot_setup(opf_h, opf_slot, d_Re, d_alpha, tim_dt, d_implicit, cf), 'vel_matrices')
      opf_slot_set(opf_slot) = .true.
      if(opf_put_pending) call vel_put_state()    ! main.f90 loads the state before vel_matrices
      LDp(:)%slot = opf_slot;  LDp(:)%set = 0
      LDm(:)%slot = opf_slot;  LDm(:)%set = 1
      LDz(:)%slot = opf_slot;  LDz(:)%set = 2
      LNp(:)%slot = opf_slot;  LNp(:)%set = 3
   end subroutine vel_matrices

   ! vel_imposesym (velocity.f90:61-74)
   subroutine vel_imposesym()
      call opf_need(opf_vel_imposesym(opf_h, opf_slot), 'vel_imposesym')
   end subroutine vel_imposesym

   ! vel_transform (velocity.f90:304-317)
   subroutine vel_transform()
      call opf_need(opf_vel_transform(opf_h, opf_slot), 'vel_transform')
   end subroutine vel_transform

   ! vel_nonlinear (velocity.f90:323-335), with vel_addHPF; vel_Pr0 stays on
   ! the device (io_write_friction reads it there)
   subroutine vel_nonlinear()
      call opf_need(opf_vel_nonlinear(opf_h, opf_slot), 'vel_nonlinear')
   end subroutine vel_nonlinear

   ! vel_predictor (velocity.f90:400-410)
   subroutine vel_predictor()
      call opf_need(opf_vel_predictor(opf_h, opf_slot), 'vel_predictor')
   end subroutine vel_predictor

   ! vel_corrector (velocity.f90:416-432) with tim_measurecorr
   ! (timestep.f90:190-207): tim_dterr = max(tim_dterr, sqrt(max |du|^2))
   subroutine vel_corrector()
      real(c_double) :: e
      call opf_need(opf_vel_corrector(opf_h, opf_slot, e), 'vel_corrector')
      tim_dterr = max(tim_dterr, e)
   end subroutine vel_corrector

   ! vel_maxtstep (velocity.f90:438-481): tim_cfl_dt, tim_cfl_dir from the
   ! physical fields of the last vel_transform
   subroutine vel_maxtstep()
      real(c_double) :: cfl
      integer(c_int) :: dir
      call opf_need(opf_vel_maxtstep(opf_h, opf_slot, cfl, dir), 'vel_maxtstep')
      tim_cfl_dt  = cfl
      tim_cfl_dir = dir
   end subroutine vel_maxtstep

   ! host copies -> device state (after loading vel_ur, vel_ut, vel_uz); before
   ! the slot's first vel_matrices the copy waits for it
   subroutine vel_put_state()
      if(.not. opf_slot_set(opf_slot)) then
         opf_put_pending = .true.
         return
      end if
      opf_put_pending = .false.
      call opf_need(opf_put_state(opf_h, opf_slot, vel_ur%Re, vel_ur%Im, vel_ut%Re, vel_ut%Im,  &
                                  vel_uz%Re, vel_uz%Im, tim_t), 'vel_put_state')
   end subroutine vel_put_state

   ! device state -> host copies
   subroutine vel_get_state()
      real(c_double) :: t
      call opf_need(opf_get_state(opf_h, opf_slot, vel_ur%Re, vel_ur%Im, vel_ut%Re, vel_ut%Im,  &
                                  vel_uz%Re, vel_uz%Im, t), 'vel_get_state')
   end subroutine vel_get_state

   ! vel_r, vel_t, vel_z, vel_curlr, vel_curlt, vel_curlz of the current state
   subroutine vel_get_phys()
      call opf_need(opf_get_phys(opf_h, opf_slot, 0_c_int, vel_r%Re), 'vel_get_phys')
      call opf_need(opf_get_phys(opf_h, opf_slot, 1_c_int, vel_t%Re), 'vel_get_phys')
      call opf_need(opf_get_phys(opf_h, opf_slot, 2_c_int, vel_z%Re), 'vel_get_phys')
      call opf_need(opf_get_phys(opf_h, opf_slot, 3_c_int, vel_curlr%Re), 'vel_get_phys')
      call opf_need(opf_get_phys(opf_h, opf_slot, 4_c_int, vel_curlt%Re), 'vel_get_phys')
      call opf_need(opf_get_phys(opf_h, opf_slot, 5_c_int, vel_curlz%Re), 'vel_get_phys')
   end subroutine vel_get_phys

   ! vel_Nr, vel_Nt, vel_Nz of the last vel_nonlinear, base-flow terms included
   subroutine vel_get_nonlinear()
      call opf_need(opf_get_nonlinear(opf_h, opf_slot, vel_Nr%Re, vel_Nr%Im, vel_Nt%Re, vel_Nt%Im,  &
                                      vel_Nz%Re, vel_Nz%Im), 'vel_get_nonlinear')
   end subroutine vel_get_nonlinear

 end module velocity
