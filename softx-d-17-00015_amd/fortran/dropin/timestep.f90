!*************************************************************************
! timestep.f90 -- module timestep of the drop-in set: the clock and
! corrector control of timestep.f90:12-278 and the banded operators.
! tim_lumesh_init / tim_lumesh_invert build and solve on the device
! (opf_tim_lumesh_init / opf_tim_lumesh_invert); a lumesh variable is a
! handle to its factor set there.  tim_mesh_init / tim_meshmult work on
! host arrays.  tim_check_cgce and tim_new_tstep take the reference's
! decisions on the host scalars.
!*************************************************************************
 module timestep
   use iso_c_binding
   use variables
   implicit none
   save

   double precision :: tim_t
   double precision :: tim_dt
   double precision :: tim_dterr
   integer          :: tim_it
   integer          :: tim_step
   double precision :: tim_corr_dt
   double precision :: tim_cfl_dt
   integer          :: tim_cfl_dir
   logical          :: tim_new_dt
   integer          :: tim_pcn

 contains

   ! tim_precompute (timestep.f90:30-41)
   subroutine tim_precompute()
      tim_t       = d_time
      tim_dt      = d_timestep
      tim_dterr   = 0d0
      tim_it      = 0
      tim_step    = 0
      tim_corr_dt = 0d0
      tim_cfl_dt  = 0d0
      tim_cfl_dir = 0
      tim_new_dt  = .false.
      tim_pcn     = -1
   end subroutine tim_precompute

   ! A = c1 I + c2 Lap per harmonic, the modified Laplacian for PM = +-1
   ! (tim_mesh_init, timestep.f90:46-60)
   subroutine tim_mesh_init(PM, c1, c2, A)
      integer,          intent(in)  :: PM
      double precision, intent(in)  :: c1, c2
      type (mesh),      intent(out) :: A(0:i_pH1)
      double precision :: d(i_N)
      integer :: nh, k, m
      do nh = 0, i_H1
         call harmonic_km(nh, k, m)
         d = -mes_D%r(:,-2)*i_Mp*m*i_Mp*m - d_alpha*k*d_alpha*k
         if(PM/=0) d = d - mes_D%r(:,-2) - 2d0*PM*i_Mp*m*mes_D%r(:,-2)
         A(nh)%M = c2 * mes_D%radLap%M
         A(nh)%M(i_KL+1,1:) = A(nh)%M(i_KL+1,1:) + c2*d + c1
      end do
   end subroutine tim_mesh_init

   ! d = A b + c per harmonic (tim_meshmult, timestep.f90:130-154); S is the
   ! parity argument of the reference's interface, unused by the product
   subroutine tim_meshmult(S, A, b, c, d)
      integer,     intent(in)  :: S
      type (mesh), intent(in)  :: A(0:i_pH1)
      type (coll), intent(in)  :: b, c
      type (coll), intent(out) :: d
      double precision :: re(i_N), im(i_N)
      integer :: nh, j, n
      do nh = 0, i_H1
         re = c%Re(:,nh)
         im = c%Im(:,nh)
         do j = 1, i_N
            do n = max(1, j-i_KL), min(j+i_KL, i_N)
               re(n) = re(n) + A(nh)%M(i_KL+1+n-j, j) * b%Re(j,nh)
               im(n) = im(n) + A(nh)%M(i_KL+1+n-j, j) * b%Im(j,nh)
            end do
         end do
         d%Re(:,nh) = re
         d%Im(:,nh) = im
      end do
   end subroutine tim_meshmult

   ! (k, m) of harmonic nh in the order of parallel.h:47-56
   subroutine harmonic_km(nh, k, m)
      integer, intent(in)  :: nh
      integer, intent(out) :: k, m
      if(nh < i_K) then
         m = 0
         k = nh
      else
         m = 1 + (nh - i_K) / (2*i_K-1)
         k = modulo(nh - i_K, 2*i_K-1) - i_K1
      end if
   end subroutine harmonic_km

   ! tim_lumesh_init (timestep.f90:67-96) on the device.  A handle that does
   ! not name a factor set yet gets a free one (slots 1.. of the context);
   ! the reference's stop 'tim_lumesh_init' on a singular factor is kept.
   subroutine tim_lumesh_init(PM, BC, c1, c2, A)
      integer,          intent(in)    :: PM, BC
      double precision, intent(in)    :: c1, c2
      type (lumesh),    intent(inout) :: A(0:i_pH1)
      integer(c_int) :: rc, slot, set
      call opf_context()
      if(A(0)%slot < 0) then
         if(opf_sets_used >= 4*(OPF_NSLOTS-1)) stop 'tim_lumesh_init: no free factor set'
         A(:)%slot = 1 + opf_sets_used / 4
         A(:)%set  = modulo(opf_sets_used, 4)
         opf_sets_used = opf_sets_used + 1
      end if
      slot = A(0)%slot
      set  = A(0)%set
      if(.not. opf_slot_set(slot)) then            ! d_alpha of the run for the factor sets
         call opf_need(opf_slot_setup(opf_h, slot, d_Re, d_alpha, max(tim_dt, 1d-3), d_implicit, 0_c_int),  &
                       'tim_lumesh_init')
         opf_slot_set(slot) = .true.
      end if
      rc = opf_tim_lumesh_init(opf_h, slot, set, int(PM, c_int), int(BC, c_int), c1, c2)
      if(rc == OPF_ERR_SINGULAR) stop 'tim_lumesh_init'
      call opf_need(rc, 'tim_lumesh_init')
   end subroutine tim_lumesh_init

   ! tim_lumesh_invert (timestep.f90:102-123): b := A^-1 b, on the device
   subroutine tim_lumesh_invert(BC, A, b)
      integer,       intent(in)    :: BC
      type (lumesh), intent(in)    :: A(0:i_pH1)
      type (coll),   intent(inout) :: b
      if(A(0)%slot < 0) stop 'tim_lumesh_invert: factor set not built'
      call opf_need(opf_tim_lumesh_invert(opf_h, int(A(0)%slot, c_int), int(A(0)%set, c_int), int(BC, c_int),  &
                                          b%Re, b%Im), 'mes_lumesh_invert')
   end subroutine tim_lumesh_invert

   ! tim_new_tstep (timestep.f90:259-278): the new tim_dt and tim_new_dt
   subroutine tim_new_tstep()
      double precision :: dt
      integer, save :: i = 0
      dt = min(tim_dt * 1.11d0, d_maxdt)
      if(tim_step == 0 .and. tim_corr_dt == 0d0) dt = min(dt, tim_cfl_dt * 0.1d0)
      if(tim_cfl_dt > 0d0) dt = min(dt, tim_cfl_dt * d_courant)
      if(tim_corr_dt > 0d0) dt = min(dt, tim_corr_dt * 0.95d0)
      i = i - 1
      tim_new_dt = (dt < tim_dt * 0.95d0 .or. (dt > tim_dt * 1.10d0 .and. i < 0)  &
                    .or. (dt == d_maxdt .and. tim_dt < dt .and. i < 0))
      if(tim_new_dt) then
         if(dt >= tim_dt) i = 99
         tim_dt = dt
      end if
   end subroutine tim_new_tstep

   ! tim_check_cgce (timestep.f90:214-253): the corrector iteration control
   subroutine tim_check_cgce()
      double precision, save :: lasterr
      if(tim_it == 1) then
         tim_corr_dt = tim_dt * dsqrt(d_dterr / tim_dterr)
         lasterr = 1d99
      end if
      if(tim_dt < 1d-9 .and. tim_step > 30) then
         print*, 'tim_check_cgce: dt --> 0 !!!?'
         tim_step = i_maxtstep - 1
         tim_it = 0
      else if(tim_it == tim_pcn) then
         tim_it = 0
      else if(tim_it > 10) then
         print*, 'tim_check_cgce: too many its!!!'
         tim_step = i_maxtstep - 1
         tim_it = 0
      else if(tim_dterr > lasterr) then
         print*, 'tim_check_cgce: increasing error!!!'
         if(tim_dterr > 2d0 * d_dterr) tim_step = i_maxtstep - 1
         if(tim_dterr < 2d0 * d_dterr) tim_corr_dt = tim_dt / (1d1 * d_courant)
         tim_it = 0
      else if(tim_dterr > d_dterr) then
         lasterr = tim_dterr
         tim_it = tim_it + 1
      else
         tim_it = 0
      end if
      tim_dterr = 0d0
   end subroutine tim_check_cgce

 end module timestep
