!*************************************************************************
! opf_dropin.f90 -- the device context behind the drop-in modules (mpif,
! meshs, variables, transform, timestep, velocity, io).  One context holds
! the resolution of parameters.f90 (i_N, i_K, i_M, i_Mp): slot 0 is the
! state of the time loop (vel_ur, vel_ut, vel_uz and the factor sets of
! vel_matrices); slots 1 .. OPF_NSLOTS-1 hold factor sets that
! tim_lumesh_init builds for other lumesh variables (four per slot).
!*************************************************************************
 module opf_dropin
   use iso_c_binding
   use opf_c
   use parameters
   implicit none
   save

   integer(c_int), parameter :: OPF_NSLOTS = 3
   integer(c_int), parameter :: opf_slot = 0
   type(c_ptr)   :: opf_h = c_null_ptr
   logical       :: opf_slot_set(0:OPF_NSLOTS-1) = .false.
   integer       :: opf_sets_used = 0          ! factor sets handed out in slots >= 1
   logical       :: opf_put_pending = .false.  ! host state loaded before the slot's first vel_matrices

 contains

   ! a failing call stops with the reference procedure's name, as the
   ! reference stops on LAPACK failures (timestep.f90:93,117,120)
   subroutine opf_need(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if(rc == OPF_OK) return
      print*, trim(where)//': '//opf_error_message()
      stop 'opf: device call failed'
   end subroutine opf_need

   ! the context of parameters.f90's resolution, created on first use
   subroutine opf_context()
      if(c_associated(opf_h)) return
      if(i_pN /= i_N) stop 'opf_context: the device path runs as one process (_Nr = _Ns = 1)'
      call opf_need(opf_ctx_create(opf_h, int(i_N, c_int), int(i_K, c_int), int(i_M, c_int),  &
                                   int(i_Mp, c_int), OPF_NSLOTS, c_null_ptr), 'opf_context')
   end subroutine opf_context

 end module opf_dropin
