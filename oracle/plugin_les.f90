!*************************************************************************
! plugin_les.f90 -- the var_null(2) point of main.f90:23,38 with the LES
! plugin: the FLAG = 2 branch of its util (LES.f90:39-64), whose first call
! reads params_model.in (init_params_model, LES.f90:69-96); the reference
! viscosity is revised at predictor evaluations (tim_it = 0).  The plugin's
! own module and procedures are compiled from LES.f90 by
! oracle/build_ref.sh.  Test infrastructure only.
!*************************************************************************
 subroutine plugin_hook()
   use timestep
   implicit none
   logical, save :: set = .false.
   if(.not.set) then
      set = .true.
      call init_params_model()
   end if
   call get_nu()
   if(tim_it==0) call update_nu_r()
   call update_nonlin_terms()
 end subroutine plugin_hook
