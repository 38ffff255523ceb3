!*************************************************************************
! plugin_nonnewtonian.f90 -- the var_null(2) point of main.f90:23,38 with
! the non-Newtonian plugin: the FLAG = 2 branch of its util
! (nonnewtonian.f90:32-57), whose first call reads params_model.in
! (init_params_model, nonnewtonian.f90:62-86).  The plugin's own module and
! procedures are compiled from nonnewtonian.f90 by oracle/build_ref.sh.
! Test infrastructure only.
!*************************************************************************
 subroutine plugin_hook()
   implicit none
   logical, save :: set = .false.
   if(.not.set) then
      set = .true.
      call init_params_model()
   end if
   call get_nu()
   call update_nonlin_terms()
 end subroutine plugin_hook
