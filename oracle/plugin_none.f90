!*************************************************************************
! plugin_none.f90 -- the var_null(2) point of main.f90:23,38 without a
! plugin: var_null does nothing (variables.f90:110-112).  Test
! infrastructure: linked into the reference oracle by oracle/build_ref.sh.
!*************************************************************************
 subroutine plugin_hook()
 end subroutine plugin_hook
