!*************************************************************************
! parameters_example.f90 -- a parameters module with the names and meaning
! of the reference's parameters.f90:12-63 (the file a user edits for a
! run), here for (N, K, M) = (16, 6, 4), Re 2400, alpha 1.25, dt 0.01.  It
! lets the drop-in modules build without the user's file; a run builds them
! against its own parameters.f90 (build_dropin.sh).
!*************************************************************************
 module parameters
   implicit none
   save

   integer,          parameter :: i_N           = 16
   integer,          parameter :: i_K           = 6
   integer,          parameter :: i_M           = 4
   integer,          parameter :: i_Mp          = 1

   double precision            :: d_Re          = 2400d0
   double precision            :: d_alpha       = 1.25d0
   logical,          parameter :: b_const_flux  = .true.
   logical,          parameter :: b_mirrorsym   = .false.
   logical,          parameter :: b_shiftrefl   = .false.
   logical,          parameter :: b_shiftrott   = .false.
   double precision, parameter :: d_minE3d      = 1d-5

   integer,          parameter :: i_save_rate1  = 100
   integer,          parameter :: i_save_rate2  = 10
   integer,          parameter :: i_maxtstep    = 100
   double precision, parameter :: d_maxt        = -1d0
   double precision, parameter :: d_cpuhours    = 1d99
   double precision, parameter :: d_time        = 0d0
   double precision, parameter :: d_timestep    = 0.01d0
   double precision, parameter :: d_maxdt       = 1d99
   double precision, parameter :: d_dterr       = 1d-5
   double precision, parameter :: d_courant     = 0.5d0
   double precision, parameter :: d_implicit    = 0.51d0

   integer,          parameter :: i_KL  = 4
   integer,          parameter :: i_K1  = i_K-1
   integer,          parameter :: i_M1  = i_M-1
   integer,          parameter :: i_Z   = 3*i_K
   integer,          parameter :: i_Th  = 3*i_M
   integer,          parameter :: i_H1  = (2*i_K1+1)*i_M-i_K1-1
   integer,          parameter :: i_pN  = i_N
   integer,          parameter :: i_pH1 = (2*i_K-1)*i_M-1
   integer,          parameter :: i_pZ  = i_Z
   double precision, parameter :: d_PI  = 3.1415926535897931d0

 contains

   subroutine par_precompute()
   end subroutine par_precompute

 end module parameters
