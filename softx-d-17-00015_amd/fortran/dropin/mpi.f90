!*************************************************************************
! mpi.f90 -- module mpif of the drop-in set: the process layout of
! mpi.f90:6-35 for the device path, which runs as one process (the GPU
! takes the place of the ranks; parallel.h must set _Nr = _Ns = 1).
!*************************************************************************
 module mpif
   implicit none
   save
   integer :: mpi_er = 0
   integer :: mpi_rnk = 0, mpi_sze = 1

 contains

   subroutine mpi_precompute()
      mpi_rnk = 0
      mpi_sze = 1
   end subroutine mpi_precompute

 end module mpif
