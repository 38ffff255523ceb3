!*************************************************************************
! meshs.f90 -- module meshs of the drop-in set: the types and the radial
! mesh of meshs.f90:28-103.  mes_precompute takes the default mesh,
! finite-difference bands and quadrature weights from libopf
! (opf_mesh_build, the same arithmetic as mes_precompute / mes_rdom_init).
! type lumesh is a handle to a factor set held on the device (the
! reference's holds the LAPACK band and pivots, meshs.f90:41-45).
!*************************************************************************
 module meshs
   use iso_c_binding
   use mpif
   use parameters
   use opf_c
   implicit none
   save
                                ! M(KL+1+n-j, j) = A(n,j)
   type mesh
      double precision :: M(2*i_KL+1, i_N)
   end type mesh
                                ! a factor set on the device: context slot and set (0..3)
   type lumesh
      integer :: slot = -1, set = -1
   end type lumesh

   type rdom
      integer          :: N, pNi, pN, pNi_(0:0), pN_(0:0)
      double precision :: r(i_N,-3:3)
      double precision :: intrdr(i_N)
      double precision :: dr0(1:1+i_KL,0:i_KL), dr1(1:1+i_KL,0:i_KL)
      type (mesh)      :: dr(i_KL)
      type (mesh)      :: radLap
   end type rdom

   type (rdom) :: mes_D

 contains

   ! mes_precompute (meshs.f90:66-103) for the default mesh.  dr(1) and
   ! radLap are filled; the higher derivative bands dr(2:) are not used by
   ! the time step and stay zero.
   subroutine mes_precompute()
      real(c_double), target :: r(i_N), intrdr(i_N), dr0(1+i_KL,0:i_KL), dr1(1+i_KL,0:i_KL)
      real(c_double), target :: d1(2*i_KL+1,i_N), lap(2*i_KL+1,i_N)
      integer :: p
      integer(c_int) :: rc
      rc = opf_mesh_build(int(i_N, c_int), c_null_ptr, c_loc(r), c_loc(intrdr), c_loc(dr0), c_loc(dr1),  &
                          c_loc(d1), c_loc(lap))
      if(rc /= OPF_OK) stop 'mes_precompute'
      mes_D%N   = i_N
      mes_D%pNi = 1
      mes_D%pN  = i_N
      mes_D%pNi_(0) = 1
      mes_D%pN_(0)  = i_N
      mes_D%r(:,1) = r
      mes_D%r(:,0) = 1d0
      do p = 2, 3
         mes_D%r(:,p) = mes_D%r(:,p-1) * r
      end do
      do p = -1, -3, -1
         mes_D%r(:,p) = mes_D%r(:,p+1) / r
      end do
      mes_D%intrdr = intrdr
      mes_D%dr0 = dr0
      mes_D%dr1 = dr1
      do p = 1, i_KL
         mes_D%dr(p)%M = 0d0
      end do
      mes_D%dr(1)%M = d1
      mes_D%radLap%M = lap
   end subroutine mes_precompute

 end module meshs
