!*************************************************************************
! Harness driver for the reference's Newton-Krylov-hookstep solver
! (NewtonHook.f90, GMRESm.f90), compiled from the sources in place by
! oracle/build_newton_ref.sh.  Test infrastructure only.
!
! It solves a small nonlinear system with the reference's newtonhook, using
! a Jacobian-vector product by finite differences in the form of multJ
! (newton.f90:399-437: eps = epsJ |x_n| / |dx|), the Euclidean dot product
! and no preconditioner, and writes every accepted iterate:
!
! usage: newton_ref out.txt n m nits gtol tol del mndl mxdl epsJ [amp]
!   f_i(x) = x_i + 0.4 sin(2 x_{i+1}) + 0.2 x_i^3 - b_i   (cyclic in i)
!   with b = f(x*) for x*_i = cos(i), started from x0_i = x*_i + amp sin(3 i)
!   (amp = 1.5 by default).  newtonhook prints its progress (info = 1).
!   out.txt: one line per call of sub(): new_nits new_gits new_tol new_del
!   x(1:n); then a line "info <info>".
!*************************************************************************
 module testsys
   implicit none
   integer :: ns
   double precision, allocatable :: bvec(:)
   double precision :: epsJ
   integer :: uout
 contains
   subroutine fsys(n, x, y)
      integer, intent(in) :: n
      double precision, intent(in) :: x(n)
      double precision, intent(out) :: y(n)
      integer :: i, j
      do i = 1, n
         j = modulo(i, n) + 1
         y(i) = x(i) + 0.4d0 * dsin(2d0 * x(j)) + 0.2d0 * x(i)**3 - bvec(i)
      end do
   end subroutine fsys
 end module testsys

 subroutine getrhs(n, x, y)
   use testsys
   implicit none
   integer, intent(in) :: n
   double precision, intent(in) :: x(n)
   double precision, intent(out) :: y(n)
   call fsys(n, x, y)
 end subroutine getrhs

 subroutine multJ(n, x, y)
   use newton
   use testsys, only : epsJ
   implicit none
   integer, intent(in) :: n
   double precision, intent(in) :: x(n)
   double precision, intent(out) :: y(n)
   double precision :: eps, s(n)
   eps = dsqrt(dot_product(x, x))
   if(eps == 0d0) stop 'multJ: eps=0'
   eps = epsJ * dsqrt(dot_product(new_x, new_x)) / eps
   y = new_x + eps * x
   call getrhs(n, y, s)
   y = (s - new_fx) / eps
 end subroutine multJ

 subroutine multJp(n, x)
   implicit none
   integer, intent(in) :: n
   double precision, intent(inout) :: x(n)
 end subroutine multJp

 subroutine saveit()
   use newton
   use testsys, only : uout
   implicit none
   write(uout, '(2I6,2ES26.17E3)', advance='no') new_nits, new_gits, new_tol, new_del
   write(uout, '(*(ES26.17E3))') new_x
 end subroutine saveit

 double precision function dotp(n, a, b)
   implicit none
   integer, intent(in) :: n
   double precision, intent(in) :: a(n), b(n)
   dotp = dot_product(a, b)
 end function dotp

 program newton_ref
   use newton
   use testsys
   implicit none
   external :: getrhs, multJ, multJp, saveit
   double precision, external :: dotp
   character(len=512) :: fout, arg
   integer :: n, m, nits, info, i
   double precision :: gtol, tol, del, mndl, mxdl, amp
   double precision, allocatable :: xs(:)

   if(command_argument_count() < 10) stop 'usage: out n m nits gtol tol del mndl mxdl epsJ'
   call get_command_argument(1, fout)
   call get_command_argument(2, arg); read(arg, *) n
   call get_command_argument(3, arg); read(arg, *) m
   call get_command_argument(4, arg); read(arg, *) nits
   call get_command_argument(5, arg); read(arg, *) gtol
   call get_command_argument(6, arg); read(arg, *) tol
   call get_command_argument(7, arg); read(arg, *) del
   call get_command_argument(8, arg); read(arg, *) mndl
   call get_command_argument(9, arg); read(arg, *) mxdl
   call get_command_argument(10, arg); read(arg, *) epsJ
   amp = 1.5d0
   if(command_argument_count() >= 11) then
      call get_command_argument(11, arg); read(arg, *) amp
   end if

   ns = n
   allocate(bvec(n), xs(n), new_x(n), new_fx(n))
   do i = 1, n
      xs(i) = dcos(dble(i))
   end do
   bvec = 0d0
   call fsys(n, xs, bvec)          ! f(x*) with b = 0
   do i = 1, n
      new_x(i) = xs(i) + amp * dsin(3d0 * dble(i))
   end do

   open(newunit=uout, file=trim(fout), status='replace')
   info = 1
   call newtonhook(getrhs, multJ, multJp, saveit, dotp, m, n, gtol, tol, del, mndl, mxdl, nits, info)
   write(uout, '(A,I4)') 'info ', info
   close(uout)
 end program newton_ref
