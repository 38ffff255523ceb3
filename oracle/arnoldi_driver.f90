!*************************************************************************
! Harness driver for the reference's Arnoldi eigenvalue routine (arnold and
! logeig of arnoldi.f), compiled from the source in place by
! oracle/build_newton_ref.sh.  Test infrastructure only.
!
! It runs the reverse-communication loop of getEigen (newton.f90:885-905)
! with the action of a fixed matrix: sv <- A sv until arnold converges.
!
! usage: arnoldi_ref A.bin n kmax ncgd tol out.txt vecs.bin
!   A.bin: n x n float64, column-major; start vector sv_i = 1 + sin(i).
!   out.txt: "k ifail" then ncgd lines "wr2 wi2 sigma_r sigma_i" with the
!   eigenvalues and logeig's exponents for t = 1; vecs.bin: b(1:n, 1:2 ncgd).
!*************************************************************************
 double precision function dotp(n, a, b)
   implicit none
   integer, intent(in) :: n
   double precision, intent(in) :: a(n), b(n)
   dotp = dot_product(a, b)
 end function dotp

 program arnoldi_ref
   implicit none
   double precision, external :: dotp
   character(len=512) :: fa, fo, fv, arg
   integer :: n, kmax, ncgd, k, ifail, i, u
   double precision :: tol
   double precision, allocatable :: A(:,:), sv(:), h(:,:), q(:,:), b(:,:), wr(:), wi(:), sr(:), si(:)

   if(command_argument_count() < 7) stop 'usage: A.bin n kmax ncgd tol out.txt vecs.bin'
   call get_command_argument(1, fa)
   call get_command_argument(2, arg); read(arg, *) n
   call get_command_argument(3, arg); read(arg, *) kmax
   call get_command_argument(4, arg); read(arg, *) ncgd
   call get_command_argument(5, arg); read(arg, *) tol
   call get_command_argument(6, fo)
   call get_command_argument(7, fv)
   allocate(A(n,n), sv(n), h(kmax,kmax), q(n,kmax), b(n,2*ncgd), wr(ncgd), wi(ncgd), sr(ncgd), si(ncgd))
   open(newunit=u, file=trim(fa), access='stream', form='unformatted', status='old')
   read(u) A
   close(u)
   do i = 1, n
      sv(i) = 1d0 + dsin(dble(i))
   end do
   wr = 0d0
   wi = 0d0
   b = 0d0
   k = 0
   do while(.true.)
      call arnold(n, k, kmax, ncgd, dotp, tol, sv, h, q, b, wr, wi, ifail)
      if(ifail /= 0) exit
      sv = matmul(A, sv)
   end do
   sr = wr
   si = wi
   call logeig(n, ncgd, b, sr, si, 1d0, i)
   open(newunit=u, file=trim(fo), status='replace')
   write(u, '(2I6)') k, ifail
   do i = 1, ncgd
      write(u, '(4ES26.17E3)') wr(i), wi(i), sr(i), si(i)
   end do
   close(u)
   open(newunit=u, file=trim(fv), access='stream', form='unformatted', status='replace')
   write(u) b
   close(u)
 end program arnoldi_ref
