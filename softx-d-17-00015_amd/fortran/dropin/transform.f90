!*************************************************************************
! transform.f90 -- module transform of the drop-in set, for the TRANSFORM
! slot of the reference's Makefile (transform.$(TRANSFORM).f90,
! Makefile:6,70-72): tra_precompute, tra_spec2phys and tra_phys2spec with
! the conventions of transform.fftw3.f90:35-147 (3/2-padded grid, no
! normalisation forward, 1/(3K 3M) back, m = 0 kept Hermitian), run on the
! device by libopf's plan-specialised LDS transforms.
!*************************************************************************
 module transform
   use iso_c_binding
   use variables
   implicit none
   save

   type (coll), private :: tc

 contains

   ! tra_precompute (transform.fftw3.f90:35-59): the device context holds the plans
   subroutine tra_precompute()
      call opf_context()
   end subroutine tra_precompute

   ! spectral -> physical (transform.fftw3.f90:65-102)
   subroutine tra_spec2phys(s, p)
      type (spec), intent(in)  :: s
      type (phys), intent(out) :: p
      call var_spec2coll(s, tc)
      call tra_coll2phys(tc, p)
   end subroutine tra_spec2phys

   ! physical -> spectral (transform.fftw3.f90:108-147)
   subroutine tra_phys2spec(p, s)
      type (phys), intent(in)  :: p
      type (spec), intent(out) :: s
      call tra_phys2coll(p, tc)
      call var_coll2spec(tc, s)
   end subroutine tra_phys2spec

   ! coll -> phys directly (var_coll2spec + tra_spec2phys)
   subroutine tra_coll2phys(c, p)
      type (coll), intent(in)  :: c
      type (phys), intent(out) :: p
      call opf_context()
      call opf_need(opf_tra_coll2phys(opf_h, c%Re, c%Im, p%Re), 'tra_spec2phys')
   end subroutine tra_coll2phys

   ! phys -> coll directly (tra_phys2spec + var_spec2coll)
   subroutine tra_phys2coll(p, c)
      type (phys), intent(in)  :: p
      type (coll), intent(out) :: c
      call opf_context()
      call opf_need(opf_tra_phys2coll(opf_h, p%Re, c%Re, c%Im), 'tra_phys2spec')
   end subroutine tra_phys2coll

 end module transform
