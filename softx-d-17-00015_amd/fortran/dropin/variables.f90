!*************************************************************************
! variables.f90 -- module variables of the drop-in set: the field types of
! variables.f90:28-47 (coll: radial index fastest; spec: harmonic fastest;
! phys: z fastest, then theta, then r) and the procedures a host or a
! plugin calls on them.  One process holds every harmonic and plane, so the
! coll <-> spec changes of layout are plain transposes.  var_coll_norm runs
! on the device (opf_var_coll_norm).
!*************************************************************************
 module variables
   use iso_c_binding
   use mpif
   use parameters
   use meshs
   use opf_dropin
   implicit none
   save

   type harm
      integer              :: pH0, pH1, pH0_(0:0), pH1_(0:0)
   end type harm

   type spec
      double precision     :: Re(0:i_pH1, i_pN)
      double precision     :: Im(0:i_pH1, i_pN)
   end type spec

   type coll
      double precision     :: Re(i_N, 0:i_pH1)
      double precision     :: Im(i_N, 0:i_pH1)
   end type coll

   type phys
      double precision     :: Re(0:i_pZ-1, 0:i_Th-1, i_pN)
   end type phys

   type (harm)             :: var_H

 contains

   ! var_precompute (variables.f90:60-88) for one process: every harmonic
   subroutine var_precompute()
      var_H%pH0 = 0
      var_H%pH1 = i_H1
      var_H%pH0_(0) = 0
      var_H%pH1_(0) = i_H1
   end subroutine var_precompute

   ! the hook of main.f90:23,30,38 (variables.f90:110-112); plugins replace it
   subroutine var_null(flag)
      integer, intent(in), optional :: flag
   end subroutine var_null

   subroutine var_coll_init(a)
      type (coll), intent(out) :: a
      a%Re = 0d0
      a%Im = 0d0
   end subroutine var_coll_init

   subroutine var_coll_copy(in, out)
      type (coll), intent(in)  :: in
      type (coll), intent(out) :: out
      out%Re(:,0:i_H1) = in%Re(:,0:i_H1)
      out%Im(:,0:i_H1) = in%Im(:,0:i_H1)
   end subroutine var_coll_copy

   ! a := a + ac
   subroutine var_coll_add(ac, a)
      type (coll), intent(in)    :: ac
      type (coll), intent(inout) :: a
      a%Re(:,0:i_H1) = a%Re(:,0:i_H1) + ac%Re(:,0:i_H1)
      a%Im(:,0:i_H1) = a%Im(:,0:i_H1) + ac%Im(:,0:i_H1)
   end subroutine var_coll_add

   ! a := a - ac
   subroutine var_coll_sub(ac, a)
      type (coll), intent(in)    :: ac
      type (coll), intent(inout) :: a
      a%Re(:,0:i_H1) = a%Re(:,0:i_H1) - ac%Re(:,0:i_H1)
      a%Im(:,0:i_H1) = a%Im(:,0:i_H1) - ac%Im(:,0:i_H1)
   end subroutine var_coll_sub

   ! a := d a
   subroutine var_coll_dmult(d, a)
      double precision, intent(in) :: d
      type (coll), intent(inout)   :: a
      a%Re(:,0:i_H1) = a%Re(:,0:i_H1) * d
      a%Im(:,0:i_H1) = a%Im(:,0:i_H1) * d
   end subroutine var_coll_dmult

   ! a := -ac
   subroutine var_coll_neg(ac, a)
      type (coll), intent(in)    :: ac
      type (coll), intent(inout) :: a
      a%Re(:,0:i_H1) = - ac%Re(:,0:i_H1)
      a%Im(:,0:i_H1) = - ac%Im(:,0:i_H1)
   end subroutine var_coll_neg

   ! coll -> spec of up to three fields (variables.f90:184-197)
   subroutine var_coll2spec(c,s, c2,s2, c3,s3)
      type (coll), intent(in)  :: c
      type (spec), intent(out) :: s
      type (coll), intent(in),  optional :: c2,c3
      type (spec), intent(out), optional :: s2,s3
      s%Re = transpose(c%Re)
      s%Im = transpose(c%Im)
      if(.not. present(c2)) return
      s2%Re = transpose(c2%Re)
      s2%Im = transpose(c2%Im)
      if(.not. present(c3)) return
      s3%Re = transpose(c3%Re)
      s3%Im = transpose(c3%Im)
   end subroutine var_coll2spec

   ! spec -> coll of up to three fields (variables.f90:281-294)
   subroutine var_spec2coll(s,c, s2,c2, s3,c3)
      type (spec), intent(in)  :: s
      type (coll), intent(out) :: c
      type (spec), intent(in),  optional :: s2,s3
      type (coll), intent(out), optional :: c2,c3
      c%Re = transpose(s%Re)
      c%Im = transpose(s%Im)
      if(.not. present(s2)) return
      c2%Re = transpose(s2%Re)
      c2%Im = transpose(s2%Im)
      if(.not. present(s3)) return
      c3%Re = transpose(s3%Re)
      c3%Im = transpose(s3%Im)
   end subroutine var_spec2coll

   ! var_coll_norm (variables.f90:718-751) of any coll field, on the device,
   ! with d_alpha of the state's parameters (set by vel_matrices)
   subroutine var_coll_norm(a, E, Ek, Em)
      type (coll),      intent(in)  :: a
      double precision, intent(out) :: E, Ek(0:i_K1), Em(0:i_M1)
      if(d_alpha==0d0) stop 'var_coll_norm: alpha=0'
      call opf_need(opf_var_coll_norm(opf_h, opf_slot, a%Re, a%Im, E, Ek, Em), 'var_coll_norm')
   end subroutine var_coll_norm

 end module variables
