!*************************************************************************
! opf_c.f90 -- Fortran interfaces (iso_c_binding) of the C ABI in
! include/opf.h (libopf.so).  Pointer arguments that may be NULL are
! type(c_ptr), value; fields are passed as the reference's coll planes
! Re(1:N,0:H-1) / Im(1:N,0:H-1) (variables.f90:37-40), radial index fastest.
!*************************************************************************
module opf_c
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: OPF_OK = 0, OPF_ERR_ARG = 1, OPF_ERR_HIP = 2, &
                                OPF_ERR_SINGULAR = 3, OPF_ERR_CORRECTOR = 4, &
                                OPF_ERR_STATE = 5, OPF_ERR_NOMEM = 6

   interface
      function opf_last_error() bind(C, name="opf_last_error") result(msg)
         import :: c_ptr
         type(c_ptr) :: msg
      end function

      function opf_version() bind(C, name="opf_version") result(v)
         import :: c_int
         integer(c_int) :: v
      end function

      ! mes_precompute + var_precompute + tra_precompute (meshs.f90:66-103,
      ! variables.f90:60-88, transform.fftw3.f90:35-59); r = NULL: default mesh
      function opf_ctx_create(ctx, N, K, M, Mp, capacity, r) bind(C, name="opf_ctx_create") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), intent(out) :: ctx
         integer(c_int), value :: N, K, M, Mp, capacity
         type(c_ptr), value :: r
         integer(c_int) :: rc
      end function

      function opf_ctx_destroy(ctx) bind(C, name="opf_ctx_destroy") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int) :: rc
      end function

      function opf_ctx_dims(ctx, N, K, M, Mp, H, capacity, bytes) bind(C, name="opf_ctx_dims") result(rc)
         import :: c_ptr, c_int, c_size_t
         type(c_ptr), value :: ctx
         integer(c_int), intent(out) :: N, K, M, Mp, H, capacity
         integer(c_size_t), intent(out) :: bytes
         integer(c_int) :: rc
      end function

      function opf_ctx_bytes_needed(N, K, M, capacity) bind(C, name="opf_ctx_bytes_needed") result(b)
         import :: c_int, c_size_t
         integer(c_int), value :: N, K, M, capacity
         integer(c_size_t) :: b
      end function

      ! mes_D (meshs.f90:48-58); NULL pointers are skipped
      function opf_mesh_get(ctx, r, intrdr, dr0, dr1, dr1band, radlap) bind(C, name="opf_mesh_get") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, r, intrdr, dr0, dr1, dr1band, radlap
         integer(c_int) :: rc
      end function

      function opf_mesh_build(N, r_in, r, intrdr, dr0, dr1, dr1band, radlap) bind(C, name="opf_mesh_build") result(rc)
         import :: c_ptr, c_int
         integer(c_int), value :: N
         type(c_ptr), value :: r_in, r, intrdr, dr0, dr1, dr1band, radlap
         integer(c_int) :: rc
      end function

      ! vel_matrices (velocity.f90:120-146) for one state slot
      function opf_slot_setup(ctx, slot, Re, alpha, dt, implicit, const_flux) bind(C, name="opf_slot_setup") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot, const_flux
         real(c_double), value :: Re, alpha, dt, implicit
         integer(c_int) :: rc
      end function

      function opf_put_state(ctx, slot, ur_re, ur_im, ut_re, ut_im, uz_re, uz_im, t) &
            bind(C, name="opf_put_state") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         real(c_double), intent(in) :: ur_re(*), ur_im(*), ut_re(*), ut_im(*), uz_re(*), uz_im(*)
         real(c_double), value :: t
         integer(c_int) :: rc
      end function

      function opf_get_state(ctx, slot, ur_re, ur_im, ut_re, ut_im, uz_re, uz_im, t) &
            bind(C, name="opf_get_state") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         real(c_double), intent(out) :: ur_re(*), ur_im(*), ut_re(*), ut_im(*), uz_re(*), uz_im(*)
         real(c_double), intent(out) :: t
         integer(c_int) :: rc
      end function

      ! main.f90:18-46 on the device
      function opf_run(ctx, nsteps, corr_fixed, diag_every) bind(C, name="opf_run") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: nsteps, corr_fixed, diag_every
         integer(c_int) :: rc
      end function

      function opf_run_slots(ctx, slots, nslots, nsteps, corr_fixed, diag_every) &
            bind(C, name="opf_run_slots") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), intent(in) :: slots(*)
         integer(c_int), value :: nslots, nsteps, corr_fixed, diag_every
         integer(c_int) :: rc
      end function

      function opf_slot_status(ctx, slot, step, t, status, last_ncorr, last_dterr, Pr0) &
            bind(C, name="opf_slot_status") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         integer(c_int), intent(out) :: step, status, last_ncorr
         real(c_double), intent(out) :: t, last_dterr, Pr0
         integer(c_int) :: rc
      end function

      function opf_diag_get(ctx, slot, recs, maxrec, nrec) bind(C, name="opf_diag_get") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot, maxrec
         real(c_double), intent(out) :: recs(*)
         integer(c_int), intent(out) :: nrec
         integer(c_int) :: rc
      end function

      ! reference procedures, one call each
      function opf_vel_imposesym(ctx, slot) bind(C, name="opf_vel_imposesym") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         integer(c_int) :: rc
      end function

      function opf_vel_transform(ctx, slot) bind(C, name="opf_vel_transform") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         integer(c_int) :: rc
      end function

      function opf_vel_nonlinear(ctx, slot) bind(C, name="opf_vel_nonlinear") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         integer(c_int) :: rc
      end function

      function opf_vel_predictor(ctx, slot) bind(C, name="opf_vel_predictor") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         integer(c_int) :: rc
      end function

      function opf_vel_corrector(ctx, slot, dterr) bind(C, name="opf_vel_corrector") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         real(c_double), intent(out) :: dterr
         integer(c_int) :: rc
      end function

      ! vel_maxtstep (velocity.f90:438-481) of the last evaluation
      function opf_vel_maxtstep(ctx, slot, cfl_dt, cfl_dir) bind(C, name="opf_vel_maxtstep") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         real(c_double), intent(out) :: cfl_dt
         integer(c_int), intent(out) :: cfl_dir
         integer(c_int) :: rc
      end function

      ! vel_matrices (velocity.f90:120-146) for a new tim_dt
      function opf_vel_matrices(ctx, slot, dt) bind(C, name="opf_vel_matrices") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         real(c_double), value :: dt
         integer(c_int) :: rc
      end function

      ! d_timestep < 0 in the device-resident loop (opf_run)
      function opf_slot_adaptive(ctx, slot, enable, maxdt, courant) bind(C, name="opf_slot_adaptive") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot, enable
         real(c_double), value :: maxdt, courant
         integer(c_int) :: rc
      end function

      ! b_mirrorsym, b_shiftrefl, b_shiftrott (parameters.f90:26-28)
      function opf_slot_symmetry(ctx, slot, mirror, shiftrefl, shiftrott) bind(C, name="opf_slot_symmetry") &
            result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot, mirror, shiftrefl, shiftrott
         integer(c_int) :: rc
      end function

      ! tim_dt, tim_corr_dt, tim_cfl_dt, tim_cfl_dir (io_write_timestep)
      function opf_slot_timestep(ctx, slot, dt, corr_dt, cfl_dt, cfl_dir) bind(C, name="opf_slot_timestep") &
            result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         real(c_double), intent(out) :: dt, corr_dt, cfl_dt
         integer(c_int), intent(out) :: cfl_dir
         integer(c_int) :: rc
      end function

      function opf_get_nonlinear(ctx, slot, nr_re, nr_im, nt_re, nt_im, nz_re, nz_im) &
            bind(C, name="opf_get_nonlinear") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         real(c_double), intent(out) :: nr_re(*), nr_im(*), nt_re(*), nt_im(*), nz_re(*), nz_im(*)
         integer(c_int) :: rc
      end function

      ! var_coll_norm (variables.f90:718-751) of one component of the state
      function opf_var_norm(ctx, slot, comp, E, Ek, Em) bind(C, name="opf_var_norm") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot, comp
         real(c_double), intent(out) :: E, Ek(*), Em(*)
         integer(c_int) :: rc
      end function

      ! io_write_friction (io.f90:627-639)
      function opf_friction(ctx, slot, Ub, Uc, Ufr) bind(C, name="opf_friction") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         real(c_double), intent(out) :: Ub, Uc, Ufr
         integer(c_int) :: rc
      end function

      ! vel_r/t/z, vel_curlr/t/z for var_null(2) hooks
      function opf_get_phys(ctx, slot, which, phys) bind(C, name="opf_get_phys") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot, which
         real(c_double), intent(out) :: phys(*)
         integer(c_int) :: rc
      end function

      ! vel_N* += dN (forcing of var_null(2) hooks)
      function opf_add_nonlinear(ctx, slot, nr_re, nr_im, nt_re, nt_im, nz_re, nz_im) &
            bind(C, name="opf_add_nonlinear") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         real(c_double), intent(in) :: nr_re(*), nr_im(*), nt_re(*), nt_im(*), nz_re(*), nz_im(*)
         integer(c_int) :: rc
      end function

      function opf_tra_coll2phys(ctx, c_re, c_im, phys) bind(C, name="opf_tra_coll2phys") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: c_re(*), c_im(*)
         real(c_double), intent(out) :: phys(*)
         integer(c_int) :: rc
      end function

      function opf_tra_phys2coll(ctx, phys, c_re, c_im) bind(C, name="opf_tra_phys2coll") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: phys(*)
         real(c_double), intent(out) :: c_re(*), c_im(*)
         integer(c_int) :: rc
      end function

      function opf_tim_lumesh_get(ctx, slot, which, ab, ipiv) bind(C, name="opf_tim_lumesh_get") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot, which
         type(c_ptr), value :: ab, ipiv
         integer(c_int) :: rc
      end function

      ! tim_lumesh_invert(BC, A, b) (timestep.f90:102-123) with factor set `which`
      function opf_tim_lumesh_invert(ctx, slot, which, BC, b_re, b_im) bind(C, name="opf_tim_lumesh_invert") &
            result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot, which, BC
         real(c_double), intent(inout) :: b_re(*), b_im(*)
         integer(c_int) :: rc
      end function

      ! tim_lumesh_init(PM, BC, c1, c2, A) (timestep.f90:67-96) into factor set `which`
      function opf_tim_lumesh_init(ctx, slot, which, PM, BC, c1, c2) bind(C, name="opf_tim_lumesh_init") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot, which, PM, BC
         real(c_double), value :: c1, c2
         integer(c_int) :: rc
      end function

      ! var_coll_norm(a, E, Ek, Em) (variables.f90:718-751) of any coll field
      function opf_var_coll_norm(ctx, slot, a_re, a_im, E, Ek, Em) bind(C, name="opf_var_coll_norm") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         real(c_double), intent(in) :: a_re(*), a_im(*)
         real(c_double), intent(out) :: E, Ek(*), Em(*)
         integer(c_int) :: rc
      end function

      function opf_kernel_timing(ctx, enable) bind(C, name="opf_kernel_timing") result(rc)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: enable
         integer(c_int) :: rc
      end function

      function opf_kernel_time(ctx, name, ms, launches) bind(C, name="opf_kernel_time") result(rc)
         import :: c_ptr, c_int, c_double, c_char
         type(c_ptr), value :: ctx
         character(kind=c_char), intent(in) :: name(*)
         real(c_double), intent(out) :: ms
         integer(c_int), intent(out) :: launches
         integer(c_int) :: rc
      end function

      ! forcing plugins at var_null(2) (Coriolis.f90, nonnewtonian.f90, LES.f90)
      function opf_slot_coriolis(ctx, slot, enable, Ek0, phi) bind(C, name="opf_slot_coriolis") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot, enable
         real(c_double), value :: Ek0, phi
         integer(c_int) :: rc
      end function

      function opf_slot_nonnewtonian(ctx, slot, enable, rho, radius, mu_0, mu_inf, lam_CY, n_CY, a_CY) &
            bind(C, name="opf_slot_nonnewtonian") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot, enable
         real(c_double), value :: rho, radius, mu_0, mu_inf, lam_CY, n_CY, a_CY
         integer(c_int) :: rc
      end function

      function opf_slot_les(ctx, slot, enable, model, C_s, filter, A_plus) bind(C, name="opf_slot_les") &
            result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot, enable, model, filter
         real(c_double), value :: C_s, A_plus
         integer(c_int) :: rc
      end function

      ! steporbit (newton.f90:721-775) of several slots, var_coll_shift
      function opf_steporbit(ctx, nslots, slots, ndts, T, sth, sz) bind(C, name="opf_steporbit") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: nslots
         integer(c_int), intent(in) :: slots(*), ndts(*)
         real(c_double), intent(in) :: T(*), sth(*), sz(*)
         integer(c_int) :: rc
      end function

      function opf_vel_shift(ctx, slot, dth, dz) bind(C, name="opf_vel_shift") result(rc)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         real(c_double), value :: dth, dz
         integer(c_int) :: rc
      end function

      ! state files: io_save_state / io_load_state (io.f90:98-392)
      function opf_state_file_info(path, N, K, M, Mp, t, Re, alpha, dt) bind(C, name="opf_state_file_info") &
            result(rc)
         import :: c_int, c_double, c_char
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), intent(out) :: N, K, M, Mp
         real(c_double), intent(out) :: t, Re, alpha, dt
         integer(c_int) :: rc
      end function

      function opf_state_file_write(path, N, K, M, Mp, r, t, Re, alpha, dt, dtcor, dtcfl, &
            ur_re, ur_im, ut_re, ut_im, uz_re, uz_im) bind(C, name="opf_state_file_write") result(rc)
         import :: c_int, c_double, c_char
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value :: N, K, M, Mp
         real(c_double), intent(in) :: r(*)
         real(c_double), value :: t, Re, alpha, dt, dtcor, dtcfl
         real(c_double), intent(in) :: ur_re(*), ur_im(*), ut_re(*), ut_im(*), uz_re(*), uz_im(*)
         integer(c_int) :: rc
      end function

      function opf_state_file_read(path, N, K, M, Mp, r, t, Re, alpha, dt, dtcor, dtcfl, &
            ur_re, ur_im, ut_re, ut_im, uz_re, uz_im) bind(C, name="opf_state_file_read") result(rc)
         import :: c_int, c_double, c_char
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value :: N, K, M, Mp
         real(c_double), intent(in) :: r(*)
         real(c_double), intent(out) :: t, Re, alpha, dt, dtcor, dtcfl
         real(c_double), intent(out) :: ur_re(*), ur_im(*), ut_re(*), ut_im(*), uz_re(*), uz_im(*)
         integer(c_int) :: rc
      end function

      function opf_state_save(ctx, slot, path) bind(C, name="opf_state_save") result(rc)
         import :: c_ptr, c_int, c_char
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int) :: rc
      end function

      function opf_state_load(ctx, slot, path, t) bind(C, name="opf_state_load") result(rc)
         import :: c_ptr, c_int, c_double, c_char
         type(c_ptr), value :: ctx
         integer(c_int), value :: slot
         character(kind=c_char), intent(in) :: path(*)
         real(c_double), value :: t
         integer(c_int) :: rc
      end function
   end interface

contains

   ! message of the last failing call, as a Fortran string
   function opf_error_message() result(s)
      character(len=:), allocatable :: s
      character(kind=c_char), pointer :: p(:)
      type(c_ptr) :: cp
      integer :: n
      interface
         function c_strlen(str) bind(C, name="strlen") result(len)
            import :: c_ptr, c_size_t
            type(c_ptr), value :: str
            integer(c_size_t) :: len
         end function
      end interface
      cp = opf_last_error()
      if (.not. c_associated(cp)) then
         s = ''
         return
      end if
      n = int(c_strlen(cp))
      call c_f_pointer(cp, p, [n])
      allocate(character(len=n) :: s)
      s = transfer(p(1:n), s)
   end function opf_error_message

end module opf_c
