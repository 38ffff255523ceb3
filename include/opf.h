/*
 * opf.h -- C ABI of the MI355X Openpipeflow time-step library (libopf.so).
 *
 * The reference's interface for this path is Fortran module procedures over
 * module state (velocity.f90, timestep.f90, transform.fftw3.f90,
 * variables.f90); these entry points are what a Fortran host binds with
 * iso_c_binding to keep those procedures as thin wrappers.  Each comment
 * names the reference procedure the entry point replaces.
 *
 * Conventions
 *  - A context (opf_ctx) is one resolution class (N, K, M, Mp): the radial
 *    mesh, FFT plans and device arrays for `capacity` independent states
 *    ("slots").  A plain run uses one slot.
 *  - Spectral fields cross the boundary in the reference's coll layout:
 *    separate real and imaginary arrays Re(1:N, 0:H-1), Im(1:N, 0:H-1),
 *    radial index fastest (variables.f90:37-40), H = (2K-1)M - (K-1),
 *    harmonics ordered as parallel.h:47-56.
 *  - Physical fields use the phys layout Re(0:3K-1, 0:3M-1, 1:N), z fastest
 *    (variables.f90:42-44).
 *  - Every call returns OPF_OK (0) or an error code; opf_last_error() gives a
 *    message.  Calls on one context are not re-entrant.
 *  - State stays in device memory between calls; only the explicit put/get
 *    calls move fields across PCIe.
 */
#ifndef OPF_H
#define OPF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct opf_ctx opf_ctx;

enum {
  OPF_OK = 0,
  OPF_ERR_ARG = 1,        /* bad argument */
  OPF_ERR_HIP = 2,        /* HIP runtime error / no device */
  OPF_ERR_SINGULAR = 3,   /* 'tim_lumesh_init' / 'matrix inversion error' */
  OPF_ERR_CORRECTOR = 4,  /* tim_check_cgce failure (too many its, ...) */
  OPF_ERR_STATE = 5,      /* slot not set up / no state */
  OPF_ERR_NOMEM = 6       /* device memory exhausted / admission refused */
};

/* Message of the last failing call on this thread. */
const char* opf_last_error(void);
int opf_version(void);

/* ---- resolution class -------------------------------------------------- */

/* Create a context: mes_precompute + var_precompute + tra_precompute
 * (meshs.f90:66-103, variables.f90:60-88, transform.fftw3.f90:35-59) for a
 * radial grid of N points (default mesh when r == NULL, else the points of
 * mesh.in) and K axial / M azimuthal modes, with device storage for
 * `capacity` states. */
int opf_ctx_create(opf_ctx** out, int N, int K, int M, int Mp, int capacity, const double* r);
/* opf_ctx_create with fixed launch shapes (no reference counterpart; the
 * autotuner's candidates): `pairs` (|k|, m) pairs per workgroup of the fused
 * step kernel (1, 2 or 4), `tm` m columns per z-transform workgroup, `tz`
 * z-lines per theta workgroup (each at most the default width; tz even,
 * which keeps the pairing of z-lines in the theta transforms).  0 takes,
 * in order, OPF_STEP_GP (pairs only), the row of the tuning table
 * (tuning.tsv beside libopf.so, or the file OPF_TUNING names; OPF_TUNING=none
 * skips it) for this (N, K, M) and capacity, then the built-in choice.
 * opf_ctx_create is this call with 0, 0, 0.  Results do not depend on the
 * shapes. */
int opf_ctx_create_tiles(opf_ctx** out, int N, int K, int M, int Mp, int capacity, const double* r, int pairs,
                         int tm, int tz);
int opf_ctx_destroy(opf_ctx* ctx);
/* H = number of harmonics; bytes = device memory held by the context now
 * (including buffers allocated on first use of opf_vel_shift, opf_steporbit
 * and the variable-viscosity plugins). */
int opf_ctx_dims(const opf_ctx* ctx, int* N, int* K, int* M, int* Mp, int* H, int* capacity,
                 size_t* bytes);
/* Device bytes a context of this shape holds after opf_ctx_create
 * (admission control).  Allocated later, on first use: the shift tables
 * ((2K + M) * 16 bytes per slot), and for opf_slot_nonnewtonian /
 * opf_slot_les 16 physical fields of N*3K*3M doubles, 4 spectral fields and
 * 3 fields of N*3K*M complex values per context, and 4N + 1 words per slot. */
size_t opf_ctx_bytes_needed(int N, int K, int M, int capacity);

/* mes_D of meshs.f90:48-58: r(1:N), intrdr(1:N), dr0(1:5,0:4), dr1(1:5,0:4)
 * (column-major), dr(1)%M and radLap%M in mesh band storage M(9, N)
 * (M(KL+1+n-j, j) = A(n, j), column-major).  Any pointer may be NULL. */
int opf_mesh_get(const opf_ctx* ctx, double* r, double* intrdr, double* dr0, double* dr1,
                 double* dr1band, double* radlap);

/* Host-only mesh precompute (mes_precompute + mes_rdom_init, meshs.f90:66-194)
 * without a context or device, same outputs as opf_mesh_get. */
int opf_mesh_build(int N, const double* r_in, double* r, double* intrdr, double* dr0, double* dr1,
                   double* dr1band, double* radlap);

/* ---- per-state set-up ---------------------------------------------------- */

/* vel_matrices (velocity.f90:120-146) for one slot: tim_lumesh_init of LNp,
 * LDp, LDm, LDz (timestep.f90:67-96), the influence vectors and 4x4
 * inverses of vel_adjPPE(0) (velocity.f90:191-225) and vel_adjustFlux(0)
 * (:160-168), for Re, alpha, dt and the implicitness (d_implicit).  Also
 * sets tim_dt.  Factorisation runs on the device. */
int opf_slot_setup(opf_ctx* ctx, int slot, double Re, double alpha, double dt, double implicit,
                   int const_flux);

/* ---- adaptive time step (d_timestep < 0) ------------------------------------ */

/* The adaptive time step of main.f90:24-29 for one slot: before every
 * predictor of opf_run, vel_maxtstep (velocity.f90:438-481) and
 * tim_new_tstep (timestep.f90:259-278) run on the device and, when tim_dt
 * changes, vel_matrices.  enable = 0 restores the fixed tim_dt.  maxdt and
 * courant are d_maxdt and d_courant (parameters.f90:39, 41; 1e99 and 0.5 in
 * the reference); courant also sets the corrector's fallback
 * tim_corr_dt = tim_dt / (10 courant) (timestep.f90:239). */
int opf_slot_adaptive(opf_ctx* ctx, int slot, int enable, double maxdt, double courant);
/* ---- symmetric subspaces ----------------------------------------------------- */

/* The symmetries vel_imposesym imposes at the start of every step
 * (velocity.f90:61-74): b_mirrorsym, b_shiftrefl, b_shiftrott of
 * parameters.f90:26-28, i.e. var_imposesym modes 1, 2, 3
 * (variables.f90:538-618), applied in that order after mode 0, in opf_run
 * and in opf_vel_imposesym. */
int opf_slot_symmetry(opf_ctx* ctx, int slot, int mirror, int shiftrefl, int shiftrott);

/* ---- forcing plugins (var_null(2) after vel_nonlinear, main.f90:23,38) ------ */

/* The Coriolis force of the Coriolis plugin (Coriolis.f90:54-71, parameters of
 * coriolis.in): N += -1/(4 Ro) Omega^ x (U + u) with Ro = Re Ek0 and Omega^ =
 * (sin(phi) cos(theta), -sin(phi) sin(theta), cos(phi)), evaluated on the
 * physical grid inside the nonlinear products of every evaluation.  Needs
 * Mp = 1.  enable = 0 removes it. */
int opf_slot_coriolis(opf_ctx* ctx, int slot, int enable, double Ek0, double phi);

/* The non-Newtonian plugin (nonnewtonian.f90, params_model.in with model 1,
 * Carreau-Yasuda): after every evaluation the strain rate of u + U, the
 * viscosity nu = mu_inf/mu_0 + (1 - mu_inf/mu_0)(1 + (lam sr)^a)^(-n/a) with
 * lam = Re lam_CY mu_0 / (rho radius^2), and N += [(nu - 1) Lap(u + U) +
 * grad(nu) . 2S] / Re (get_nu, update_nonlin_terms, :93-244); the slot then
 * takes exactly one corrector per step (tim_pcn = 1, :70).  enable = 0
 * removes it. */
int opf_slot_nonnewtonian(opf_ctx* ctx, int slot, int enable, double rho, double radius,
                          double mu_0, double mu_inf, double lam_CY, double n_CY, double a_CY);

/* The LES plugin (LES.f90, params_model.in): eddy viscosity nu_t by
 * Smagorinsky (model 1, filters 1-3, van Driest damping with A_plus) or
 * Vreman (model 2, filters 1-2, damping when A_plus > 0) from the strain of
 * u + U after every evaluation; before each predictor vel_nu = 1.1 (1 + Re
 * max nu_t) when it left 7 % of that value, the matrices refactored on the
 * device (update_nu_r, :264-273); N += (1/Re) nu' Lap(u + U) - 4 nu_t z^ +
 * grad(nu_t) . 2S with nu' = Re nu_t + 1 - vel_nu (:280-320); one corrector
 * per step (tim_pcn = 1).  enable = 0 removes it (and restores vel_nu = 1). */
int opf_slot_les(opf_ctx* ctx, int slot, int enable, int model, double C_s, int filter,
                 double A_plus);

/* tim_dt, tim_corr_dt, tim_cfl_dt, tim_cfl_dir of the slot (the columns of
 * io_write_timestep, io.f90:645-657). */
int opf_slot_timestep(opf_ctx* ctx, int slot, double* dt, double* corr_dt, double* cfl_dt, int* cfl_dir);
/* vel_maxtstep (velocity.f90:438-481) for a host-driven loop: the CFL limit
 * and its direction from the physical fields of the last
 * opf_vel_transform + opf_vel_nonlinear. */
int opf_vel_maxtstep(opf_ctx* ctx, int slot, double* cfl_dt, int* cfl_dir);
/* vel_matrices (velocity.f90:120-146) for a new tim_dt, the slot's other
 * parameters unchanged (main.f90:27-28 after tim_new_tstep). */
int opf_vel_matrices(opf_ctx* ctx, int slot, double dt);

/* Load / read a state (vel_ur, vel_ut, vel_uz and tim_t); the layout of
 * io_load_state after radial interpolation (io.f90:98-165).  put applies
 * vel_imposesym (velocity.f90:61-74) and resets the step counter. */
int opf_put_state(opf_ctx* ctx, int slot, const double* ur_re, const double* ur_im,
                  const double* ut_re, const double* ut_im, const double* uz_re,
                  const double* uz_im, double t);
int opf_get_state(opf_ctx* ctx, int slot, double* ur_re, double* ur_im, double* ut_re,
                  double* ut_im, double* uz_re, double* uz_im, double* t);

/* ---- state files (netCDF classic, io.f90) ---------------------------------
 * The file of io_save_state (io.f90:295-392): global attributes t, Re,
 * alpha; variables r(N), dt, dtcor, dtcfl and Ur, Ut, Uz over (r, H, ReIm)
 * with attributes K, M, Mp.  Reading follows io_load_state / io_load_coll
 * (io.f90:98-232): harmonics the grid does not hold are dropped, harmonics
 * the file does not hold are zero, and a file on other radial points is
 * interpolated by local polynomials through 2*KL+1 points (io_interp_wts,
 * io_interp, io.f90:238-289).  The library needs no netCDF installation. */
/* The file's grid and scalars (N from r; K, M, Mp from Ur); any pointer may
 * be NULL. */
int opf_state_file_info(const char* path, int* N, int* K, int* M, int* Mp, double* t,
                        double* Re, double* alpha, double* dt);
/* io_save_state on host arrays in the coll layout; r = the N radial points. */
int opf_state_file_write(const char* path, int N, int K, int M, int Mp, const double* r,
                         double t, double Re, double alpha, double dt, double dtcor,
                         double dtcfl, const double* ur_re, const double* ur_im,
                         const double* ut_re, const double* ut_im, const double* uz_re,
                         const double* uz_im);
/* io_load_state onto the grid (N, K, M, Mp, r) on the host (r == NULL: the
 * default mesh of mes_precompute).  Scalar pointers may be NULL. */
int opf_state_file_read(const char* path, int N, int K, int M, int Mp, const double* r,
                        double* t, double* Re, double* alpha, double* dt, double* dtcor,
                        double* dtcfl, double* ur_re, double* ur_im, double* ut_re,
                        double* ut_im, double* uz_re, double* uz_im);
/* io_save_state of a slot: its state, tim_t, Re, alpha, tim_dt, tim_corr_dt
 * and tim_cfl_dt. */
int opf_state_save(opf_ctx* ctx, int slot, const char* path);
/* io_load_state into a set-up slot, then opf_put_state.  t < 0 takes tim_t
 * from the file (d_time < 0), else t.  An adaptive slot takes tim_dt from
 * the file (d_timestep < 0) and refactors its matrices; tim_corr_dt and
 * tim_cfl_dt come from the file. */
int opf_state_load(opf_ctx* ctx, int slot, const char* path, double t);

/* ---- the time loop (main.f90:18-46), device resident ------------------------
 * Advance every slot that holds a state by `nsteps` steps.  corr_fixed = 0
 * iterates the corrector to convergence as main.f90:35-41 does
 * (tim_check_cgce, timestep.f90:214-253); corr_fixed = n runs exactly n
 * correctors per step (steporbit, newton.f90:743-787, uses 1).  diag_every
 * > 0 records io_write_energy / io_write_friction quantities at the start of
 * every diag_every-th step (io.f90:76-83 uses 10).  Predictor/corrector
 * decisions are taken on the device; the host only enqueues iterations. */
int opf_run(opf_ctx* ctx, int nsteps, int corr_fixed, int diag_every);
/* Same for a subset of slots (slots[0..nslots)). */
int opf_run_slots(opf_ctx* ctx, const int* slots, int nslots, int nsteps, int corr_fixed,
                  int diag_every);

/* tim_step, tim_t, failure status (0 ok), correctors of the last step,
 * tim_dterr of the last corrector, vel_Pr0. */
int opf_slot_status(opf_ctx* ctx, int slot, int* step, double* t, int* status, int* last_ncorr,
                    double* last_dterr, double* Pr0);
/* Diagnostics records, oldest first, 10 doubles each:
 * step, t, E, Ek0, Em0, Ub (= vel_Pr0 with constant flux), Uc, u_tau, ncorr, dterr
 * (vel_energy.dat / vel_friction.dat, io.f90:517,637). */
int opf_diag_get(opf_ctx* ctx, int slot, double* recs, int maxrec, int* nrec);

/* ---- reference procedures, one call each (Fortran drop-in) --------------- */

int opf_vel_imposesym(opf_ctx* ctx, int slot);      /* vel_imposesym, velocity.f90:61-74 */
int opf_vel_transform(opf_ctx* ctx, int slot);      /* vel_transform, velocity.f90:304-317 */
int opf_vel_nonlinear(opf_ctx* ctx, int slot);      /* vel_nonlinear, velocity.f90:323-335 */
int opf_vel_predictor(opf_ctx* ctx, int slot);      /* vel_predictor, velocity.f90:400-410 */
/* vel_corrector (velocity.f90:416-432); returns sqrt(max|du|^2) as
 * tim_measurecorr computes it (timestep.f90:190-207). */
int opf_vel_corrector(opf_ctx* ctx, int slot, double* dterr);
/* vel_Nr, vel_Nt, vel_Nz after vel_nonlinear, base-flow terms included
 * (vel_addHPF, velocity.f90:338-361); coll layout. */
int opf_get_nonlinear(opf_ctx* ctx, int slot, double* nr_re, double* nr_im, double* nt_re,
                      double* nt_im, double* nz_re, double* nz_im);

/* var_coll_norm (variables.f90:718-751) of component comp (0 u_r, 1 u_th,
 * 2 u_z) of the slot's current state: E, Ek(0:K-1), Em(0:M-1) (NULL
 * skips).  io_write_energy sums E over the three components. */
int opf_var_norm(opf_ctx* ctx, int slot, int comp, double* E, double* Ek, double* Em);
/* var_coll_norm (variables.f90:718-751) of any field a in the coll layout
 * (host arrays re, im), with d_alpha of the slot's parameters: E,
 * Ek(0:K-1), Em(0:M-1) (NULL skips).  Runs on the device. */
int opf_var_coll_norm(opf_ctx* ctx, int slot, const double* a_re, const double* a_im, double* E, double* Ek,
                      double* Em);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): opf_var_coll_norm
 * on device arrays, allocating nothing.  a_re, a_im: device arrays of N*H
 * doubles in the coll layout; out: device array of 1 + K + M doubles that
 * receives E, Ek(0:K-1), Em(0:M-1).  The per-harmonic partials use a static
 * per-device buffer of 65536 harmonics (512 KiB, held until the process
 * ends); H above that is rejected with OPF_ERR_ARG.  The two kernels are
 * enqueued on `stream` (a hipStream_t of the context's device), or, when
 * stream is NULL, on the context's stream, which is then waited for.  All
 * calls on one GPU share the buffer: calls that overlap in time on one GPU
 * (different streams or threads) race. */
int opf_var_coll_norm_noalloc(opf_ctx* ctx, int slot, const double* a_re, const double* a_im, double* out,
                              void* stream);
/* io_write_friction (io.f90:627-639): Ub, Uc, u_tau of the current state. */
int opf_friction(opf_ctx* ctx, int slot, double* Ub, double* Uc, double* Ufr);
/* vel_transform's physical fields (velocity.f90:304-317) of the current
 * state for var_null(2) hooks: which = 0..2 vel_r, vel_t, vel_z; 3..5
 * vel_curlr, vel_curlt, vel_curlz; phys layout. */
int opf_get_phys(opf_ctx* ctx, int slot, int which, double* phys);
/* vel_N* += dN for the current evaluation (forcing of var_null(2) hooks,
 * e.g. Coriolis.f90:54-71); coll layout. */
int opf_add_nonlinear(opf_ctx* ctx, int slot, const double* nr_re, const double* nr_im, const double* nt_re,
                      const double* nt_im, const double* nz_re, const double* nz_im);

/* tra_coll2phys = var_coll2spec + tra_spec2phys (variables.f90:184-197,
 * transform.fftw3.f90:65-102) and tra_phys2coll = tra_phys2spec +
 * var_spec2coll (transform.fftw3.f90:108-147, variables.f90:281-294), on
 * host arrays of one field. */
int opf_tra_coll2phys(opf_ctx* ctx, const double* c_re, const double* c_im, double* phys);
int opf_tra_phys2coll(opf_ctx* ctx, const double* phys, double* c_re, double* c_im);

/* tim_lumesh_init result of a slot (timestep.f90:67-96): which = 0 LDp,
 * 1 LDm, 2 LDz, 3 LNp.  ab: lumesh%M(13, N) per harmonic (13*N*H doubles,
 * column-major per harmonic), ipiv: lumesh%ipiv(N) per harmonic (1-based).
 * Rows 1..KL (fill-in workspace) are returned as the factorisation left them. */
int opf_tim_lumesh_get(opf_ctx* ctx, int slot, int which, double* ab, int* ipiv);
/* tim_lumesh_init (timestep.f90:67-96) into factor set `which` (0..3) of a
 * set-up slot: A = c1 I + c2 Lap for every harmonic (PM = +1 / -1: the
 * modified Laplacian of u+ / u-), the axis row of parity
 * S = (m Mp + |PM|) mod 2 and the wall row dr1(:, BC), factored by DGBTF2
 * (the (0,0) matrix is skipped when BC = 1).  d_alpha is the slot's.  The
 * set replaces what vel_matrices put there; OPF_ERR_SINGULAR is the
 * reference's stop 'tim_lumesh_init'. */
int opf_tim_lumesh_init(opf_ctx* ctx, int slot, int which, int PM, int BC, double c1, double c2);
/* tim_lumesh_invert(BC, A, b) (timestep.f90:102-123) with the slot's factor
 * set `which` as A, on a host coll field (in place); BC = 1 zeroes the
 * (0,0) harmonic, BC < 0 takes the BC the set was built with. */
int opf_tim_lumesh_invert(opf_ctx* ctx, int slot, int which, int BC, double* b_re, double* b_im);

/* ---- Newton-Krylov clients (newton.f90) -------------------------------------
 * var_coll_shift (variables.f90:487-528) of a slot's state: harmonic (k, m)
 * times e^{-i(alpha k dz + m Mp dth)} (dz first, then dth, each rounded as
 * the reference does), the vorticity recomputed for the next evaluation. */
int opf_vel_shift(opf_ctx* ctx, int slot, double dth, double dz);
/* steporbit (newton.f90:721-775) for `nslots` slots at once: for ndts[i] > 1,
 * tim_dt = T[i]/ndts[i] (vel_matrices when it changes), ndts[i] steps of a
 * predictor and exactly one corrector, vel_imposesym, then the state shifted
 * back by the orbit's shifts, var_coll_shift(-sth[i], -sz[i]); for
 * ndts[i] == 1, one such step with the slot's current tim_dt and no shift
 * (the trajectory direction of multJ).  Slots must hold states and use a
 * fixed time step.  Independent segments (multiple shooting, several
 * searches, the perturbed and base integrations of a Jacobian product)
 * advance together in one launch sequence. */
int opf_steporbit(opf_ctx* ctx, int nslots, const int* slots, const int* ndts, const double* T,
                  const double* sth, const double* sz);

/* ---- request server (continuous batching of integrations) ----------------
 * A request integrates one state for `nsteps` steps with its own Re, alpha
 * and dt (the time integration steporbit hands to the stepper,
 * newton.f90:743-787, with corr_fixed = 1 correctors per step there;
 * 0 = to convergence).  No counterpart in the reference (its states are
 * integrated one per process).
 *
 * Memory: all classes (N, K, M of nkm[3i..3i+2]) draw on one device pool of
 * pool_bytes.  A class's slots come in blocks of block_slots[i] (one context
 * and stream each, opf_ctx_bytes_needed(N, K, M, block_slots[i]) bytes),
 * created when a queued request finds no free slot and the pool has room,
 * released when idle and another class needs the room.  A request's
 * footprint is its block's share.
 * Scheduling: a scheduler thread keeps `depth` (1..8) iterations queued on
 * every block with running slots; between iterations it admits queued
 * requests into free slots (they join at the next iteration) and retires
 * finished ones (final state copied behind the queued iterations).  Slots
 * advance independently, so results do not depend on what else is served.
 * Thread-safe. */
typedef struct opf_server opf_server;
const char* opf_server_last_error(void);
int opf_server_create(opf_server** out, int nclass, const int* nkm, const int* block_slots, size_t pool_bytes,
                      int depth);
int opf_server_destroy(opf_server* srv);
int opf_server_submit(opf_server* srv, int cls, double Re, double alpha, double dt, int nsteps, int corr_fixed,
                      const double* ur_re, const double* ur_im, const double* ut_re, const double* ut_im,
                      const double* uz_re, const double* uz_im, long long* id);
/* done = 1 once finished; status 0 = ok; times in seconds since creation. */
int opf_server_poll(opf_server* srv, long long id, int* done, int* status, double* t_submit, double* t_admit,
                    double* t_done);
/* Final state of a finished request (coll planes), then forgets the request. */
int opf_server_result(opf_server* srv, long long id, double* ur_re, double* ur_im, double* ut_re,
                      double* ut_im, double* uz_re, double* uz_im, double* t);
/* queued requests, requests in slots, state-steps of finished requests */
int opf_server_stats(opf_server* srv, int* queued, int* running, long long* state_steps);
/* device pool bytes in use and its limit, blocks (contexts) alive */
int opf_server_pool(opf_server* srv, size_t* used, size_t* limit, int* blocks);

/* ---- timing hooks --------------------------------------------------------- */

/* Record HIP events around every launch of the named kernel on the
 * context's stream ("step", "zfwd", "theta", "zbwd", "curl", "control");
 * 0 disables.  opf_kernel_time returns the summed milliseconds and the
 * launch count since the last reset. */
int opf_kernel_timing(opf_ctx* ctx, int enable);
int opf_kernel_time(opf_ctx* ctx, const char* name, double* ms, int* launches);

/* Launch modes of the device-resident time loop (no reference counterpart).
 * graph = n > 0: n iterations of the loop are captured as one HIP graph and
 * replayed (not with the variable-viscosity plugins or kernel timing on,
 * which launch eagerly; n <= 64); graph = 0: eager launches; the default
 * comes from the tuning table's row of the context, else 0.  compact = 1
 * (default): once at most half of a run's slots are still iterating,
 * iterations launch over the list of those slots only.  streams = 2: a run
 * over four or more slots is split into two halves on two streams, so one
 * half's step kernels overlap the other half's transforms (default from
 * the tuning table, else 2).  A negative value (0 for streams) leaves a
 * mode as it is.  Outputs (any pointer may be NULL): the modes now, and the
 * graph replays, graph captures and compacted iterations since the context
 * was created. */
int opf_launch_mode(opf_ctx* ctx, int graph, int compact, int streams, int* graph_now, int* compact_now,
                    int* streams_now, long long* graph_launches, long long* graph_captures,
                    long long* compact_iterations);

/* Launch shapes in use (no reference counterpart): pairs per workgroup of the
 * step kernel, m columns per z-transform workgroup, z-lines per theta
 * workgroup. */
int opf_tile_config(const opf_ctx* ctx, int* pairs, int* tm, int* tz);

/* Launch shape of the fused step kernel chosen at opf_create: (|k|, m)
 * pairs per workgroup (1, 2 or 4; environment OPF_STEP_GP overrides the
 * default), threads per workgroup and dynamic LDS bytes.  No reference
 * counterpart (introspection for tests and benchmarks). */
int opf_step_config(const opf_ctx* ctx, int* pairs_per_group, int* threads, long long* lds_bytes);

#ifdef __cplusplus
}
#endif
#endif /* OPF_H */
