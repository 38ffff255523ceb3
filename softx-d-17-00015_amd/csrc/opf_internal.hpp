// Internal declarations shared by the host orchestration and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace opf {

constexpr int KL = 4;          // parameters.f90:45 (half band width)
constexpr int LDA = 2 * KL + 1;  // 'mesh' band rows (meshs.f90:37-39)
constexpr int LDAB = 3 * KL + 1; // 'lumesh' rows (meshs.f90:42-45)
constexpr int NLU = 13;        // rows per factor column: 9 of U + 4 multipliers
constexpr int LUS = 16;        // stored doubles per factor column (128-byte aligned)
constexpr int MAXN = 128;      // largest radial resolution supported
constexpr double D_PI = 3.1415926535897931;  // parameters.f90:54

// Host radial mesh: the rdom type of meshs.f90:48-55 for a serial run.
struct HostMesh {
  int N = 0;
  std::vector<double> r, rinv, rinv2, r2;   // r^1, r^-1, r^-2, r^2
  std::vector<double> dr1b, dr2b, lapb;      // bands [row][col], row = KL+n-j
  std::vector<double> dr0, dr1;              // [(KL+1)*(KL+1)], (i, c) -> i + (KL+1)*c
  std::vector<double> intrdr;
};

void build_mesh(int N, const double* r_in, HostMesh& out);  // mes_precompute

// Twiddles and radix plan of one FFT length.
struct FftPlan {
  int L = 0;
  int nrad = 0;
  int rad[16] = {0};
};
FftPlan make_plan(int L);

// Factor matrices per pair, in order of storage.
enum Mat { MAT_LDP = 0, MAT_LDM = 1, MAT_LDZ = 2, MAT_LNP = 3 };

// Per-slot physical parameters (device resident, read by the kernels).
struct SlotParams {
  double Re, alpha, dt, implicit, nu;
  double dterr_tol, courant;
  int const_flux;
  int adaptive;   // d_timestep < 0: vel_maxtstep + tim_new_tstep before every predictor
  double maxdt;   // d_maxdt (parameters.f90:39)
  int symm;       // bits 0, 1, 2: b_mirrorsym, b_shiftrefl, b_shiftrott (parameters.f90:26-28)
  int coriolis;   // 1: the Coriolis force of Coriolis.f90:54-71 is added to N
  double cor_d1;  // -1 / (4 Ro), Ro = Re Ek0
  double cor_sphi, cor_cphi;   // sin(phi), cos(phi): the rotation axis in the (x, z) plane
};

// Slot modes for the next step kernel.
enum { MODE_PRED = 0, MODE_CORR = 1 };

// Per-slot control block (device resident).  The step and control kernels
// advance it; the host reads it back to follow progress.
struct SlotCtl {
  int active;          // slot integrates
  int mode;            // MODE_PRED / MODE_CORR for the next step kernel
  int it;              // tim_it
  int step;            // tim_step
  int ncorr;           // correctors done in the current step
  int ubase, ucur, unext;  // u buffer roles (0..2): u_, current iterate, output
  int ncur, npred;     // Nraw buffer roles (0..1)
  int status;          // 0 ok, else a failure code (see OPF_FAIL_*)
  int target_steps;    // deactivate when step reaches this (<0: never)
  int corr_fixed;      // >0: exactly this many correctors per step (steporbit)
  int diag_every;      // record diagnostics when step % diag_every == 0 (0: never)
  int ndiag;           // diagnostics records written
  int diag_cap;        // capacity of the slot's record array
  int fresh;           // imposesym + curl needed before the next evaluation
  int last_ncorr;      // correctors used by the last completed step
  int host_driven;     // 1: the host runs tim_check_cgce and the clock (drop-in calls)
  int ntc;             // the saved counter i of tim_new_tstep (timestep.f90:261)
  int refactor;        // 1: vel_matrices runs before this step's predictor (tim_new_dt)
  int cfl_dir;         // tim_cfl_dir
  double cfl_dt;       // tim_cfl_dt
  double t, dterr, lasterr, corr_dt, Pr0, Pr0_hpf, last_dterr;
  unsigned long long dterr_bits;  // max |du|^2 of the running corrector (bit pattern)
  double flux_d1;      // vel_adjustFlux(0) d1
};

// One diagnostics record, the content of io_write_energy / io_write_friction.
struct DiagRec {
  double step, t, E, Ek0, Em0, Ub, Uc, Ufr, ncorr, dterr;
};

enum {
  OPF_FAIL_NONE = 0,
  OPF_FAIL_DT0 = 1,        // tim_check_cgce: dt --> 0
  OPF_FAIL_ITS = 2,        // too many its
  OPF_FAIL_INCERR = 3,     // increasing error
  OPF_FAIL_NAN = 4,
};

// Geometry of a resolution class as seen by the kernels (passed by value).
struct Geo {
  int N, K, M, Mp;
  int H;          // harmonics (parallel.h:47-56 order)
  int Hp;         // row stride of a spectral field (H rounded up to 64)
  int P;          // unique (|k|, m) pairs = K*M
  int Z, T;       // physical grid 3K x 3M
  size_t coll;    // doubles2 per component of one slot: N*Hp
  const double* mesh;   // device mesh block, see MeshOff
};

// Offsets (in doubles) of the pieces of the device mesh block.
struct MeshOff {
  int r, rinv, rinv2, Ubase, twor, dr1b, lapb, intrdr, dr0, dr1, total;
  __host__ __device__ static MeshOff make(int N) {
    MeshOff o{};
    int p = 0;
    o.r = p; p += N;
    o.rinv = p; p += N;
    o.rinv2 = p; p += N;
    o.Ubase = p; p += N;
    o.twor = p; p += N;
    o.dr1b = p; p += LDA * N;
    o.lapb = p; p += LDA * N;
    o.intrdr = p; p += N;
    o.dr0 = p; p += (KL + 1) * (KL + 1);
    o.dr1 = p; p += (KL + 1) * (KL + 1);
    o.total = p;
    return o;
  }
};

// DGETF2 + DGETRI (unblocked) on a row-major n x n matrix, in place
// (mes_mat_invert, meshs.f90:231-243).
void mat_invert(std::vector<double>& A, int n);

// State files (io_save_state / io_load_state, io.f90:98-392).  Fields are the
// six coll planes Ur.Re, Ur.Im, Ut.Re, Ut.Im, Uz.Re, Uz.Im, each N x H with r
// fastest.  read_state_file maps the file onto the grid (N, K, M, Mp, r);
// any scalar pointer may be null.  Both throw std::runtime_error.
struct StateInfo {
  int N, K, M, Mp;
  double t, Re, alpha, dt;
};
StateInfo read_state_info(const std::string& path);   // the file's grid (from Ur) and scalars
void write_state_file(const std::string& path, int N, int K, int M, int Mp, double t, double Re, double alpha,
                      double dt, double dtcor, double dtcfl, const double* r, const double* const f[6]);
void read_state_file(const std::string& path, int N, int K, int M, int Mp, const double* r, double* t,
                     double* Re, double* alpha, double* dt, double* dtcor, double* dtcfl, double* const out[6]);

}  // namespace opf
