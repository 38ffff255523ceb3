// Host orchestration and the C ABI (include/opf.h).
//
// One opf_ctx holds a resolution class: the host mesh, FFT plans, and the
// device arrays of `capacity` slots.  The time loop of main.f90:18-46 runs as
// a sequence of "iterations", each one nonlinear evaluation plus one
// predictor or corrector (per slot, as the slot's control block says):
//   k_zfwd -> k_theta -> k_zbwd -> k_step -> k_control
// (k_step ends with the curl for the next evaluation).  The corrector test
// (tim_check_cgce) runs in k_control on the device, so the host only enqueues
// iterations and checks progress between chunks.
#include <hip/hip_runtime.h>
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/opf.h"
#include "kernels.hpp"
#include "engine_internal.hpp"

using namespace opf;

namespace {

thread_local std::string g_err;

struct OpfError : std::runtime_error {
  int code;
  OpfError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess)                                                                \
      throw OpfError(OPF_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));       \
  } while (0)

template <typename F> int guard(F&& f) {
  try {
    f();
    return OPF_OK;
  } catch (const OpfError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return OPF_ERR_ARG;
  }
}

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct Timing {
  bool on = false;
  struct Rec { std::string name; hipEvent_t a, b; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  }
};

}  // namespace

struct opf_ctx {
  int N, K, M, Mp, H, Hp, P, Z, T, cap;
  int diag_cap = 256;
  HostMesh mesh;
  hipStream_t stream = nullptr;
  double* d_mesh = nullptr;
  Bufs B{};
  Geo G{};
  FftTiles tiles{};
  StepTiles st{};
  size_t smem_z = 0, smem_t = 0;
  FftKerns fk{};
  PlaneKern plane = nullptr;   // whole-plane evaluation, where the grid has one
  size_t smem_plane = 0;
  std::vector<void*> allocs;
  size_t bytes = 0;
  std::vector<int> occupied;   // slot holds a state
  std::vector<int> setup;      // slot matrices built
  struct Adapt { int on = 0; double maxdt = 1e99, courant = 0.5; };
  std::vector<Adapt> adapt;    // d_timestep < 0, d_maxdt, d_courant per slot
  int n_adaptive = 0;          // slots with the adaptive time step on
  std::vector<int> symm;       // symmetry bits per slot (SlotParams::symm)
  int n_symm = 0;              // slots with a symmetry imposed
  std::vector<double> alpha_h; // d_alpha per slot (host copy of SlotParams::alpha)
  struct Cor { int on = 0; double Ek0 = 0.0, phi = 0.0; };
  std::vector<Cor> cor;        // coriolis.in per slot (Coriolis.f90:36-39)
  struct NN { int on = 0; double rho = 0, radius = 0, mu0 = 0, muinf = 0, lam = 0, n = 0, a = 0; };
  std::vector<NN> nn;          // params_model.in per slot (nonnewtonian.f90:67-86)
  int n_nn = 0;                // slots with the non-Newtonian viscosity on
  std::vector<double> Re_h;    // d_Re per slot (host copy of SlotParams::Re)
  struct LES { int on = 0, model = 1, filter = 1; double Cs = 0, Aplus = 0; };
  std::vector<LES> les;        // params_model.in per slot (LES.f90:70-95)
  int n_les = 0;               // slots with the LES eddy viscosity on
  double* lesD = nullptr;      // per slot [Delta, Delta_r, L_theta, coef] x N
  unsigned long long* lesMax = nullptr;   // per slot nu_max (bit pattern)
  double* nnP = nullptr;       // non-Newtonian physical fields (16 of N*Z*T), first use
  double2* nnC = nullptr;      // non-Newtonian coll temporaries (4 of coll), first use
  double2* nnY = nullptr;      // their z-transform intermediates (3 fields), first use
  // granular transform temporaries (allocated on first use)
  double2* tY = nullptr;
  double* tP = nullptr;
  double2* tC = nullptr;
  double2* shiftTab = nullptr;   // var_coll_shift tables, (2K + M) per slot (first use)
  double* normBuf = nullptr;     // opf_var_coll_norm: field, partials and sums (first use)
  SlotParams* prmHost = nullptr; // pinned parameters of asynchronous admissions (first use)
  std::vector<int> lu_bc;        // BC of each factor set [slot][4] (tim_lumesh_invert's (0,0) rule)
  int device = 0;                // HIP device of the context
  // launch modes of the time loop: one iteration captured as a HIP graph
  // and replayed; iterations of a run's tail over the list of slots still
  // active (compaction)
  int use_graph = 0, use_compact = 1;   // use_graph: iterations per captured graph (0: eager)
  int use_streams = 2;           // streams a run's slots are split over (1 or 2; table or 2)
  hipStream_t stream2 = nullptr; // the second stream of a split run (first use)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int* smapDev = nullptr;        // compacted slot list on the device [cap]
  std::vector<int> smapHost;     // its host copy (what the device holds)
  struct GraphKey {
    int s0 = -1, ns = 0, symm = 0, adapt = 0, iters = 0;
    const int* smap = nullptr;
    bool operator==(const GraphKey& o) const {
      return s0 == o.s0 && ns == o.ns && symm == o.symm && adapt == o.adapt && iters == o.iters &&
             smap == o.smap;
    }
  } gkey;
  hipGraphExec_t gexec = nullptr;
  long graph_launches = 0, graph_captures = 0, compact_iterations = 0;
  Timing tm;

  template <typename T> T* alloc(size_t n) {
    void* p = nullptr;
    const size_t b = round_up(std::max<size_t>(n, 1) * sizeof(T), 256);
    HIPCHK(hipMalloc(&p, b));
    HIPCHK(hipMemset(p, 0, b));
    allocs.push_back(p);
    bytes += b;
    return (T*)p;
  }
};

namespace {

// -------------------------------------------------------------------------
// shapes
// -------------------------------------------------------------------------
struct Shape {
  int H, Hp, P, Z, T;
  size_t coll;
};
Shape shape_of(int N, int K, int M) {
  Shape s;
  s.H = (2 * K - 1) * M - (K - 1);
  s.Hp = (int)round_up((size_t)s.H, 64);
  s.P = K * M;
  s.Z = 3 * K;
  s.T = 3 * M;
  s.coll = (size_t)N * s.Hp;
  return s;
}

size_t bytes_needed(int N, int K, int M, int cap, int diag_cap) {
  const Shape s = shape_of(N, K, M);
  size_t b = 0;
  const size_t c3 = (size_t)cap * 3 * s.coll * sizeof(double2);
  b += c3 * 7;                                               // U[3], Nraw[2], om, R
  b += (size_t)cap * s.coll * sizeof(double2);               // S
  b += (size_t)cap * 4 * N * LUS * s.P * sizeof(double);     // LU
  b += (size_t)cap * 4 * N * s.P;                            // PV
  b += (size_t)cap * 4 * N * s.P * (FW_Q + BW_Q) * sizeof(double2);  // FW, BW
  b += (size_t)cap * 4 * s.P * 2;                            // META
  b += (size_t)cap * 6 * N * s.P * sizeof(double);           // Uinf
  b += (size_t)cap * 16 * s.H * sizeof(double);              // Ainv
  b += (size_t)cap * 3 * s.Hp * sizeof(double);              // diagb
  b += (size_t)cap * 3 * N * sizeof(unsigned long long);     // cflmax
  b += (size_t)cap * diag_cap * sizeof(DiagRec);
  b += (size_t)cap * N * s.Z * 9 * M * sizeof(double2);      // Yf + Yb
  return b;
}

FftTiles choose_tiles(int Z, int T, int M) {
  FftTiles t{};
  t.pz = make_plan(Z);
  t.pt = make_plan(T);
  // a workgroup transforms at most FFT_TILE complex values: tm columns of
  // length Z, or 3*tz packed lines of length T
  if (Z > FFT_TILE || 3 * T > FFT_TILE) throw OpfError(OPF_ERR_ARG, "transform length too large for the FFT tiles");
  t.tm = std::max(1, std::min(M, tile_m(Z)));
  t.tz = std::max(1, std::min(Z, FFT_TILE / (3 * T)));
  if (t.tz == 1 && Z >= 2 && 2 * T <= THETA_ITEMS) t.tz = 2;   // long theta lines
  return t;
}

// Tuned launch shapes.  The autotuner (opf/autotune.py) times the fused step
// kernel's pairs per workgroup and the transforms' tile widths per (N, K, M,
// capacity) on the device and writes the fastest into a table, tuning.tsv
// beside libopf.so, or the file OPF_TUNING names (OPF_TUNING=none: no table).
// Rows: N K M min_capacity pairs tm tz graph [measured figures]; '#' starts
// a comment; graph = iterations per captured HIP graph (0: eager launches).
// A context takes the row of its (N, K, M) with the largest min_capacity not
// above its capacity.
struct Tuned {
  int gp = 0, tm = 0, tz = 0, graph = 0, streams = 0;
};
std::string tuning_path() {
  if (const char* e = std::getenv("OPF_TUNING")) return std::string(e);
  Dl_info info;
  if (dladdr((void*)&tuning_path, &info) && info.dli_fname) {
    const std::string lib(info.dli_fname);
    const size_t k = lib.rfind('/');
    return (k == std::string::npos ? std::string(".") : lib.substr(0, k)) + "/tuning.tsv";
  }
  return std::string();
}
Tuned lookup_tuning(int N, int K, int M, int cap) {
  Tuned t;
  const std::string path = tuning_path();
  if (path.empty() || path == "none") return t;
  std::ifstream f(path);
  std::string line;
  int best = -1;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    int n, k, m, cmin, gp, tm, tz, graph;
    if (!(is >> n >> k >> m >> cmin >> gp >> tm >> tz >> graph)) continue;
    // an optional ninth integer column: streams (a token with a '.' is the
    // first timing column of a row without it)
    std::string tok;
    int streams = 0;
    if (is >> tok && tok.find('.') == std::string::npos) streams = std::atoi(tok.c_str());
    if (n == N && k == K && m == M && cmin <= cap && cmin > best) {
      best = cmin;
      t.gp = gp;
      t.tm = tm;
      t.tz = tz;
      t.graph = std::max(0, std::min(64, graph));
      t.streams = streams == 1 || streams == 2 ? streams : 0;
    }
  }
  return t;
}

size_t step_smem(int N, int gp, int bufd2) {
  const int GH = 2 * gp;
  size_t b = (size_t)4 * N * GH * sizeof(double2);           // V[4]
  b += std::max(((size_t)2 * bufd2 + FC_TAIL) * sizeof(double2), (size_t)3 * N * GH * sizeof(double));
  b += (size_t)GH * 8 * sizeof(double) + 8 * sizeof(double);  // AC, RED
  b += 4 * sizeof(int);                                       // META
  return round_up(b, 16);
}

// Pairs per workgroup of the fused step kernel and its factor ring.  A
// chunk block must hold one window of the widest U sweep (9 columns of 5
// double2) plus its pad; 48 double2 per Helmholtz block gives 15-column
// chunks for the common band widths.  Four pairs per workgroup (N <= 96)
// put 48 of the solver wave's 64 lanes to work in the Helmholtz solves
// instead of 24 and halve the solve instructions per harmonic; they are the
// default when the slots give at least three such workgroups per CU (51 KiB
// of LDS each at N = 64), otherwise two pairs (26 KiB, six per CU).  Above
// N = 80 four pairs need more than 64 KiB of LDS (69 KiB at N = 96, two
// workgroups per CU), which the kernel requests explicitly; at
// (96,256,128) they take the step kernel from 1.55 to 1.35 ms.
StepTiles choose_step_tiles(int N, int P, int cap, int ncu, int gp_req, size_t lds_max) {
  StepTiles t{};
  int gp = gp_req;
  if (gp <= 0)
    gp = (N <= 96 && step_smem(N, 4, step_bufd2(4)) <= lds_max && (long long)((P + 3) / 4) * cap >= 3LL * ncu)
             ? 4 : 2;
  if (gp != 1 && gp != 2 && gp != 4) throw OpfError(OPF_ERR_ARG, "step pairs per workgroup must be 1, 2 or 4");
  if (gp == 4 && N > 96) throw OpfError(OPF_ERR_ARG, "four pairs per workgroup need N <= 96");
  t.gp = gp;
  t.bufd2 = step_bufd2(gp);
  t.ngroups = (P + gp - 1) / gp;
  t.smem = step_smem(N, gp, t.bufd2);
  if (t.smem > lds_max) throw OpfError(OPF_ERR_ARG, "step kernel tiles do not fit in LDS");
  return t;
}

// the instantiation of the fused step kernel for the context's pairs per
// workgroup and N (items per thread cover N * 2 * pairs)
using StepKern = void (*)(Geo, Bufs, StepTiles, int);
StepKern step_kernel(const opf_ctx* c) {
  if (c->st.gp == 4) return c->N <= 64 ? k_step_g4 : k_step_g4x;
  if (c->st.gp == 1) return k_step_g1;
  return c->N <= 64 ? k_step_g2 : k_step_g2x;
}

void check_slot(opf_ctx* c, int slot) {
  if (!c) throw OpfError(OPF_ERR_ARG, "null context");
  if (slot < 0 || slot >= c->cap) throw OpfError(OPF_ERR_ARG, "slot out of range");
}

std::vector<SlotCtl> read_ctl(opf_ctx* c) {
  std::vector<SlotCtl> v(c->cap);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(v.data(), c->B.ctl, sizeof(SlotCtl) * c->cap, hipMemcpyDeviceToHost));
  return v;
}
void write_ctl(opf_ctx* c, const std::vector<SlotCtl>& v) {
  HIPCHK(hipMemcpy(c->B.ctl, v.data(), sizeof(SlotCtl) * c->cap, hipMemcpyHostToDevice));
}

// Launch helper with optional event timing.
template <typename Kern, typename... Args>
void launch(opf_ctx* c, const char* name, Kern k, dim3 grid, dim3 block, size_t smem, Args... args) {
  hipEvent_t a = nullptr, b = nullptr;
  if (c->tm.on) {
    a = c->tm.get();
    b = c->tm.get();
    HIPCHK(hipEventRecord(a, c->stream));
  }
  hipLaunchKernelGGL(k, grid, block, smem, c->stream, args...);
  HIPCHK(hipGetLastError());
  if (c->tm.on) {
    HIPCHK(hipEventRecord(b, c->stream));
    c->tm.recs.push_back({name, a, b});
  }
}

dim3 grid_h(opf_ctx* c, int nslots) { return dim3((c->H + 63) / 64, 1, nslots); }
// k_curl's blocks: (64-harmonic block, n) tiles rounded up to 8 XCDs
dim3 grid_hn(opf_ctx* c, int nslots) {
  const int ntile = ((c->H + 63) / 64) * c->N;
  return dim3(8 * ((ntile + 7) / 8), 1, nslots);
}
// blocks of k_zfwd_t / k_zbwd_t: their (m tile, n, field) tiles, 6 or 3
// fields, rounded up to 8 XCDs
unsigned ztile_blocks(const opf_ctx* c, int nfield) {
  const int gx = (c->M + c->tiles.tm - 1) / c->tiles.tm;
  return (unsigned)(8 * ((gx * nfield * c->N + 7) / 8));
}

// One evaluation (vel_transform + vel_nonlinear without the base-flow
// terms) for slots [s0, s0+ns).
void launch_eval(opf_ctx* c, int s0, int ns) {
  const FftTiles& t = c->tiles;
  if (c->plane) {   // one workgroup per radial plane, the intermediates in LDS
    launch(c, "plane", c->plane, dim3(c->N, 1, ns), dim3(PLANE_THREADS), c->smem_plane, c->G, c->B, t, s0);
    return;
  }
  launch(c, "zfwd", c->fk.zfwd, dim3(ztile_blocks(c, 6), 1, ns), dim3(FFT_THREADS), c->smem_z, c->G, c->B, t,
         s0);
  launch(c, "theta", c->fk.theta, dim3((c->Z + t.tz - 1) / t.tz, c->N, ns), dim3(FFT_THREADS), c->smem_t,
         c->G, c->B, t, s0);
  launch(c, "zbwd", c->fk.zbwd, dim3(ztile_blocks(c, 3), 1, ns), dim3(FFT_THREADS), c->smem_z,
         c->G, c->B, t, s0);
}

void launch_step(opf_ctx* c, int s0, int ns) {
  const int per = (c->st.ngroups + 7) / 8;   // XCD-aware order inside the kernel
  launch(c, "step", step_kernel(c), dim3(8 * per, 1, ns), dim3(step_threads(c->st.gp)), c->st.smem, c->G, c->B,
         c->st, s0);
  launch(c, "control", k_control, dim3(1, 1, ns), dim3(256), 0, c->G, c->B, s0);
}

// vel_matrices for slots s0 .. s0+ns-1 (gated: only slots whose control
// block asks for it, the tim_new_dt of this step)
void launch_matrices(opf_ctx* c, int s0, int ns, int gated) {
  launch(c, "factor", k_lu_factor, dim3((c->P + 15) / 16, 4, ns), dim3(64), 0, c->G, c->B, s0, gated,
         LuSpec{-1, 0, 0, 0.0, 0.0});
  launch(c, "influence", k_influence, dim3((c->P + 63) / 64, 5, ns), dim3(64), 0, c->G, c->B, s0, gated);
  launch(c, "ainv", k_ainv, grid_h(c, ns), dim3(64), 0, c->G, c->B, s0, gated);
}

// One iteration.  With adaptive slots, the predictor's evaluation is
// followed by vel_maxtstep + tim_new_tstep and, where dt changed,
// vel_matrices (main.f90:24-29), all decided on the device.
void launch_nonnewtonian(opf_ctx* c, int slot);
void launch_les(opf_ctx* c, int slot);

// slots [s0, s0 + ns) of the launch, or, in a compacted launch, entries
// s0 .. s0 + ns - 1 of the host copy of the slot list
void launch_iteration(opf_ctx* c, int s0, int ns) {
  if (c->n_symm > 0) {   // vel_imposesym at the start of a step (main.f90:21), then the new curl
    launch(c, "symmetry", k_symmetry, dim3((3 * c->N * c->H + 255) / 256, 1, ns), dim3(256), 0, c->G, c->B, s0, 1);
    launch(c, "curl", k_curl, grid_hn(c, ns), dim3(64), 0, c->G, c->B, s0, 1);
  }
  launch_eval(c, s0, ns);
  if (c->n_nn > 0 || c->n_les > 0)   // var_null(2) after vel_nonlinear (main.f90:23,38)
    for (int i = s0; i < s0 + ns; ++i) {
      const int s = c->B.smap ? c->smapHost[i] : i;
      if (c->nn[s].on) launch_nonnewtonian(c, s);
      if (c->les[s].on) launch_les(c, s);
    }
  if (c->n_adaptive > 0) {
    launch(c, "newdt", k_newdt, dim3(1, 1, ns), dim3(64), 0, c->G, c->B, s0, 0);
    launch_matrices(c, s0, ns, 1);
  }
  launch_step(c, s0, ns);
}

// coll (reference layout) <-> device rows
void coll_to_dev(opf_ctx* c, const double* re, const double* im, double2* dst) {
  std::vector<double2> h((size_t)c->N * c->Hp, make_double2(0.0, 0.0));
  for (int hh = 0; hh < c->H; ++hh)
    for (int n = 0; n < c->N; ++n)
      h[(size_t)n * c->Hp + hh] = make_double2(re[(size_t)hh * c->N + n], im ? im[(size_t)hh * c->N + n] : 0.0);
  HIPCHK(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}
void dev_to_coll(opf_ctx* c, const double2* src, double* re, double* im) {
  std::vector<double2> h((size_t)c->N * c->Hp);
  HIPCHK(hipMemcpyAsync(h.data(), src, h.size() * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int hh = 0; hh < c->H; ++hh)
    for (int n = 0; n < c->N; ++n) {
      const double2 v = h[(size_t)n * c->Hp + hh];
      if (re) re[(size_t)hh * c->N + n] = v.x;
      if (im) im[(size_t)hh * c->N + n] = v.y;
    }
}

void ensure_temps(opf_ctx* c) {
  if (c->tY) return;
  c->tY = c->alloc<double2>((size_t)c->N * c->Z * c->M);
  c->tP = c->alloc<double>((size_t)c->N * c->Z * c->T);
  c->tC = c->alloc<double2>(c->G.coll);
}

// Transforms of nf coll fields (stride coll) to / from physical fields
// (stride N*Z*T) on device buffers for the plugins (tra_spec2phys /
// tra_phys2spec, transform.fftw3.f90:65-147), one launch per stage with the
// field as grid z; physical fields in the plugins' layout [n][z][theta].
int lines_per_block(const opf_ctx* c) { return std::max(1, FFT_TILE / padded_len(c->T)); }
void coll2phys_dev(opf_ctx* c, const double2* src, double* dst, int nf) {
  const FftTiles& t = c->tiles;
  const int BT = lines_per_block(c);
  const size_t fy = (size_t)c->N * c->Z * c->M, fp = (size_t)c->N * c->Z * c->T;
  launch(c, "zfwd1", k_zfwd_one, dim3((c->M + t.tm - 1) / t.tm, c->N, nf), dim3(FFT_THREADS), c->smem_z, c->G,
         c->B, t, src, c->nnY, (size_t)c->G.coll, fy);
  launch(c, "c2r_lines", k_c2r_lines, dim3((c->Z + 2 * BT - 1) / (2 * BT), c->N, nf), dim3(FFT_THREADS),
         ((size_t)BT * padded_len(c->T) + c->T) * sizeof(double2), c->G, c->B, t, BT, (const double2*)c->nnY, dst,
         fy, fp);
}
void phys2coll_dev(opf_ctx* c, const double* src, double2* dst, int nf) {
  const FftTiles& t = c->tiles;
  const int BT = lines_per_block(c);
  const size_t fy = (size_t)c->N * c->Z * c->M, fp = (size_t)c->N * c->Z * c->T;
  HIPCHK(hipMemsetAsync(dst, 0, sizeof(double2) * c->G.coll * nf, c->stream));
  launch(c, "r2c_lines", k_r2c_lines, dim3((c->Z + 2 * BT - 1) / (2 * BT), c->N, nf), dim3(FFT_THREADS),
         ((size_t)BT * padded_len(c->T) + c->T) * sizeof(double2), c->G, c->B, t, BT, src, c->nnY, fp, fy);
  launch(c, "zbwd1", k_zbwd_one, dim3((c->M + t.tm - 1) / t.tm, c->N, nf), dim3(FFT_THREADS), c->smem_z, c->G,
         c->B, t, (const double2*)c->nnY, dst, fy, (size_t)c->G.coll);
}

// The variable-viscosity plugins share their transforms: the gradient
// tensor of u + U on the physical grid (into P.S), then, once nu is known on
// the grid, grad(nu) (P.g) and curl curl (u + U) (P.q), and finally the
// forcing in P.q added to the slot's raw N.
NnPhys visc_fields(opf_ctx* c) {
  ensure_temps(c);
  const size_t np = (size_t)c->N * c->Z * c->T;
  if (!c->nnP) c->nnP = c->alloc<double>(16 * np);
  if (!c->nnC) c->nnC = c->alloc<double2>(4 * c->G.coll);
  if (!c->nnY) c->nnY = c->alloc<double2>(3 * (size_t)c->N * c->Z * c->M);
  NnPhys P{};
  for (int i = 0; i < 9; ++i) P.S[i] = c->nnP + (size_t)i * np;
  P.nu = c->nnP + 9 * np;
  for (int i = 0; i < 3; ++i) P.g[i] = c->nnP + (size_t)(10 + i) * np;
  for (int i = 0; i < 3; ++i) P.q[i] = c->nnP + (size_t)(13 + i) * np;
  return P;
}
void visc_gradient(opf_ctx* c, int slot, const NnPhys& P) {
  double2* c3 = c->nnC;
  for (int row = 0; row < 3; ++row) {          // A(row, j), velocity-gradient rows
    launch(c, "nn_grad", k_nn_grad, dim3((c->H + 63) / 64, c->N), dim3(64), 0, c->G, c->B, slot, row, c3);
    coll2phys_dev(c, c3, P.S[row * 3], 3);
  }
}
void visc_derivatives(opf_ctx* c, int slot, const NnPhys& P) {
  double2* c3 = c->nnC;
  double2* nuc = c->nnC + 3 * c->G.coll;
  const dim3 gh((c->H + 63) / 64, c->N), b64(64);
  phys2coll_dev(c, P.nu, nuc, 1);
  launch(c, "nn_gradp", k_nn_gradp, gh, b64, 0, c->G, c->B, slot, (const double2*)nuc, c3);
  coll2phys_dev(c, c3, P.g[0], 3);
  launch(c, "nn_curlcurl", k_nn_curlcurl, gh, b64, 0, c->G, c->B, slot, c3);
  coll2phys_dev(c, c3, P.q[0], 3);
}
void visc_add(opf_ctx* c, int slot, const NnPhys& P) {
  double2* c3 = c->nnC;
  phys2coll_dev(c, P.q[0], c3, 3);
  for (int j = 0; j < 3; ++j)
    launch(c, "nn_addN", k_nn_addN, dim3((c->H + 63) / 64, c->N, 1), dim3(64), 0, c->G, c->B, slot, j,
           (const double2*)(c3 + (size_t)j * c->G.coll));
}

// util(2) of the non-Newtonian plugin for one slot: get_nu and
// update_nonlin_terms (nonnewtonian.f90:93-244) on the slot's current
// evaluation, vel_nu = 1.
void launch_nonnewtonian(opf_ctx* c, int slot) {
  const NnPhys P = visc_fields(c);
  const size_t np = (size_t)c->N * c->Z * c->T;
  visc_gradient(c, slot, P);
  const opf_ctx::NN& q = c->nn[slot];
  const double lam = c->Re_h[slot] * q.lam * q.mu0 / (q.rho * (q.radius * q.radius));   // lam_ (:77)
  const dim3 gp((unsigned)((np + 255) / 256)), b256(256);
  launch(c, "nn_visc", k_nn_visc, gp, b256, 0, P, np, lam, q.muinf / q.mu0, -q.n / q.a, q.a);
  visc_derivatives(c, slot, P);
  launch(c, "nn_force", k_nn_force, gp, b256, 0, P, np, 1.0, c->Re_h[slot]);
  visc_add(c, slot, P);
}

void launch_matrices(opf_ctx* c, int s0, int ns, int gated);

// util(2) of the LES plugin for one slot (LES.f90:100-320): get_nu with the
// slot's eddy-viscosity model, update_nu_r before a predictor (vel_matrices
// on the device where vel_nu changed), update_nonlin_terms.
void launch_les(opf_ctx* c, int slot) {
  const NnPhys P = visc_fields(c);
  const size_t np = (size_t)c->N * c->Z * c->T;
  if (!c->lesD) c->lesD = c->alloc<double>((size_t)c->cap * 4 * c->N);
  if (!c->lesMax) c->lesMax = c->alloc<unsigned long long>(c->cap);
  double* D = c->lesD + (size_t)slot * 4 * c->N;
  const opf_ctx::LES& q = c->les[slot];
  visc_gradient(c, slot, P);
  launch(c, "les_prep", k_les_prep, dim3(1), dim3(64), 0, c->G, c->B, slot, q.model, q.Cs, q.Aplus, D, c->lesMax);
  const double dz = (2.0 * D_PI / (c->alpha_h[slot] * (double)c->Z)) * (3.0 / 2.0);
  const dim3 gp((unsigned)((np + 255) / 256)), b256(256);
  launch(c, "les_visc", k_les_visc, gp, b256, 0, P, np, c->N, c->Z * c->T, q.model, q.Cs, dz, (const double*)D,
         c->lesMax + slot);
  launch(c, "les_nur", k_les_nur, dim3(1), dim3(1), 0, c->B, slot, (const unsigned long long*)c->lesMax);
  launch_matrices(c, slot, 1, 1);
  visc_derivatives(c, slot, P);
  launch(c, "les_force", k_les_force, gp, b256, 0, P, np, c->B, slot);
  visc_add(c, slot, P);
}

// The filter lengths of init_params_model and Smagorinsky (LES.f90:81-89,
// 199-206) for the slot's alpha, uploaded once.
void les_geometry(opf_ctx* c, int slot) {
  if (!c->lesD) c->lesD = c->alloc<double>((size_t)c->cap * 4 * c->N);
  const int N = c->N;
  const std::vector<double>& r = c->mesh.r;
  const opf_ctx::LES& q = c->les[slot];
  const double Rdth = (2.0 * D_PI / (double)(c->Mp * c->T)) * (3.0 / 2.0);
  const double dz = (2.0 * D_PI / (c->alpha_h[slot] * (double)c->Z)) * (3.0 / 2.0);
  std::vector<double> h((size_t)4 * N, 0.0);
  double* Delta = h.data();
  double* Dr = h.data() + N;
  double* Lt = h.data() + 2 * N;
  Dr[0] = (r[1] - 0.0) / 2.0;
  for (int n = 1; n < N - 1; ++n) Dr[n] = (r[n + 1] - r[n - 1]) / 2.0;
  Dr[N - 1] = 1.0 - r[N - 2];
  for (int n = 0; n < N; ++n) {
    if (q.filter == 1) Delta[n] = std::pow(Dr[n] * Rdth * dz, 1.0 / 3.0);
    else if (q.filter == 2) Delta[n] = std::pow(Dr[n] * r[n] * Rdth * dz, 1.0 / 3.0);
    else Delta[n] = std::sqrt((Dr[n] * Dr[n] + (r[n] * Rdth) * (r[n] * Rdth) + dz * dz) / 3.0);
    Lt[n] = q.filter == 2 ? r[n] * Rdth : Rdth;   // Vreman's azimuthal length (:229-230)
  }
  HIPCHK(hipMemcpy(c->lesD + (size_t)slot * 4 * N, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
}

std::vector<double2> twiddles(int L) {
  std::vector<double2> tw(L);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int q = 0; q < L; ++q) {
    const long double a = 2.0L * pi * (long double)q / (long double)L;
    tw[q] = make_double2((double)cosl(a), (double)sinl(a));
  }
  return tw;
}

// gp_req, tm_req, tz_req > 0 fix the launch shapes (the autotuner's
// candidates); 0 takes OPF_STEP_GP, then the tuning table, then the defaults.
void create(opf_ctx* c, int N, int K, int M, int Mp, int cap, const double* r, int gp_req, int tm_req,
            int tz_req) {
  if (N < 9 || N > MAXN) throw OpfError(OPF_ERR_ARG, "N must be in [9, 128]");
  if (K < 2 || M < 1 || Mp < 1 || cap < 1) throw OpfError(OPF_ERR_ARG, "bad K, M, Mp or capacity");
  if (M > 256) throw OpfError(OPF_ERR_ARG, "M must be <= 256");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw OpfError(OPF_ERR_HIP, "no HIP device");
  c->N = N; c->K = K; c->M = M; c->Mp = Mp; c->cap = cap;
  const Shape s = shape_of(N, K, M);
  c->H = s.H; c->Hp = s.Hp; c->P = s.P; c->Z = s.Z; c->T = s.T;
  build_mesh(N, r, c->mesh);
  c->tiles = choose_tiles(c->Z, c->T, M);
  c->fk = fft_kernels(c->tiles.pz, c->tiles.pt);
  const Tuned tuned = lookup_tuning(N, K, M, cap);
  c->use_graph = tuned.graph;
  c->use_streams = tuned.streams > 0 ? tuned.streams : 2;
  {
    // narrower transform tiles than the defaults (the kernels hold at most
    // the default width) from the caller or the table
    const int tm = tm_req > 0 ? tm_req : tuned.tm;
    const int tz = tz_req > 0 ? tz_req : tuned.tz;
    if (tm > 0) {
      if (tm > c->tiles.tm) throw OpfError(OPF_ERR_ARG, "z tile wider than the kernels hold");
      c->tiles.tm = tm;
    }
    if (tz > 0 && tz != c->tiles.tz) {
      // the theta kernel transforms N_z of z-lines 2q and 2q + 1 together; an
      // even tile keeps those partners (and so the rounding) of the default
      if (tz > c->tiles.tz) throw OpfError(OPF_ERR_ARG, "theta tile wider than the kernels hold");
      if (tz % 2 != 0 || c->tiles.tz % 2 != 0) throw OpfError(OPF_ERR_ARG, "theta tile must be even");
      c->tiles.tz = tz;
    }
  }
  {
    const char* e1 = std::getenv("OPF_STEP_GP");
    int dev = 0, ncu = 0, lds = 0;
    HIPCHK(hipGetDevice(&dev));
    c->device = dev;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const int gp = gp_req > 0 ? gp_req : (e1 ? std::atoi(e1) : tuned.gp);
    c->st = choose_step_tiles(N, c->P, cap, ncu, gp, (size_t)lds);
    if (c->st.smem > 64 * 1024)   // dynamic LDS beyond 64 KiB is requested per kernel
      HIPCHK(hipFuncSetAttribute((const void*)step_kernel(c), hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)c->st.smem));
    // the whole-plane evaluation where the grid has one and its LDS fits
    // (OPF_PLANE=0: the three transform kernels instead)
    const char* e2 = std::getenv("OPF_PLANE");
    c->plane = plane_kernel(c->tiles.pz, c->tiles.pt, &c->smem_plane);
    if (c->plane && ((e2 && std::atoi(e2) == 0) || c->smem_plane > (size_t)lds)) c->plane = nullptr;
    if (c->plane && c->smem_plane > 64 * 1024)
      HIPCHK(hipFuncSetAttribute((const void*)c->plane, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)c->smem_plane));
  }
  c->smem_z = ((size_t)c->tiles.tm * padded_len(c->Z) + c->Z) * sizeof(double2);       // data + twiddles
  c->smem_t = ((size_t)3 * c->tiles.tz * padded_len(c->T) + c->T) * sizeof(double2);
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  // device mesh block
  const MeshOff mo = MeshOff::make(N);
  std::vector<double> hm(mo.total, 0.0);
  const HostMesh& me = c->mesh;
  for (int n = 0; n < N; ++n) {
    hm[mo.r + n] = me.r[n];
    hm[mo.rinv + n] = me.rinv[n];
    hm[mo.rinv2 + n] = me.rinv2[n];
    hm[mo.Ubase + n] = 1.0 - me.r2[n];        // vel_U = 1 - r^2 (velocity.f90:53)
    hm[mo.twor + n] = -(-2.0 * me.r[n]);      // -vel_Up, vel_Up = -2r (velocity.f90:54, 342)
    hm[mo.intrdr + n] = me.intrdr[n];
  }
  std::copy(me.dr1b.begin(), me.dr1b.end(), hm.begin() + mo.dr1b);
  std::copy(me.lapb.begin(), me.lapb.end(), hm.begin() + mo.lapb);
  std::copy(me.dr0.begin(), me.dr0.end(), hm.begin() + mo.dr0);
  std::copy(me.dr1.begin(), me.dr1.end(), hm.begin() + mo.dr1);
  c->d_mesh = c->alloc<double>(mo.total);
  HIPCHK(hipMemcpy(c->d_mesh, hm.data(), hm.size() * sizeof(double), hipMemcpyHostToDevice));

  Bufs& B = c->B;
  const size_t c3 = (size_t)cap * 3 * s.coll;
  for (int i = 0; i < 3; ++i) B.U[i] = c->alloc<double2>(c3);
  for (int i = 0; i < 2; ++i) B.Nraw[i] = c->alloc<double2>(c3);
  B.om = c->alloc<double2>(c3);
  B.R = c->alloc<double2>(c3);
  B.S = c->alloc<double2>((size_t)cap * s.coll);
  B.LU = c->alloc<double>((size_t)cap * 4 * N * LUS * s.P);
  B.PV = c->alloc<unsigned char>((size_t)cap * 4 * N * s.P);
  B.FW = c->alloc<double2>((size_t)cap * 4 * N * s.P * FW_Q);
  B.BW = c->alloc<double2>((size_t)cap * 4 * N * s.P * BW_Q);
  B.META = c->alloc<unsigned char>((size_t)cap * 4 * s.P * 2);
  B.Uinf = c->alloc<double>((size_t)cap * 6 * N * s.P);
  B.Ainv = c->alloc<double>((size_t)cap * 16 * s.H);
  B.Ui = c->alloc<double>((size_t)cap * N);
  B.diagb = c->alloc<double>((size_t)cap * 3 * s.Hp);
  B.cflmax = c->alloc<unsigned long long>((size_t)cap * 3 * N);
  B.diag = c->alloc<DiagRec>((size_t)cap * c->diag_cap);
  B.prm = c->alloc<SlotParams>(cap);
  B.ctl = c->alloc<SlotCtl>(cap);
  B.Yf = c->alloc<double2>((size_t)cap * N * s.Z * 6 * M);
  B.Yb = c->alloc<double2>((size_t)cap * N * s.Z * 3 * M);
  B.twZ = c->alloc<double2>(s.Z);
  B.twT = c->alloc<double2>(s.T);
  B.smap = nullptr;
  c->smapDev = c->alloc<int>(cap);
  auto tz = twiddles(s.Z), tt = twiddles(s.T);
  HIPCHK(hipMemcpy(B.twZ, tz.data(), tz.size() * sizeof(double2), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(B.twT, tt.data(), tt.size() * sizeof(double2), hipMemcpyHostToDevice));

  Geo& G = c->G;
  G.N = N; G.K = K; G.M = M; G.Mp = Mp; G.H = s.H; G.Hp = s.Hp; G.P = s.P; G.Z = s.Z; G.T = s.T;
  G.coll = s.coll;
  G.mesh = c->d_mesh;

  std::vector<SlotCtl> ctl(cap);
  for (int i = 0; i < cap; ++i) {
    SlotCtl& k = ctl[i];
    std::memset(&k, 0, sizeof(k));
    k.ubase = 2; k.ucur = 0; k.unext = 1;
    k.ncur = 0; k.npred = 1;
    k.target_steps = -1;
    k.diag_cap = c->diag_cap;
    k.lasterr = 1e99;
  }
  write_ctl(c, ctl);
  c->occupied.assign(cap, 0);
  c->setup.assign(cap, 0);
  c->lu_bc.assign((size_t)cap * 4, 0);
  for (int i = 0; i < cap; ++i) c->lu_bc[(size_t)i * 4 + MAT_LNP] = 1;
  c->adapt.assign(cap, opf_ctx::Adapt{});
  c->symm.assign(cap, 0);
  c->alpha_h.assign(cap, 0.0);
  c->cor.assign(cap, opf_ctx::Cor{});
  c->nn.assign(cap, opf_ctx::NN{});
  c->les.assign(cap, opf_ctx::LES{});
  c->Re_h.assign(cap, 0.0);
}

// The Coriolis parameters of Coriolis.f90:49-50, 58: Ro = Re Ek0, d1 = -1/(4 Ro).
void set_coriolis(SlotParams& p, const opf_ctx::Cor& k) {
  p.coriolis = k.on;
  p.cor_d1 = k.on ? -1.0 / (4.0 * (p.Re * k.Ek0)) : 0.0;
  p.cor_sphi = k.on ? std::sin(k.phi) : 0.0;
  p.cor_cphi = k.on ? std::cos(k.phi) : 0.0;
}

void slot_setup(opf_ctx* c, int slot, double Re, double alpha, double dt, double implicit, int cf) {
  check_slot(c, slot);
  if (!(dt > 0.0) || !(Re > 0.0) || alpha == 0.0) throw OpfError(OPF_ERR_ARG, "Re, dt must be > 0 and alpha != 0");
  SlotParams p{};
  p.Re = Re; p.alpha = alpha; p.dt = dt; p.implicit = implicit; p.nu = 1.0;   // vel_nu (velocity.f90:51)
  if (c->setup[slot] && c->les[slot].on) {   // vel_nu as update_nu_r left it
    HIPCHK(hipStreamSynchronize(c->stream));
    SlotParams old{};
    HIPCHK(hipMemcpy(&old, c->B.prm + slot, sizeof(old), hipMemcpyDeviceToHost));
    p.nu = old.nu;
  }
  p.dterr_tol = 1e-5;   // d_dterr (parameters.f90:40)
  p.courant = c->adapt[slot].courant;   // d_courant (parameters.f90:41)
  p.const_flux = cf;
  p.adaptive = c->adapt[slot].on;        // d_timestep < 0 (parameters.f90:38)
  p.maxdt = c->adapt[slot].maxdt;        // d_maxdt (parameters.f90:39)
  p.symm = c->symm[slot];                // b_mirrorsym, b_shiftrefl, b_shiftrott
  set_coriolis(p, c->cor[slot]);
  HIPCHK(hipMemcpy(c->B.prm + slot, &p, sizeof(p), hipMemcpyHostToDevice));
  c->alpha_h[slot] = alpha;
  c->Re_h[slot] = Re;
  if (c->les[slot].on) les_geometry(c, slot);
  auto ctl = read_ctl(c);
  ctl[slot].status = 0;
  write_ctl(c, ctl);
  launch_matrices(c, slot, 1, 0);
  ctl = read_ctl(c);
  if (ctl[slot].status != 0) {
    const int st = ctl[slot].status;
    ctl[slot].status = 0;
    write_ctl(c, ctl);
    throw OpfError(OPF_ERR_SINGULAR, st < 110 ? "tim_lumesh_init: singular factor"
                                              : (st == 110 ? "vel_adjustFlux: err 2" : "matrix inversion error"));
  }
  c->setup[slot] = 1;
}

void put_state(opf_ctx* c, int slot, const double* const* f, double t) {
  check_slot(c, slot);
  if (!c->setup[slot]) throw OpfError(OPF_ERR_STATE, "slot not set up (opf_slot_setup)");
  auto ctl = read_ctl(c);
  SlotCtl& k = ctl[slot];
  double2* u = c->B.U[k.ucur] + (size_t)slot * 3 * c->G.coll;
  for (int comp = 0; comp < 3; ++comp) coll_to_dev(c, f[2 * comp], f[2 * comp + 1], u + comp * c->G.coll);
  k.step = 0;
  k.t = t;
  k.mode = MODE_PRED;
  k.it = 0;
  k.ncorr = 0;
  k.status = 0;
  k.ndiag = 0;
  k.active = 1;
  k.dterr_bits = 0;
  k.corr_dt = 0.0;   // tim_corr_dt, tim_cfl_dt and tim_new_tstep's counter of a fresh run
  k.cfl_dt = 0.0;
  k.cfl_dir = 0;
  k.ntc = 0;
  k.refactor = 0;
  write_ctl(c, ctl);
  launch(c, "imposesym", k_imposesym, dim3((c->N + 63) / 64, 1, 1), dim3(64), 0, c->G, c->B, slot);
  launch(c, "curl0", k_curl, grid_hn(c, 1), dim3(64), 0, c->G, c->B, slot, 0);
  HIPCHK(hipStreamSynchronize(c->stream));
  ctl = read_ctl(c);
  ctl[slot].active = 0;
  write_ctl(c, ctl);
  c->occupied[slot] = 1;
}

}  // namespace

// ---- the device-resident loop in phases (also used by the server) --------
namespace opf {

RunPlan run_begin(opf_ctx* c, const std::vector<int>& slots, const std::vector<int>& nsteps,
                  const std::vector<int>& corr_fixed, int diag_every, bool strict) {
  if (slots.size() != nsteps.size() || slots.size() != corr_fixed.size())
    throw OpfError(OPF_ERR_ARG, "run plan: size mismatch");
  auto ctl = read_ctl(c);
  RunPlan p;
  int lo = c->cap, hi = -1;
  for (size_t i = 0; i < slots.size(); ++i) {
    const int s = slots[i];
    check_slot(c, s);
    if (!c->occupied[s]) throw OpfError(OPF_ERR_STATE, "slot holds no state");
    if (nsteps[i] < 0) throw OpfError(OPF_ERR_ARG, "nsteps < 0");
    SlotCtl& k = ctl[s];
    if (k.status != 0) {
      if (strict) throw OpfError(OPF_ERR_CORRECTOR, "slot failed earlier");
      continue;
    }
    if (nsteps[i] == 0) continue;
    k.active = 1;
    k.mode = MODE_PRED;
    k.target_steps = k.step + nsteps[i];
    // the variable-viscosity plugins set tim_pcn = 1 (nonnewtonian.f90:70, LES.f90:76)
    k.corr_fixed = (c->nn[s].on || c->les[s].on) ? 1 : corr_fixed[i];
    k.diag_every = diag_every;
    k.host_driven = 0;
    p.slots.push_back(s);
    p.max_steps = std::max(p.max_steps, nsteps[i]);
    lo = std::min(lo, s);
    hi = std::max(hi, s);
  }
  if (hi < 0) return p;
  write_ctl(c, ctl);
  p.lo = lo;
  p.ns = hi - lo + 1;
  return p;
}

// Iterations over [s0, s0 + ns) (of the slot list when c->B.smap is set):
// replays of one captured iteration where the launch sequence allows it
// (no per-slot plugin launches, no event timing), else eager launches.
void enqueue_iterations(opf_ctx* c, int s0, int ns, long iterations) {
  if (iterations <= 0 || ns <= 0) return;
  const int gi = c->use_graph;   // iterations per captured graph (0: eager launches)
  const bool graph = gi > 0 && iterations >= gi && !c->tm.on && c->n_nn == 0 && c->n_les == 0;
  if (!graph) {
    for (long i = 0; i < iterations; ++i) launch_iteration(c, s0, ns);
    return;
  }
  opf_ctx::GraphKey key;
  key.s0 = s0;
  key.ns = ns;
  key.symm = c->n_symm > 0;
  key.adapt = c->n_adaptive > 0;
  key.smap = c->B.smap;
  key.iters = gi;
  if (!c->gexec || !(key == c->gkey)) {
    if (c->gexec) {
      HIPCHK(hipGraphExecDestroy(c->gexec));
      c->gexec = nullptr;
    }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
      for (int i = 0; i < gi; ++i) launch_iteration(c, s0, ns);
    } catch (...) {
      (void)hipStreamEndCapture(c->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    const hipError_t e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      c->gexec = nullptr;
      throw OpfError(OPF_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    c->gkey = key;
    ++c->graph_captures;
  }
  const long reps = iterations / gi;
  for (long i = 0; i < reps; ++i) HIPCHK(hipGraphLaunch(c->gexec, c->stream));
  for (long i = reps * gi; i < iterations; ++i) launch_iteration(c, s0, ns);
  c->graph_launches += reps;
}

// Iterations over two halves of [s0, s0 + ns) on two streams, so that one
// half's step kernels overlap the other half's transforms (the slots are
// independent).  The second stream first waits for the context stream's
// earlier work, and the context stream waits for the second stream's at the
// end, so every later call on the context sees both.
void enqueue_split(opf_ctx* c, int s0, int ns, long iterations) {
  if (!c->stream2) {
    HIPCHK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  }
  const int na = ns / 2;
  HIPCHK(hipEventRecord(c->ev_fork, c->stream));
  HIPCHK(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
  hipStream_t main = c->stream;
  c->stream = c->stream2;
  try {
    for (long i = 0; i < iterations; ++i) launch_iteration(c, s0 + na, ns - na);
  } catch (...) {
    c->stream = main;
    throw;
  }
  c->stream = main;
  for (long i = 0; i < iterations; ++i) launch_iteration(c, s0, na);
  HIPCHK(hipEventRecord(c->ev_join, c->stream2));
  HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
}

void run_enqueue(opf_ctx* c, const RunPlan& p, long iterations) {
  if (p.ns == 0 || iterations <= 0) return;
  const int nl = (int)p.live.size();
  if (c->use_compact && nl > 0 && 2 * nl <= p.ns) {
    // the run's tail: only the listed slots are still active
    if (c->smapHost != p.live) {
      c->smapHost = p.live;
      HIPCHK(hipMemcpyAsync(c->smapDev, c->smapHost.data(), sizeof(int) * nl, hipMemcpyHostToDevice, c->stream));
    }
    c->B.smap = c->smapDev;
    try {
      enqueue_iterations(c, 0, nl, iterations);
    } catch (...) {
      c->B.smap = nullptr;
      throw;
    }
    c->B.smap = nullptr;
    c->compact_iterations += iterations;
    return;
  }
  // two streams for runs over many slots without per-slot plugin launches
  if (c->use_streams > 1 && p.ns >= 4 && !c->tm.on && c->n_nn == 0 && c->n_les == 0) {
    enqueue_split(c, p.lo, p.ns, iterations);
    return;
  }
  enqueue_iterations(c, p.lo, p.ns, iterations);
}

hipStream_t ctx_stream(opf_ctx* c) { return c->stream; }

size_t state_rows(const opf_ctx* c) { return 3 * c->G.coll; }

void coll_to_rows(const opf_ctx* c, const double* const planes[6], double2* rows) {
  for (int comp = 0; comp < 3; ++comp) {
    double2* d = rows + (size_t)comp * c->G.coll;
    const double* re = planes[2 * comp];
    const double* im = planes[2 * comp + 1];
    for (int n = 0; n < c->N; ++n) {
      double2* row = d + (size_t)n * c->Hp;
      for (int hh = 0; hh < c->H; ++hh) row[hh] = make_double2(re[(size_t)hh * c->N + n], im[(size_t)hh * c->N + n]);
      for (int hh = c->H; hh < c->Hp; ++hh) row[hh] = make_double2(0.0, 0.0);
    }
  }
}

void rows_to_coll(const opf_ctx* c, const double2* rows, double* const planes[6]) {
  for (int comp = 0; comp < 3; ++comp) {
    const double2* d = rows + (size_t)comp * c->G.coll;
    for (int hh = 0; hh < c->H; ++hh)
      for (int n = 0; n < c->N; ++n) {
        const double2 v = d[(size_t)n * c->Hp + hh];
        planes[2 * comp][(size_t)hh * c->N + n] = v.x;
        planes[2 * comp + 1][(size_t)hh * c->N + n] = v.y;
      }
  }
}

void admit_async(opf_ctx* c, int slot, double Re, double alpha, double dt, int ucur, const double2* rows,
                 int nsteps, int corr_fixed) {
  check_slot(c, slot);
  if (!(dt > 0.0) || !(Re > 0.0) || alpha == 0.0) throw OpfError(OPF_ERR_ARG, "Re, dt must be > 0 and alpha != 0");
  if (ucur < 0 || ucur > 2 || nsteps < 1 || corr_fixed < 0) throw OpfError(OPF_ERR_ARG, "bad admission");
  if (c->n_nn || c->n_les || c->n_adaptive || c->n_symm) throw OpfError(OPF_ERR_STATE, "plain slots only");
  if (!c->prmHost) HIPCHK(hipHostMalloc((void**)&c->prmHost, sizeof(SlotParams) * c->cap, hipHostMallocDefault));
  SlotParams& p = c->prmHost[slot];
  std::memset(&p, 0, sizeof p);
  p.Re = Re; p.alpha = alpha; p.dt = dt; p.implicit = 0.51; p.nu = 1.0;   // velocity.f90:51
  p.dterr_tol = 1e-5;
  p.courant = 0.5;
  p.const_flux = 1;
  p.maxdt = 1e99;
  c->alpha_h[slot] = alpha;
  c->Re_h[slot] = Re;
  hipLaunchKernelGGL(k_admit, dim3(1), dim3(1), 0, c->stream, c->B, slot, 0, 0, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->B.prm + slot, &p, sizeof p, hipMemcpyHostToDevice, c->stream));
  launch_matrices(c, slot, 1, 0);
  HIPCHK(hipMemcpyAsync(c->B.U[ucur] + (size_t)slot * 3 * c->G.coll, rows, sizeof(double2) * 3 * c->G.coll,
                        hipMemcpyHostToDevice, c->stream));
  launch(c, "imposesym", k_imposesym, dim3((c->N + 63) / 64, 1, 1), dim3(64), 0, c->G, c->B, slot);
  hipLaunchKernelGGL(k_admit, dim3(1), dim3(1), 0, c->stream, c->B, slot, 1, nsteps, corr_fixed);
  HIPCHK(hipGetLastError());
  launch(c, "curl0", k_curl, grid_hn(c, 1), dim3(64), 0, c->G, c->B, slot, 0);
  c->setup[slot] = 1;
  c->occupied[slot] = 1;
}

void enqueue_iteration_range(opf_ctx* c, int s0, int ns) { launch_iteration(c, s0, ns); }

void ctl_snapshot_async(opf_ctx* c, SlotCtl* out) {
  HIPCHK(hipMemcpyAsync(out, c->B.ctl, sizeof(SlotCtl) * c->cap, hipMemcpyDeviceToHost, c->stream));
}

void fetch_async(opf_ctx* c, int slot, int ucur, double2* rows) {
  check_slot(c, slot);
  HIPCHK(hipMemcpyAsync(rows, c->B.U[ucur] + (size_t)slot * 3 * c->G.coll, sizeof(double2) * 3 * c->G.coll,
                        hipMemcpyDeviceToHost, c->stream));
}

int run_remaining(opf_ctx* c, RunPlan& p) {
  p.live.clear();
  if (p.ns == 0) return 0;
  auto ctl = read_ctl(c);
  int remaining = 0;
  for (int s : p.slots)
    if (ctl[s].active) {
      remaining = std::max(remaining, std::max(1, ctl[s].target_steps - ctl[s].step));
      p.live.push_back(s);
    }
  return remaining;
}

}  // namespace opf

namespace {

void run_plan(opf_ctx* c, const std::vector<int>& slots, const std::vector<int>& nsteps,
              const std::vector<int>& corr_fixed, int diag_every) {
  RunPlan p = run_begin(c, slots, nsteps, corr_fixed, diag_every, true);
  if (p.ns == 0) return;
  // Most steps take one corrector; enqueue 2 iterations per step, then top
  // up until every slot has reached its target.  A step takes at most 11
  // iterations (a predictor and tim_check_cgce's 10 correctors), or fails.
  const long budget = 11L * p.max_steps + 2;
  long enqueued = 0;
  long chunk = std::min(budget, 2L * p.max_steps);
  int remaining = 1;
  while (chunk > 0) {
    run_enqueue(c, p, chunk);
    enqueued += chunk;
    remaining = run_remaining(c, p);
    chunk = remaining > 0 ? std::min(budget - enqueued, std::max(2L, 2L * remaining)) : 0;
  }
  auto ctl = read_ctl(c);
  for (int s : slots)
    if (ctl[s].status != 0) {
      char msg[128];
      std::snprintf(msg, sizeof msg, "tim_check_cgce failure %d in slot %d at step %d", ctl[s].status, s, ctl[s].step);
      throw OpfError(OPF_ERR_CORRECTOR, msg);
    }
  if (remaining > 0) {   // cannot happen while k_control keeps its corrector limit; never leave slots running
    for (int s : p.slots) ctl[s].active = 0;
    write_ctl(c, ctl);
    throw OpfError(OPF_ERR_CORRECTOR, "time loop: steps left after the iteration budget");
  }
}

void run_slots(opf_ctx* c, const std::vector<int>& slots, int nsteps, int corr_fixed, int diag_every) {
  if (nsteps < 0) throw OpfError(OPF_ERR_ARG, "nsteps < 0");
  run_plan(c, slots, std::vector<int>(slots.size(), nsteps), std::vector<int>(slots.size(), corr_fixed),
           diag_every);
}

// var_coll_shift tables of slots [s0, s0 + ns) (k_shift's layout); a slot
// not in `sel` gets none.  The host's cos / sin stand for dcos / dsin.
void launch_shift(opf_ctx* c, int s0, int ns, const std::vector<int>& sel, const std::vector<double>& dth,
                  const std::vector<double>& dz) {
  const int per = 2 * c->K + c->M;
  if (!c->shiftTab) c->shiftTab = c->alloc<double2>((size_t)c->cap * per);
  std::vector<double2> tab((size_t)ns * per, make_double2(0.0, 0.0));
  bool any = false;
  for (size_t i = 0; i < sel.size(); ++i) {
    const int j = sel[i] - s0;
    double2* t = &tab[(size_t)j * per];
    t[0] = make_double2(dz[i] != 0.0 ? 1.0 : 0.0, dth[i] != 0.0 ? 1.0 : 0.0);
    any = any || dz[i] != 0.0 || dth[i] != 0.0;
    for (int k = -(c->K - 1); k <= c->K - 1; ++k) {
      const double a = dz[i] * (double)k * c->alpha_h[sel[i]];
      t[k + c->K] = make_double2(std::cos(a), std::sin(a));
    }
    for (int m = 0; m < c->M; ++m) {
      const double a = dth[i] * (double)m * (double)c->Mp;
      t[2 * c->K + m] = make_double2(std::cos(a), std::sin(a));
    }
  }
  if (!any) return;
  HIPCHK(hipStreamSynchronize(c->stream));   // no kernel still reads the previous tables
  HIPCHK(hipMemcpy(c->shiftTab, tab.data(), tab.size() * sizeof(double2), hipMemcpyHostToDevice));
  launch(c, "shift", k_shift, dim3((3 * c->N * c->H + 255) / 256, 1, ns), dim3(256), 0, c->G, c->B, s0,
         (const double2*)c->shiftTab);
}

// The end of steporbit (newton.f90:721-775) for slots `sel`: vel_imposesym
// after the last corrector, var_coll_shift by (dth, dz), and the vorticity
// of the result for the next evaluation.
void orbit_finish(opf_ctx* c, const std::vector<int>& sel, const std::vector<double>& dth,
                  const std::vector<double>& dz) {
  if (sel.empty()) return;
  int lo = c->cap, hi = -1;
  for (int s : sel) { lo = std::min(lo, s); hi = std::max(hi, s); }
  const int ns = hi - lo + 1;
  auto ctl = read_ctl(c);
  for (int s : sel) {
    ctl[s].active = 1;
    ctl[s].target_steps = -1;
  }
  write_ctl(c, ctl);
  launch(c, "symmetry", k_symmetry, dim3((3 * c->N * c->H + 255) / 256, 1, ns), dim3(256), 0, c->G, c->B, lo, 0);
  launch_shift(c, lo, ns, sel, dth, dz);
  launch(c, "curl0", k_curl, grid_hn(c, ns), dim3(64), 0, c->G, c->B, lo, 0);
  HIPCHK(hipStreamSynchronize(c->stream));
  ctl = read_ctl(c);
  for (int s : sel) ctl[s].active = 0;
  write_ctl(c, ctl);
}

}  // namespace

namespace {
// run `fn` with the slot marked active (kernels skip inactive slots)
template <typename F> void with_active(opf_ctx* c, int slot, int mode, F&& fn) {
  check_slot(c, slot);
  if (!c->occupied[slot]) throw OpfError(OPF_ERR_STATE, "slot holds no state");
  auto ctl = read_ctl(c);
  ctl[slot].active = 1;
  ctl[slot].target_steps = -1;
  if (mode >= 0) ctl[slot].mode = mode;
  ctl[slot].diag_every = 0;
  ctl[slot].corr_fixed = 0;
  ctl[slot].host_driven = 1;
  write_ctl(c, ctl);
  fn();
  HIPCHK(hipStreamSynchronize(c->stream));
  ctl = read_ctl(c);
  ctl[slot].active = 0;
  write_ctl(c, ctl);
}
}  // namespace

// =========================================================================
// C ABI
// =========================================================================
extern "C" {

const char* opf_last_error(void) { return g_err.c_str(); }
int opf_version(void) { return 100; }

static double* scratch_on(int dev);

int opf_ctx_create_tiles(opf_ctx** out, int N, int K, int M, int Mp, int capacity, const double* r, int pairs,
                         int tm, int tz) {
  return guard([&] {
    if (!out) throw OpfError(OPF_ERR_ARG, "null out");
    *out = nullptr;
    if (pairs < 0 || tm < 0 || tz < 0) throw OpfError(OPF_ERR_ARG, "negative tile size");
    auto* c = new opf_ctx();
    try {
      create(c, N, K, M, Mp, capacity, r, pairs, tm, tz);
      (void)scratch_on(c->device);   // opf_var_coll_norm_noalloc's buffer on this device
    } catch (...) {
      opf_ctx_destroy(c);
      throw;
    }
    *out = c;
  });
}

int opf_ctx_create(opf_ctx** out, int N, int K, int M, int Mp, int capacity, const double* r) {
  return opf_ctx_create_tiles(out, N, K, M, Mp, capacity, r, 0, 0, 0);
}

int opf_ctx_destroy(opf_ctx* c) {
  if (!c) return OPF_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->prmHost) (void)hipHostFree(c->prmHost);
  for (void* p : c->allocs) (void)hipFree(p);
  for (auto& r : c->tm.recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : c->tm.pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return OPF_OK;
}

int opf_ctx_dims(const opf_ctx* c, int* N, int* K, int* M, int* Mp, int* H, int* cap, size_t* bytes) {
  return guard([&] {
    if (!c) throw OpfError(OPF_ERR_ARG, "null context");
    if (N) *N = c->N;
    if (K) *K = c->K;
    if (M) *M = c->M;
    if (Mp) *Mp = c->Mp;
    if (H) *H = c->H;
    if (cap) *cap = c->cap;
    if (bytes) *bytes = c->bytes;
  });
}

size_t opf_ctx_bytes_needed(int N, int K, int M, int capacity) { return bytes_needed(N, K, M, capacity, 256); }

namespace {
void mesh_out(const HostMesh& me, double* r, double* intrdr, double* dr0, double* dr1, double* dr1band,
              double* radlap) {
  const int N = me.N;
    if (r) std::copy(me.r.begin(), me.r.end(), r);
    if (intrdr) std::copy(me.intrdr.begin(), me.intrdr.end(), intrdr);
    if (dr0) std::copy(me.dr0.begin(), me.dr0.end(), dr0);
    if (dr1) std::copy(me.dr1.begin(), me.dr1.end(), dr1);
    // band [row][col] -> Fortran M(row+1, col+1) column-major: M[row + 9*col]
    for (int row = 0; row < LDA; ++row)
      for (int col = 0; col < N; ++col) {
        if (dr1band) dr1band[row + LDA * col] = me.dr1b[(size_t)row * N + col];
        if (radlap) radlap[row + LDA * col] = me.lapb[(size_t)row * N + col];
      }
}
}  // namespace

extern "C" int opf_mesh_get(const opf_ctx* c, double* r, double* intrdr, double* dr0, double* dr1,
                            double* dr1band, double* radlap) {
  return guard([&] {
    if (!c) throw OpfError(OPF_ERR_ARG, "null context");
    mesh_out(c->mesh, r, intrdr, dr0, dr1, dr1band, radlap);
  });
}

extern "C" int opf_mesh_build(int N, const double* r_in, double* r, double* intrdr, double* dr0, double* dr1,
                              double* dr1band, double* radlap) {
  return guard([&] {
    HostMesh me;
    try {
      build_mesh(N, r_in, me);
    } catch (const std::runtime_error& e) {
      throw OpfError(OPF_ERR_ARG, e.what());
    }
    mesh_out(me, r, intrdr, dr0, dr1, dr1band, radlap);
  });
}

int opf_slot_setup(opf_ctx* c, int slot, double Re, double alpha, double dt, double implicit, int cf) {
  return guard([&] { slot_setup(c, slot, Re, alpha, dt, implicit, cf); });
}

int opf_slot_adaptive(opf_ctx* c, int slot, int enable, double maxdt, double courant) {
  return guard([&] {
    check_slot(c, slot);
    if (enable && !(maxdt > 0.0 && courant > 0.0)) throw OpfError(OPF_ERR_ARG, "maxdt and courant must be > 0");
    opf_ctx::Adapt& a = c->adapt[slot];
    c->n_adaptive += (enable ? 1 : 0) - a.on;
    a.on = enable ? 1 : 0;
    if (enable) {
      a.maxdt = maxdt;
      a.courant = courant;
    }
    if (c->setup[slot]) {   // the slot's parameters on the device follow at once
      SlotParams p{};
      HIPCHK(hipMemcpy(&p, c->B.prm + slot, sizeof(p), hipMemcpyDeviceToHost));
      p.adaptive = a.on;
      p.maxdt = a.maxdt;
      p.courant = a.courant;
      HIPCHK(hipMemcpy(c->B.prm + slot, &p, sizeof(p), hipMemcpyHostToDevice));
    }
  });
}

int opf_slot_symmetry(opf_ctx* c, int slot, int mirror, int shiftrefl, int shiftrott) {
  return guard([&] {
    check_slot(c, slot);
    const int bits = (mirror ? 1 : 0) | (shiftrefl ? 2 : 0) | (shiftrott ? 4 : 0);
    c->n_symm += (bits != 0) - (c->symm[slot] != 0);
    c->symm[slot] = bits;
    if (c->setup[slot]) {
      SlotParams p{};
      HIPCHK(hipMemcpy(&p, c->B.prm + slot, sizeof(p), hipMemcpyDeviceToHost));
      p.symm = bits;
      HIPCHK(hipMemcpy(c->B.prm + slot, &p, sizeof(p), hipMemcpyHostToDevice));
    }
  });
}

int opf_slot_coriolis(opf_ctx* c, int slot, int enable, double Ek0, double phi) {
  return guard([&] {
    check_slot(c, slot);
    if (enable && c->Mp != 1) throw OpfError(OPF_ERR_ARG, "add_coriolis: i_Mp/=1");
    if (enable && !(Ek0 > 0.0)) throw OpfError(OPF_ERR_ARG, "Ek0 must be > 0");
    opf_ctx::Cor& k = c->cor[slot];
    k.on = enable ? 1 : 0;
    k.Ek0 = enable ? Ek0 : 0.0;
    k.phi = enable ? phi : 0.0;
    if (c->setup[slot]) {
      SlotParams p{};
      HIPCHK(hipMemcpy(&p, c->B.prm + slot, sizeof(p), hipMemcpyDeviceToHost));
      set_coriolis(p, k);
      HIPCHK(hipMemcpy(c->B.prm + slot, &p, sizeof(p), hipMemcpyHostToDevice));
    }
  });
}

int opf_slot_les(opf_ctx* c, int slot, int enable, int model, double C_s, int filter, double A_plus) {
  return guard([&] {
    check_slot(c, slot);
    if (enable && model != 1 && model != 2) throw OpfError(OPF_ERR_ARG, "load_params_model: unknown model");
    if (enable && (filter < 1 || filter > (model == 1 ? 3 : 2))) throw OpfError(OPF_ERR_ARG, "bad LES filter");
    if (enable && !(C_s > 0.0)) throw OpfError(OPF_ERR_ARG, "C_s must be > 0");
    opf_ctx::LES& q = c->les[slot];
    c->n_les += (enable ? 1 : 0) - q.on;
    q = opf_ctx::LES{};
    if (enable) q = opf_ctx::LES{1, model, filter, C_s, A_plus};
    if (enable && c->setup[slot]) les_geometry(c, slot);
    if (!enable && c->setup[slot]) {   // vel_nu back to 1 and the matrices with it
      SlotParams p{};
      HIPCHK(hipMemcpy(&p, c->B.prm + slot, sizeof(p), hipMemcpyDeviceToHost));
      if (p.nu != 1.0) slot_setup(c, slot, p.Re, p.alpha, p.dt, p.implicit, p.const_flux);
    }
  });
}

int opf_slot_nonnewtonian(opf_ctx* c, int slot, int enable, double rho, double radius, double mu_0,
                          double mu_inf, double lam_CY, double n_CY, double a_CY) {
  return guard([&] {
    check_slot(c, slot);
    if (enable && !(rho > 0.0 && radius > 0.0 && mu_0 > 0.0 && mu_inf >= 0.0 && lam_CY >= 0.0 && a_CY > 0.0))
      throw OpfError(OPF_ERR_ARG, "bad Carreau-Yasuda parameters");
    opf_ctx::NN& q = c->nn[slot];
    c->n_nn += (enable ? 1 : 0) - q.on;
    q = opf_ctx::NN{};
    if (enable) q = opf_ctx::NN{1, rho, radius, mu_0, mu_inf, lam_CY, n_CY, a_CY};
  });
}

int opf_slot_timestep(opf_ctx* c, int slot, double* dt, double* corr_dt, double* cfl_dt, int* cfl_dir) {
  return guard([&] {
    check_slot(c, slot);
    HIPCHK(hipStreamSynchronize(c->stream));
    SlotParams p{};
    HIPCHK(hipMemcpy(&p, c->B.prm + slot, sizeof(p), hipMemcpyDeviceToHost));
    auto ctl = read_ctl(c);
    if (dt) *dt = p.dt;
    if (corr_dt) *corr_dt = ctl[slot].corr_dt;
    if (cfl_dt) *cfl_dt = ctl[slot].cfl_dt;
    if (cfl_dir) *cfl_dir = ctl[slot].cfl_dir;
  });
}

int opf_vel_maxtstep(opf_ctx* c, int slot, double* cfl_dt, int* cfl_dir) {
  return guard([&] {
    check_slot(c, slot);
    if (!c->occupied[slot]) throw OpfError(OPF_ERR_STATE, "slot holds no state");
    launch(c, "newdt", k_newdt, dim3(1, 1, 1), dim3(64), 0, c->G, c->B, slot, 1);
    HIPCHK(hipStreamSynchronize(c->stream));
    auto ctl = read_ctl(c);
    if (cfl_dt) *cfl_dt = ctl[slot].cfl_dt;
    if (cfl_dir) *cfl_dir = ctl[slot].cfl_dir;
  });
}

int opf_vel_matrices(opf_ctx* c, int slot, double dt) {
  return guard([&] {
    check_slot(c, slot);
    if (!c->setup[slot]) throw OpfError(OPF_ERR_STATE, "slot not set up (opf_slot_setup)");
    if (!(dt > 0.0)) throw OpfError(OPF_ERR_ARG, "dt must be > 0");
    SlotParams p{};
    HIPCHK(hipMemcpy(&p, c->B.prm + slot, sizeof(p), hipMemcpyDeviceToHost));
    slot_setup(c, slot, p.Re, p.alpha, dt, p.implicit, p.const_flux);
  });
}

int opf_put_state(opf_ctx* c, int slot, const double* urr, const double* uri, const double* utr,
                  const double* uti, const double* uzr, const double* uzi, double t) {
  return guard([&] {
    const double* f[6] = {urr, uri, utr, uti, uzr, uzi};
    for (auto p : f)
      if (!p) throw OpfError(OPF_ERR_ARG, "null field");
    put_state(c, slot, f, t);
  });
}

int opf_get_state(opf_ctx* c, int slot, double* urr, double* uri, double* utr, double* uti, double* uzr,
                  double* uzi, double* t) {
  return guard([&] {
    check_slot(c, slot);
    auto ctl = read_ctl(c);
    const double2* u = c->B.U[ctl[slot].ucur] + (size_t)slot * 3 * c->G.coll;
    double* f[6] = {urr, uri, utr, uti, uzr, uzi};
    for (int comp = 0; comp < 3; ++comp) dev_to_coll(c, u + comp * c->G.coll, f[2 * comp], f[2 * comp + 1]);
    if (t) *t = ctl[slot].t;
  });
}

int opf_run(opf_ctx* c, int nsteps, int corr_fixed, int diag_every) {
  return guard([&] {
    if (!c) throw OpfError(OPF_ERR_ARG, "null context");
    std::vector<int> slots;
    for (int s = 0; s < c->cap; ++s)
      if (c->occupied[s]) slots.push_back(s);
    run_slots(c, slots, nsteps, corr_fixed, diag_every);
  });
}

int opf_run_slots(opf_ctx* c, const int* slots, int nslots, int nsteps, int corr_fixed, int diag_every) {
  return guard([&] {
    if (!c || (!slots && nslots > 0)) throw OpfError(OPF_ERR_ARG, "null argument");
    run_slots(c, std::vector<int>(slots, slots + nslots), nsteps, corr_fixed, diag_every);
  });
}

int opf_slot_status(opf_ctx* c, int slot, int* step, double* t, int* status, int* last_ncorr, double* last_dterr,
                    double* Pr0) {
  return guard([&] {
    check_slot(c, slot);
    auto ctl = read_ctl(c);
    const SlotCtl& k = ctl[slot];
    if (step) *step = k.step;
    if (t) *t = k.t;
    if (status) *status = k.status;
    if (last_ncorr) *last_ncorr = k.last_ncorr;
    if (last_dterr) *last_dterr = k.last_dterr;
    if (Pr0) *Pr0 = k.Pr0;
  });
}

int opf_diag_get(opf_ctx* c, int slot, double* recs, int maxrec, int* nrec) {
  return guard([&] {
    check_slot(c, slot);
    auto ctl = read_ctl(c);
    const int nd = ctl[slot].ndiag;
    const int have = std::min(nd, c->diag_cap);
    std::vector<DiagRec> all(c->diag_cap);
    HIPCHK(hipMemcpy(all.data(), c->B.diag + (size_t)slot * c->diag_cap, sizeof(DiagRec) * c->diag_cap,
                     hipMemcpyDeviceToHost));
    const int n = std::min(have, maxrec);
    const int first = nd - have + (have - n);
    for (int i = 0; i < n; ++i) {
      const DiagRec& r = all[(first + i) % c->diag_cap];
      std::memcpy(recs + 10 * i, &r, sizeof(DiagRec));
    }
    if (nrec) *nrec = n;
  });
}

// ---- granular reference procedures -----------------------------------------

int opf_vel_imposesym(opf_ctx* c, int slot) {
  return guard([&] {
    check_slot(c, slot);
    if (!c->occupied[slot]) throw OpfError(OPF_ERR_STATE, "slot holds no state");
    with_active(c, slot, -1, [&] {
      launch(c, "symmetry", k_symmetry, dim3((3 * c->N * c->H + 255) / 256, 1, 1), dim3(256), 0, c->G, c->B, slot,
             0);
      launch(c, "curl", k_curl, grid_hn(c, 1), dim3(64), 0, c->G, c->B, slot, 0);
    });
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}


int opf_vel_transform(opf_ctx* c, int slot) {
  return guard([&] {
    with_active(c, slot, -1, [&] {
      const FftTiles& t = c->tiles;
      launch(c, "zfwd", c->fk.zfwd, dim3(ztile_blocks(c, 6), 1, 1), dim3(FFT_THREADS), c->smem_z, c->G, c->B, t,
             slot);
    });
  });
}

int opf_vel_nonlinear(opf_ctx* c, int slot) {
  return guard([&] {
    with_active(c, slot, -1, [&] {
      const FftTiles& t = c->tiles;
      launch(c, "theta", c->fk.theta, dim3((c->Z + t.tz - 1) / t.tz, c->N, 1), dim3(FFT_THREADS), c->smem_t, c->G,
             c->B, t, slot);
      launch(c, "zbwd", c->fk.zbwd, dim3(ztile_blocks(c, 3), 1, 1), dim3(FFT_THREADS), c->smem_z, c->G,
             c->B, t, slot);
      if (c->nn[slot].on) launch_nonnewtonian(c, slot);   // var_null(2)
      if (c->les[slot].on) launch_les(c, slot);
    });
  });
}

int opf_vel_predictor(opf_ctx* c, int slot) {
  return guard([&] { with_active(c, slot, MODE_PRED, [&] { launch_step(c, slot, 1); }); });
}

int opf_vel_corrector(opf_ctx* c, int slot, double* dterr) {
  return guard([&] {
    with_active(c, slot, MODE_CORR, [&] { launch_step(c, slot, 1); });
    auto ctl = read_ctl(c);
    if (dterr) *dterr = ctl[slot].last_dterr;

  });
}

int opf_get_nonlinear(opf_ctx* c, int slot, double* nrr, double* nri, double* ntr, double* nti, double* nzr,
                      double* nzi) {
  // Nraw of the last evaluation plus the base-flow terms of vel_addHPF,
  // added on the host from the current iterate (velocity.f90:338-361).
  return guard([&] {
    check_slot(c, slot);
    auto ctl = read_ctl(c);
    SlotParams prm;
    HIPCHK(hipMemcpy(&prm, c->B.prm + slot, sizeof prm, hipMemcpyDeviceToHost));
    const int N = c->N, H = c->H;
    std::vector<double> f[12];
    for (auto& v : f) v.assign((size_t)N * H, 0.0);
    const double2* nr = c->B.Nraw[ctl[slot].ncur] + (size_t)slot * 3 * c->G.coll;
    const double2* u = c->B.U[ctl[slot].ucur] + (size_t)slot * 3 * c->G.coll;
    for (int comp = 0; comp < 3; ++comp) {
      dev_to_coll(c, nr + comp * c->G.coll, f[2 * comp].data(), f[2 * comp + 1].data());
      dev_to_coll(c, u + comp * c->G.coll, f[6 + 2 * comp].data(), f[6 + 2 * comp + 1].data());
    }
    const HostMesh& me = c->mesh;
    double Pr0 = 0.0;
    for (int hh = 0; hh < H; ++hh) {
      int k, m;
      if (hh < c->K) { m = 0; k = hh; }
      else { const int q = hh - c->K; m = 1 + q / (2 * c->K - 1); k = q % (2 * c->K - 1) - (c->K - 1); }
      (void)m;
      for (int n = 0; n < N; ++n) {
        const size_t o = (size_t)hh * N + n;
        const double a = (prm.alpha * (double)k) * (1.0 - me.r2[n]);
        const double b = 2.0 * me.r[n];
        f[0][o] = f[0][o] + a * f[7][o];
        f[1][o] = f[1][o] - a * f[6][o];
        f[2][o] = f[2][o] + a * f[9][o];
        f[3][o] = f[3][o] - a * f[8][o];
        f[4][o] = (f[4][o] + a * f[11][o]) + b * f[6][o];
        f[5][o] = (f[5][o] - a * f[10][o]) + b * f[7][o];
      }
    }
    if (prm.const_flux) {
      double s = 0.0;
      for (int i = 0; i <= KL; ++i) s = s + f[10][(size_t)(N - 1 - KL + i)] * me.dr1[i + (KL + 1)];
      Pr0 = s / (-2.0);
      for (int n = 0; n < N; ++n) f[4][n] = f[4][n] + (4.0 * Pr0) / prm.Re;
    }
    double* outs[6] = {nrr, nri, ntr, nti, nzr, nzi};
    for (int i = 0; i < 6; ++i)
      if (outs[i]) std::copy(f[i].begin(), f[i].end(), outs[i]);
  });
}

// the slot's current iterate, one component, on the host (coll planes)
static void slot_field_to_host(opf_ctx* c, int slot, int comp, std::vector<double>& re, std::vector<double>& im) {
  auto ctl = read_ctl(c);
  re.assign((size_t)c->N * c->H, 0.0);
  im.assign((size_t)c->N * c->H, 0.0);
  const double2* u = c->B.U[ctl[slot].ucur] + (size_t)slot * 3 * c->G.coll + (size_t)comp * c->G.coll;
  dev_to_coll(c, u, re.data(), im.data());
}

int opf_var_norm(opf_ctx* c, int slot, int comp, double* E, double* Ek, double* Em) {
  // var_coll_norm (variables.f90:718-751): per harmonic b = (w f) . intrdr,
  // halved for (0,0); Em(m) and Ek(|k|) accumulate in harmonic order;
  // E = sum(Em)
  return guard([&] {
    check_slot(c, slot);
    if (comp < 0 || comp > 2) throw OpfError(OPF_ERR_ARG, "component must be 0, 1 or 2");
    if (!c->occupied[slot]) throw OpfError(OPF_ERR_STATE, "slot holds no state");
    SlotParams prm;
    HIPCHK(hipMemcpy(&prm, c->B.prm + slot, sizeof prm, hipMemcpyDeviceToHost));
    std::vector<double> re, im;
    slot_field_to_host(c, slot, comp, re, im);
    const int N = c->N;
    std::vector<double> ek(c->K, 0.0), em(c->M, 0.0);
    const double w = 4.0 * D_PI * D_PI / prm.alpha;
    for (int hh = 0; hh < c->H; ++hh) {
      int k, m;
      if (hh < c->K) { m = 0; k = hh; }
      else { const int q = hh - c->K; m = 1 + q / (2 * c->K - 1); k = q % (2 * c->K - 1) - (c->K - 1); }
      double b = 0.0;
      for (int n = 0; n < N; ++n) {
        const size_t o = (size_t)hh * N + n;
        const double f = w * (re[o] * re[o] + im[o] * im[o]);
        b = b + f * c->mesh.intrdr[n];
      }
      if (k == 0 && m == 0) b = 0.5 * b;
      em[m] = em[m] + b;
      ek[k < 0 ? -k : k] = ek[k < 0 ? -k : k] + b;
    }
    double e = 0.0;
    for (double v : em) e = e + v;
    if (E) *E = e;
    if (Ek) std::copy(ek.begin(), ek.end(), Ek);
    if (Em) std::copy(em.begin(), em.end(), Em);
  });
}

// var_coll_norm's two passes (k_norm_harm, k_norm_sum) on device arrays:
// the field (re, im in the coll layout), partials b[H], out[1 + K + M].
static void launch_norm(opf_ctx* c, int slot, const double* re, const double* im, double* b, double* out,
                 hipStream_t stream) {
  hipLaunchKernelGGL(k_norm_harm, dim3((c->H + 255) / 256), dim3(256), 0, stream, c->G,
                     (const SlotParams*)(c->B.prm + slot), re, im, b);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_norm_sum, dim3(1), dim3(256), 0, stream, c->G, (const double*)b, out);
  HIPCHK(hipGetLastError());
}

int opf_var_coll_norm(opf_ctx* c, int slot, const double* re, const double* im, double* E, double* Ek,
                      double* Em) {
  // var_coll_norm (variables.f90:718-751) of any coll field, with d_alpha of
  // the slot's parameters, on the device
  return guard([&] {
    check_slot(c, slot);
    if (!re || !im) throw OpfError(OPF_ERR_ARG, "null field");
    if (!c->setup[slot]) throw OpfError(OPF_ERR_STATE, "slot not set up (alpha unknown)");
    const size_t nh = (size_t)c->N * c->H, nout = 1 + (size_t)c->K + c->M;
    if (!c->normBuf) c->normBuf = c->alloc<double>(2 * nh + c->H + nout);
    double* dre = c->normBuf;
    double* dim = dre + nh;
    double* b = dim + nh;
    double* out = b + c->H;
    HIPCHK(hipMemcpyAsync(dre, re, nh * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dim, im, nh * sizeof(double), hipMemcpyHostToDevice, c->stream));
    launch_norm(c, slot, dre, dim, b, out, c->stream);
    std::vector<double> h(nout);
    HIPCHK(hipMemcpyAsync(h.data(), out, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (E) *E = h[0];
    if (Ek) std::copy(h.begin() + 1, h.begin() + 1 + c->K, Ek);
    if (Em) std::copy(h.begin() + 1 + c->K, h.end(), Em);
  });
}

// ---- var_coll_norm without allocation --------------------------------------
// The per-harmonic partials of opf_var_coll_norm_noalloc live in a file-scope
// device array, one copy per GPU, held until the process ends.  Host code
// never passes `scratch` itself to a kernel: its device address, looked up
// once per device with hipGetSymbolAddress (at context creation), is kept in
// g_scratch and used instead.
constexpr int NORM_SCRATCH_MAX = 65536;   // harmonics (512 KiB of doubles)
static __device__ __align__(256) double scratch[NORM_SCRATCH_MAX];

static std::mutex g_scratch_mu;
static std::map<int, double*> g_scratch;   // device -> address of its scratch

// Address of `scratch` on device `dev` (looked up with dev current).
static double* scratch_on(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) return it->second;
  int cur = 0;
  HIPCHK(hipGetDevice(&cur));
  if (cur != dev) HIPCHK(hipSetDevice(dev));
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (cur != dev) (void)hipSetDevice(cur);
  if (e != hipSuccess || !ptr)
    throw OpfError(OPF_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[dev] = (double*)ptr;
  return (double*)ptr;
}

int opf_var_coll_norm_noalloc(opf_ctx* c, int slot, const double* re, const double* im, double* out,
                              void* stream) {
  // opf_var_coll_norm on device arrays; no device memory is allocated
  return guard([&] {
    check_slot(c, slot);
    if (!re || !im || !out) throw OpfError(OPF_ERR_ARG, "null device pointer");
    if (c->H > NORM_SCRATCH_MAX)
      throw OpfError(OPF_ERR_ARG, "H = " + std::to_string(c->H) + " exceeds the " +
                                      std::to_string(NORM_SCRATCH_MAX) + " harmonics of the norm scratch buffer");
    if (!c->setup[slot]) throw OpfError(OPF_ERR_STATE, "slot not set up (alpha unknown)");
    double* b = scratch_on(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    launch_norm(c, slot, re, im, b, out, s);
    if (!stream) HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int opf_friction(opf_ctx* c, int slot, double* Ub, double* Uc, double* Ufr) {
  // io_write_friction (io.f90:627-639) on the (0,0) mode of u_z; with
  // constant flux Ub is vel_Pr0 of the current state (velocity.f90:356-357)
  return guard([&] {
    check_slot(c, slot);
    if (!c->occupied[slot]) throw OpfError(OPF_ERR_STATE, "slot holds no state");
    SlotParams prm;
    HIPCHK(hipMemcpy(&prm, c->B.prm + slot, sizeof prm, hipMemcpyDeviceToHost));
    std::vector<double> re, im;
    slot_field_to_host(c, slot, 2, re, im);
    const int N = c->N;
    const HostMesh& me = c->mesh;
    const double* uz = re.data();   // harmonic 0 occupies the first N values
    double ub = 0.0;
    if (prm.const_flux) {
      double s = 0.0;
      for (int i = 0; i <= KL; ++i) s = s + uz[N - 1 - KL + i] * me.dr1[i + (KL + 1) * 1];
      ub = s / (-2.0);
    } else {
      double s = 0.0;
      for (int n = 0; n < N; ++n) s = s + uz[n] * me.intrdr[n];
      ub = 0.5 + 2.0 * s;
    }
    double uc = 0.0;
    for (int i = 0; i <= KL; ++i) uc = uc + uz[i] * me.dr0[i];
    double f = 0.0;
    for (int i = 0; i <= KL; ++i) f = f + uz[N - 1 - KL + i] * me.dr1[i + (KL + 1) * 1];
    if (Ub) *Ub = ub;
    if (Uc) *Uc = 1.0 + uc;
    if (Ufr) *Ufr = std::sqrt(std::fabs((f - 2.0) / prm.Re));
  });
}

int opf_get_phys(opf_ctx* c, int slot, int which, double* phys) {
  // vel_r/t/z (which 0..2) and vel_curlr/t/z (3..5) of vel_transform
  // (velocity.f90:304-317) for the slot's current iterate, phys layout
  return guard([&] {
    check_slot(c, slot);
    if (which < 0 || which > 5 || !phys) throw OpfError(OPF_ERR_ARG, "which must be 0..5");
    if (!c->occupied[slot]) throw OpfError(OPF_ERR_STATE, "slot holds no state");
    ensure_temps(c);
    auto ctl = read_ctl(c);
    const double2* src = which < 3 ? c->B.U[ctl[slot].ucur] + (size_t)slot * 3 * c->G.coll + (size_t)which * c->G.coll
                                   : c->B.om + (size_t)slot * 3 * c->G.coll + (size_t)(which - 3) * c->G.coll;
    const FftTiles& t = c->tiles;
    launch(c, "zfwd1", k_zfwd_one, dim3((c->M + t.tm - 1) / t.tm, c->N, 1), dim3(FFT_THREADS), c->smem_z, c->G,
           c->B, t, src, c->tY, (size_t)0, (size_t)0);
    launch(c, "c2r", k_theta_c2r, dim3((c->Z + 1) / 2, c->N, 1), dim3(FFT_THREADS),
           (padded_len(c->T) + c->T) * sizeof(double2), c->G, c->B, t, (const double2*)c->tY, c->tP);
    HIPCHK(hipMemcpyAsync(phys, c->tP, sizeof(double) * c->N * c->Z * c->T, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int opf_add_nonlinear(opf_ctx* c, int slot, const double* nrr, const double* nri, const double* ntr,
                      const double* nti, const double* nzr, const double* nzi) {
  // var_null(2) hooks add forcing to vel_N* after vel_nonlinear
  // (Coriolis.f90:54-71); added here to the slot's current evaluation
  return guard([&] {
    check_slot(c, slot);
    if (!nrr || !ntr || !nzr) throw OpfError(OPF_ERR_ARG, "null argument");
    ensure_temps(c);
    auto ctl = read_ctl(c);
    double2* nraw = c->B.Nraw[ctl[slot].ncur] + (size_t)slot * 3 * c->G.coll;
    const double* re[3] = {nrr, ntr, nzr};
    const double* im[3] = {nri, nti, nzi};
    for (int comp = 0; comp < 3; ++comp) {
      coll_to_dev(c, re[comp], im[comp], c->tC);
      launch(c, "add_coll", k_add_coll, dim3((c->H + 63) / 64, c->N, 1), dim3(64), 0, c->G,
             nraw + (size_t)comp * c->G.coll, (const double2*)c->tC);
      HIPCHK(hipStreamSynchronize(c->stream));
    }
  });
}

int opf_tra_coll2phys(opf_ctx* c, const double* cre, const double* cim, double* phys) {
  return guard([&] {
    if (!c || !cre || !phys) throw OpfError(OPF_ERR_ARG, "null argument");
    ensure_temps(c);
    coll_to_dev(c, cre, cim, c->tC);
    const FftTiles& t = c->tiles;
    launch(c, "zfwd1", k_zfwd_one, dim3((c->M + t.tm - 1) / t.tm, c->N, 1), dim3(FFT_THREADS), c->smem_z, c->G,
           c->B, t, (const double2*)c->tC, c->tY, (size_t)0, (size_t)0);
    launch(c, "c2r", k_theta_c2r, dim3((c->Z + 1) / 2, c->N, 1), dim3(FFT_THREADS), (padded_len(c->T) + c->T) * sizeof(double2), c->G,
           c->B, t, (const double2*)c->tY, c->tP);
    HIPCHK(hipMemcpyAsync(phys, c->tP, sizeof(double) * c->N * c->Z * c->T, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int opf_tra_phys2coll(opf_ctx* c, const double* phys, double* cre, double* cim) {
  return guard([&] {
    if (!c || !phys) throw OpfError(OPF_ERR_ARG, "null argument");
    ensure_temps(c);
    HIPCHK(hipMemcpyAsync(c->tP, phys, sizeof(double) * c->N * c->Z * c->T, hipMemcpyHostToDevice, c->stream));
    const FftTiles& t = c->tiles;
    HIPCHK(hipMemsetAsync(c->tC, 0, sizeof(double2) * c->G.coll, c->stream));
    launch(c, "r2c", k_theta_r2c, dim3((c->Z + 1) / 2, c->N, 1), dim3(FFT_THREADS), (padded_len(c->T) + c->T) * sizeof(double2), c->G,
           c->B, t, (const double*)c->tP, c->tY);
    launch(c, "zbwd1", k_zbwd_one, dim3((c->M + t.tm - 1) / t.tm, c->N, 1), dim3(FFT_THREADS), c->smem_z, c->G,
           c->B, t, (const double2*)c->tY, c->tC, (size_t)0, (size_t)0);
    dev_to_coll(c, c->tC, cre, cim);
  });
}

int opf_tim_lumesh_get(opf_ctx* c, int slot, int which, double* ab, int* ipiv) {
  return guard([&] {
    check_slot(c, slot);
    if (which < 0 || which > 3) throw OpfError(OPF_ERR_ARG, "which must be 0..3");
    if (!c->setup[slot]) throw OpfError(OPF_ERR_STATE, "slot not set up");
    const int N = c->N, P = c->P, H = c->H;
    // per-pair blocks [pair][n][LUS] and [pair][n]
    std::vector<double> F((size_t)P * N * LUS);
    std::vector<unsigned char> pv((size_t)P * N);
    HIPCHK(hipMemcpy(F.data(), lu_block(c->B, slot, which, 0, P, N), F.size() * sizeof(double),
                     hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(pv.data(), pv_block(c->B, slot, which, 0, P, N), pv.size(), hipMemcpyDeviceToHost));
    for (int hh = 0; hh < H; ++hh) {
      int k, m;
      if (hh < c->K) { m = 0; k = hh; }
      else { const int q = hh - c->K; m = 1 + q / (2 * c->K - 1); k = q % (2 * c->K - 1) - (c->K - 1); }
      const int p = m * c->K + std::abs(k);
      for (int n = 0; n < N; ++n) {
        for (int row = 0; row < NLU; ++row)
          if (ab) ab[((size_t)hh * N + n) * NLU + row] = F[((size_t)p * N + n) * LUS + row];
        if (ipiv) ipiv[(size_t)hh * N + n] = n + 1 + pv[(size_t)p * N + n];
      }
    }
  });
}

}  // extern "C"

// tim_lumesh_invert on a host field: solved on the device, lane per harmonic.
namespace opf {
__global__ void k_lumesh_invert(Geo G, Bufs B, int slot, int which, int bc, double2* x);
}

extern "C" int opf_tim_lumesh_invert(opf_ctx* c, int slot, int which, int BC, double* bre, double* bim) {
  return guard([&] {
    check_slot(c, slot);
    if (which < 0 || which > 3) throw OpfError(OPF_ERR_ARG, "which must be 0..3");
    if (BC > 1) throw OpfError(OPF_ERR_ARG, "BC must be 0 or 1 (negative: the matrix's own)");
    if (!c->setup[slot]) throw OpfError(OPF_ERR_STATE, "slot not set up");
    if (!bre || !bim) throw OpfError(OPF_ERR_ARG, "null field");
    const int bc = BC >= 0 ? BC : c->lu_bc[(size_t)slot * 4 + which];
    ensure_temps(c);
    coll_to_dev(c, bre, bim, c->tC);
    launch(c, "lumesh_invert", k_lumesh_invert, grid_h(c, 1), dim3(64), 0, c->G, c->B, slot, which, bc, c->tC);
    dev_to_coll(c, c->tC, bre, bim);
  });
}

extern "C" int opf_tim_lumesh_init(opf_ctx* c, int slot, int which, int PM, int BC, double c1, double c2) {
  // tim_lumesh_init (timestep.f90:67-96): c1 I + c2 Lap (PM = +-1: the
  // modified Laplacian of u+-), the axis parity row and the BC row, then
  // DGBTF2, for every (|k|, m) pair, into factor set `which` of the slot
  return guard([&] {
    check_slot(c, slot);
    if (which < 0 || which > 3) throw OpfError(OPF_ERR_ARG, "which must be 0..3");
    if (PM < -1 || PM > 1 || BC < 0 || BC > 1) throw OpfError(OPF_ERR_ARG, "PM must be -1, 0, 1 and BC 0 or 1");
    if (!c->setup[slot]) throw OpfError(OPF_ERR_STATE, "slot not set up (alpha unknown)");
    auto ctl = read_ctl(c);
    ctl[slot].status = 0;
    write_ctl(c, ctl);
    launch(c, "factor", k_lu_factor, dim3((c->P + 15) / 16, 4, 1), dim3(64), 0, c->G, c->B, slot, 0,
           LuSpec{which, PM, BC, c1, c2});
    ctl = read_ctl(c);
    c->lu_bc[(size_t)slot * 4 + which] = BC;
    if (ctl[slot].status != 0) {
      ctl[slot].status = 0;
      write_ctl(c, ctl);
      throw OpfError(OPF_ERR_SINGULAR, "tim_lumesh_init");
    }
  });
}

extern "C" int opf_kernel_timing(opf_ctx* c, int enable) {
  return guard([&] {
    if (!c) throw OpfError(OPF_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto& r : c->tm.recs) { c->tm.pool.push_back(r.a); c->tm.pool.push_back(r.b); }
    c->tm.recs.clear();
    c->tm.on = enable != 0;
  });
}

extern "C" int opf_launch_mode(opf_ctx* c, int graph, int compact, int streams, int* graph_now, int* compact_now,
                               int* streams_now, long long* graph_launches, long long* graph_captures,
                               long long* compact_iterations) {
  return guard([&] {
    if (!c) throw OpfError(OPF_ERR_ARG, "null context");
    if (graph > 64) throw OpfError(OPF_ERR_ARG, "at most 64 iterations per captured graph");
    if (streams > 2) throw OpfError(OPF_ERR_ARG, "at most 2 streams");
    if (graph >= 0) c->use_graph = graph;
    if (compact >= 0) c->use_compact = compact != 0;
    if (streams >= 1) c->use_streams = streams;
    if (graph_now) *graph_now = c->use_graph;
    if (compact_now) *compact_now = c->use_compact;
    if (streams_now) *streams_now = c->use_streams;
    if (graph_launches) *graph_launches = c->graph_launches;
    if (graph_captures) *graph_captures = c->graph_captures;
    if (compact_iterations) *compact_iterations = c->compact_iterations;
  });
}

extern "C" int opf_tile_config(const opf_ctx* c, int* pairs, int* tm, int* tz) {
  return guard([&] {
    if (!c) throw OpfError(OPF_ERR_ARG, "null context");
    if (pairs) *pairs = c->st.gp;
    if (tm) *tm = c->tiles.tm;
    if (tz) *tz = c->tiles.tz;
  });
}

extern "C" int opf_step_config(const opf_ctx* c, int* gp, int* threads, long long* lds) {
  return guard([&] {
    if (!c || !gp || !threads || !lds) throw OpfError(OPF_ERR_ARG, "null argument");
    *gp = c->st.gp;
    *threads = step_threads(c->st.gp);
    *lds = (long long)c->st.smem;
  });
}

extern "C" int opf_kernel_time(opf_ctx* c, const char* name, double* ms, int* launches) {
  return guard([&] {
    if (!c || !name) throw OpfError(OPF_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    int cnt = 0;
    for (auto& r : c->tm.recs)
      if (r.name == name) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
        tot += t;
        ++cnt;
      }
    if (ms) *ms = tot;
    if (launches) *launches = cnt;
  });
}

// ---- state files (io.f90) -------------------------------------------------
extern "C" int opf_state_file_info(const char* path, int* N, int* K, int* M, int* Mp, double* t, double* Re,
                                   double* alpha, double* dt) {
  return guard([&] {
    if (!path) throw OpfError(OPF_ERR_ARG, "null path");
    StateInfo s = read_state_info(path);
    if (N) *N = s.N;
    if (K) *K = s.K;
    if (M) *M = s.M;
    if (Mp) *Mp = s.Mp;
    if (t) *t = s.t;
    if (Re) *Re = s.Re;
    if (alpha) *alpha = s.alpha;
    if (dt) *dt = s.dt;
  });
}

extern "C" int opf_state_file_write(const char* path, int N, int K, int M, int Mp, const double* r, double t,
                                    double Re, double alpha, double dt, double dtcor, double dtcfl,
                                    const double* urr, const double* uri, const double* utr, const double* uti,
                                    const double* uzr, const double* uzi) {
  return guard([&] {
    const double* f[6] = {urr, uri, utr, uti, uzr, uzi};
    if (!path || !r) throw OpfError(OPF_ERR_ARG, "null argument");
    for (auto p : f)
      if (!p) throw OpfError(OPF_ERR_ARG, "null field");
    if (N <= KL || N > MAXN || K < 1 || M < 1 || Mp < 1) throw OpfError(OPF_ERR_ARG, "bad N, K, M or Mp");
    write_state_file(path, N, K, M, Mp, t, Re, alpha, dt, dtcor, dtcfl, r, f);
  });
}

extern "C" int opf_state_file_read(const char* path, int N, int K, int M, int Mp, const double* r, double* t,
                                   double* Re, double* alpha, double* dt, double* dtcor, double* dtcfl, double* urr,
                                   double* uri, double* utr, double* uti, double* uzr, double* uzi) {
  return guard([&] {
    double* f[6] = {urr, uri, utr, uti, uzr, uzi};
    if (!path) throw OpfError(OPF_ERR_ARG, "null path");
    for (auto p : f)
      if (!p) throw OpfError(OPF_ERR_ARG, "null field");
    if (N <= KL || N > MAXN || K < 1 || M < 1 || Mp < 1) throw OpfError(OPF_ERR_ARG, "bad N, K, M or Mp");
    HostMesh me;
    if (!r) build_mesh(N, nullptr, me);   // the default mesh of mes_precompute
    read_state_file(path, N, K, M, Mp, r ? r : me.r.data(), t, Re, alpha, dt, dtcor, dtcfl, f);
  });
}

extern "C" int opf_state_save(opf_ctx* c, int slot, const char* path) {
  return guard([&] {
    check_slot(c, slot);
    if (!path) throw OpfError(OPF_ERR_ARG, "null path");
    if (!c->occupied[slot]) throw OpfError(OPF_ERR_STATE, "slot holds no state");
    HIPCHK(hipStreamSynchronize(c->stream));
    SlotParams p{};
    HIPCHK(hipMemcpy(&p, c->B.prm + slot, sizeof(p), hipMemcpyDeviceToHost));
    auto ctl = read_ctl(c);
    const SlotCtl& k = ctl[slot];
    const size_t nh = (size_t)c->N * c->H;
    std::vector<double> buf(6 * nh);
    const double2* u = c->B.U[k.ucur] + (size_t)slot * 3 * c->G.coll;
    for (int comp = 0; comp < 3; ++comp)
      dev_to_coll(c, u + comp * c->G.coll, &buf[2 * comp * nh], &buf[(2 * comp + 1) * nh]);
    const double* f[6];
    for (int i = 0; i < 6; ++i) f[i] = &buf[i * nh];
    write_state_file(path, c->N, c->K, c->M, c->Mp, k.t, p.Re, p.alpha, p.dt, k.corr_dt, k.cfl_dt,
                     c->mesh.r.data(), f);
  });
}

extern "C" int opf_state_load(opf_ctx* c, int slot, const char* path, double t) {
  return guard([&] {
    check_slot(c, slot);
    if (!path) throw OpfError(OPF_ERR_ARG, "null path");
    if (!c->setup[slot]) throw OpfError(OPF_ERR_STATE, "slot not set up (opf_slot_setup)");
    const size_t nh = (size_t)c->N * c->H;
    std::vector<double> buf(6 * nh);
    double* f[6];
    for (int i = 0; i < 6; ++i) f[i] = &buf[i * nh];
    double tf = 0.0, dtf = 0.0, dtcor = 0.0, dtcfl = 0.0;
    read_state_file(path, c->N, c->K, c->M, c->Mp, c->mesh.r.data(), &tf, nullptr, nullptr, &dtf, &dtcor, &dtcfl,
                    f);
    if (c->adapt[slot].on && dtf > 0.0) {   // d_timestep < 0: tim_dt from the file (io.f90:145)
      SlotParams p{};
      HIPCHK(hipMemcpy(&p, c->B.prm + slot, sizeof(p), hipMemcpyDeviceToHost));
      if (p.dt != dtf) slot_setup(c, slot, p.Re, p.alpha, dtf, p.implicit, p.const_flux);
    }
    const double* cf[6] = {f[0], f[1], f[2], f[3], f[4], f[5]};
    put_state(c, slot, cf, t < 0.0 ? tf : t);   // d_time < 0: tim_t from the file (io.f90:120)
    auto ctl = read_ctl(c);
    ctl[slot].corr_dt = dtcor;   // tim_corr_dt, tim_cfl_dt (io.f90:147-150)
    ctl[slot].cfl_dt = dtcfl;
    write_ctl(c, ctl);
  });
}

// ---- Newton-Krylov clients (newton.f90) ------------------------------------
extern "C" int opf_vel_shift(opf_ctx* c, int slot, double dth, double dz) {
  return guard([&] {
    check_slot(c, slot);
    if (!c->occupied[slot]) throw OpfError(OPF_ERR_STATE, "slot holds no state");
    HIPCHK(hipStreamSynchronize(c->stream));
    auto ctl = read_ctl(c);
    ctl[slot].active = 1;
    write_ctl(c, ctl);
    launch_shift(c, slot, 1, {slot}, {dth}, {dz});
    launch(c, "curl0", k_curl, grid_hn(c, 1), dim3(64), 0, c->G, c->B, slot, 0);
    HIPCHK(hipStreamSynchronize(c->stream));
    ctl = read_ctl(c);
    ctl[slot].active = 0;
    write_ctl(c, ctl);
  });
}

extern "C" int opf_steporbit(opf_ctx* c, int nslots, const int* slots, const int* ndts, const double* T,
                             const double* sth, const double* sz) {
  return guard([&] {
    if (!c || nslots < 0 || (nslots > 0 && (!slots || !ndts || !T || !sth || !sz)))
      throw OpfError(OPF_ERR_ARG, "null argument");
    std::vector<int> sel, steps, fixed;
    std::vector<double> dth, dz;
    std::vector<int> seen(c->cap, 0);
    for (int i = 0; i < nslots; ++i) {
      const int s = slots[i];
      check_slot(c, s);
      if (seen[s]++) throw OpfError(OPF_ERR_ARG, "slot listed twice");
      if (!c->occupied[s]) throw OpfError(OPF_ERR_STATE, "slot holds no state");
      if (c->adapt[s].on) throw OpfError(OPF_ERR_ARG, "steporbit needs a fixed time step (adaptive slot)");
      if (ndts[i] < 1) throw OpfError(OPF_ERR_ARG, "ndts must be >= 1");
      if (ndts[i] != 1) {   // tim_dt = T / ndts and vel_matrices (newton.f90:736-743)
        if (!(T[i] > 0.0)) throw OpfError(OPF_ERR_ARG, "T must be > 0");
        SlotParams p{};
        HIPCHK(hipMemcpy(&p, c->B.prm + s, sizeof(p), hipMemcpyDeviceToHost));
        const double dt = T[i] / (double)ndts[i];
        if (p.dt != dt) slot_setup(c, s, p.Re, p.alpha, dt, p.implicit, p.const_flux);
      }
      sel.push_back(s);
      steps.push_back(ndts[i]);
      fixed.push_back(1);   // predictor and one corrector per step (newton.f90:745-757)
      // the shift back by the orbit's shifts after the last step (:765-770)
      dth.push_back(ndts[i] != 1 ? -sth[i] : 0.0);
      dz.push_back(ndts[i] != 1 ? -sz[i] : 0.0);
    }
    // multJ's trajectory direction (ndts = 1) leaves tim_t and tim_step as
    // they were (newton.f90:759-762 advances them only for ndts /= 1)
    std::vector<std::pair<int, double>> keep_t;
    std::vector<int> keep_step;
    {
      const auto ctl = read_ctl(c);
      for (int i = 0; i < nslots; ++i)
        if (ndts[i] == 1) {
          keep_t.push_back({slots[i], ctl[slots[i]].t});
          keep_step.push_back(ctl[slots[i]].step);
        }
    }
    run_plan(c, sel, steps, fixed, 0);
    orbit_finish(c, sel, dth, dz);
    if (!keep_t.empty()) {
      auto ctl = read_ctl(c);
      for (size_t i = 0; i < keep_t.size(); ++i) {
        ctl[keep_t[i].first].t = keep_t[i].second;
        ctl[keep_t[i].first].step = keep_step[i];
      }
      write_ctl(c, ctl);
    }
  });
}
