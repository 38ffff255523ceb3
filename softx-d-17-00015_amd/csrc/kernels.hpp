// Device buffers of a resolution class and the kernel entry points.
#pragma once
#include <hip/hip_runtime.h>

#include "opf_internal.hpp"

namespace opf {

// All device arrays of one resolution class (slot-major).  Passed by value.
struct Bufs {
  double2* U[3];        // velocity buffers [slot][3][N][Hp]; roles per slot in SlotCtl
  double2* Nraw[2];     // nonlinear terms without base-flow terms [slot][3][N][Hp]
  double2* om;          // curl of the current iterate [slot][3][N][Hp]
  double2* R;           // step scratch: right-hand sides [slot][3][N][Hp]
  double2* S;           // step scratch: pressure [slot][N][Hp]
  double* LU;           // factors [slot][4 mats][P pairs][N columns][LUS]
  unsigned char* PV;    // pivot offsets ipiv(j)-j [slot][4][P][N]
  double2* FW;          // forward-sweep copy [slot][4][P][N][FW_Q]: (offset, l1) (l2, l3) (l4, 0)
  double2* BW;          // backward-sweep copy [slot][4][P][BW_Q][N]: (u0, 1/u0) (u1, u2) .. (u7, u8)
  unsigned char* META;  // [slot][4][P][2]: widest used superdiagonal, largest pivot offset
  double* Uinf;         // influence vectors [slot][6][N][P]
  double* Ainv;         // inverted influence matrices [slot][16][H]
  double* Ui;           // flux correction vector [slot][N]
  double* diagb;        // per-harmonic energy partials [slot][3][Hp]
  unsigned long long* cflmax;  // [slot][3][N] per-plane max of |u_r|, |u_t|, |u_z + 1 - r^2| on the
                               // physical grid of the last evaluation (bit patterns, vel_maxtstep)
  DiagRec* diag;        // diagnostics records [slot][diag_cap]
  SlotParams* prm;      // [slot]
  SlotCtl* ctl;         // [slot]
  double2* Yf;          // forward intermediates [slot][N][Z][6][M]
  double2* Yb;          // backward intermediates [slot][N][Z][3][M]
  double2* twZ;         // twiddles exp(2 pi i q / Z), q < Z
  double2* twT;         // twiddles exp(2 pi i q / T), q < T
  const int* smap;      // null: grid z counts slots from slot0; else the slots of a
                        // compacted launch, smap[slot0 + z]
};

// The slot a workgroup serves: slot0 + blockIdx.z, or its entry of the
// compacted slot list.
__device__ __forceinline__ int slot_of(const Bufs& B, int slot0) {
  return B.smap ? B.smap[slot0 + (int)blockIdx.z] : slot0 + (int)blockIdx.z;
}

// Tile shapes of the transform kernels (chosen by the host / autotuner).
struct FftTiles {
  int tm;     // m columns per workgroup in the z kernels
  int tz;     // z lines per workgroup in the theta kernel
  FftPlan pz, pt;
};

// Factor blocks: matrix `mat` of pair p, column n, LAPACK band row r at
// lu_block(...)[n*LUS + r]; pivot offset at pv_block(...)[n].
__host__ __device__ inline double* lu_block(const Bufs& B, int slot, int mat, int p, int P, int N) {
  return B.LU + (((size_t)(slot * 4 + mat) * P + p) * N) * LUS;
}
__host__ __device__ inline unsigned char* pv_block(const Bufs& B, int slot, int mat, int p, int P, int N) {
  return B.PV + ((size_t)(slot * 4 + mat) * P + p) * N;
}

// Solve-ready copy of the factors used by the fused step kernel.  Column j of
// the forward block holds the pivot offset (as a double) and the multipliers
// l1..l4 = AB(10..13, j); plane q of the backward block holds, per column j,
// q = 0: (u0, 1/u0) with u0 = AB(9, j), q >= 1: (u_{2q-1}, u_{2q}) with
// u_d = AB(9-d, j) = U(j-d, j).
constexpr int FW_Q = 3;
constexpr int BW_Q = 5;
__host__ __device__ inline double2* fw_block(const Bufs& B, int slot, int mat, int p, int P, int N) {
  return B.FW + (((size_t)(slot * 4 + mat) * P + p) * N) * FW_Q;
}
__host__ __device__ inline double2* bw_block(const Bufs& B, int slot, int mat, int p, int P, int N) {
  return B.BW + (((size_t)(slot * 4 + mat) * P + p) * N) * BW_Q;
}
__host__ __device__ inline unsigned char* meta_block(const Bufs& B, int slot, int mat, int p, int P) {
  return B.META + ((size_t)(slot * 4 + mat) * P + p) * 2;
}

// Shape of the fused step kernel.
struct StepTiles {
  int gp;        // (|k|, m) pairs per workgroup: 1, 2 or 4
  int bufd2;     // double2 per factor chunk buffer (two buffers)
  int ngroups;   // workgroups per slot
  size_t smem;   // dynamic LDS bytes
};
// threads of a step workgroup: wave 0 solves, the other waves stream factor chunks
inline int step_threads(int gp) { return gp == 4 ? 256 : 128; }
constexpr int step_bufd2(int gp) { return 3 * gp * 49; }   // double2 per factor chunk buffer
constexpr int STEP_BLKCAP = 64;    // double2 per (system, pair) block of a chunk (one per loader lane)
constexpr int FC_TAIL = 16;        // double2 after the factor ring: the solvers' prefetch reads up to two
                                   // columns (at most 10 double2) past a chunk

constexpr int FFT_THREADS = 256;
constexpr int FFT_TILE = 2048;   // complex values per transform workgroup (all its lines)
// items (line, theta point) one theta workgroup holds in registers: FFT_TILE / 3
// rounded up to whole rounds of FFT_THREADS (768).  Lines of 3M > FFT_TILE / 2
// would leave one line per workgroup under FFT_TILE; they take two lines
// when 2 * 3M fits in these items (up to 3M = 384).
constexpr int THETA_ITEMS = ((FFT_TILE / 3 + FFT_THREADS - 1) / FFT_THREADS) * FFT_THREADS;

// LDS stride of one transform of length L in the transform kernels: one pad
// slot per 16 elements, rounded up to an odd count.
// m columns per workgroup of the z transforms for length Z (as many as fit
// in FFT_TILE values; rounding down to whole 128-byte lines of the
// intermediates measured slower)
__host__ __device__ constexpr int tile_m(int Z) { return FFT_TILE / Z; }

__host__ __device__ constexpr int padded_len(int L) { return (L + (L - 1) / 16 + 1) | 1; }

// The matrix tim_lumesh_init builds: mat < 0 the four of vel_matrices
// (LDp, LDm, LDz with c1 = 1/dt, c2 = -implicit nu/Re; LNp with 0, 1, BC = 1);
// mat >= 0 only that one, with the given PM, BC, c1, c2.
struct LuSpec {
  int mat, PM, BC;
  double c1, c2;
};
__global__ void k_lu_factor(Geo G, Bufs B, int slot0, int gated, LuSpec spec);
__global__ void k_influence(Geo G, Bufs B, int slot0, int gated);
__global__ void k_ainv(Geo G, Bufs B, int slot0, int gated);
__global__ void k_curl(Geo G, Bufs B, int slot0, int gated);
__global__ void k_symmetry(Geo G, Bufs B, int slot0, int gated);
__global__ void k_newdt(Geo G, Bufs B, int slot0, int what);
__global__ void k_imposesym(Geo G, Bufs B, int slot0);
__global__ void k_admit(Bufs B, int slot, int phase, int nsteps, int corr_fixed);
__global__ void k_shift(Geo G, Bufs B, int slot0, const double2* __restrict__ tab);
// non-Newtonian plugin: physical-grid fields of one slot (row-major S(i,j) at S[3i+j])
struct NnPhys {
  double* S[9];
  double* nu;
  double* g[3];
  double* q[3];
};
__global__ void k_nn_grad(Geo G, Bufs B, int slot, int row, double2* __restrict__ out);
__global__ void k_nn_gradp(Geo G, Bufs B, int slot, const double2* __restrict__ p, double2* __restrict__ out);
__global__ void k_nn_curlcurl(Geo G, Bufs B, int slot, double2* __restrict__ out);
__global__ void k_nn_visc(NnPhys P, size_t np, double lam, double d1, double d2, double a);
__global__ void k_nn_force(NnPhys P, size_t np, double nu_r, double Re);
__global__ void k_nn_addN(Geo G, Bufs B, int slot, int comp, const double2* __restrict__ src);
__global__ void k_les_prep(Geo G, Bufs B, int slot, int model, double Cs, double Aplus, double* D,
                           unsigned long long* numax);
__global__ void k_les_visc(NnPhys P, size_t np, int N, int ZT, int model, double Cs, double dz, const double* D,
                           unsigned long long* numax_slot);
__global__ void k_les_nur(Bufs B, int slot, const unsigned long long* numax);
__global__ void k_les_force(NnPhys P, size_t np, Bufs B, int slot);
__global__ void k_c2r_lines(Geo G, Bufs B, FftTiles T, int BT, const double2* Y, double* phys, size_t fy,
                            size_t fp);
__global__ void k_r2c_lines(Geo G, Bufs B, FftTiles T, int BT, const double* phys, double2* Y, size_t fp,
                            size_t fy);
__global__ void k_add_coll(Geo G, double2* dst, const double2* src);
// var_coll_norm of a field in the coll layout: per-harmonic partials b[H], then [E, Ek, Em]
__global__ void k_norm_harm(Geo G, const SlotParams* prm, const double* re, const double* im, double* b);
__global__ void k_norm_sum(Geo G, const double* b, double* out);
__global__ void k_step_g1(Geo G, Bufs B, StepTiles ST, int slot0);
__global__ void k_step_g2(Geo G, Bufs B, StepTiles ST, int slot0);    // N <= 64
__global__ void k_step_g2x(Geo G, Bufs B, StepTiles ST, int slot0);   // N <= 128
__global__ void k_step_g4(Geo G, Bufs B, StepTiles ST, int slot0);    // N <= 64, 256 threads
__global__ void k_step_g4x(Geo G, Bufs B, StepTiles ST, int slot0);   // N <= 96, 256 threads
__global__ void k_control(Geo G, Bufs B, int slot0);

// transforms: kernels specialised to the radix plans of (Z, T), chosen once
// per context by fft_kernels (generic kernels for other lengths)
using FftKern = void (*)(Geo, Bufs, FftTiles, int);
struct FftKerns {
  FftKern zfwd, theta, zbwd;
};
FftKerns fft_kernels(const FftPlan& pz, const FftPlan& pt);
// whole-plane evaluation (z, theta, products, theta, z in one workgroup per
// radial plane) for the plan pairs that have one; nullptr otherwise.
// smem: its dynamic LDS bytes.
constexpr int PLANE_THREADS = 1024;
using PlaneKern = void (*)(Geo, Bufs, FftTiles, int);
PlaneKern plane_kernel(const FftPlan& pz, const FftPlan& pt, size_t* smem);
// materialise physical fields / transform arbitrary coll fields (granular API)
__global__ void k_zfwd_one(Geo G, Bufs B, FftTiles T, const double2* a, double2* Y, size_t fa, size_t fy);
__global__ void k_theta_c2r(Geo G, Bufs B, FftTiles T, const double2* Y, double* phys);
__global__ void k_theta_r2c(Geo G, Bufs B, FftTiles T, const double* phys, double2* Y);
__global__ void k_zbwd_one(Geo G, Bufs B, FftTiles T, const double2* Y, double2* a, size_t fy, size_t fa);

}  // namespace opf
