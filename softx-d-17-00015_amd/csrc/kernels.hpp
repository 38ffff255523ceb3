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
  double* Uinf;         // influence vectors [slot][6][N][P]
  double* Ainv;         // inverted influence matrices [slot][16][H]
  double* Ui;           // flux correction vector [slot][N]
  double* diagb;        // per-harmonic energy partials [slot][3][Hp]
  DiagRec* diag;        // diagnostics records [slot][diag_cap]
  SlotParams* prm;      // [slot]
  SlotCtl* ctl;         // [slot]
  double2* Yf;          // forward intermediates [slot][N][Z][6][M]
  double2* Yb;          // backward intermediates [slot][N][Z][3][M]
  double2* twZ;         // twiddles exp(2 pi i q / Z), q < Z
  double2* twT;         // twiddles exp(2 pi i q / T), q < T
};

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

// Shape of the fused step kernel: pairs per workgroup and factor chunk.
struct StepTiles {
  int gp;        // (|k|, m) pairs per workgroup (<= 8)
  int cs;        // factor columns per LDS chunk
  int ngroups;   // workgroups per slot
  size_t smem;   // dynamic LDS bytes
};
constexpr int STEP_THREADS = 256;
constexpr int STEP_LOADMAX = 4;  // factor double2 loaded per loader thread per chunk

constexpr int FFT_THREADS = 256;
constexpr int FFT_MAXB = 4;      // butterflies held per thread per pass

__global__ void k_lu_factor(Geo G, Bufs B, int slot0);
__global__ void k_influence(Geo G, Bufs B, int slot0);
__global__ void k_ainv(Geo G, Bufs B, int slot0);
__global__ void k_curl(Geo G, Bufs B, int slot0);
__global__ void k_imposesym(Geo G, Bufs B, int slot0);
__global__ void k_step(Geo G, Bufs B, StepTiles ST, int slot0);
__global__ void k_control(Geo G, Bufs B, int slot0);

// transforms
__global__ void k_zfwd(Geo G, Bufs B, FftTiles T, int slot0);
__global__ void k_theta(Geo G, Bufs B, FftTiles T, int slot0);
__global__ void k_zbwd(Geo G, Bufs B, FftTiles T, int slot0);
// materialise physical fields / transform arbitrary coll fields (granular API)
__global__ void k_zfwd_one(Geo G, Bufs B, FftTiles T, const double2* a, double2* Y);
__global__ void k_theta_c2r(Geo G, Bufs B, FftTiles T, const double2* Y, double* phys);
__global__ void k_theta_r2c(Geo G, Bufs B, FftTiles T, const double* phys, double2* Y);
__global__ void k_zbwd_one(Geo G, Bufs B, FftTiles T, const double2* Y, double2* a);

}  // namespace opf
