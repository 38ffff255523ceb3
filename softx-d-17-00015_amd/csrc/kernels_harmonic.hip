// The fused per-harmonic step kernel: vel_predictor / vel_corrector
// (velocity.f90:400-432) with vel_addHPF (:338-361), tim_nlincorr
// (timestep.f90:174-184), vel_step (velocity.f90:367-394), tim_measurecorr
// (timestep.f90:190-207) and the curl of the new velocity for the next
// nonlinear evaluation (var_coll_curl, variables.f90:408-436).
//
// One workgroup owns a group of GP consecutive (|k|, m) pairs, i.e. the
// harmonics +k and -k of each pair (2*GP "harmonic slots"): two waves for
// GP = 1, 2 and four for GP = 4.
// Everything a harmonic needs between its input and output stays in LDS:
//   V[4][N][2GP]  the three right-hand sides (u+, u-, uz) and the pressure.
// Both waves map threads to (radial point, harmonic) for the stages that are
// parallel in r (right-hand side, divergence, gradient, influence correction,
// curl).  The banded LU solves are serial in r: wave 0 runs them, one lane
// per (system, harmonic, real or imaginary part), while the other waves
// stream the factor columns of the group into a two-buffer LDS ring with direct
// global-to-LDS copies, one chunk ahead of the solver.  Factors are stored
// once per pair and shared by the +k / -k harmonics and their real and
// imaginary parts.
//
// The solve loops are unrolled over the length of their register window (5
// rows for the L sweep, kue+1 for the U sweep, kue being the widest
// superdiagonal of U that holds a non-zero in the group's matrices), so the
// window rotates by renaming instead of moving registers.
//
// Arithmetic follows the reference operation by operation (see the cited
// routines) with products and sums rounded separately (no contraction),
// except in the banded solves: there each update b - l*t is one fused
// multiply-add and the division by the pivot is a product with its
// reciprocal (correctly rounded), which halves the solves' FP64
// instructions; a solve then differs from the separately rounded one by a
// few ulps per row (relative 1e-15 on the state).
#include <hip/hip_runtime.h>

#include "kernels.hpp"

#pragma clang fp contract(off)

namespace opf {

namespace {

__device__ __forceinline__ double diag_d(double r2inv, int Mp, int m, double alpha, int k, int PM) {
  // tim_mesh_init / tim_lumesh_init diagonal term (timestep.f90:54-55, 76-77)
  double d = -((((r2inv * (double)Mp) * (double)m) * (double)Mp) * (double)m) -
             (((alpha * (double)k) * alpha) * (double)k);
  if (PM != 0) d = (d - r2inv) - ((((2.0 * (double)PM) * (double)Mp) * (double)m) * r2inv);
  return d;
}

struct Slot {
  bool valid;
  bool h00;
  int pl, s, p, m, ka, k, h;
};

__device__ __forceinline__ Slot slot_of(int lh, int gp, int np, int p0, int K) {
  Slot r;
  r.pl = lh % gp;
  r.s = lh / gp;
  r.p = p0 + r.pl;
  r.m = r.p / K;
  r.ka = r.p - r.m * K;
  r.k = r.s ? -r.ka : r.ka;
  r.valid = r.pl < np && (r.s == 0 || (r.ka > 0 && r.m > 0));
  r.h = r.m == 0 ? r.k : K + (r.m - 1) * (2 * K - 1) + r.k + K - 1;
  r.h00 = r.valid && r.h == 0;
  return r;
}

// wall boundary residuals (vel_evalBC, velocity.f90:262-298)
__device__ void eval_bc5(const double2* up, const double2* um, const double2* uz, const double* dr1c,
                         double d, double m_, double a_, double BRe[4], double BIm[4]) {
  const double2 upN = up[KL], umN = um[KL], uzN = uz[KL];
  const double urRe = 0.5 * (upN.x + umN.x);
  const double urIm = 0.5 * (upN.y + umN.y);
  const double utRe = 0.5 * (upN.y - umN.y);
  const double utIm = -0.5 * (upN.x - umN.x);
  double drRe = 0.0, drIm = 0.0;
#pragma unroll
  for (int i = 0; i <= KL; ++i) {
    drRe = drRe + dr1c[i] * (up[i].x + um[i].x);
    drIm = drIm + dr1c[i] * (up[i].y + um[i].y);
  }
  drRe = 0.5 * drRe;
  drIm = 0.5 * drIm;
  BRe[0] = upN.x;
  BIm[0] = upN.y;
  BRe[1] = umN.x;
  BIm[1] = umN.y;
  BRe[2] = -uzN.y;
  BIm[2] = uzN.x;
  BRe[3] = ((urRe * d + drRe) - (d * m_) * utIm) - a_ * uzN.y;
  BIm[3] = ((urIm * d + drIm) + (d * m_) * utRe) + a_ * uzN.x;
}

struct Lds {
  double2* V;          // [4][N][GH]
  double2* FC;         // factor chunk ring: 2 buffers of bufd2 double2
  double* E;           // energy terms [3][N][GH] (aliases FC; used before the solves)
  double* AC;          // [GH][8]
  double* RED;         // [8]
  int* META;           // [4]: (kue, largest pivot offset) of the pressure and Helmholtz matrices
  int GH, bufd2;
};

__device__ __forceinline__ double2& Vat(const Lds& L, int c, int n, int lh, int N) {
  return L.V[((size_t)c * N + n) * L.GH + lh];
}

// ---------------------------------------------------------------------------
// The factor stream: the solves consume, in this order, the chunks of the
// pressure L sweep, pressure U sweep, Helmholtz L sweep and Helmholtz U
// sweep ("segments").  A chunk holds, per (system, pair) block, `cs` columns
// of `qn` double2 each; blocks are padded by one double2 (bank spread).
// ---------------------------------------------------------------------------
struct Seg {
  int m0;     // first factor matrix of the phase
  int nsys;   // systems (matrices) of the phase
  int fwd;    // 1: L sweep (columns 0..N-2 ascending), 0: U sweep (N-1..0)
  int qn;     // double2 per column
  int cs;     // columns per chunk (a multiple of the window length)
  int nch;    // chunks
};

__device__ __forceinline__ Seg make_seg(int m0, int nsys, int fwd, int kue, int N, int gp, int bufd2) {
  Seg s;
  s.m0 = m0;
  s.nsys = nsys;
  s.fwd = fwd;
  s.qn = fwd ? FW_Q : 1 + (kue + 1) / 2;
  const int W = fwd ? 5 : kue + 1;
  int cap = bufd2 / (nsys * gp) - 1;
  cap = cap < STEP_BLKCAP ? cap : STEP_BLKCAP;
  s.cs = (cap / s.qn) / W * W;
  const int tot = fwd ? N - 1 : N;
  s.nch = (tot + s.cs - 1) / s.cs;
  return s;
}

__device__ __forceinline__ int seg_ncol(const Seg& S, int c, int N) {
  const int tot = S.fwd ? N - 1 : N;
  const int rem = tot - c * S.cs;
  return rem < S.cs ? rem : S.cs;
}

struct Stream {
  Seg s[4];
  int total;
};

__device__ __forceinline__ Seg seg_of(const Stream& st, int g, int& c) {
  c = g;
  Seg S = st.s[0];
  if (c >= S.nch) {
    c -= S.nch;
    S = st.s[1];
    if (c >= S.nch) {
      c -= S.nch;
      S = st.s[2];
      if (c >= S.nch) {
        c -= S.nch;
        S = st.s[3];
      }
    }
  }
  return S;
}

// Loader waves (every wave but the first): chunk g of the stream is copied
// from global memory straight into ring buffer g & 1
// (global_load_lds_dwordx4, no registers); loader wave w takes the
// (system, pair) blocks b with b % NLW == w and its lane l moves element l
// of each.  The copy is issued while the solver works on chunk g-1 and is
// complete at the workgroup barrier that ends that chunk.
typedef __attribute__((address_space(3))) void* lds_ptr_t;

template <int GP, int NLW>
__device__ __forceinline__ void ld_dma(const Lds& L, const Bufs& B, const Stream& st, int g, int slot, int p0,
                                       int np, int P, int N, int lane) {
  if (g >= st.total) return;
  int c;
  const Seg S = seg_of(st, g, c);
  const int lw = lane >> 6;   // loader wave: blocks lw, lw + NLW, ...
  lane &= 63;
  const int per = seg_ncol(S, c, N) * S.qn;
  if (lane >= per) return;
  const int col = lane / S.qn, q = lane - col * S.qn;
  const int BS = S.cs * S.qn + 1;
  double2* buf = L.FC + (size_t)(g & 1) * L.bufd2;
#pragma unroll
  for (int sys = 0; sys < 3; ++sys)
#pragma unroll
    for (int pl = 0; pl < GP; ++pl) {
      if ((sys * GP + pl) % NLW == lw && sys < S.nsys && pl < np) {
        const int mat = S.m0 + sys;
        const double2* src = S.fwd ? fw_block(B, slot, mat, p0 + pl, P, N) + (size_t)c * S.cs * FW_Q + lane
                                   : bw_block(B, slot, mat, p0 + pl, P, N) + (size_t)q * N + (N - 1 - c * S.cs - col);
        __builtin_amdgcn_global_load_lds((const void*)src, (lds_ptr_t)(buf + (sys * GP + pl) * BS), 16, 0, 0);
      }
    }
}

// ---------------------------------------------------------------------------
// DGBTRS 'N' (KL = KU = 4) for one solver lane, split into chunks.
// L sweep with the row interchanges: w[i] holds b(j+i), the rows that the
// multipliers of column j touch.  F: the lane's block of the chunk, column
// cl at F[3*cl]: (pivot offset, l1) (l2, l3) (l4, 0).  MO: largest pivot
// offset in the block's matrices.
//
// The whole-window loop keeps the factors of column cl + 2 and the row that
// enters the window two columns later in flight while column cl is
// eliminated (register rings of five named slots, so the prefetch carries
// across windows without moves).  The last two prefetches of a chunk read
// up to two columns past it (the next block, or the ring's tail pad of
// FC_TAIL double2) and are discarded.  Rows at or beyond N enter the window as a
// copy of row N-1: such a row is never a pivot (pivot rows lie inside the
// matrix), its multipliers are exact zeros and it is never stored, so its
// value cannot reach a result.
// ---------------------------------------------------------------------------
template <int MO>
__device__ __forceinline__ void fwd_chunk(const double2* __restrict__ F, double* __restrict__ x, int xs, int j0, int ncol, int N,
                                          double (&w)[5]) {
  int cl = 0;
  if (ncol >= 5) {
    double2 f[5][3];
    double xi[5];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
      for (int q = 0; q < 3; ++q) f[p][q] = F[p * 3 + q];
      const int jn = j0 + p + 5;
      xi[p] = x[(jn < N ? jn : N - 1) * xs];
    }
    for (; cl + 5 <= ncol; cl += 5) {
      const int j = j0 + cl;
#pragma unroll
      for (int s = 0; s < 5; ++s) {
        {
#pragma unroll
          for (int q = 0; q < 3; ++q) f[(s + 2) % 5][q] = F[(cl + s + 2) * 3 + q];
          const int jn = j + s + 7;
          xi[(s + 2) % 5] = x[(jn < N ? jn : N - 1) * xs];
        }
        const double2 f0 = f[s][0], f1 = f[s][1], f2 = f[s][2];
        if (MO > 0) {
          const double offd = f0.x;
          const double a0 = w[s % 5];
          double sv = a0;
#pragma unroll
          for (int i = 1; i <= MO; ++i) sv = offd == (double)i ? w[(s + i) % 5] : sv;
#pragma unroll
          for (int i = 1; i <= MO; ++i) w[(s + i) % 5] = offd == (double)i ? a0 : w[(s + i) % 5];
          w[s % 5] = sv;
        }
        const double t = w[s % 5];
        const double nt = -t;
        x[(j + s) * xs] = t;
        w[(s + 1) % 5] = fma(f0.y, nt, w[(s + 1) % 5]);
        w[(s + 2) % 5] = fma(f1.x, nt, w[(s + 2) % 5]);
        w[(s + 3) % 5] = fma(f1.y, nt, w[(s + 3) % 5]);
        w[(s + 4) % 5] = fma(f2.x, nt, w[(s + 4) % 5]);
        w[s % 5] = xi[s];
      }
    }
  }
  for (; cl < ncol; ++cl) {   // remainder of the last chunk: explicit shifts
    const int j = j0 + cl;
    const double2 f0 = F[cl * 3], f1 = F[cl * 3 + 1], f2 = F[cl * 3 + 2];
    if (MO > 0) {
      const double offd = f0.x;
      const double a0 = w[0];
      double sv = a0;
#pragma unroll
      for (int i = 1; i <= MO; ++i) sv = offd == (double)i ? w[i] : sv;
#pragma unroll
      for (int i = 1; i <= MO; ++i) w[i] = offd == (double)i ? a0 : w[i];
      w[0] = sv;
    }
    const double t = w[0];
    const double nt = -t;
    x[j * xs] = t;
    w[1] = fma(f0.y, nt, w[1]);
    w[2] = fma(f1.x, nt, w[2]);
    w[3] = fma(f1.y, nt, w[3]);
    w[4] = fma(f2.x, nt, w[4]);
    w[0] = w[1];
    w[1] = w[2];
    w[2] = w[3];
    w[3] = w[4];
    w[4] = j + 5 < N ? x[(j + 5) * xs] : 0.0;
  }
}

// U sweep (DTBSV upper, no transpose) for rows jtop .. jtop-ncol+1: v[d]
// holds the pending row jtop-d for d = 0..KUE.  F: column cl (row jtop-cl)
// at F[QN*cl]: (u0, 1/u0) (u1, u2) ...  Superdiagonals beyond KUE are exact
// zeros and are skipped.
template <int KUE>
__device__ __forceinline__ void bwd_chunk(const double2* __restrict__ F, double* __restrict__ x, int xs, int jtop, int ncol,
                                          double (&v)[9]) {
  constexpr int W = KUE + 1;
  constexpr int QN = 1 + (KUE + 1) / 2;
  // windows wider than 7 rows run the shifting loop only (register budget)
  constexpr bool ROT = KUE <= 6;
  int cl = 0;
  if (ROT && ncol >= W) {
    // register rings of W named slots: the factors of row jtop-cl-2 and the
    // row entering the window two rows later are in flight while row
    // jtop-cl is solved (over-reads past the chunk as in fwd_chunk).  Rows below 0 enter as copies of row 0: they are
    // only ever updated, never solved or stored.
    double2 u[W][QN];
    double xi[W];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
      for (int q = 0; q < QN; ++q) u[p][q] = F[p * QN + q];
      const int jn = jtop - p - W;
      xi[p] = x[(jn >= 0 ? jn : 0) * xs];
    }
    for (; cl + W <= ncol; cl += W) {
      const int j = jtop - cl;
#pragma unroll
      for (int s = 0; s < W; ++s) {
        {
#pragma unroll
          for (int q = 0; q < QN; ++q) u[(s + 2) % W][q] = F[(cl + s + 2) * QN + q];
          const int jn = j - s - 2 - W;
          xi[(s + 2) % W] = x[(jn >= 0 ? jn : 0) * xs];
        }
        const double t = v[s % W] * u[s][0].y;
        x[(j - s) * xs] = t;
        const double nt = -t;
#pragma unroll
        for (int d = 1; d <= KUE; ++d) {
          const double ud = (d & 1) ? u[s][(d + 1) / 2].x : u[s][d / 2].y;
          v[(s + d) % W] = fma(nt, ud, v[(s + d) % W]);
        }
        v[s % W] = xi[s];
      }
    }
  }
  for (; cl < ncol; ++cl) {   // remainder of the last chunk: explicit shifts
    const int j = jtop - cl;
    double2 u[QN];
#pragma unroll
    for (int q = 0; q < QN; ++q) u[q] = F[cl * QN + q];
    const double t = v[0] * u[0].y;
    x[j * xs] = t;
    const double nt = -t;
#pragma unroll
    for (int d = 1; d <= KUE; ++d) {
      const double ud = (d & 1) ? u[(d + 1) / 2].x : u[d / 2].y;
      v[d] = fma(nt, ud, v[d]);
    }
#pragma unroll
    for (int d = 0; d < KUE; ++d) v[d] = v[d + 1];
    v[KUE] = j - W >= 0 ? x[(j - W) * xs] : 0.0;
  }
}

// Straight-line forms of the two sweeps for a whole chunk of CS columns
// (rows), CS a multiple of the window length: with no loop back-edge and no
// clamped row indices the compiler is free to issue every LDS read as early
// as the registers allow.  Callers use them only where every row entering
// the window lies inside the matrix (fwd: j0 + CS + 4 <= N - 1; bwd:
// jtop - CS + 1 - (KUE + 1) >= 0); the chunk loops above cover the rest.
template <int MO, int CS>
__device__ __forceinline__ void fwd_chunk_s(const double2* __restrict__ F, double* __restrict__ x, int xs, int j0,
                                            double (&w)[5]) {
  static_assert(CS % 5 == 0, "whole windows only");
  constexpr int D = 3;   // steps of lookahead for the LDS reads
  double* xb = x + (size_t)j0 * xs;
  double2 f[CS][3];
  double xi[CS];
#pragma unroll
  for (int p = 0; p < D && p < CS; ++p) {
#pragma unroll
    for (int q = 0; q < 3; ++q) f[p][q] = F[p * 3 + q];
    xi[p] = xb[(p + 5) * xs];
  }
#pragma unroll
  for (int s = 0; s < CS; ++s) {
    if (s + D < CS) {
#pragma unroll
      for (int q = 0; q < 3; ++q) f[s + D][q] = F[(s + D) * 3 + q];
      xi[s + D] = xb[(s + D + 5) * xs];
    }
    // keep the reads issued here, ahead of the arithmetic that consumes
    // older ones
    __builtin_amdgcn_sched_barrier(0);
    const double2 f0 = f[s][0], f1 = f[s][1], f2 = f[s][2];
    if (MO > 0) {
      const double offd = f0.x;
      const double a0 = w[s % 5];
      double sv = a0;
#pragma unroll
      for (int i = 1; i <= MO; ++i) sv = offd == (double)i ? w[(s + i) % 5] : sv;
#pragma unroll
      for (int i = 1; i <= MO; ++i) w[(s + i) % 5] = offd == (double)i ? a0 : w[(s + i) % 5];
      w[s % 5] = sv;
    }
    const double t = w[s % 5];
    const double nt = -t;
    xb[s * xs] = t;
    w[(s + 1) % 5] = fma(f0.y, nt, w[(s + 1) % 5]);
    w[(s + 2) % 5] = fma(f1.x, nt, w[(s + 2) % 5]);
    w[(s + 3) % 5] = fma(f1.y, nt, w[(s + 3) % 5]);
    w[(s + 4) % 5] = fma(f2.x, nt, w[(s + 4) % 5]);
    w[s % 5] = xi[s];
  }
}

template <int KUE, int CS>
__device__ __forceinline__ void bwd_chunk_s(const double2* __restrict__ F, double* __restrict__ x, int xs, int jtop,
                                            double (&v)[9]) {
  constexpr int W = KUE + 1;
  constexpr int QN = 1 + (KUE + 1) / 2;
  static_assert(CS % W == 0, "whole windows only");
  constexpr int D = 3;
  double* xb = x + (size_t)jtop * xs;
  double2 u[CS][QN];
  double xi[CS];
#pragma unroll
  for (int p = 0; p < D && p < CS; ++p) {
#pragma unroll
    for (int q = 0; q < QN; ++q) u[p][q] = F[p * QN + q];
    xi[p] = xb[-(p + W) * xs];
  }
#pragma unroll
  for (int s = 0; s < CS; ++s) {
    if (s + D < CS) {
#pragma unroll
      for (int q = 0; q < QN; ++q) u[s + D][q] = F[(s + D) * QN + q];
      xi[s + D] = xb[-(s + D + W) * xs];
    }
    __builtin_amdgcn_sched_barrier(0);
    const double t = v[s % W] * u[s][0].y;
    xb[-s * xs] = t;
    const double nt = -t;
#pragma unroll
    for (int d = 1; d <= KUE; ++d) {
      const double ud = (d & 1) ? u[s][(d + 1) / 2].x : u[s][d / 2].y;
      v[(s + d) % W] = fma(nt, ud, v[(s + d) % W]);
    }
    v[s % W] = xi[s];
  }
}

// chunk lengths of a segment (make_seg) as compile-time values
constexpr int seg_cap(int nsys, int gp) {
  return step_bufd2(gp) / (nsys * gp) - 1 < STEP_BLKCAP ? step_bufd2(gp) / (nsys * gp) - 1 : STEP_BLKCAP;
}
constexpr int seg_cs_fwd(int cap) { return (cap / FW_Q) / 5 * 5; }
constexpr int seg_cs_bwd(int cap, int kue) { return (cap / (1 + (kue + 1) / 2)) / (kue + 1) * (kue + 1); }

// a chunk of either sweep: the straight-line form where it applies
template <int MO, int CAP>
__device__ __forceinline__ void fwd_any(const double2* F, double* x, int xs, int j0, int ncol, int cs, int N,
                                        double (&w)[5]) {
  constexpr int CS = seg_cs_fwd(CAP);
  if constexpr (MO <= 2 && CS >= 5) {
    if (ncol == CS && cs == CS && j0 + CS + 4 <= N - 1) {
      fwd_chunk_s<MO, CS>(F, x, xs, j0, w);
      return;
    }
  }
  fwd_chunk<MO>(F, x, xs, j0, ncol, N, w);
}

template <int KUE, int CAP>
__device__ __forceinline__ void bwd_any(const double2* F, double* x, int xs, int jtop, int ncol, int cs,
                                        double (&v)[9]) {
  constexpr int CS = seg_cs_bwd(CAP, KUE);
  if constexpr (KUE >= 3 && KUE <= 6 && CS >= KUE + 1) {
    if (ncol == CS && cs == CS && jtop - CS + 1 - (KUE + 1) >= 0) {
      bwd_chunk_s<KUE, CS>(F, x, xs, jtop, v);
      return;
    }
  }
  bwd_chunk<KUE>(F, x, xs, jtop, ncol, v);
}

// One solve phase (pressure: LNp on component 3; Helmholtz: LDp, LDm, LDz on
// components 0, 1, 2): L sweep then U sweep, chunk by chunk, with wave 1
// streaming the factor ring.  gbase: index of the phase's first chunk.
template <int GP, int NLW, bool PRESSURE>
__device__ __forceinline__ void solve_phase(const Lds& L, const Bufs& B, const Stream& st, const Seg& SF,
                                            const Seg& SB, int kue, int mo, int gbase, int slot, int p0, int np,
                                            int P, int K, int N) {
  constexpr bool pressure = PRESSURE;
  constexpr int CAP = seg_cap(PRESSURE ? 1 : 3, GP);
  const int tid = threadIdx.x;
  const int nsolv = 2 * L.GH * SF.nsys;
  const bool solver = tid < nsolv;
  const bool loader = tid >= 64;
  const int lane = tid - 64;
  const int sys = solver ? tid / (2 * L.GH) : 0;
  const int rem = tid - sys * 2 * L.GH;
  const int lh = rem >> 1, part = rem & 1;
  const Slot S = slot_of(lh, GP, np, p0, K);
  const bool act = solver && S.valid && !(pressure && S.h00);
  const int comp = pressure ? 3 : sys;
  double* x = (double*)&Vat(L, comp, 0, lh, N) + part;
  const int xs = 2 * L.GH;
  const int blk = sys * GP + S.pl;

  double w[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) w[i] = 0.0;
  if (act) {
#pragma unroll
    for (int i = 0; i < 5; ++i) w[i] = i < N ? x[i * xs] : 0.0;
  }
  const int BSF = SF.cs * SF.qn + 1;
  for (int c = 0; c < SF.nch; ++c) {
    const int g = gbase + c;
    if (loader) ld_dma<GP, NLW>(L, B, st, g + 1, slot, p0, np, P, N, lane);
    if (act) {
      const double2* F = L.FC + (size_t)(g & 1) * L.bufd2 + (size_t)blk * BSF;
      const int j0 = c * SF.cs, ncol = seg_ncol(SF, c, N);
      switch (mo) {
        case 0: fwd_any<0, CAP>(F, x, xs, j0, ncol, SF.cs, N, w); break;
        case 1: fwd_any<1, CAP>(F, x, xs, j0, ncol, SF.cs, N, w); break;
        case 2: fwd_any<2, CAP>(F, x, xs, j0, ncol, SF.cs, N, w); break;
        case 3: fwd_any<3, CAP>(F, x, xs, j0, ncol, SF.cs, N, w); break;
        default: fwd_any<4, CAP>(F, x, xs, j0, ncol, SF.cs, N, w); break;
      }
    }
    __syncthreads();
  }
  double v[9];
#pragma unroll
  for (int d = 0; d < 9; ++d) v[d] = 0.0;
  if (act) {
    v[0] = w[0];
#pragma unroll
    for (int d = 1; d < 9; ++d) {
      const int jn = N - 1 - d;
      const double xv = x[(jn >= 0 ? jn : 0) * xs];
      v[d] = (d <= kue && jn >= 0) ? xv : 0.0;
    }
  }
  const int BSB = SB.cs * SB.qn + 1;
  for (int c = 0; c < SB.nch; ++c) {
    const int g = gbase + SF.nch + c;
    if (loader) ld_dma<GP, NLW>(L, B, st, g + 1, slot, p0, np, P, N, lane);
    if (act) {
      const double2* F = L.FC + (size_t)(g & 1) * L.bufd2 + (size_t)blk * BSB;
      const int jtop = N - 1 - c * SB.cs, ncol = seg_ncol(SB, c, N);
      switch (kue) {
        case 1: bwd_any<1, CAP>(F, x, xs, jtop, ncol, SB.cs, v); break;
        case 2: bwd_any<2, CAP>(F, x, xs, jtop, ncol, SB.cs, v); break;
        case 3: bwd_any<3, CAP>(F, x, xs, jtop, ncol, SB.cs, v); break;
        case 4: bwd_any<4, CAP>(F, x, xs, jtop, ncol, SB.cs, v); break;
        case 5: bwd_any<5, CAP>(F, x, xs, jtop, ncol, SB.cs, v); break;
        case 6: bwd_any<6, CAP>(F, x, xs, jtop, ncol, SB.cs, v); break;
        case 7: bwd_any<7, CAP>(F, x, xs, jtop, ncol, SB.cs, v); break;
        default: bwd_any<8, CAP>(F, x, xs, jtop, ncol, SB.cs, v); break;
      }
    }
    __syncthreads();
  }
}

__device__ __forceinline__ double pr0_of(const double2* uz0, size_t hs, int N, const double* dr1c) {
  // vel_Pr0 = dot_product(uz%Re(N-KL:N,0), dr1(:,1)) / (-2)  (velocity.f90:356-357)
  double s = 0.0;
#pragma unroll
  for (int i = 0; i <= KL; ++i) s = s + uz0[(size_t)(N - 1 - KL + i) * hs].x * dr1c[i];
  return s / (-2.0);
}

// IT: (radial point, harmonic) items per thread in the parallel stages,
// N * 2 * GP <= IT * THR; THR: threads of the workgroup (wave 0 solves,
// the others stream factors during the solves).
template <int GP, int IT, int THR>
__device__ __forceinline__ void step_body(const Geo& G, const Bufs& B, const StepTiles& ST, int slot0) {
  const int slot = slot_of(B, slot0);
  SlotCtl* ctl = &B.ctl[slot];
  if (!ctl->active) return;
  // XCD-aware group order: consecutive workgroup ids go to different XCDs,
  // so give each XCD a contiguous range of pairs (neighbouring harmonics
  // share cache lines of the spectral fields).
  const int per = (ST.ngroups + 7) / 8;
  const int grp = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (grp >= ST.ngroups) return;
  const int N = G.N, P = G.P, K = G.K;
  const int p0 = grp * GP;
  const int np = min(GP, P - p0);
  const int tid = threadIdx.x;
  const int lane = tid - 64;
  const bool loader = tid >= 64;
  const size_t hs = G.Hp, csz = G.coll;
  const MeshOff mo = MeshOff::make(N);
  const double* __restrict__ me = G.mesh;
  const double* dr1c = me + mo.dr1 + (KL + 1);
  const SlotParams prm = B.prm[slot];
  const bool corr = ctl->mode == MODE_CORR;
  const size_t s3 = (size_t)slot * 3 * csz;
  // distinct buffers (the ring of velocity roles; uc and ub coincide only in
  // the predictor, where both are read-only), so restrict-qualified: loads
  // may move past the stores of un and om
  const double2* __restrict__ uc = B.U[ctl->ucur] + s3;                 // current iterate
  const double2* __restrict__ ub = B.U[corr ? ctl->ubase : ctl->ucur] + s3;   // u_ (vel_predictor copies u -> u_)
  double2* __restrict__ un = B.U[ctl->unext] + s3;
  const double2* __restrict__ Nc = B.Nraw[ctl->ncur] + s3;
  const double2* __restrict__ Np = B.Nraw[ctl->npred] + s3;
  double2* __restrict__ om = B.om + s3;
  const bool want_diag = !corr && ctl->diag_every > 0 && (ctl->step % ctl->diag_every) == 0;

  extern __shared__ double2 lds2[];
  Lds L;
  L.GH = 2 * GP;
  L.bufd2 = ST.bufd2;
  unsigned char* base = (unsigned char*)lds2;
  L.V = (double2*)base;
  base += (size_t)4 * N * L.GH * sizeof(double2);
  L.FC = (double2*)base;
  L.E = (double*)base;
  const size_t fc_bytes = ((size_t)2 * L.bufd2 + FC_TAIL) * sizeof(double2);
  const size_t e_bytes = (size_t)3 * N * L.GH * sizeof(double);
  base += (fc_bytes > e_bytes ? fc_bytes : e_bytes);
  L.AC = (double*)base;
  base += (size_t)L.GH * 8 * sizeof(double);
  L.RED = (double*)base;
  base += 8 * sizeof(double);
  L.META = (int*)base;
  // Stage 1's mesh weights (the radLap band, Ubase, 2r and 1/r^2 of every
  // row: 12 N doubles) staged in LDS behind E, in the factor ring's second
  // buffer, which the first solve fills only after stage 2.  Read from
  // global memory instead, the stencil's nine weight loads per item each
  // wait a full cache round trip behind the other workgroups' HBM traffic.
  // Four pairs at N <= 64 always have the room (E + weights = 288 N bytes
  // against the ring's 19 KiB).
  constexpr bool MWL = GP == 4 && IT == 2;
  double* MW = (double*)((unsigned char*)L.FC + e_bytes);

  const double th = prm.implicit, th1 = 1.0 - prm.implicit;
  const double c1 = 1.0 / prm.dt;
  const double c2 = (1.0 - prm.implicit) * prm.nu / prm.Re;
  const double wnorm = 4.0 * D_PI * D_PI / prm.alpha;
  const int NI = N * L.GH;
  // items it = tid + q * THR of the parallel stages: radial point it / GH,
  // harmonic slot it % GH, the same slot for all of a thread's items
  static_assert(THR % (2 * GP) == 0, "threads must cover whole radial rows");
  const int lh0 = tid % (2 * GP);
  const Slot S0 = slot_of(lh0, GP, np, p0, K);

  // ---- stage 0: u_ of the group into V[0..2]; band metadata of the group --
  if (tid < 2) {
    const int m0 = tid == 0 ? MAT_LNP : MAT_LDP, nm = tid == 0 ? 1 : 3;
    int kue = 1, mxo = 0;
    for (int mi = 0; mi < nm; ++mi)
      for (int pl = 0; pl < np; ++pl) {
        const unsigned char* mt = meta_block(B, slot, m0 + mi, p0 + pl, P);
        kue = max(kue, (int)mt[0]);
        mxo = max(mxo, (int)mt[1]);
      }
    L.META[2 * tid] = kue;
    L.META[2 * tid + 1] = mxo;
  }
  // all global inputs of this thread's items in one batch: u_ (staged into
  // V[0..2] for the stencils), the current iterate (a separate buffer only
  // for the corrector), N and, for the corrector, N_
  // (unconditional loads: in the predictor the current iterate is u_ and
  // N_ is not used, and the repeated addresses hit the cache)
  // the weights first (one batch, ahead of the state loads)
  constexpr int MWT = (12 * 64 + THR - 1) / THR;
  double mwv[MWT];
  if constexpr (MWL) {
#pragma unroll
    for (int u = 0; u < MWT; ++u) {
      const int i = tid + u * THR, r = i - 9 * N;
      const int src = i < 9 * N ? mo.lapb + i
                    : r < N ? mo.Ubase + r : r < 2 * N ? mo.twor + (r - N) : mo.rinv2 + (r - 2 * N);
      mwv[u] = i < 12 * N ? me[src] : 0.0;
    }
  }
  double2 vb[IT][3], vc[IT][3], vn[IT][3], vp[IT][3];
  const double2* ucl = corr ? uc : ub;
  const double2* npl = corr ? Np : Nc;
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int it = tid + q * THR;
    const int n = it / L.GH, lh = lh0;
    const Slot& S = S0;
    const bool ok = it < NI && S.valid;
    const size_t o = ok ? (size_t)n * hs + S.h : 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      vb[q][c] = ub[c * csz + o];
      vn[q][c] = Nc[c * csz + o];
      vc[q][c] = ucl[c * csz + o];
      vp[q][c] = npl[c * csz + o];
    }
    if (!ok) {
#pragma unroll
      for (int c = 0; c < 3; ++c) vb[q][c] = make_double2(0.0, 0.0);
    }
  }
  if constexpr (MWL) {
#pragma unroll
    for (int u = 0; u < MWT; ++u)
      if (tid + u * THR < 12 * N) MW[tid + u * THR] = mwv[u];
  }
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int it = tid + q * THR;
    if (it >= NI) continue;
    const int n = it / L.GH, lh = lh0;
#pragma unroll
    for (int c = 0; c < 3; ++c) Vat(L, c, n, lh, N) = vb[q][c];
  }
  __syncthreads();

  // the factor stream of the two solve phases
  const int kueP = L.META[0], moP = L.META[1], kueH = L.META[2], moH = L.META[3];
  Stream st;
  st.s[0] = make_seg(MAT_LNP, 1, 1, kueP, N, GP, L.bufd2);
  st.s[1] = make_seg(MAT_LNP, 1, 0, kueP, N, GP, L.bufd2);
  st.s[2] = make_seg(MAT_LDP, 3, 1, kueH, N, GP, L.bufd2);
  st.s[3] = make_seg(MAT_LDP, 3, 0, kueH, N, GP, L.bufd2);
  st.total = st.s[0].nch + st.s[1].nch + st.s[2].nch + st.s[3].nch;

  // ---- stage 1: effective N and RHS = Lt u_ + N (tim_meshmult) -----------
  // Results are held in registers until every thread has read its stencil
  // from u_ (V[0..2]), then written over it.
  double2 res[IT][3];
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int it = tid + q * THR;
    res[q][0] = res[q][1] = res[q][2] = make_double2(0.0, 0.0);
    if (it >= NI) continue;
    const int n = it / L.GH, lh = lh0;
    const Slot& S = S0;
    if (!S.valid) {
      if (want_diag)
        for (int c = 0; c < 3; ++c) L.E[((size_t)c * N + n) * L.GH + lh] = 0.0;
      continue;
    }
    const size_t o = (size_t)n * hs + S.h;
    const double ak = prm.alpha * (double)S.k;
    const double a = ak * (MWL ? MW[9 * N + n] : me[mo.Ubase + n]);
    const double b = MWL ? MW[10 * N + n] : me[mo.twor + n];
    const double2 ucr = vc[q][0], uct = vc[q][1], ucz = vc[q][2];
    double2 nr = vn[q][0], nt = vn[q][1], nz = vn[q][2];
    nr = make_double2(nr.x + a * ucr.y, nr.y - a * ucr.x);
    nt = make_double2(nt.x + a * uct.y, nt.y - a * uct.x);
    nz = make_double2((nz.x + a * ucz.y) + b * ucr.x, (nz.y - a * ucz.x) + b * ucr.y);
    double pr_c = 0.0;
    if (S.h00 && prm.const_flux) {
      pr_c = pr0_of(uc + 2 * csz, hs, N, dr1c);
      nz.x = nz.x + (4.0 * pr_c) / prm.Re;
    }
    if (corr) {
      const double2 bur = Vat(L, 0, n, lh, N);
      const double2 but = Vat(L, 1, n, lh, N);
      const double2 buz = Vat(L, 2, n, lh, N);
      double2 pr = vp[q][0], pt = vp[q][1], pz = vp[q][2];
      pr = make_double2(pr.x + a * bur.y, pr.y - a * bur.x);
      pt = make_double2(pt.x + a * but.y, pt.y - a * but.x);
      pz = make_double2((pz.x + a * buz.y) + b * bur.x, (pz.y - a * buz.x) + b * bur.y);
      if (S.h00 && prm.const_flux) pz.x = pz.x + (4.0 * pr0_of(ub + 2 * csz, hs, N, dr1c)) / prm.Re;
      nr = make_double2(th * nr.x + th1 * pr.x, th * nr.y + th1 * pr.y);
      nt = make_double2(th * nt.x + th1 * pt.x, th * nt.y + th1 * pt.y);
      nz = make_double2(th * nz.x + th1 * pz.x, th * nz.y + th1 * pz.y);
    }
    double2 rp = make_double2(nr.x - nt.y, nr.y + nt.x);
    double2 rm = make_double2(nr.x + nt.y, nr.y - nt.x);
    double2 rz = nz;
    const double r2i = MWL ? MW[11 * N + n] : me[mo.rinv2 + n];
#pragma unroll
    for (int dd = -KL; dd <= KL; ++dd) {
      // rows outside the matrix contribute an exact zero (weight 0 on a
      // clamped row), so the loads need no branch
      const int jr = n + dd;
      const bool in = jr >= 0 && jr <= N - 1;
      const int j = in ? jr : n;
      const double lap_ = MWL ? MW[(KL - dd) * N + j] : me[mo.lapb + (KL - dd) * N + j];
      const double lap = in ? lap_ : 0.0;
      double ap = c2 * lap, am = c2 * lap, az = c2 * lap;
      if (dd == 0) {
        ap = (ap + c2 * diag_d(r2i, G.Mp, S.m, prm.alpha, S.k, 1)) + c1;
        am = (am + c2 * diag_d(r2i, G.Mp, S.m, prm.alpha, S.k, -1)) + c1;
        az = (az + c2 * diag_d(r2i, G.Mp, S.m, prm.alpha, S.k, 0)) + c1;
      }
      const double2 br = Vat(L, 0, j, lh, N);
      const double2 bt = Vat(L, 1, j, lh, N);
      const double2 bz = Vat(L, 2, j, lh, N);
      rp.x = rp.x + ap * (br.x - bt.y);
      rp.y = rp.y + ap * (br.y + bt.x);
      rm.x = rm.x + am * (br.x + bt.y);
      rm.y = rm.y + am * (br.y - bt.x);
      rz.x = rz.x + az * bz.x;
      rz.y = rz.y + az * bz.y;
    }
    res[q][0] = rp;
    res[q][1] = rm;
    res[q][2] = rz;
    if (want_diag) {
      L.E[((size_t)0 * N + n) * L.GH + lh] = wnorm * (ucr.x * ucr.x + ucr.y * ucr.y);
      L.E[((size_t)1 * N + n) * L.GH + lh] = wnorm * (uct.x * uct.x + uct.y * uct.y);
      L.E[((size_t)2 * N + n) * L.GH + lh] = wnorm * (ucz.x * ucz.x + ucz.y * ucz.y);
      if (S.h00 && n == 0) {
        double s = 0.0;
        for (int i = 0; i <= KL; ++i) s = s + uc[2 * csz + (size_t)i * hs].x * me[mo.dr0 + i];
        double f = 0.0;
        for (int i = 0; i <= KL; ++i) f = f + uc[2 * csz + (size_t)(N - 1 - KL + i) * hs].x * dr1c[i];
        DiagRec* rec = B.diag + (size_t)slot * ctl->diag_cap + (ctl->ndiag % ctl->diag_cap);
        rec->step = ctl->step;
        rec->t = ctl->t;
        if (prm.const_flux) {
          rec->Ub = pr0_of(uc + 2 * csz, hs, N, dr1c);
        } else {   // io.f90:630
          double q = 0.0;
          for (int i = 0; i < N; ++i) q = q + uc[2 * csz + (size_t)i * hs].x * me[mo.intrdr + i];
          rec->Ub = 0.5 + 2.0 * q;
        }
        rec->Uc = 1.0 + s;
        rec->Ufr = sqrt(fabs((f - 2.0) / prm.Re));
      }
    }
    if (S.h00 && n == 0) ctl->Pr0_hpf = pr_c;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < IT; ++q) {
    const int it = tid + q * THR;
    if (it >= NI) continue;
    const int n = it / L.GH, lh = lh0;
    Vat(L, 0, n, lh, N) = res[q][0];
    Vat(L, 1, n, lh, N) = res[q][1];
    Vat(L, 2, n, lh, N) = res[q][2];
  }
  // energy partials (var_coll_norm, variables.f90:732-735), n ascending
  if (want_diag && tid < 3 * L.GH) {
    const int c = tid / L.GH, lh = tid - c * L.GH;
    const Slot S = slot_of(lh, GP, np, p0, K);
    if (S.valid) {
      double bsum = 0.0;
      for (int n = 0; n < N; ++n) bsum = bsum + L.E[((size_t)c * N + n) * L.GH + lh] * me[mo.intrdr + n];
      if (S.h00) bsum = 0.5 * bsum;
      B.diagb[(size_t)slot * 3 * hs + (size_t)c * hs + S.h] = bsum;
    }
  }
  __syncthreads();

  // ---- stage 2: pressure RHS = div(pm2rt(R+, R-), Rz) (vel_adjPPE(1)) ----
  // the first factor chunk streams in meanwhile (the energy terms that
  // shared its memory have been summed)
  if (loader) ld_dma<GP, THR / 64 - 1>(L, B, st, 0, slot, p0, np, P, N, lane);
  for (int it = tid; it < NI; it += THR) {
    const int n = it / L.GH, lh = lh0;
    const Slot& S = S0;
    double2 dv = make_double2(0.0, 0.0);
    if (S.valid && !S.h00 && n != 0 && n != N - 1) {
      const double mr = (me[mo.rinv + n] * (double)G.Mp) * (double)S.m;
      const double ak = prm.alpha * (double)S.k;
      double2 d = make_double2(0.0, 0.0);
#pragma unroll
      for (int dd = -KL; dd <= KL; ++dd) {
        const int jr = n + dd;
        const bool in = jr >= 0 && jr <= N - 1;
        const int j = in ? jr : n;
        const double2 a = Vat(L, 0, j, lh, N), bm = Vat(L, 1, j, lh, N);
        const double w_ = me[mo.dr1b + (KL - dd) * N + j];
        const double wgt = in ? w_ : 0.0;
        d.x = d.x + wgt * (0.5 * (a.x + bm.x));
        d.y = d.y + wgt * (0.5 * (a.y + bm.y));
      }
      const double2 a = Vat(L, 0, n, lh, N), bm = Vat(L, 1, n, lh, N), z = Vat(L, 2, n, lh, N);
      const double2 rr = make_double2(0.5 * (a.x + bm.x), 0.5 * (a.y + bm.y));
      const double2 tt = make_double2(0.5 * (a.y - bm.y), -0.5 * (a.x - bm.x));
      const double ri = me[mo.rinv + n];
      dv.x = ((rr.x * ri + d.x) - tt.y * mr) - z.y * ak;
      dv.y = ((rr.y * ri + d.y) + tt.x * mr) + z.x * ak;
    }
    Vat(L, 3, n, lh, N) = dv;
  }
  __syncthreads();

  // ---- stage 3: pressure solve with LNp (tim_lumesh_invert(1,...)) -------
  solve_phase<GP, THR / 64 - 1, true>(L, B, st, st.s[0], st.s[1], kueP, moP, 0, slot, p0, np, P, K, N);

  // ---- stage 4: subtract grad p (rt2pm), zero the BC rows ----------------
  for (int it = tid; it < NI; it += THR) {
    const int n = it / L.GH, lh = lh0;
    const Slot& S = S0;
    if (!S.valid) continue;
    if (n == 0 || n == N - 1) {
      for (int c = 0; c < 3; ++c) Vat(L, c, n, lh, N) = make_double2(0.0, 0.0);
      continue;
    }
    const double mr = (me[mo.rinv + n] * (double)G.Mp) * (double)S.m;
    const double ak = prm.alpha * (double)S.k;
    double2 gr = make_double2(0.0, 0.0);
#pragma unroll
    for (int dd = -KL; dd <= KL; ++dd) {
      const int jr = n + dd;
      const bool in = jr >= 0 && jr <= N - 1;
      const int j = in ? jr : n;
      const double w_ = me[mo.dr1b + (KL - dd) * N + j];
      const double wgt = in ? w_ : 0.0;
      const double2 pj = Vat(L, 3, j, lh, N);
      gr.x = gr.x + wgt * pj.x;
      gr.y = gr.y + wgt * pj.y;
    }
    const double2 pv = Vat(L, 3, n, lh, N);
    const double2 gt = make_double2(-(pv.y * mr), pv.x * mr);
    const double2 gz = make_double2(-(pv.y * ak), pv.x * ak);
    double2& r0 = Vat(L, 0, n, lh, N);
    double2& r1 = Vat(L, 1, n, lh, N);
    double2& r2 = Vat(L, 2, n, lh, N);
    r0 = make_double2(r0.x - (gr.x - gt.y), r0.y - (gr.y + gt.x));
    r1 = make_double2(r1.x - (gr.x + gt.y), r1.y - (gr.y - gt.x));
    r2 = make_double2(r2.x - gz.x, r2.y - gz.y);
  }
  __syncthreads();

  // ---- stage 5: Helmholtz solves LDp, LDm, LDz ---------------------------
  solve_phase<GP, THR / 64 - 1, false>(L, B, st, st.s[2], st.s[3], kueH, moH, st.s[0].nch + st.s[1].nch, slot, p0,
                                       np, P, K, N);

  // ---- stage 6: influence-matrix correction (vel_adjPPE(2)) --------------
  // the 4x4 inverses of the group's harmonics into LDS, one entry per thread
  // (the factor ring is free after the solves)
  double* AIl = (double*)L.FC;
  for (int e = tid; e < 16 * L.GH; e += THR) {
    const int lh = e >> 4, ij = e & 15;
    const Slot S = slot_of(lh, GP, np, p0, K);
    AIl[e] = (S.valid && !S.h00) ? B.Ainv[(size_t)slot * 16 * G.H + (size_t)ij * G.H + S.h] : 0.0;
  }
  __syncthreads();
  if (tid < L.GH) {
    const int lh = tid;
    const Slot S = slot_of(lh, GP, np, p0, K);
    double aR[4] = {0, 0, 0, 0}, aI[4] = {0, 0, 0, 0};
    if (S.valid && !S.h00) {
      double2 wp[KL + 1], wm[KL + 1], wz[KL + 1];
      for (int i = 0; i <= KL; ++i) {
        wp[i] = Vat(L, 0, N - 1 - KL + i, lh, N);
        wm[i] = Vat(L, 1, N - 1 - KL + i, lh, N);
        wz[i] = Vat(L, 2, N - 1 - KL + i, lh, N);
      }
      double BRe[4], BIm[4];
      eval_bc5(wp, wm, wz, dr1c, me[mo.rinv + N - 1], (double)(S.m * G.Mp), prm.alpha * (double)S.k, BRe, BIm);
      const double* Ai = AIl + lh * 16;
      for (int i = 0; i < 4; ++i) {
        double sR = 0.0, sI = 0.0;
        for (int j = 0; j < 4; ++j) {
          const double aij = Ai[i * 4 + j];
          sR = sR + aij * BRe[j];
          sI = sI + aij * BIm[j];
        }
        aR[i] = -sR;
        aI[i] = -sI;
      }
    }
    for (int i = 0; i < 4; ++i) {
      L.AC[lh * 8 + i] = aR[i];
      L.AC[lh * 8 + 4 + i] = aI[i];
    }
  }
  __syncthreads();
  for (int it = tid; it < NI; it += THR) {
    const int n = it / L.GH, lh = lh0;
    const Slot& S = S0;
    if (!S.valid) continue;
    double2 up = Vat(L, 0, n, lh, N), um = Vat(L, 1, n, lh, N), uz = Vat(L, 2, n, lh, N);
    if (!S.h00) {
      const double* aR = L.AC + lh * 8;
      const double* aI = aR + 4;
      const double* U = B.Uinf + (size_t)slot * 6 * N * P + (size_t)n * P + S.p;
      const size_t us = (size_t)N * P;
      const double U1 = U[0], U2 = U[us], U3 = U[2 * us], U4 = U[3 * us], U5 = U[4 * us];
      const double U6 = U[5 * us] * (prm.alpha * (double)S.k);
      up.x = up.x + aR[0] * U1; up.y = up.y + aI[0] * U1;
      um.x = um.x + aR[1] * U2; um.y = um.y + aI[1] * U2;
      uz.x = uz.x - aI[2] * U3; uz.y = uz.y + aR[2] * U3;
      up.x = up.x + aR[3] * U4; up.y = up.y + aI[3] * U4;
      um.x = um.x + aR[3] * U5; um.y = um.y + aI[3] * U5;
      uz.x = uz.x - aI[3] * U6; uz.y = uz.y + aR[3] * U6;
    }
    // vel_pm2rt
    Vat(L, 0, n, lh, N) = make_double2(0.5 * (up.x + um.x), 0.5 * (up.y + um.y));
    Vat(L, 1, n, lh, N) = make_double2(0.5 * (up.y - um.y), -0.5 * (up.x - um.x));
    Vat(L, 2, n, lh, N) = uz;
  }
  // vel_adjustFlux(1) (velocity.f90:170-175) on harmonic (0,0): n ascending,
  // the quadrature weights staged in LDS first
  const bool has00 = p0 == 0;
  double* QW = AIl + 16 * L.GH;
  if (has00)
    for (int n = tid; n < N; n += THR) QW[n] = me[mo.intrdr + n];
  __syncthreads();
  if (has00 && tid == 0) {
    double d3 = 0.0, Pr0 = ctl->Pr0_hpf;
    if (prm.const_flux) {
      double s = 0.0;
      for (int n = 0; n < N; ++n) s = s + Vat(L, 2, n, 0, N).x * QW[n];
      const double d2 = 2.0 * s;
      d3 = -d2 / ctl->flux_d1;
      Pr0 = Pr0 + (d3 * prm.Re) / 4.0;
    }
    L.RED[0] = d3;
    ctl->Pr0 = Pr0;
  }
  __syncthreads();
  double dmax = 0.0;
  for (int it = tid; it < NI; it += THR) {
    const int n = it / L.GH, lh = lh0;
    const Slot& S = S0;
    if (!S.valid) continue;
    double2 ur = Vat(L, 0, n, lh, N), ut = Vat(L, 1, n, lh, N), uz = Vat(L, 2, n, lh, N);
    if (S.h00) {
      if (prm.const_flux) uz.x = uz.x + L.RED[0] * B.Ui[(size_t)slot * N + n];
      ur = make_double2(0.0, 0.0);
      ut.y = 0.0;
      uz.y = 0.0;
      Vat(L, 0, n, lh, N) = ur;
      Vat(L, 1, n, lh, N) = ut;
      Vat(L, 2, n, lh, N) = uz;
    }
    const size_t o = (size_t)n * hs + S.h;
    un[o] = ur;
    un[csz + o] = ut;
    un[2 * csz + o] = uz;
    if (corr) {
      const double2 a = uc[o], bt = uc[csz + o], cz = uc[2 * csz + o];
      const double d0x = a.x - ur.x, d0y = a.y - ur.y;
      const double d1x = bt.x - ut.x, d1y = bt.y - ut.y;
      const double d2x = cz.x - uz.x, d2y = cz.y - uz.y;
      dmax = fmax(dmax, fmax(d0x * d0x + d0y * d0y, fmax(d1x * d1x + d1y * d1y, d2x * d2x + d2y * d2y)));
    }
  }
  __syncthreads();

  // ---- stage 7: curl of the new velocity for the next evaluation ---------
  for (int it = tid; it < NI; it += THR) {
    const int n = it / L.GH, lh = lh0;
    const Slot& S = S0;
    if (!S.valid) continue;
    const double mr = (me[mo.rinv + n] * (double)G.Mp) * (double)S.m;
    const double ak = prm.alpha * (double)S.k;
    const double2 vr = Vat(L, 0, n, lh, N), vt = Vat(L, 1, n, lh, N), vz = Vat(L, 2, n, lh, N);
    double2 d1 = make_double2(0.0, 0.0), d2 = make_double2(0.0, 0.0);
#pragma unroll
    for (int dd = -KL; dd <= KL; ++dd) {
      const int jr = n + dd;
      const bool in = jr >= 0 && jr <= N - 1;
      const int j = in ? jr : n;
      const double w_ = me[mo.dr1b + (KL - dd) * N + j];
      const double wgt = in ? w_ : 0.0;
      const double2 zj = Vat(L, 2, j, lh, N), tj = Vat(L, 1, j, lh, N);
      d1.x = d1.x + wgt * zj.x;
      d1.y = d1.y + wgt * zj.y;
      d2.x = d2.x + wgt * tj.x;
      d2.y = d2.y + wgt * tj.y;
    }
    const double ri = me[mo.rinv + n];
    const size_t o = (size_t)n * hs + S.h;
    om[o] = make_double2(-(vz.y * mr) + vt.y * ak, vz.x * mr - vt.x * ak);
    om[csz + o] = make_double2(-(vr.y * ak) - d1.x, vr.x * ak - d1.y);
    om[2 * csz + o] = make_double2((vt.x * ri + d2.x) + vr.y * mr, (vt.y * ri + d2.y) - vr.x * mr);
  }
  if (corr) {
    // max is order independent and the bit pattern of a non-negative double
    // orders like its value: an integer atomic max is exact.  One atomic per
    // workgroup (all of them target the slot's one word).
    for (int off = 32; off >= 1; off >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, off));
    if ((tid & 63) == 0) L.RED[1 + (tid >> 6)] = dmax;
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < THR / 64; ++w) dmax = fmax(dmax, L.RED[1 + w]);
      atomicMax(&ctl->dterr_bits, (unsigned long long)__double_as_longlong(dmax));
    }
  }
}

}  // namespace

__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3)))
k_step_g1(Geo G, Bufs B, StepTiles ST, int slot0) {
  step_body<1, 2, 128>(G, B, ST, slot0);
}
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3)))
k_step_g2(Geo G, Bufs B, StepTiles ST, int slot0) {
  step_body<2, 2, 128>(G, B, ST, slot0);
}
// 64 < N <= 128: at most four workgroups (two waves each) fit in a CU's LDS,
// so two waves per SIMD; the register budget that allows keeps the four
// items per thread out of scratch
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2)))
k_step_g2x(Geo G, Bufs B, StepTiles ST, int slot0) {
  step_body<2, 4, 128>(G, B, ST, slot0);
}
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_step_g4(Geo G, Bufs B, StepTiles ST, int slot0) {
  step_body<4, 2, 256>(G, B, ST, slot0);
}
// 64 < N <= 96: about 69 KiB of LDS at N = 96, two workgroups per CU
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_step_g4x(Geo G, Bufs B, StepTiles ST, int slot0) {
  step_body<4, 3, 256>(G, B, ST, slot0);
}

}  // namespace opf
