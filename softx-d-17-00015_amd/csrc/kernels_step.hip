// Per-state set-up kernels and the small kernels around the fused step:
// banded LU factorisation of the four matrices per (|k|, m) pair and its
// solve-ready copy, the influence vectors and 4x4 inverses of the
// influence-matrix method, the curl, imposesym, on-device corrector control
// and diagnostics sums, the granular tim_lumesh_invert, and forcing added
// to the nonlinear terms.
//
// Layout: a spectral (coll) field is stored harmonic-fastest, one complex
// double2 per (n, h): element (slot, comp, n, h) at ((slot*3+comp)*N + n)*Hp + h.
// The LU factors are stored once per (|k|, m) pair (the matrices depend on k
// only through (alpha k)^2, timestep.f90:76-77) in LAPACK band form,
// [slot][mat][pair][n][16] (see lu_block).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

// Products and sums are rounded separately, as in the reference's Fortran
// (built for x86-64 without FMA): the ill-conditioned Neumann pressure solve
// then reproduces the reference's rounding instead of amplifying the
// difference a fused multiply-add makes.
#pragma clang fp contract(off)

namespace opf {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ void harmonic_km(int h, int K, int& k, int& m) {
  // parallel.h:47-56: m = 0 holds k = 0..K-1, each m >= 1 holds k = -(K-1)..K-1
  if (h < K) {
    m = 0;
    k = h;
  } else {
    const int hh = h - K;
    m = 1 + hh / (2 * K - 1);
    k = hh % (2 * K - 1) - (K - 1);
  }
}

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// Banded matrix-vector row: sum_j A(n,j) x(j), j ascending, accumulated from
// `init` (var_coll_meshmult starts from 0, tim_meshmult from c(n)).
__device__ __forceinline__ double2 band_row(const double* band, int N, int n,
                                            const double2* x, size_t xs, double2 init) {
  double2 acc = init;
  const int j0 = n - KL < 0 ? 0 : n - KL;
  const int j1 = n + KL > N - 1 ? N - 1 : n + KL;
  for (int j = j0; j <= j1; ++j) {
    const double a = band[(KL + n - j) * N + j];
    const double2 v = x[j * xs];
    acc.x = acc.x + a * v.x;
    acc.y = acc.y + a * v.y;
  }
  return acc;
}

// ---------------------------------------------------------------------------
// banded LU solve (LAPACK DGBTRS 'N', KL = KU = 4), one system per lane.
// F: factor of this lane's matrix, column n at F[n*LUS] with the 13 rows of
// LAPACK band storage AB(1:13, n) (rows 9..12 hold the multipliers); pv:
// pivot offsets ipiv(j)-j; x: vector with stride xs, solved in place.  T is
// double or double2 (real and imaginary parts solved with the same factor,
// as tim_lumesh_invert does with two dgbtrs calls).
// ---------------------------------------------------------------------------
template <typename T> struct VT;
template <> struct VT<double> {
  __device__ static double zero() { return 0.0; }
  __device__ static void axpy(double& y, double a, double t) { y = y + a * t; }
  __device__ static double neg(double v) { return -v; }
  __device__ static double div(double v, double d) { return v / d; }
  __device__ static void sub_mul(double& y, double t, double a) { y = y - t * a; }
};
template <> struct VT<double2> {
  __device__ static double2 zero() { return make_double2(0.0, 0.0); }
  __device__ static void axpy(double2& y, double a, double2 t) {
    y.x = y.x + a * t.x;
    y.y = y.y + a * t.y;
  }
  __device__ static double2 neg(double2 v) { return make_double2(-v.x, -v.y); }
  __device__ static double2 div(double2 v, double d) { return make_double2(v.x / d, v.y / d); }
  __device__ static void sub_mul(double2& y, double2 t, double a) {
    y.x = y.x - t.x * a;
    y.y = y.y - t.y * a;
  }
};

template <typename T>
__device__ void lu_solve(const double* F, const unsigned char* pv, T* x, size_t xs, int N) {
  using V = VT<T>;
  // L solve with row interchanges: window w[i] = b(j+i)
  T w[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) w[i] = i < N ? x[i * xs] : V::zero();
  for (int j = 0; j < N - 1; ++j) {
    const int off = pv[j];
    const T a0 = w[0];
    T s = w[0];
    s = off == 1 ? w[1] : s;
    s = off == 2 ? w[2] : s;
    s = off == 3 ? w[3] : s;
    s = off == 4 ? w[4] : s;
    w[0] = s;
    w[1] = off == 1 ? a0 : w[1];
    w[2] = off == 2 ? a0 : w[2];
    w[3] = off == 3 ? a0 : w[3];
    w[4] = off == 4 ? a0 : w[4];
    const T t = V::neg(w[0]);
    const int lm = N - 1 - j < KL ? N - 1 - j : KL;
    const double* Fj = F + (size_t)j * LUS;
#pragma unroll
    for (int i = 1; i <= KL; ++i)
      if (i <= lm) V::axpy(w[i], Fj[8 + i], t);
    x[j * xs] = w[0];
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = w[i + 1];
    w[4] = j + 5 < N ? x[(j + 5) * xs] : V::zero();
  }
  // U solve (DTBSV upper, 8 superdiagonals): window v[q] = x(j-8+q)
  T v[9];
  const int jN = N - 1;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int idx = jN - 8 + q;
    v[q] = idx >= 0 ? x[idx * xs] : V::zero();
  }
  v[8] = w[0];
  for (int j = N - 1; j >= 0; --j) {
    const double* Fj = F + (size_t)j * LUS;
    v[8] = V::div(v[8], Fj[8]);
    const T t = v[8];
    x[j * xs] = t;
#pragma unroll
    for (int d = 1; d <= 8; ++d)
      if (j - d >= 0) V::sub_mul(v[8 - d], t, Fj[8 - d]);
#pragma unroll
    for (int q = 8; q >= 1; --q) v[q] = v[q - 1];
    v[0] = j - 9 >= 0 ? x[(j - 9) * xs] : V::zero();
  }
}

// ---------------------------------------------------------------------------
// wall boundary residuals of the influence method (vel_evalBC,
// velocity.f90:262-298) for one harmonic; inputs are the last KL+1 radial
// values of u+, u-, uz.
// ---------------------------------------------------------------------------
__device__ void eval_bc(const double2* up, const double2* um, const double2* uz,
                        const double* dr1c, double d, double m_, double a_,
                        double BRe[4], double BIm[4]) {
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

// 4x4 inverse as mes_mat_invert (meshs.f90:231-243): DGETF2 + DGETRI.
// A row-major, in place.  Returns false if singular.
__device__ bool inv4(double A[16]) {
  int ipiv[4];
  for (int j = 0; j < 4; ++j) {
    int p = j;
    double best = fabs(A[j * 4 + j]);
    for (int i = j + 1; i < 4; ++i) {
      const double v = fabs(A[i * 4 + j]);
      if (v > best) { best = v; p = i; }
    }
    ipiv[j] = p;
    if (A[p * 4 + j] == 0.0) return false;
    if (p != j)
      for (int c = 0; c < 4; ++c) { const double t = A[p * 4 + c]; A[p * 4 + c] = A[j * 4 + c]; A[j * 4 + c] = t; }
    if (j < 3) {
      const double rc = 1.0 / A[j * 4 + j];
      for (int i = j + 1; i < 4; ++i) A[i * 4 + j] = A[i * 4 + j] * rc;
      for (int i = j + 1; i < 4; ++i) {
        const double lij = A[i * 4 + j];
        for (int c = j + 1; c < 4; ++c) A[i * 4 + c] = A[i * 4 + c] - lij * A[j * 4 + c];
      }
    }
  }
  double x[4];
  for (int j = 0; j < 4; ++j) {
    A[j * 4 + j] = 1.0 / A[j * 4 + j];
    const double ajj = -A[j * 4 + j];
    for (int i = 0; i < j; ++i) x[i] = A[i * 4 + j];
    for (int jj = 0; jj < j; ++jj) {
      if (x[jj] != 0.0) {
        const double temp = x[jj];
        for (int i = 0; i < jj; ++i) x[i] = x[i] + temp * A[i * 4 + jj];
        x[jj] = x[jj] * A[jj * 4 + jj];
      }
    }
    for (int i = 0; i < j; ++i) A[i * 4 + j] = x[i] * ajj;
  }
  double work[4];
  for (int j = 3; j >= 0; --j) {
    for (int i = j + 1; i < 4; ++i) { work[i] = A[i * 4 + j]; A[i * 4 + j] = 0.0; }
    for (int jj = j + 1; jj < 4; ++jj) {
      const double temp = -work[jj];
      if (temp != 0.0)
        for (int i = 0; i < 4; ++i) A[i * 4 + j] = A[i * 4 + j] + temp * A[i * 4 + jj];
    }
  }
  for (int j = 2; j >= 0; --j) {
    const int jp = ipiv[j];
    if (jp != j)
      for (int i = 0; i < 4; ++i) { const double t = A[i * 4 + j]; A[i * 4 + j] = A[i * 4 + jp]; A[i * 4 + jp] = t; }
  }
  return true;
}

// Diagonal term d(n) of tim_mesh_init / tim_lumesh_init (timestep.f90:54-55,
// 76-77) for azimuthal index m, axial index k and PM in {-1, 0, 1}.
__device__ __forceinline__ double diag_d(double r2inv, int Mp, int m, double alpha, int k, int PM) {
  double d = -((((r2inv * (double)Mp) * (double)m) * (double)Mp) * (double)m) -
             (((alpha * (double)k) * alpha) * (double)k);
  if (PM != 0) d = (d - r2inv) - ((((2.0 * (double)PM) * (double)Mp) * (double)m) * r2inv);
  return d;
}

// ---------------------------------------------------------------------------
// tim_lumesh_init (timestep.f90:67-96) + DGBTF2, four lanes per (pair,
// matrix), writing the factors in LAPACK band layout and the solve-ready
// copy (see fw_block / bw_block, meta_block).
// grid: x = pairs/16, y = 4 matrices, z = slot; 64 threads = 16 systems x 4 lanes.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_lu_factor(Geo G, Bufs B, int slot0, int gated, LuSpec spec) {
  constexpr int LPS = 4;              // lanes per system
  constexpr int SPW = 64 / LPS;       // systems per workgroup
  const int q = threadIdx.x % LPS, sl = threadIdx.x / LPS;
  const int p = blockIdx.x * SPW + sl;
  const int mat = blockIdx.y;
  const int slot = slot_of(B, slot0);
  // no barriers below: a lane may leave on its own
  if (p >= G.P || (gated && !B.ctl[slot].refactor) || (spec.mat >= 0 && mat != spec.mat)) return;
  const int N = G.N, P = G.P;
  const int m = p / G.K, k = p % G.K;  // pair (|k|, m)
  const SlotParams prm = B.prm[slot];
  const MeshOff mo = MeshOff::make(N);
  const double* me = G.mesh;
  int PM, BC;
  double c1, c2;
  if (spec.mat >= 0) { PM = spec.PM; BC = spec.BC; c1 = spec.c1; c2 = spec.c2; }
  else if (mat == MAT_LNP) { PM = 0; BC = 1; c1 = 0.0; c2 = 1.0; }
  else {
    PM = mat == MAT_LDP ? 1 : (mat == MAT_LDM ? -1 : 0);
    BC = 0;
    c1 = 1.0 / prm.dt;
    c2 = -prm.implicit * prm.nu / prm.Re;
  }
  double* F = lu_block(B, slot, mat, p, P, N);
  unsigned char* pv = pv_block(B, slot, mat, p, P, N);
  // The elimination at column j touches columns j .. j + KV only, so each
  // system keeps a window of KV + 1 band columns in LDS (column c in slot
  // c mod 9, the systems' values interleaved) instead of working on the
  // factor array in global memory, where every access of the serial
  // elimination waited a full cache round trip.  A column leaves the window
  // finished (stored once) and the next enters built from the mesh.  The
  // four lanes of a system share a step's work, each on its own data with
  // the same instructions: lane q takes the window columns j + c with
  // c = q, q + 4 (and 8 for lane 0), the rows q, q + 4, q + 8 (and 12) of a
  // new column, and a quarter of a finished one; all four find the pivot.
  // Lanes pass values to each other through the window: a wave's LDS
  // operations complete in the order issued, and lds_fence() keeps the
  // compiler from moving an access across a hand-over point.  The
  // arithmetic is DGBTF2's, operation for operation.
  constexpr int WC = 2 * KL + 1;   // window columns
  __shared__ double Wd[WC * NLU * SPW];
  auto W = [&](int ws, int row) -> double& { return Wd[(ws * NLU + row) * SPW + sl]; };
  auto lds_fence = [] { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); };
  const int S = (m * G.Mp + (PM < 0 ? -PM : PM)) % 2;
  // column `col` of tim_lumesh_init's band (timestep.f90:67-96): the mesh
  // operator, then the boundary rows of the first and last KL + 1 columns.
  // Lane q builds rows brow(t) = q + 4t (t < 3) and row 12 (t = 3, all four
  // lanes, the same value).  fetch() reads the mesh values of a column into
  // registers a column ahead of place(), so the loads overlap an
  // elimination step.  Both run on every step (with the column clamped,
  // place() writing only when `put`): a load left pending on one path would
  // make every later step wait for all the stores before it.  Each lane
  // loads its own rows' values and reads its own rows of a finished column,
  // so no lane ever selects among values by q (the compiler turns such
  // selects into an indexed stack array).
  auto brow = [&](int t) { return t < 3 ? LPS * t + q : NLU - 1; };
  struct Src {
    double lap[4], r2, b0, b1;
  };
  auto fetch = [&](int col, Src& c) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int a = brow(t) - KL;
      c.lap[t] = me[mo.lapb + (a < 0 ? 0 : a) * N + col];
    }
    c.r2 = me[mo.rinv2 + col];
    const int i0 = col <= KL ? col : KL, i1 = col >= N - KL - 1 ? col - (N - KL - 1) : 0;
    c.b0 = me[mo.dr0 + i0 + (KL + 1) * (1 - S)];
    c.b1 = me[mo.dr1 + i1 + (KL + 1) * BC];
  };
  auto place = [&](int ws, int col, const Src& c, bool put) {
    const double d = diag_d(c.r2, G.Mp, m, prm.alpha, k, PM);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int row = brow(t);
      double x = row < KL ? 0.0 : c2 * c.lap[t];
      x = row == 2 * KL ? (x + c2 * d) + c1 : x;
      x = (col <= KL && row == 2 * KL - col) ? c.b0 : x;
      x = (col >= N - KL - 1 && row == 2 * KL + N - 1 - col) ? c.b1 : x;
      if (put) W(ws, row) = x;
    }
  };
  // A finished column leaves as the LAPACK band column (F, pv) and as the
  // solve-ready copy the fused step kernel reads (fw, bw); meanwhile the
  // widest superdiagonal of U holding a non-zero and the largest pivot
  // offset are tracked for the meta block.  The step kernel skips the
  // superdiagonals that are exactly zero (x - t*0 = x) and multiplies by
  // r = 1/u0 (correctly rounded) instead of dividing by the diagonal.
  double2* fw = fw_block(B, slot, mat, p, P, N);
  double2* bw = bw_block(B, slot, mat, p, P, N);
  int kue = 1, mxo = 0;   // this lane's share of the meta block
  // The backward copy's planes hold one 16-byte pair per column (the pivot
  // rows one byte), so writing each column as it is finished would fill a
  // 128-byte line in eight pieces far apart in time, and lines left partly
  // written in the cache cost write bandwidth.  Each lane collects its own
  // entries for 8 columns (planes q and, lane 0, 4 in LDS; lane 0 the pivot
  // offsets in a register) and writes them together: flush(c0, n, npv).
  constexpr int SG = 8;
  __shared__ double2 Sb[SG * BW_Q * SPW];
  auto sb = [&](int cc, int qq) -> double2& { return Sb[(cc * BW_Q + qq) * SPW + sl]; };
  unsigned long long pvs = 0;
  auto flush = [&](int c0, int n, int npv) {
#pragma unroll
    for (int cc = 0; cc < SG; ++cc)
      if (cc < n) {
        bw[(size_t)q * N + c0 + cc] = sb(cc, q);
        if (q == 0) bw[(size_t)LPS * N + c0 + cc] = sb(cc, LPS);
      }
    if (q == 0)
      for (int cc = 0; cc < npv; ++cc) pv[c0 + cc] = (unsigned char)(pvs >> (8 * cc));
    pvs = 0;
  };
  auto rd = [&](int ws, int row) { return W(ws, row < NLU ? row : NLU - 1); };
  auto store = [&](int ws, int col, int pj) {
    // LAPACK column: rows 4q .. 4q + 3 from lane q (rows 13 .. 15 zero)
    double x[LPS];
#pragma unroll
    for (int r = 0; r < LPS; ++r) {
      const int row = LPS * q + r;
      const double v = rd(ws, row);
      x[r] = row < NLU ? v : 0.0;
      // superdiagonal d = 8 - row of U in column col
      if (row < 2 * KL && col - (2 * KL - row) >= 0 && v != 0.0) kue = 2 * KL - row > kue ? 2 * KL - row : kue;
    }
    reinterpret_cast<double4*>(F + (size_t)col * LUS)[q] = make_double4(x[0], x[1], x[2], x[3]);
    // solve-ready copy, in window rows r: fw pair q (q < 3) = (pj, r9),
    // (r10, r11), (r12, 0); bw pair q = (r8, 1/r8), (r7, r6), (r5, r4),
    // (r3, r2) and, from lane 0, pair 4 = (r1, r0)
    const double fa = rd(ws, q == 1 ? 10 : 12), fb = rd(ws, q == 0 ? 9 : 11);
    if (q < FW_Q) fw[col * FW_Q + q] = make_double2(q == 0 ? (double)pj : fa, q == 2 ? 0.0 : fb);
    static_assert(BW_Q == LPS + 1, "bw pairs: one per lane and a fifth from lane 0");
    const double ba = rd(ws, q == 0 ? 2 * KL : 2 * KL + 1 - 2 * q), bb = rd(ws, 2 * KL - 2 * q);
    sb(col % SG, q) = make_double2(ba, q == 0 ? (ba != 0.0 ? 1.0 / ba : 0.0) : bb);
    if (q == 0) sb(col % SG, LPS) = make_double2(rd(ws, 1), rd(ws, 0));
    mxo = pj > mxo ? pj : mxo;
  };
  // the pressure matrix of (0,0) is never factored: its lanes store the
  // columns as built, in the same loop as the others
  const bool raw = BC == 1 && k == 0 && m == 0;
  for (int col = 0; col < WC && col < N; ++col) {
    Src c;
    fetch(col, c);
    place(col, col, c, true);
  }
  // DGBTF2, KL = KU = 4, KV = 8 (0-based rows).  A step touches columns
  // j .. j + 8 (the window) and rows KV - 8 .. KV + 4 of them; each lane
  // reads the elements of its columns that the step may change at once and
  // writes every one back, changed where DGBTF2 changes it and unchanged
  // elsewhere, so the step is straight-line code (selects instead of
  // branches on the run-time limits km, ju and the pivot).  Window slots
  // past the matrix's last column hold columns already stored; rewriting
  // them with their own values is harmless.
  constexpr int KV = 2 * KL;
  int ju = 0;
  Src nx;
  for (int j = 0, wj = 0; j < N; ++j, wj = wj + 1 == WC ? 0 : wj + 1) {
    const int jn = j + WC < N ? j + WC : N - 1;
    fetch(jn, nx);
    lds_fence();   // the previous step's columns and new column are written
    const int km = raw ? 0 : (N - 1 - j < KL ? N - 1 - j : KL);
    double pc[KL + 1];   // rows KV .. KV + KL of column j
#pragma unroll
    for (int i = 0; i <= KL; ++i) pc[i] = W(wj, KV + i);
    int jp = 0;
    double best = fabs(pc[0]), piv = pc[0];
#pragma unroll
    for (int i = 1; i <= KL; ++i) {
      const double v = fabs(pc[i]);
      const bool better = i <= km && v > best;
      best = better ? v : best;
      jp = better ? i : jp;
      piv = better ? pc[i] : piv;
    }
    pvs |= (unsigned long long)jp << (8 * (j % SG));
    if (!raw && piv == 0.0) {
      flush(j - j % SG, j % SG, j % SG + 1);
      if (q == 0) atomicMax(&B.ctl[slot].status, 100 + mat);  // singular factor
      return;
    }
    const int lim = j + KL + jp < N - 1 ? j + KL + jp : N - 1;
    ju = raw ? ju : (ju > lim ? ju : lim);
    const int nc = raw ? -1 : ju - j;   // columns j .. j + nc take part (nc <= KV)
    // multipliers: the pivot column below the diagonal after the
    // interchange, scaled by 1/pivot (rows 1 .. km)
    const double rc = 1.0 / (raw ? 1.0 : piv);
    double l[KL + 1];
#pragma unroll
    for (int i = 1; i <= KL; ++i) {
      const double li = (jp != 0 && i == jp) ? pc[0] : pc[i];
      l[i] = i <= km ? li * rc : li;
    }
    // this lane's columns j + c: the row interchange of rows j and j + jp,
    // then, for c = 0, the multipliers and, for c > 0, the rank-1 update of
    // the rows below the diagonal
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const int c = q + LPS * t;
      if (t == 2 && q != 0) break;   // c = 8 is lane 0's alone
      const int ws = wj + c >= WC ? wj + c - WC : wj + c;
      const double ra = W(ws, KV + jp - c), rb = W(ws, KV - c);
      const bool sw = jp != 0 && c <= nc;
      W(ws, KV + jp - c) = sw ? rb : ra;
      W(ws, KV - c) = sw ? ra : rb;
      const double temp = -(sw ? ra : rb);   // row KV - c after the interchange
      const bool col_on = c <= nc && temp != 0.0;
      double cv[KL + 1];
#pragma unroll
      for (int i = 1; i <= KL; ++i) cv[i] = W(ws, KV + i - c);
#pragma unroll
      for (int i = 1; i <= KL; ++i)
        W(ws, KV + i - c) = c == 0 ? l[i] : ((col_on && i <= km) ? cv[i] + l[i] * temp : cv[i]);
    }
    lds_fence();   // column j is final
    store(wj, j, jp);
    lds_fence();   // every lane has read column j before its slot is refilled
    place(wj, jn, nx, j + WC < N);
    // after place(): its wait for the mesh loads then does not include
    // these stores (loads and stores share one in-order counter)
    if (j % SG == SG - 1 || j == N - 1) flush(j - j % SG, j % SG + 1, j % SG + 1);
  }
  // the four lanes of a system are neighbours in the wave
  kue = max(kue, __shfl_xor(kue, 1));
  kue = max(kue, __shfl_xor(kue, 2));
  if (q == 0) {
    unsigned char* mt = meta_block(B, slot, mat, p, P);
    mt[0] = (unsigned char)kue;
    mt[1] = (unsigned char)mxo;
  }
}

// ---------------------------------------------------------------------------
// dst += src on one coll component (harmonics 0..H-1 of every radial row):
// forcing terms added to vel_N* by var_null(2) hooks.  grid: x = H/64, y = N.
// ---------------------------------------------------------------------------
__global__ void k_add_coll(Geo G, double2* dst, const double2* src) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = blockIdx.y;
  if (h >= G.H || n >= G.N) return;
  const size_t o = (size_t)n * G.Hp + h;
  const double2 a = dst[o], b = src[o];
  dst[o] = make_double2(a.x + b.x, a.y + b.y);
}

// ---------------------------------------------------------------------------
// var_coll_norm (variables.f90:718-751) of a field in the reference's coll
// layout (re, im: N x H, r fastest).  Pass 1, one thread per harmonic:
// b(h) = sum_n f(n) intrdr(n) with f = w (re^2 + im^2), w = 4 pi^2 / alpha,
// in n order, halved for (0,0).  Pass 2, one workgroup: Em(m) and Ek(|k|)
// accumulate b in harmonic order (parallel.h:47-56), then E = sum(Em) in m
// order.  out = [E, Ek(0:K-1), Em(0:M-1)].
// ---------------------------------------------------------------------------
__global__ void k_norm_harm(Geo G, const SlotParams* __restrict__ prm, const double* __restrict__ re,
                            const double* __restrict__ im, double* __restrict__ b) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= G.H) return;
  const double w = 4.0 * D_PI * D_PI / prm->alpha;
  const double* intrdr = G.mesh + MeshOff::make(G.N).intrdr;
  const double* a = re + (size_t)h * G.N;
  const double* c = im + (size_t)h * G.N;
  double s = 0.0;
  for (int n = 0; n < G.N; ++n) {
    const double f = w * (a[n] * a[n] + c[n] * c[n]);
    s = s + f * intrdr[n];
  }
  b[h] = h == 0 ? 0.5 * s : s;
}

__global__ void k_norm_sum(Geo G, const double* __restrict__ b, double* __restrict__ out) {
  const int t = threadIdx.x;
  const int K = G.K, M = G.M, W = 2 * G.K - 1;
  double* Ek = out + 1;
  double* Em = out + 1 + K;
  for (int m = t; m < M; m += blockDim.x) {   // the harmonics of m, k ascending
    double s = 0.0;
    if (m == 0) {
      for (int k = 0; k < K; ++k) s = s + b[k];
    } else {
      const double* bm = b + K + (size_t)(m - 1) * W;
      for (int i = 0; i < W; ++i) s = s + bm[i];
    }
    Em[m] = s;
  }
  for (int q = t; q < K; q += blockDim.x) {   // |k| = q: (q, 0), then (-q, m), (q, m) for m >= 1
    double s = b[q];
    for (int m = 1; m < M; ++m) {
      const double* bm = b + K + (size_t)(m - 1) * W + (K - 1);
      if (q > 0) s = s + bm[-q];
      s = s + bm[q];
    }
    Ek[q] = s;
  }
  __syncthreads();
  if (t == 0) {
    double e = 0.0;
    for (int m = 0; m < M; ++m) e = e + Em[m];
    out[0] = e;
  }
}

// ---------------------------------------------------------------------------
// lu_solve<double> for one lane whose vector lives in LDS (element n at
// X[n * 64], the workgroup's lanes interleaved) and whose factor columns and
// pivot offsets are read D columns ahead into a register ring, so a solve
// waits on LDS rather than on a cache round trip per column.  The factors
// come from the solve-ready copy (FW: pivot offset and multipliers, 48
// contiguous bytes per column; BW: U's diagonal and superdiagonals in five
// planes), half the bytes of the LAPACK columns' two 128-byte lines.  The
// operations and their order are lu_solve's (the diagonal divides).
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void lu_solve_lds(const double2* FW, const double2* BW, double* X, int N) {
  // Every global load is unconditional (column indices clamped into the
  // matrix; values past the last column are never used), and a ring slot is
  // refilled only after its last use: a load issued on one path only, or a
  // slot refilled while its old value is still live (the two then need
  // different registers and a copy at the loop's back edge), makes each pass
  // round the loop wait for every load in flight.
  double w[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) w[i] = i < N ? X[i * 64] : 0.0;
  double fl[D][KL];
  int pr[D];
  // column c of L: (pivot offset, l1) (l2, l3) (l4, 0)
  auto lcol = [&](int c, double (&f)[KL], int& o) {
    const double2 a = FW[(size_t)c * FW_Q], b = FW[(size_t)c * FW_Q + 1], e = FW[(size_t)c * FW_Q + 2];
    f[0] = a.y;
    f[1] = b.x;
    f[2] = b.y;
    f[3] = e.x;
    o = (int)a.x;
  };
  // column c of U, band rows 0..8 (row 8 - d: superdiagonal d): plane 0
  // holds (u0, 1/u0), plane q >= 1 (u_{2q-1}, u_{2q})
  auto ucol = [&](int c, double (&f)[2 * KL + 1]) {
    f[2 * KL] = ((const double*)(BW + c))[0];
#pragma unroll
    for (int q = 1; q < BW_Q; ++q) {
      const double2 v = BW[(size_t)q * N + c];
      f[2 * KL + 1 - 2 * q] = v.x;
      f[2 * KL - 2 * q] = v.y;
    }
  };
#pragma unroll
  for (int u = 0; u < D; ++u) lcol(u < N - 1 ? u : N - 2, fl[u], pr[u]);
  for (int j0 = 0; j0 < N - 1; j0 += D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int j = j0 + u;
      const int off = pr[u];
      if (j < N - 1) {
        const double a0 = w[0];
        double s = w[0];
        s = off == 1 ? w[1] : s;
        s = off == 2 ? w[2] : s;
        s = off == 3 ? w[3] : s;
        s = off == 4 ? w[4] : s;
        w[0] = s;
        w[1] = off == 1 ? a0 : w[1];
        w[2] = off == 2 ? a0 : w[2];
        w[3] = off == 3 ? a0 : w[3];
        w[4] = off == 4 ? a0 : w[4];
        const double t = -w[0];
        const int lm = N - 1 - j < KL ? N - 1 - j : KL;
#pragma unroll
        for (int i = 1; i <= KL; ++i)
          if (i <= lm) w[i] = w[i] + fl[u][i - 1] * t;
        X[j * 64] = w[0];
#pragma unroll
        for (int i = 0; i < 4; ++i) w[i] = w[i + 1];
        const double xn = X[(j + 5 < N ? j + 5 : N - 1) * 64];
        w[4] = j + 5 < N ? xn : 0.0;
      }
      lcol(j + D < N - 1 ? j + D : N - 2, fl[u], pr[u]);
    }
  }
  double v[9];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int idx = N - 9 + q;
    v[q] = idx >= 0 ? X[idx * 64] : 0.0;
  }
  v[8] = w[0];
  double fu[D][2 * KL + 1];
#pragma unroll
  for (int u = 0; u < D; ++u) ucol(N - 1 - u >= 0 ? N - 1 - u : 0, fu[u]);
  for (int j0 = N - 1; j0 >= 0; j0 -= D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int j = j0 - u;
      if (j >= 0) {
        v[8] = v[8] / fu[u][8];
        const double t = v[8];
        X[j * 64] = t;
#pragma unroll
        for (int d = 1; d <= 8; ++d)
          if (j - d >= 0) v[8 - d] = v[8 - d] - t * fu[u][8 - d];
#pragma unroll
        for (int q = 8; q >= 1; --q) v[q] = v[q - 1];
        const double xn = X[(j - 9 >= 0 ? j - 9 : 0) * 64];
        v[0] = j - 9 >= 0 ? xn : 0.0;
      }
      ucol(j - D >= 0 ? j - D : 0, fu[u]);
    }
  }
}

// ---------------------------------------------------------------------------
// vel_adjPPE(0) influence vectors (velocity.f90:191-218) per pair, and
// vel_adjustFlux(0) (velocity.f90:160-168) on pair 0.
// Uinf slots: 0 U1 (LDp), 1 U2 (LDm), 2 U3 (LDz), 3 U4, 4 U5, 5 p (U6 = p*alpha*k).
// The solves are independent, so each gets its own lane: grid y = 0..2 the
// three Helmholtz solves, 3 the pressure solve and its gradient, 4 the flux
// solve (lane of pair 0 only).  grid: x = pairs/64, y = 5, z = slot.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_influence(Geo G, Bufs B, int slot0, int gated) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int which = blockIdx.y;
  const int slot = slot_of(B, slot0);
  if (p >= G.P || (gated && !B.ctl[slot].refactor) || (which == 4 && p != 0)) return;
  const int N = G.N, P = G.P;
  const int m = p / G.K, k = p % G.K;
  const MeshOff mo = MeshOff::make(N);
  const double* me = G.mesh;
  double* U = B.Uinf + (size_t)slot * 6 * N * P + p;
  __shared__ double Xs[MAXN * 64];
  double* x = Xs + threadIdx.x;
  auto FWp = [&](int mat) { return (const double2*)fw_block(B, slot, mat, p, P, N); };
  auto BWp = [&](int mat) { return (const double2*)bw_block(B, slot, mat, p, P, N); };
  // right-hand side and factor of this lane's solve (one inlined solve for
  // every case, so the vector stays in LDS and the factor loads in flight)
  const SlotParams prm = B.prm[slot];
  int mat;
  bool solve = true;
  if (which < 3) mat = which;
  else if (which == 3) { mat = MAT_LNP; solve = !(k == 0 && m == 0); }
  else { mat = MAT_LDZ; solve = prm.const_flux != 0; }
  for (int n = 0; n < N; ++n) {
    double v;
    if (which < 3) v = n == N - 1 ? 1.0 : 0.0;
    else if (which == 3) v = solve && n == N - 1 ? -1.0 : 0.0;
    else v = (n == 0 || n == N - 1) ? 0.0 : 1.0;
    x[n * 64] = v;
  }
  if (solve) lu_solve_lds<8>(FWp(mat), BWp(mat), x, N);
  if (which < 3) {
    double* u = U + (size_t)which * N * P;
    for (int n = 0; n < N; ++n) u[(size_t)n * P] = x[n * 64];
  } else if (which == 3) {
    double* pp = U + (size_t)5 * N * P;
    double* U4 = U + (size_t)3 * N * P;
    double* U5 = U + (size_t)4 * N * P;
    for (int n = 0; n < N; ++n) {
      // grad: dr1 . p (meshmult from 0), t.Im = p * m/r; rt2pm -> U4 = gr - tIm, U5 = gr + tIm
      double gr = 0.0;
      const int j0 = n - KL < 0 ? 0 : n - KL, j1 = n + KL > N - 1 ? N - 1 : n + KL;
      for (int j = j0; j <= j1; ++j) gr = gr + me[mo.dr1b + (KL + n - j) * N + j] * x[j * 64];
      const double mr = (me[mo.rinv + n] * (double)G.Mp) * (double)m;
      const double tIm = x[n * 64] * mr;
      pp[(size_t)n * P] = x[n * 64];
      U4[(size_t)n * P] = gr - tIm;
      U5[(size_t)n * P] = gr + tIm;
    }
  } else {
    double* Ui = B.Ui + (size_t)slot * N;
    for (int n = 0; n < N; ++n) Ui[n] = x[n * 64];
    if (prm.const_flux) {
      double s = 0.0;
      for (int n = 0; n < N; ++n) s = s + x[n * 64] * me[mo.intrdr + n];
      B.ctl[slot].flux_d1 = 2.0 * s;
      if (2.0 * s == 0.0) atomicMax(&B.ctl[slot].status, 110);
    }
  }
}

// Inverted 4x4 influence matrices per harmonic (velocity.f90:219-225).
__global__ void k_ainv(Geo G, Bufs B, int slot0, int gated) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int slot = slot_of(B, slot0);
  if (h >= G.H || (gated && !B.ctl[slot].refactor)) return;
  int k, m;
  harmonic_km(h, G.K, k, m);
  double* Ai = B.Ainv + (size_t)slot * 16 * G.H + h;
  if (k == 0 && m == 0) {
    for (int q = 0; q < 16; ++q) Ai[(size_t)q * G.H] = 0.0;
    return;
  }
  const int N = G.N, P = G.P;
  const int p = m * G.K + (k < 0 ? -k : k);
  const SlotParams prm = B.prm[slot];
  const MeshOff mo = MeshOff::make(N);
  const double* me = G.mesh;
  const double* U = B.Uinf + (size_t)slot * 6 * N * P + p;
  const double ak = prm.alpha * (double)k;
  const double d = me[mo.rinv + N - 1];
  const double m_ = (double)(m * G.Mp);
  double A[16];
  for (int j = 0; j < 4; ++j) {
    double2 up[KL + 1], um[KL + 1], uz[KL + 1];
    for (int i = 0; i <= KL; ++i) {
      const size_t n = (size_t)(N - 1 - KL + i) * P;
      up[i] = um[i] = uz[i] = make_double2(0.0, 0.0);
      if (j == 0) up[i].x = U[n];
      if (j == 1) um[i].x = U[(size_t)N * P + n];
      if (j == 2) uz[i].y = U[(size_t)2 * N * P + n];
      if (j == 3) {
        up[i].x = U[(size_t)3 * N * P + n];
        um[i].x = U[(size_t)4 * N * P + n];
        uz[i].y = U[(size_t)5 * N * P + n] * ak;
      }
    }
    double BRe[4], BIm[4];
    eval_bc(up, um, uz, me + mo.dr1 + (KL + 1), d, m_, ak, BRe, BIm);
    for (int i = 0; i < 4; ++i) A[i * 4 + j] = BRe[i];
  }
  if (!inv4(A)) atomicMax(&B.ctl[slot].status, 120);
  for (int q = 0; q < 16; ++q) Ai[(size_t)q * G.H] = A[q];
}

// ---------------------------------------------------------------------------
// curl (var_coll_curl, variables.f90:408-436) of the current iterate -> om
// ---------------------------------------------------------------------------
__device__ __forceinline__ void curl_point(const Geo& G, int h, int n, int k, int m, double alpha,
                                           const double2* ur, const double2* ut, const double2* uz, double2* om) {
  const int N = G.N;
  const size_t hs = G.Hp;
  const MeshOff mo = MeshOff::make(N);
  const double* me = G.mesh;
  const double ak = alpha * (double)k;
  double2* orr = om;
  double2* ot = om + G.coll;
  double2* oz = om + 2 * G.coll;
  const double mr = (me[mo.rinv + n] * (double)G.Mp) * (double)m;
  const size_t o = (size_t)n * hs + h;
  const double2 vr = ur[o], vt = ut[o], vz = uz[o];
  orr[o] = make_double2(-(vz.y * mr) + vt.y * ak, vz.x * mr - vt.x * ak);
  const double2 d1 = band_row(me + mo.dr1b, N, n, uz + h, hs, make_double2(0.0, 0.0));
  ot[o] = make_double2(-(vr.y * ak) - d1.x, vr.x * ak - d1.y);
  const double2 d2 = band_row(me + mo.dr1b, N, n, ut + h, hs, make_double2(0.0, 0.0));
  const double ri = me[mo.rinv + n];
  oz[o] = make_double2((vt.x * ri + d2.x) + vr.y * mr, (vt.y * ri + d2.y) - vr.x * mr);
}

// which = 0: curl of U[ucur]; used after put_state and after every step.
// One thread per (h, n).  grid: x = 8 * per blocks over the (64-harmonic
// block, n) tiles, z = slot; tiles in XCD-aware order (block b takes tile
// (b % 8) per + b / 8, tiles ordered harmonic block first), so the nine rows
// of a stencil are read by workgroups on one XCD and come from its L2.
__global__ void k_curl(Geo G, Bufs B, int slot0, int gated) {
  const int nhb = (G.H + 63) / 64, ntile = nhb * G.N, per = (ntile + 7) / 8;
  const int tile = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (tile >= ntile) return;
  const int hb = tile / G.N, n = tile - hb * G.N;
  const int h = hb * 64 + threadIdx.x;
  const int slot = slot_of(B, slot0);
  const SlotCtl* c = &B.ctl[slot];
  if (!c->active || h >= G.H) return;
  if (gated && (c->mode != MODE_PRED || B.prm[slot].symm == 0)) return;
  int k, m;
  harmonic_km(h, G.K, k, m);
  const double2* u = B.U[c->ucur] + (size_t)slot * 3 * G.coll;
  curl_point(G, h, n, k, m, B.prm[slot].alpha, u, u + G.coll, u + 2 * G.coll,
             B.om + (size_t)slot * 3 * G.coll);
}

// vel_imposesym (velocity.f90:61-74) with the slot's symmetries on U[ucur]:
// var_imposesym(0), then (1) mirror, (2) shift & reflect, (3) shift &
// rotate as enabled (variables.f90:538-618; each ends with the harmonic-0
// conditions of mode 0).  Modes 1 and 2 average a harmonic (k, m) with the
// image of its partner (-k, m) (itself for m = 0): factor a = 1, or (-1)^k
// for shift & reflect; the imaginary part flips for m > 0; u_theta flips.
// Both act within a pair, mode 3 within a harmonic and every mode within a
// component, so one thread takes (n, pair, component) and applies the
// enabled modes in order in registers.  gated: only slots about to run a
// predictor (the start of a step, main.f90:21) with a symmetry enabled.
__global__ void k_symmetry(Geo G, Bufs B, int slot0, int gated) {
  const int slot = slot_of(B, slot0);
  const SlotCtl* c = &B.ctl[slot];
  const int symm = B.prm[slot].symm;
  if (!c->active || (gated && (c->mode != MODE_PRED || symm == 0))) return;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;   // (comp, n, h)
  const int h = idx % G.H, rest = idx / G.H;
  const int n = rest % G.N, comp = rest / G.N;
  if (comp >= 3) return;
  int k, m;
  harmonic_km(h, G.K, k, m);
  if (m > 0 && k < 0) return;                 // taken by the partner (k > 0)
  const int hp = m == 0 ? h : h - 2 * k;      // harmonic (-k, m)
  double2* u = B.U[c->ucur] + (size_t)slot * 3 * G.coll + (size_t)comp * G.coll + (size_t)n * G.Hp;
  double2 x = u[h], y = u[hp];
  const double b = m == 0 ? 1.0 : -1.0;
  const double sg = comp == 1 ? -1.0 : 1.0;   // u_theta changes sign
  auto fix0 = [&]() {                         // mode 0 on harmonic 0
    if (h == 0) {
      if (comp == 0) x = make_double2(0.0, 0.0);
      else x.y = 0.0;
    }
  };
  fix0();
  for (int i = 1; i <= 3; ++i) {
    if (!(symm & (1 << (i - 1)))) continue;
    if (i <= 2) {
      const double a = (i == 2 && (k & 1)) ? -1.0 : 1.0;
      const double f = sg * a;
      // image of the partner (of x for the partner's own update)
      const double2 ix = make_double2(y.x * f, y.y * (f * b));
      const double2 iy = make_double2(x.x * f, x.y * (f * b));
      x = make_double2((x.x + ix.x) / 2.0, (x.y + ix.y) / 2.0);
      if (hp != h) y = make_double2((y.x + iy.x) / 2.0, (y.y + iy.y) / 2.0);
      else y = x;
    } else if ((k + m) & 1) {
      x = make_double2(0.0, 0.0);
      y = x;
    }
    fix0();
  }
  u[h] = x;
  if (hp != h) u[hp] = y;
}

// ---------------------------------------------------------------------------
// Non-Newtonian viscosity (the plugin of nonnewtonian.f90:93-244): the
// velocity-gradient tensor of the total flow u + U(r) z^, the strain rate and
// the Carreau-Yasuda viscosity nu on the physical grid, and the forcing
// N' = [(nu - 1) Lap(u) + grad(nu) . 2S] / Re added to the raw N of the
// evaluation.  Spectral pieces run per harmonic (one thread per h, all n),
// physical pieces per grid point; the transforms between them are the
// single-field transforms of tra_coll2phys / tra_phys2coll.
// ---------------------------------------------------------------------------
// One row of the gradient tensor A of the slot's current iterate (get_nu,
// nonnewtonian.f90:108-147): row 0 the radial derivatives by
// var_coll_meshmult with dr(1), row 1 the theta terms (1/r)(d/dtheta u -
// e_theta x u), row 2 the z derivatives; out holds the three components.
// One thread per (h, n): grid x = harmonics / 64, y = N.
__global__ void k_nn_grad(Geo G, Bufs B, int slot, int row, double2* __restrict__ out) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = blockIdx.y;
  if (h >= G.H) return;
  int k, m;
  harmonic_km(h, G.K, k, m);
  const double2* u = B.U[B.ctl[slot].ucur] + (size_t)slot * 3 * G.coll;
  const MeshOff mo = MeshOff::make(G.N);
  const double* me = G.mesh;
  const size_t hs = G.Hp;
  const int N = G.N;
  auto at = [&](int comp, int j) {   // u, with vel_U added to the (0, 0) mode of u_z
    double2 v = u[(size_t)comp * G.coll + (size_t)j * hs + h];
    if (comp == 2 && h == 0) v.x = v.x + me[mo.Ubase + j];
    return v;
  };
  const size_t o = (size_t)n * hs + h;
  if (row == 0) {
    const int j0 = n - KL < 0 ? 0 : n - KL;
    const int j1 = n + KL > N - 1 ? N - 1 : n + KL;
    for (int comp = 0; comp < 3; ++comp) {
      double2 acc = make_double2(0.0, 0.0);
      for (int j = j0; j <= j1; ++j) {
        const double a = me[mo.dr1b + (KL + n - j) * N + j];
        const double2 v = at(comp, j);
        acc.x = acc.x + a * v.x;
        acc.y = acc.y + a * v.y;
      }
      out[(size_t)comp * G.coll + o] = acc;
    }
  } else if (row == 1) {
    const double ri = me[mo.rinv + n];
    const double mm = (double)(m * G.Mp);
    const double2 vr = at(0, n), vt = at(1, n), vz = at(2, n);
    out[o] = make_double2(ri * (-(vr.y * mm) - vt.x), ri * (vr.x * mm - vt.y));
    out[G.coll + o] = make_double2(ri * (-(vt.y * mm) + vr.x), ri * (vt.x * mm + vr.y));
    out[2 * G.coll + o] = make_double2(ri * (-(vz.y * mm)), ri * (vz.x * mm));
  } else {
    const double d1 = B.prm[slot].alpha * (double)k;
    for (int comp = 0; comp < 3; ++comp) {
      const double2 v = at(comp, n);
      out[(size_t)comp * G.coll + o] = make_double2(-(v.y * d1), v.x * d1);
    }
  }
}

// var_coll_grad (variables.f90:442-459) of a scalar coll field p into out;
// one thread per (h, n).
__global__ void k_nn_gradp(Geo G, Bufs B, int slot, const double2* __restrict__ p, double2* __restrict__ out) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = blockIdx.y;
  if (h >= G.H) return;
  int k, m;
  harmonic_km(h, G.K, k, m);
  const MeshOff mo = MeshOff::make(G.N);
  const double* me = G.mesh;
  const double ak = B.prm[slot].alpha * (double)k;
  const size_t o = (size_t)n * G.Hp + h;
  out[o] = band_row(me + mo.dr1b, G.N, n, p + h, G.Hp, make_double2(0.0, 0.0));
  const double mr = (me[mo.rinv + n] * (double)G.Mp) * (double)m;
  const double2 v = p[o];
  out[G.coll + o] = make_double2(-(v.y * mr), v.x * mr);
  out[2 * G.coll + o] = make_double2(-(v.y * ak), v.x * ak);
}

// curl(curl(u + U z^)) of the current iterate: the curl (var_coll_curl,
// variables.f90:408-436) of the slot's vorticity om = curl u, plus curl curl
// (U z^) = -Lap(U) z^ = 4 z^ on the (0, 0) mode (U = 1 - r^2); one thread
// per (h, n).
__global__ void k_nn_curlcurl(Geo G, Bufs B, int slot, double2* __restrict__ out) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = blockIdx.y;
  if (h >= G.H) return;
  int k, m;
  harmonic_km(h, G.K, k, m);
  const int N = G.N;
  const size_t hs = G.Hp;
  const MeshOff mo = MeshOff::make(N);
  const double* me = G.mesh;
  const double ak = B.prm[slot].alpha * (double)k;
  const double2* wr = B.om + (size_t)slot * 3 * G.coll;
  const double2* wt = wr + G.coll;
  const double2* wz = wr + 2 * G.coll;
  const double mr = (me[mo.rinv + n] * (double)G.Mp) * (double)m;
  const size_t o = (size_t)n * hs + h;
  const double2 vr = wr[o], vt = wt[o], vz = wz[o];
  out[o] = make_double2(-(vz.y * mr) + vt.y * ak, vz.x * mr - vt.x * ak);
  const double2 d1 = band_row(me + mo.dr1b, N, n, wz + h, hs, make_double2(0.0, 0.0));
  out[G.coll + o] = make_double2(-(vr.y * ak) - d1.x, vr.x * ak - d1.y);
  const double2 d2 = band_row(me + mo.dr1b, N, n, wt + h, hs, make_double2(0.0, 0.0));
  const double ri = me[mo.rinv + n];
  double2 qz = make_double2((vt.x * ri + d2.x) + vr.y * mr, (vt.y * ri + d2.y) - vr.x * mr);
  if (h == 0) qz.x += 4.0;
  out[2 * G.coll + o] = qz;
}

// get_nu's algebra per grid point (nonnewtonian.f90:148-172): S = (A + A^T)/2
// written back over A, strain rate sr = (2 S:S)^(1/2) and the Carreau-Yasuda
// viscosity nu = d1 + (1 - d1)(1 + (lam sr)^a)^d2.
__global__ void k_nn_visc(NnPhys P, size_t np, double lam, double d1, double d2, double a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double* const* S = P.S;
  const double s12 = 0.5 * (S[1][i] + S[3][i]);
  const double s13 = 0.5 * (S[2][i] + S[6][i]);
  const double s23 = 0.5 * (S[5][i] + S[7][i]);
  S[1][i] = s12; S[3][i] = s12;
  S[2][i] = s13; S[6][i] = s13;
  S[5][i] = s23; S[7][i] = s23;
  const double s11 = S[0][i], s22 = S[4][i], s33 = S[8][i];
  double sr = s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23);
  sr = sqrt(2.0 * sr);
  P.nu[i] = d1 + (1.0 - d1) * pow(1.0 + pow(lam * sr, a), d2);
}

// update_nonlin_terms' algebra per grid point (nonnewtonian.f90:227-236):
// q_i = [(nu - nu_r)(-curlcurl(u)_i) + sum_j grad(nu)_j S_ji 2 + (nu_r - 1)(-4) delta_i3] / Re.
__global__ void k_nn_force(NnPhys P, size_t np, double nu_r, double Re) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double nu_ = P.nu[i] - nu_r;
  for (int c = 0; c < 3; ++c) {
    double q = nu_ * (-P.q[c][i]);
    for (int j = 0; j < 3; ++j) q = q + P.g[j][i] * P.S[j * 3 + c][i] * 2.0;
    if (c == 2) q = q + (nu_r - 1.0) * (-4.0);
    P.q[c][i] = q / Re;
  }
}

// ---------------------------------------------------------------------------
// LES eddy viscosity (the plugin of LES.f90:100-395).  D holds per slot
// [Delta(N), Delta_r(N), L_theta(N), coef(N)]: the Smagorinsky filter width,
// the radial spacing, the azimuthal length of Vreman's filter and, written
// by k_les_prep, the per-plane model coefficient.
// ---------------------------------------------------------------------------
// u_tau, Re_tau, y+ (get_nu, LES.f90:170-177) and the per-plane coefficient:
// Smagorinsky vanDr (C_s Delta)^2 (:206-213), Vreman's damping
// d1 + (1 - d1) vanDr with d1 = 0.03/C_s, or 1 without damping (:247-256);
// clears the slot's nu_max.  One wave.
__global__ void k_les_prep(Geo G, Bufs B, int slot, int model, double Cs, double Aplus, double* D,
                           unsigned long long* numax) {
  const int N = G.N;
  const MeshOff mo = MeshOff::make(N);
  const double* me = G.mesh;
  const double2* uz = B.U[B.ctl[slot].ucur] + (size_t)slot * 3 * G.coll + 2 * G.coll;
  const double* w = me + mo.dr1 + (KL + 1);   // dr1(:,1)
  double d1 = 0.0;
  for (int i = 0; i <= KL; ++i) d1 = d1 + uz[(size_t)(N - 1 - KL + i) * G.Hp].x * w[i];
  const double Re = B.prm[slot].Re;
  const double utau = sqrt(fabs((d1 - 2.0) / Re));
  const double Retau = Re * utau;
  for (int n = threadIdx.x; n < N; n += blockDim.x) {
    const double yp = Retau * (1.0 - me[mo.r + n]);
    const double q = yp / Aplus;
    const double vanDr = 1.0 - exp(-(q * q));
    double c;
    if (model == 1) {
      const double cd = Cs * D[n];
      c = vanDr * (cd * cd);
    } else if (Aplus <= 0.0) {
      c = 1.0;
    } else {
      const double a = 0.03 / Cs;
      c = a + (1.0 - a) * vanDr;
    }
    D[3 * N + n] = c;
  }
  if (threadIdx.x == 0) numax[slot] = 0ull;
}

// The eddy viscosity per grid point from the gradient tensor A (row-major in
// P.S), which is then replaced by S = (A + A^T)/2: Smagorinsky nu = coef
// sr, sr = (2 S:S)^(1/2); Vreman nu = coef C_s (P/AA)^(1/2) with
// B = sum_m Delta_m^2 A_mi A_mj (:218-257).  The slot's nu_max by one
// atomic per workgroup on the bit pattern (nu >= 0).
__global__ void k_les_visc(NnPhys P, size_t np, int N, int ZT, int model, double Cs, double dz, const double* D,
                           unsigned long long* numax_slot) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double nu = 0.0;
  if (i < np) {
    const int n = (int)(i / (size_t)ZT);
    double a[9];
    for (int q = 0; q < 9; ++q) a[q] = P.S[q][i];
    if (model == 2) {
      double AA = 0.0;
      for (int q = 0; q < 9; ++q) AA = AA + a[q] * a[q];
      const double dl[3] = {D[N + n], D[2 * N + n], dz};
      double Bm[9];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double b = 0.0;
          for (int m = 0; m < 3; ++m) b = b + (dl[m] * dl[m]) * a[m * 3 + r] * a[m * 3 + c];
          Bm[r * 3 + c] = b;
        }
      double Pv = Bm[0] * Bm[4] - Bm[1] * Bm[1] + Bm[0] * Bm[8] - Bm[2] * Bm[2] + Bm[4] * Bm[8] - Bm[5] * Bm[5];
      Pv = AA < 1e-12 ? 0.0 : Pv;
      Pv = fmax(Pv, 0.0);
      AA = fmax(AA, 1e-12);
      nu = Cs * sqrt(Pv / AA) * D[3 * N + n];
    }
    const double s12 = 0.5 * (a[1] + a[3]);
    const double s13 = 0.5 * (a[2] + a[6]);
    const double s23 = 0.5 * (a[5] + a[7]);
    if (model == 1) {
      double sr = a[0] * a[0] + a[4] * a[4] + a[8] * a[8] + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23);
      sr = sqrt(2.0 * sr);
      nu = D[3 * N + n] * sr;
    }
    P.S[1][i] = s12; P.S[3][i] = s12;
    P.S[2][i] = s13; P.S[6][i] = s13;
    P.S[5][i] = s23; P.S[7][i] = s23;
    P.nu[i] = nu;
  }
  for (int off = 32; off >= 1; off >>= 1) nu = fmax(nu, __shfl_xor(nu, off));
  __shared__ double wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = nu;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = wm[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) v = fmax(v, wm[q]);
    atomicMax(numax_slot, (unsigned long long)__double_as_longlong(v));
  }
}

// update_nu_r (LES.f90:264-273) before a predictor: vel_nu = 1.1 (1 + Re
// nu_max) when it has left [0.93, 1.07] of that value; the change is left in
// `refactor` for the gated vel_matrices launches that follow.
__global__ void k_les_nur(Bufs B, int slot, const unsigned long long* numax) {
  SlotCtl* c = &B.ctl[slot];
  SlotParams* prm = &B.prm[slot];
  if (!(c->active && c->mode == MODE_PRED)) {
    c->refactor = 0;
    return;
  }
  const double nm = __longlong_as_double((long long)numax[slot]);
  double d1 = 1.0 + prm->Re * nm;
  d1 = 1.10 * d1;
  const bool ch = prm->nu < 0.93 * d1 || prm->nu > 1.07 * d1;
  if (ch) prm->nu = d1;
  c->refactor = ch ? 1 : 0;
}

// The LES forcing per grid point (update_nonlin_terms, LES.f90:303-311):
// q_i = (1/Re) nu' (-curlcurl(u)_i) + nu (-4) delta_i3 + sum_j grad(nu)_j S_ji 2,
// nu' = Re nu + 1 - vel_nu.
__global__ void k_les_force(NnPhys P, size_t np, Bufs B, int slot) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double Re = B.prm[slot].Re, nur = B.prm[slot].nu;
  const double nu = P.nu[i];
  const double nu_ = Re * nu + (1.0 - nur);
  for (int c = 0; c < 3; ++c) {
    double q = (1.0 / Re) * nu_ * (-P.q[c][i]);
    if (c == 2) q = q + nu * (-4.0);
    for (int j = 0; j < 3; ++j) q = q + P.g[j][i] * P.S[j * 3 + c][i] * 2.0;
    P.q[c][i] = q;
  }
}

// N += N': one component added to the raw nonlinear terms of the slot's
// current evaluation.
__global__ void k_nn_addN(Geo G, Bufs B, int slot, int comp, const double2* __restrict__ src) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = blockIdx.y;
  if (h >= G.H || n >= G.N) return;
  double2* dst = B.Nraw[B.ctl[slot].ncur] + (size_t)slot * 3 * G.coll + (size_t)comp * G.coll;
  const size_t o = (size_t)n * G.Hp + h;
  const double2 a = dst[o], b = src[o];
  dst[o] = make_double2(a.x + b.x, a.y + b.y);
}

// var_coll_shift (variables.f90:487-528) on U[ucur] of slots [slot0,
// slot0 + gridDim.z): harmonic (k, m) rotated by e^{-i alpha k dz}, then by
// e^{-i m Mp dth}, each in the reference's order of operations
// (Re' = Re c + Im s, Im' = Im c - Re s).  tab holds 2K + M entries per
// slot: [0] = (z shift on, theta shift on), [k + K] = (cos, sin)(dz k alpha)
// for k = -(K-1)..K-1, [2K + m] = (cos, sin)(dth m Mp), evaluated on the host
// with the C library's cos and sin as the reference's dcos and dsin.  One
// thread per (component, n, harmonic).
__global__ void k_shift(Geo G, Bufs B, int slot0, const double2* __restrict__ tab) {
  const int slot = slot_of(B, slot0);
  const double2* tb = tab + (size_t)blockIdx.z * (2 * G.K + G.M);
  const double2 on = tb[0];
  if (on.x == 0.0 && on.y == 0.0) return;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int h = idx % G.H, rest = idx / G.H;
  const int n = rest % G.N, comp = rest / G.N;
  if (comp >= 3) return;
  int k, m;
  harmonic_km(h, G.K, k, m);
  double2* u = B.U[B.ctl[slot].ucur] + (size_t)slot * 3 * G.coll + (size_t)comp * G.coll + (size_t)n * G.Hp + h;
  double2 x = *u;
  auto rot = [&](const double2 cs) {
    const double re = x.x;
    x.x = re * cs.x + x.y * cs.y;   // products rounded before the sum (no FMA in this file)
    x.y = x.y * cs.x - re * cs.y;
  };
  if (on.x != 0.0) rot(tb[k + G.K]);
  if (on.y != 0.0) rot(tb[2 * G.K + m]);
  *u = x;
}

// A request's admission into a free slot, in stream order between
// iterations of the other slots (one thread).  phase 0, before its factors
// are built: clear the failure status the factor kernels report into.
// phase 1, after the state is loaded: a fresh run of nsteps steps with
// corr_fixed correctors per step (put_state + run_begin of the synchronous
// path), active only if the factorisation succeeded.
__global__ void k_admit(Bufs B, int slot, int phase, int nsteps, int corr_fixed) {
  SlotCtl k = B.ctl[slot];
  if (phase == 0) {
    k.status = 0;
    k.active = 0;
  } else {
    k.step = 0;
    k.t = 0.0;
    k.mode = MODE_PRED;
    k.it = 0;
    k.ncorr = 0;
    k.ndiag = 0;
    k.dterr_bits = 0;
    k.corr_dt = 0.0;
    k.cfl_dt = 0.0;
    k.cfl_dir = 0;
    k.ntc = 0;
    k.refactor = 0;
    k.target_steps = nsteps;
    k.corr_fixed = corr_fixed;
    k.diag_every = 0;
    k.host_driven = 0;
    k.active = k.status == 0 ? 1 : 0;
  }
  B.ctl[slot] = k;
}

// var_imposesym(0) (variables.f90:611-616) on harmonic 0 of U[ucur].
__global__ void k_imposesym(Geo G, Bufs B, int slot0) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int slot = slot_of(B, slot0);
  const SlotCtl* c = &B.ctl[slot];
  if (n >= G.N) return;
  double2* u = B.U[c->ucur] + (size_t)slot * 3 * G.coll;
  const size_t o = (size_t)n * G.Hp;
  u[o] = make_double2(0.0, 0.0);
  u[G.coll + o].y = 0.0;
  u[2 * G.coll + o].y = 0.0;
}

// ---------------------------------------------------------------------------
// Control: tim_check_cgce (timestep.f90:214-253) per slot, buffer roles,
// time and step counters, and the reduction of the diagnostics partials in
// var_coll_norm's blocks (variables.f90:727-749).
// One workgroup per slot.
// ---------------------------------------------------------------------------
// The decisions of the control step on a register copy of the slot's
// control block (recs: the slot's diagnostics records).
__device__ __forceinline__ void control_decide(SlotCtl& k, const SlotParams& prm, DiagRec* recs, bool diag_now,
                                               double E, double Ek0, double Em0) {
  if (diag_now) {
    DiagRec* rec = recs + (k.ndiag % k.diag_cap);
    rec->E = E;
    rec->Ek0 = Ek0;
    rec->Em0 = Em0;
    rec->ncorr = -1.0;
    rec->dterr = 0.0;
  }
  if (k.mode == MODE_PRED) {
    // vel_predictor done: u_ := old current, N_ := this evaluation's raw N
    const int old_base = k.ubase;
    k.ubase = k.ucur;
    k.ucur = k.unext;
    k.unext = old_base;
    k.npred = k.ncur;
    k.ncur = 1 - k.ncur;
    k.it = 1;
    k.ncorr = 0;
    k.lasterr = 1e99;
    k.mode = MODE_CORR;
    if (diag_now) k.ndiag = k.ndiag + 1;
    return;
  }
  // corrector done
  const double dterr = sqrt(__longlong_as_double((long long)k.dterr_bits));
  k.dterr_bits = 0ull;
  k.ncorr = k.ncorr + 1;
  const int old_cur = k.ucur;
  k.ucur = k.unext;
  k.unext = old_cur;
  k.last_dterr = dterr;
  if (k.host_driven) return;   // the caller applies tim_check_cgce itself
  int it = k.it;
  if (it == 1) {
    k.corr_dt = prm.dt * sqrt(prm.dterr_tol / dterr);
    k.lasterr = 1e99;
  }
  int fail = OPF_FAIL_NONE;
  if (!(dterr == dterr)) {
    fail = OPF_FAIL_NAN;
    it = 0;
  } else if (prm.dt < 1e-9 && k.step > 30) {
    fail = OPF_FAIL_DT0;
    it = 0;
  } else if (it > 10) {
    fail = OPF_FAIL_ITS;
    it = 0;
  } else if (dterr > k.lasterr) {
    if (dterr > 2.0 * prm.dterr_tol) fail = OPF_FAIL_INCERR;
    else k.corr_dt = prm.dt / (10.0 * prm.courant);
    it = 0;
  } else if (dterr > prm.dterr_tol) {
    k.lasterr = dterr;
    it = it + 1;
  } else {
    it = 0;
  }
  if (k.corr_fixed > 0) it = k.ncorr >= k.corr_fixed ? 0 : (it == 0 ? 1 : it);
  if (fail != OPF_FAIL_NONE && k.corr_fixed <= 0) {
    k.status = fail;
    k.active = 0;
    return;
  }
  k.it = it;
  if (it == 0) {
    // step complete: tim_t += dt, tim_step += 1 (main.f90:43-44)
    if (k.diag_every > 0 && k.ndiag > 0) {
      DiagRec* rec = recs + ((k.ndiag - 1) % k.diag_cap);
      if ((int)rec->step == k.step) {
        rec->ncorr = k.ncorr;
        rec->dterr = dterr;
      }
    }
    k.last_ncorr = k.ncorr;
    k.t = k.t + prm.dt;
    k.step = k.step + 1;
    k.mode = MODE_PRED;
    if (k.target_steps >= 0 && k.step >= k.target_steps) k.active = 0;
  }
}

__global__ void k_control(Geo G, Bufs B, int slot0) {
  const int slot = slot_of(B, slot0);
  const SlotCtl k0 = B.ctl[slot];   // one batch of loads
  if (!k0.active) return;
  const SlotParams prm = B.prm[slot];
  __shared__ double Em[3][256];
  __shared__ double Ekm[3][256];   // the k = 0 term of each m block (the Ek(0) sum)
  __shared__ double chunk[4096];
  const bool diag_now = k0.mode == MODE_PRED && k0.diag_every > 0 && (k0.step % k0.diag_every) == 0;
  if (diag_now) {
    // Em(m) accumulates the block of m.  Whole m blocks are staged through
    // LDS by all threads (coalesced), then one wave per block sums it as a
    // tree (the reference sums in harmonic order; the two differ by
    // rounding, about 1e-16 relative).
    const int K = G.K, M = G.M, bl = 2 * K - 1;
    auto hstart = [&](int m) { return m == 0 ? 0 : K + (m - 1) * bl; };
    auto hend = [&](int m) { return m == 0 ? K : K + m * bl; };
    for (int comp = 0; comp < 3; ++comp) {
      const double* db = B.diagb + ((size_t)slot * 3 + comp) * G.Hp;
      for (int m0 = 0; m0 < M;) {
        int m1 = m0 + 1;
        while (m1 < M && hend(m1) - hstart(m0) <= 4096) ++m1;
        const int h0 = hstart(m0), nh = hend(m1 - 1) - h0;
        // all of a thread's loads of the chunk in one batch (one round trip)
        for (int b0 = 0; b0 < nh; b0 += 16 * (int)blockDim.x) {
          double v[16];
#pragma unroll
          for (int u = 0; u < 16; ++u) {
            const int i = b0 + u * (int)blockDim.x + (int)threadIdx.x;
            v[u] = i < nh ? db[h0 + i] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < 16; ++u) {
            const int i = b0 + u * (int)blockDim.x + (int)threadIdx.x;
            if (i < nh) chunk[i] = v[u];
          }
        }
        __syncthreads();
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
        for (int m = m0 + wave; m < m1; m += nw) {
          const int a = hstart(m) - h0, n = hend(m) - hstart(m);
          double sm = 0.0;
          for (int i = lane; i < n; i += 64) sm = sm + chunk[a + i];
          for (int off = 32; off >= 1; off >>= 1) sm = sm + __shfl_xor(sm, off);
          if (lane == 0 && m < 256) {
            Em[comp][m] = sm;
            Ekm[comp][m] = chunk[(m == 0 ? 0 : hstart(m) + (K - 1)) - h0];
          }
        }
        __syncthreads();
        m0 = m1;
      }
    }
  }
  if (threadIdx.x != 0) return;
  double E = 0.0, Ek0 = 0.0, Em0 = 0.0;
  if (diag_now) {
    for (int comp = 0; comp < 3; ++comp) {
      double e = 0.0, ek = 0.0;   // Ek(0): harmonics with k = 0 in order, m = 0..M-1
      for (int m = 0; m < G.M; ++m) {
        e = e + Em[comp][m];
        ek = ek + Ekm[comp][m];
      }
      E = E + e;
      Ek0 = Ek0 + ek;
      Em0 = Em0 + Em[comp][0];
    }
  }
  SlotCtl k = k0;
  control_decide(k, prm, B.diag + (size_t)slot * k.diag_cap, diag_now, E, Ek0, Em0);
  B.ctl[slot] = k;   // one batch of stores
}

// ---------------------------------------------------------------------------
// vel_maxtstep (velocity.f90:438-481) from the per-plane maxima that the
// theta kernel recorded for the last evaluation, and tim_new_tstep
// (timestep.f90:259-278).  what = 0: the device loop: for adaptive slots
// about to run a predictor, both, with tim_new_dt left in `refactor` for the
// gated vel_matrices launches that follow (every other slot: refactor = 0);
// what = 1: vel_maxtstep only.  One wave per slot: lane l takes the planes
// n = l, l+64, ...; the minima are exact, so their order is free.
// ---------------------------------------------------------------------------
__global__ void k_newdt(Geo G, Bufs B, int slot0, int what) {
  const int slot = slot_of(B, slot0);
  SlotCtl* c = &B.ctl[slot];
  SlotParams* prm = &B.prm[slot];
  const int lane = threadIdx.x;
  if (what == 0) {
    const bool act = c->active && c->mode == MODE_PRED && prm->adaptive;
    if (!act) {
      if (lane == 0) c->refactor = 0;
      return;
    }
  }
  const int N = G.N;
  const MeshOff mo = MeshOff::make(N);
  const double* r = G.mesh + mo.r;
  const unsigned long long* mx = B.cflmax + (size_t)slot * 3 * N;
  const double dth = (2.0 * D_PI) / (double)(G.T * G.Mp);
  const double dz = (2.0 * D_PI) / (prm->alpha * (double)G.Z);
  double t1 = 1e99, t2 = 1e99, t3 = 1e99;
  for (int n = lane; n < N; n += 64) {
    double d;
    if (n == 0) d = r[1] - r[0];
    else if (n == N - 1) d = r[N - 1] - r[N - 2];
    else d = fmin(r[n] - r[n - 1], r[n + 1] - r[n]);
    const double a = __longlong_as_double((long long)mx[n]);
    const double b = __longlong_as_double((long long)mx[N + n]);
    const double e = __longlong_as_double((long long)mx[2 * N + n]);
    if (a != 0.0) t1 = fmin(t1, d / a);
    if (b != 0.0) t2 = fmin(t2, dth / b);
    if (e != 0.0) t3 = fmin(t3, dz / e);
  }
  for (int off = 32; off >= 1; off >>= 1) {
    t1 = fmin(t1, __shfl_xor(t1, off));
    t2 = fmin(t2, __shfl_xor(t2, off));
    t3 = fmin(t3, __shfl_xor(t3, off));
  }
  if (lane != 0) return;
  const double cfl = fmin(t1, fmin(t2, t3));
  int dir = c->cfl_dir;
  if (cfl == t1) dir = 1;
  if (cfl == t2) dir = 2;
  if (cfl == t3) dir = 3;
  c->cfl_dt = cfl;
  c->cfl_dir = dir;
  if (what == 1) return;
  // tim_new_tstep
  const double dt0 = prm->dt;
  double dt = fmin(dt0 * 1.11, prm->maxdt);
  if (c->step == 0 && c->corr_dt == 0.0) dt = fmin(dt, cfl * 0.1);
  if (cfl > 0.0) dt = fmin(dt, cfl * prm->courant);
  if (c->corr_dt > 0.0) dt = fmin(dt, c->corr_dt * 0.95);
  const int i = c->ntc - 1;
  const bool nw = dt < dt0 * 0.95 || (dt > dt0 * 1.10 && i < 0) || (dt == prm->maxdt && dt0 < dt && i < 0);
  c->ntc = (nw && dt >= dt0) ? 99 : i;
  if (nw) prm->dt = dt;
  c->refactor = nw ? 1 : 0;
}

// tim_lumesh_invert (timestep.f90:102-123) of one coll field x[n][Hp] with
// the slot's factor `which`; (0,0) of the pressure matrix gives zero.
__global__ void k_lumesh_invert(Geo G, Bufs B, int slot, int which, int bc, double2* x) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= G.H) return;
  int k, m;
  harmonic_km(h, G.K, k, m);
  const int N = G.N, P = G.P;
  const int p = m * G.K + (k < 0 ? -k : k);
  if (bc == 1 && k == 0 && m == 0) {   // the (0,0) pressure system is singular: b = 0
    for (int n = 0; n < N; ++n) x[(size_t)n * G.Hp + h] = make_double2(0.0, 0.0);
    return;
  }
  lu_solve<double2>(lu_block(B, slot, which, p, P, N), pv_block(B, slot, which, p, P, N), x + h, G.Hp, N);
}

}  // namespace opf
