// Per-harmonic kernels of the time step: banded LU factorisation and solves,
// the pressure-Poisson projection with the influence-matrix correction, the
// predictor/corrector step, the curl, diagnostics and the on-device corrector
// control.
//
// Layout: a spectral (coll) field is stored harmonic-fastest, one complex
// double2 per (n, h): element (slot, comp, n, h) at ((slot*3+comp)*N + n)*Hp + h.
// One lane owns one harmonic, so every radial sweep reads and writes rows of
// 64 consecutive harmonics: fully coalesced, and the transposes of the
// reference's MPI code (variables.f90:184-370) disappear.  The LU factors are
// stored once per (|k|, m) pair (the matrices depend on k only through
// (alpha k)^2, timestep.f90:76-77) as [slot][mat][n][13 rows][pair].
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
// F: factor rows of this lane's matrix, element (n,row) at F[(n*13+row)*P];
// pv: pivot offsets ipiv(j)-j at pv[n*P]; x: vector with stride xs, solved
// in place.  T is double or double2 (real and imaginary parts solved with
// the same factor, as tim_lumesh_invert does with two dgbtrs calls).
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
__device__ void lu_solve(const double* F, const unsigned char* pv, int P, T* x, size_t xs, int N) {
  using V = VT<T>;
  // L solve with row interchanges: window w[i] = b(j+i)
  T w[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) w[i] = i < N ? x[i * xs] : V::zero();
  for (int j = 0; j < N - 1; ++j) {
    const int off = pv[(size_t)j * P];
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
    const double* Fj = F + (size_t)j * NLU * P;
#pragma unroll
    for (int i = 1; i <= KL; ++i)
      if (i <= lm) V::axpy(w[i], Fj[(size_t)(8 + i) * P], t);
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
    const double* Fj = F + (size_t)j * NLU * P;
    v[8] = V::div(v[8], Fj[(size_t)8 * P]);
    const T t = v[8];
    x[j * xs] = t;
#pragma unroll
    for (int d = 1; d <= 8; ++d)
      if (j - d >= 0) V::sub_mul(v[8 - d], t, Fj[(size_t)(8 - d) * P]);
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
// tim_lumesh_init (timestep.f90:67-96) + DGBTF2, one lane per (pair, matrix).
// grid: x = pairs/64, y = 4 matrices, z = slot.
// ---------------------------------------------------------------------------
__global__ void k_lu_factor(Geo G, Bufs B, int slot0) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int mat = blockIdx.y;
  const int slot = slot0 + blockIdx.z;
  if (p >= G.P) return;
  const int N = G.N, P = G.P;
  const int m = p / G.K, k = p % G.K;  // pair (|k|, m)
  const SlotParams prm = B.prm[slot];
  const MeshOff mo = MeshOff::make(N);
  const double* me = G.mesh;
  int PM, BC;
  double c1, c2;
  if (mat == MAT_LNP) { PM = 0; BC = 1; c1 = 0.0; c2 = 1.0; }
  else {
    PM = mat == MAT_LDP ? 1 : (mat == MAT_LDM ? -1 : 0);
    BC = 0;
    c1 = 1.0 / prm.dt;
    c2 = -prm.implicit * prm.nu / prm.Re;
  }
  double* F = B.LU + ((size_t)(slot * 4 + mat) * N) * NLU * P + p;
  unsigned char* pv = B.PV + ((size_t)(slot * 4 + mat) * N) * P + p;
  auto A = [&](int row, int col) -> double& { return F[((size_t)col * NLU + row) * P]; };
  for (int j = 0; j < N; ++j) {
    for (int row = 0; row < KL; ++row) A(row, j) = 0.0;
    for (int q = 0; q < LDA; ++q) A(KL + q, j) = c2 * me[mo.lapb + q * N + j];
    const double d = diag_d(me[mo.rinv2 + j], G.Mp, m, prm.alpha, k, PM);
    A(2 * KL, j) = (A(2 * KL, j) + c2 * d) + c1;
  }
  const int S = (m * G.Mp + (PM < 0 ? -PM : PM)) % 2;
  for (int j1 = 1; j1 <= KL + 1; ++j1) A(2 * KL + 1 - j1, j1 - 1) = me[mo.dr0 + (j1 - 1) + (KL + 1) * (1 - S)];
  for (int j1 = N - KL; j1 <= N; ++j1) A(2 * KL + N - j1, j1 - 1) = me[mo.dr1 + (KL - N + j1) + (KL + 1) * BC];
  for (int j = 0; j < N; ++j) pv[(size_t)j * P] = 0;
  if (BC == 1 && k == 0 && m == 0) return;   // pressure (0,0) is never factored
  // DGBTF2, KL = KU = 4, KV = 8 (0-based rows)
  const int KV = 2 * KL;
  int ju = 0;
  for (int j = 0; j < N; ++j) {
    const int km = N - 1 - j < KL ? N - 1 - j : KL;
    int jp = 0;
    double best = fabs(A(KV, j));
    for (int i = 1; i <= km; ++i) {
      const double v = fabs(A(KV + i, j));
      if (v > best) { best = v; jp = i; }
    }
    pv[(size_t)j * P] = (unsigned char)jp;
    if (A(KV + jp, j) == 0.0) {
      atomicMax(&B.ctl[slot].status, 100 + mat);  // singular factor
      return;
    }
    const int lim = j + KL + jp < N - 1 ? j + KL + jp : N - 1;
    ju = ju > lim ? ju : lim;
    if (jp != 0)
      for (int c = 0; c <= ju - j; ++c) {
        const double t = A(KV + jp - c, j + c);
        A(KV + jp - c, j + c) = A(KV - c, j + c);
        A(KV - c, j + c) = t;
      }
    if (km > 0) {
      const double rc = 1.0 / A(KV, j);
      for (int i = 1; i <= km; ++i) A(KV + i, j) = A(KV + i, j) * rc;
      for (int c = 1; c <= ju - j; ++c) {
        const double y = A(KV - c, j + c);
        if (y != 0.0) {
          const double temp = -y;
          for (int i = 1; i <= km; ++i) A(KV + i - c, j + c) = A(KV + i - c, j + c) + A(KV + i, j) * temp;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// vel_adjPPE(0) influence vectors (velocity.f90:191-218) per pair, and
// vel_adjustFlux(0) (velocity.f90:160-168) on pair 0.
// Uinf slots: 0 U1 (LDp), 1 U2 (LDm), 2 U3 (LDz), 3 U4, 4 U5, 5 p (U6 = p*alpha*k).
// ---------------------------------------------------------------------------
__global__ void k_influence(Geo G, Bufs B, int slot0) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int slot = slot0 + blockIdx.z;
  if (p >= G.P) return;
  const int N = G.N, P = G.P;
  const int m = p / G.K, k = p % G.K;
  const MeshOff mo = MeshOff::make(N);
  const double* me = G.mesh;
  double* U = B.Uinf + (size_t)slot * 6 * N * P + p;
  auto LUp = [&](int mat) { return B.LU + ((size_t)(slot * 4 + mat) * N) * NLU * P + p; };
  auto PVp = [&](int mat) { return B.PV + ((size_t)(slot * 4 + mat) * N) * P + p; };
  for (int j = 0; j < 3; ++j) {
    double* x = U + (size_t)j * N * P;
    for (int n = 0; n < N; ++n) x[(size_t)n * P] = n == N - 1 ? 1.0 : 0.0;
    lu_solve<double>(LUp(j), PVp(j), P, x, P, N);
  }
  double* pp = U + (size_t)5 * N * P;
  if (k == 0 && m == 0) {
    for (int n = 0; n < N; ++n) pp[(size_t)n * P] = 0.0;
  } else {
    for (int n = 0; n < N; ++n) pp[(size_t)n * P] = n == N - 1 ? -1.0 : 0.0;
    lu_solve<double>(LUp(MAT_LNP), PVp(MAT_LNP), P, pp, P, N);
  }
  double* U4 = U + (size_t)3 * N * P;
  double* U5 = U + (size_t)4 * N * P;
  for (int n = 0; n < N; ++n) {
    // grad: dr1 . p (meshmult from 0), t.Im = p * m/r; rt2pm -> U4 = gr - tIm, U5 = gr + tIm
    double gr = 0.0;
    const int j0 = n - KL < 0 ? 0 : n - KL, j1 = n + KL > N - 1 ? N - 1 : n + KL;
    for (int j = j0; j <= j1; ++j) gr = gr + me[mo.dr1b + (KL + n - j) * N + j] * pp[(size_t)j * P];
    const double mr = (me[mo.rinv + n] * (double)G.Mp) * (double)m;
    const double tIm = pp[(size_t)n * P] * mr;
    U4[(size_t)n * P] = gr - tIm;
    U5[(size_t)n * P] = gr + tIm;
  }
  if (p == 0) {
    const SlotParams prm = B.prm[slot];
    double* Ui = B.Ui + (size_t)slot * N;
    for (int n = 0; n < N; ++n) Ui[n] = (n == 0 || n == N - 1) ? 0.0 : 1.0;
    if (prm.const_flux) {
      lu_solve<double>(LUp(MAT_LDZ), PVp(MAT_LDZ), P, Ui, 1, N);
      double s = 0.0;
      for (int n = 0; n < N; ++n) s = s + Ui[n] * me[mo.intrdr + n];
      B.ctl[slot].flux_d1 = 2.0 * s;
      if (2.0 * s == 0.0) atomicMax(&B.ctl[slot].status, 110);
    }
  }
}

// Inverted 4x4 influence matrices per harmonic (velocity.f90:219-225).
__global__ void k_ainv(Geo G, Bufs B, int slot0) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int slot = slot0 + blockIdx.z;
  if (h >= G.H) return;
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
__device__ void curl_lane(const Geo& G, int h, int k, int m, double alpha,
                          const double2* ur, const double2* ut, const double2* uz, double2* om) {
  const int N = G.N;
  const size_t hs = G.Hp;
  const MeshOff mo = MeshOff::make(N);
  const double* me = G.mesh;
  const double ak = alpha * (double)k;
  double2* orr = om;
  double2* ot = om + G.coll;
  double2* oz = om + 2 * G.coll;
  for (int n = 0; n < N; ++n) {
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
}

// which = 0: curl of U[ucur]; used after put_state and after every step.
__global__ void k_curl(Geo G, Bufs B, int slot0) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int slot = slot0 + blockIdx.z;
  const SlotCtl* c = &B.ctl[slot];
  if (!c->active || h >= G.H) return;
  int k, m;
  harmonic_km(h, G.K, k, m);
  const double2* u = B.U[c->ucur] + (size_t)slot * 3 * G.coll;
  curl_lane(G, h, k, m, B.prm[slot].alpha, u, u + G.coll, u + 2 * G.coll,
            B.om + (size_t)slot * 3 * G.coll);
}

// var_imposesym(0) (variables.f90:611-616) on harmonic 0 of U[ucur].
__global__ void k_imposesym(Geo G, Bufs B, int slot0) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int slot = slot0 + blockIdx.z;
  const SlotCtl* c = &B.ctl[slot];
  if (n >= G.N) return;
  double2* u = B.U[c->ucur] + (size_t)slot * 3 * G.coll;
  const size_t o = (size_t)n * G.Hp;
  u[o] = make_double2(0.0, 0.0);
  u[G.coll + o].y = 0.0;
  u[2 * G.coll + o].y = 0.0;
}

// ---------------------------------------------------------------------------
// The step kernel: vel_predictor / vel_corrector (velocity.f90:400-432) with
// vel_addHPF (:338-361), tim_nlincorr (timestep.f90:174-184), vel_step
// (velocity.f90:367-394) and tim_measurecorr (timestep.f90:190-207), for all
// active slots, one lane per harmonic.  The raw nonlinear terms Nraw come
// from the transform kernels without the base-flow terms; they are added
// here from the same u the reference's vel_addHPF sees (the current iterate),
// and the corrector rebuilds the predictor's N_ from its raw terms and u_.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void hpf_point(double2& Nr, double2& Nt, double2& Nz,
                                          double2 ur, double2 ut, double2 uz, double a, double b) {
  Nr = make_double2(Nr.x + a * ur.y, Nr.y - a * ur.x);
  Nt = make_double2(Nt.x + a * ut.y, Nt.y - a * ut.x);
  Nz = make_double2((Nz.x + a * uz.y) + b * ur.x, (Nz.y - a * uz.x) + b * ur.y);
}

__device__ __forceinline__ double pr0_of(const double2* uz, size_t hs, int N, const double* dr1c) {
  // vel_Pr0 = dot_product(uz%Re(N-KL:N,0), dr1(:,1)) / (-2)
  double s = 0.0;
#pragma unroll
  for (int i = 0; i <= KL; ++i) s = s + uz[(size_t)(N - 1 - KL + i) * hs].x * dr1c[i];
  return s / (-2.0);
}

__global__ void __launch_bounds__(64) k_step(Geo G, Bufs B, int slot0) {
  const int slot = slot0 + blockIdx.z;
  SlotCtl* ctl = &B.ctl[slot];
  if (!ctl->active) return;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = h < G.H;
  const int N = G.N, P = G.P;
  const size_t hs = G.Hp, cs = G.coll;
  const MeshOff mo = MeshOff::make(N);
  const double* me = G.mesh;
  const double* dr1c = me + mo.dr1 + (KL + 1);
  const SlotParams prm = B.prm[slot];
  const int mode = ctl->mode;
  const bool corr = mode == MODE_CORR;
  const size_t s3 = (size_t)slot * 3 * cs;

  // u_ (start of step): the predictor copies u -> u_ first (velocity.f90:402-404),
  // so in predictor mode it is the current state itself
  const double2* ub = B.U[corr ? ctl->ubase : ctl->ucur] + s3;
  const double2* uc = B.U[ctl->ucur] + s3;    // current iterate (HPF source)
  double2* un = B.U[ctl->unext] + s3;         // output
  const double2* Nc = B.Nraw[ctl->ncur] + s3;
  const double2* Np = B.Nraw[ctl->npred] + s3;
  double2* R = B.R + s3;
  double2* S = B.S + (size_t)slot * cs;

  int k = 0, m = 0;
  if (live) harmonic_km(h, G.K, k, m);
  const int p = m * G.K + (k < 0 ? -k : k);
  const double ak = prm.alpha * (double)k;
  const bool h00 = live && h == 0;
  // vel_Pr0 from the u seen by each vel_addHPF
  const double pr_c = h00 ? pr0_of(uc + 2 * cs, hs, N, dr1c) : 0.0;
  const double pr_b = h00 ? pr0_of(ub + 2 * cs, hs, N, dr1c) : 0.0;
  const double th = prm.implicit, th1 = 1.0 - prm.implicit;
  double dmax = 0.0;     // max |u_new - u_cur|^2 of this lane (corrector)
  const double c1 = 1.0 / prm.dt;
  const double c2 = (1.0 - prm.implicit) * prm.nu / prm.Re;

  // diagnostics at the start of a step (io_write2files point): per-harmonic
  // energy partials of the current state and the friction quantities
  const bool want_diag = !corr && ctl->diag_every > 0 && (ctl->step % ctl->diag_every) == 0;

  if (live) {
    if (want_diag) {
      const double w = 4.0 * D_PI * D_PI / prm.alpha;
      double* db = B.diagb + (size_t)slot * 3 * hs;
      for (int c = 0; c < 3; ++c) {
        double b = 0.0;
        for (int n = 0; n < N; ++n) {
          const double2 v = uc[c * cs + (size_t)n * hs + h];
          b = b + (w * (v.x * v.x + v.y * v.y)) * me[mo.intrdr + n];
        }
        if (h == 0) b = 0.5 * b;
        db[c * hs + h] = b;
      }
      if (h00) {
        double s = 0.0;
        for (int i = 0; i <= KL; ++i) s = s + uc[2 * cs + (size_t)i * hs].x * me[mo.dr0 + i];
        double f = 0.0;
        for (int i = 0; i <= KL; ++i) f = f + uc[2 * cs + (size_t)(N - 1 - KL + i) * hs].x * dr1c[i];
        DiagRec* rec = B.diag + (size_t)slot * ctl->diag_cap + (ctl->ndiag % ctl->diag_cap);
        rec->step = ctl->step;
        rec->t = ctl->t;
        rec->Ub = pr_c;
        rec->Uc = 1.0 + s;
        rec->Ufr = sqrt(fabs((f - 2.0) / prm.Re));
      }
    }

    // -- stage 1: effective N, RHS = Lt u_ + N (tim_meshmult) --------------
    for (int n = 0; n < N; ++n) {
      const size_t o = (size_t)n * hs + h;
      const double a = ak * me[mo.Ubase + n];
      const double b = me[mo.twor + n];
      double2 nr = Nc[o], nt = Nc[cs + o], nz = Nc[2 * cs + o];
      hpf_point(nr, nt, nz, uc[o], uc[cs + o], uc[2 * cs + o], a, b);
      if (h00) nz.x = nz.x + (4.0 * pr_c) / prm.Re;
      if (corr) {
        double2 pr = Np[o], pt = Np[cs + o], pz = Np[2 * cs + o];
        hpf_point(pr, pt, pz, ub[o], ub[cs + o], ub[2 * cs + o], a, b);
        if (h00) pz.x = pz.x + (4.0 * pr_b) / prm.Re;
        nr = make_double2(th * nr.x + th1 * pr.x, th * nr.y + th1 * pr.y);
        nt = make_double2(th * nt.x + th1 * pt.x, th * nt.y + th1 * pt.y);
        nz = make_double2(th * nz.x + th1 * pz.x, th * nz.y + th1 * pz.y);
      }
      // rt2pm of N
      const double2 cp = make_double2(nr.x - nt.y, nr.y + nt.x);
      const double2 cm = make_double2(nr.x + nt.y, nr.y - nt.x);
      double2 rp = cp, rm = cm, rz = nz;
      const double r2i = me[mo.rinv2 + n];
      const int j0 = n - KL < 0 ? 0 : n - KL, j1 = n + KL > N - 1 ? N - 1 : n + KL;
      for (int j = j0; j <= j1; ++j) {
        const double lap = me[mo.lapb + (KL + n - j) * N + j];
        double ap = c2 * lap, am = c2 * lap, az = c2 * lap;
        if (j == n) {
          ap = (ap + c2 * diag_d(r2i, G.Mp, m, prm.alpha, k, 1)) + c1;
          am = (am + c2 * diag_d(r2i, G.Mp, m, prm.alpha, k, -1)) + c1;
          az = (az + c2 * diag_d(r2i, G.Mp, m, prm.alpha, k, 0)) + c1;
        }
        const size_t oj = (size_t)j * hs + h;
        const double2 br = ub[oj], bt = ub[cs + oj], bz = ub[2 * cs + oj];
        const double2 up = make_double2(br.x - bt.y, br.y + bt.x);
        const double2 um = make_double2(br.x + bt.y, br.y - bt.x);
        rp.x = rp.x + ap * up.x; rp.y = rp.y + ap * up.y;
        rm.x = rm.x + am * um.x; rm.y = rm.y + am * um.y;
        rz.x = rz.x + az * bz.x; rz.y = rz.y + az * bz.y;
      }
      R[o] = rp;
      R[cs + o] = rm;
      R[2 * cs + o] = rz;
    }

    // -- stage 2: pressure RHS = div(pm2rt(R+, R-), Rz), zero BC rows --------
    for (int n = 0; n < N; ++n) {
      const size_t o = (size_t)n * hs + h;
      double2 dv = make_double2(0.0, 0.0);
      if (n != 0 && n != N - 1) {
        const double mr = (me[mo.rinv + n] * (double)G.Mp) * (double)m;
        double2 d = make_double2(0.0, 0.0);
        const int j0 = n - KL < 0 ? 0 : n - KL, j1 = n + KL > N - 1 ? N - 1 : n + KL;
        for (int j = j0; j <= j1; ++j) {
          const size_t oj = (size_t)j * hs + h;
          const double2 a = R[oj], bm = R[cs + oj];
          const double2 rr = make_double2(0.5 * (a.x + bm.x), 0.5 * (a.y + bm.y));
          const double wgt = me[mo.dr1b + (KL + n - j) * N + j];
          d.x = d.x + wgt * rr.x;
          d.y = d.y + wgt * rr.y;
        }
        const double2 a = R[o], bm = R[cs + o], z = R[2 * cs + o];
        const double2 rr = make_double2(0.5 * (a.x + bm.x), 0.5 * (a.y + bm.y));
        const double2 tt = make_double2(0.5 * (a.y - bm.y), -0.5 * (a.x - bm.x));
        const double ri = me[mo.rinv + n];
        dv.x = ((rr.x * ri + d.x) - tt.y * mr) - z.y * ak;
        dv.y = ((rr.y * ri + d.y) + tt.x * mr) + z.x * ak;
      }
      S[o] = dv;
    }
    // pressure solve (LNp, Neumann); harmonic (0,0) has p = 0
    const size_t lu_mat = (size_t)N * NLU * P;
    const double* LUs = B.LU + (size_t)slot * 4 * lu_mat + p;
    const unsigned char* PVs = B.PV + (size_t)slot * 4 * N * P + p;
    if (h00) {
      for (int n = 0; n < N; ++n) S[(size_t)n * hs + h] = make_double2(0.0, 0.0);
    } else {
      lu_solve<double2>(LUs + MAT_LNP * lu_mat, PVs + (size_t)MAT_LNP * N * P, P, S + h, hs, N);
    }

    // -- stage 3: subtract grad p (rt2pm), zero BC rows ---------------------
    for (int n = 0; n < N; ++n) {
      const size_t o = (size_t)n * hs + h;
      if (n == 0 || n == N - 1) {
        R[o] = R[cs + o] = R[2 * cs + o] = make_double2(0.0, 0.0);
        continue;
      }
      const double mr = (me[mo.rinv + n] * (double)G.Mp) * (double)m;
      const double2 gr = band_row(me + mo.dr1b, N, n, S + h, hs, make_double2(0.0, 0.0));
      const double2 pv = S[o];
      const double2 gt = make_double2(-(pv.y * mr), pv.x * mr);
      const double2 gz = make_double2(-(pv.y * ak), pv.x * ak);
      const double2 gp = make_double2(gr.x - gt.y, gr.y + gt.x);
      const double2 gm = make_double2(gr.x + gt.y, gr.y - gt.x);
      R[o] = csub(R[o], gp);
      R[cs + o] = csub(R[cs + o], gm);
      R[2 * cs + o] = csub(R[2 * cs + o], gz);
    }

    // -- stage 4: Helmholtz solves LDp, LDm, LDz ----------------------------
    for (int c = 0; c < 3; ++c)
      lu_solve<double2>(LUs + c * lu_mat, PVs + (size_t)c * N * P, P, R + c * cs + h, hs, N);

    // -- stage 5: influence-matrix correction, pm2rt, store -----------------
    double aR[4] = {0, 0, 0, 0}, aI[4] = {0, 0, 0, 0};
    if (!h00) {
      double2 wp[KL + 1], wm[KL + 1], wz[KL + 1];
      for (int i = 0; i <= KL; ++i) {
        const size_t o = (size_t)(N - 1 - KL + i) * hs + h;
        wp[i] = R[o];
        wm[i] = R[cs + o];
        wz[i] = R[2 * cs + o];
      }
      double BRe[4], BIm[4];
      eval_bc(wp, wm, wz, dr1c, me[mo.rinv + N - 1], (double)(m * G.Mp), ak, BRe, BIm);
      const double* Ai = B.Ainv + (size_t)slot * 16 * G.H + h;
      for (int i = 0; i < 4; ++i) {
        double sR = 0.0, sI = 0.0;
        for (int j = 0; j < 4; ++j) {
          const double aij = Ai[(size_t)(i * 4 + j) * G.H];
          sR = sR + aij * BRe[j];
          sI = sI + aij * BIm[j];
        }
        aR[i] = -sR;
        aI[i] = -sI;
      }
    }
    const double* U = B.Uinf + (size_t)slot * 6 * N * P + p;
    for (int n = 0; n < N; ++n) {
      const size_t o = (size_t)n * hs + h;
      double2 up = R[o], um = R[cs + o], uz = R[2 * cs + o];
      if (!h00) {
        const size_t q = (size_t)n * P;
        const double U1 = U[q], U2 = U[(size_t)N * P + q], U3 = U[(size_t)2 * N * P + q];
        const double U4 = U[(size_t)3 * N * P + q], U5 = U[(size_t)4 * N * P + q];
        const double U6 = U[(size_t)5 * N * P + q] * ak;
        up.x = up.x + aR[0] * U1; up.y = up.y + aI[0] * U1;
        um.x = um.x + aR[1] * U2; um.y = um.y + aI[1] * U2;
        uz.x = uz.x - aI[2] * U3; uz.y = uz.y + aR[2] * U3;
        up.x = up.x + aR[3] * U4; up.y = up.y + aI[3] * U4;
        um.x = um.x + aR[3] * U5; um.y = um.y + aI[3] * U5;
        uz.x = uz.x - aI[3] * U6; uz.y = uz.y + aR[3] * U6;
      }
      const double2 ur = make_double2(0.5 * (up.x + um.x), 0.5 * (up.y + um.y));
      const double2 ut = make_double2(0.5 * (up.y - um.y), -0.5 * (up.x - um.x));
      un[o] = ur;
      un[cs + o] = ut;
      un[2 * cs + o] = uz;
      if (corr && !h00) {
        const double2 d0 = csub(uc[o], ur), d1 = csub(uc[cs + o], ut), d2 = csub(uc[2 * cs + o], uz);
        dmax = fmax(dmax, fmax(d0.x * d0.x + d0.y * d0.y, fmax(d1.x * d1.x + d1.y * d1.y, d2.x * d2.x + d2.y * d2.y)));
      }
    }
    if (h00) {
      // vel_adjustFlux(1) (velocity.f90:170-175) and the zeroing of (0,0)
      double Pr0 = pr_c;
      if (prm.const_flux) {
        double s = 0.0;
        for (int n = 0; n < N; ++n) s = s + un[2 * cs + (size_t)n * hs].x * me[mo.intrdr + n];
        const double d2 = 2.0 * s;
        const double d3 = -d2 / ctl->flux_d1;
        const double* Ui = B.Ui + (size_t)slot * N;
        for (int n = 0; n < N; ++n) {
          double2& z = un[2 * cs + (size_t)n * hs];
          z.x = z.x + d3 * Ui[n];
        }
        Pr0 = Pr0 + (d3 * prm.Re) / 4.0;
      }
      for (int n = 0; n < N; ++n) {
        const size_t o = (size_t)n * hs;
        un[o] = make_double2(0.0, 0.0);
        un[cs + o].y = 0.0;
        un[2 * cs + o].y = 0.0;
        if (corr) {
          const double2 d0 = csub(uc[o], un[o]), d1 = csub(uc[cs + o], un[cs + o]),
                        d2 = csub(uc[2 * cs + o], un[2 * cs + o]);
          dmax = fmax(dmax, fmax(d0.x * d0.x + d0.y * d0.y, fmax(d1.x * d1.x + d1.y * d1.y, d2.x * d2.x + d2.y * d2.y)));
        }
      }
      ctl->Pr0 = Pr0;
      ctl->Pr0_hpf = pr_c;
    }
  }
  if (corr) {
    // All 64 lanes take part (lanes past H carry 0).  Max is order
    // independent, and the bit pattern of a non-negative double orders like
    // its value, so an integer atomic max gives the exact global maximum.
    for (int off = 32; off >= 1; off >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, off));
    if ((threadIdx.x & 63) == 0) atomicMax(&ctl->dterr_bits, (unsigned long long)__double_as_longlong(dmax));
  }
}

// ---------------------------------------------------------------------------
// curl of the new iterate (U[unext]) -> om, for the next evaluation
// ---------------------------------------------------------------------------
__global__ void k_curl_next(Geo G, Bufs B, int slot0) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const int slot = slot0 + blockIdx.z;
  const SlotCtl* c = &B.ctl[slot];
  if (!c->active || h >= G.H) return;
  int k, m;
  harmonic_km(h, G.K, k, m);
  const double2* u = B.U[c->unext] + (size_t)slot * 3 * G.coll;
  curl_lane(G, h, k, m, B.prm[slot].alpha, u, u + G.coll, u + 2 * G.coll,
            B.om + (size_t)slot * 3 * G.coll);
}

// ---------------------------------------------------------------------------
// Control: tim_check_cgce (timestep.f90:214-253) per slot, buffer roles,
// time and step counters, and the reduction of the diagnostics partials in
// the reference's order (var_coll_norm, variables.f90:727-749).
// One workgroup per slot.
// ---------------------------------------------------------------------------
__global__ void k_control(Geo G, Bufs B, int slot0) {
  const int slot = slot0 + blockIdx.z;
  SlotCtl* c = &B.ctl[slot];
  if (!c->active) return;
  const SlotParams prm = B.prm[slot];
  __shared__ double Em[3][256];
  __shared__ double Ek0s[3];
  const bool diag_now = c->mode == MODE_PRED && c->diag_every > 0 && (c->step % c->diag_every) == 0;
  if (diag_now) {
    const double* db = B.diagb + (size_t)slot * 3 * G.Hp;
    // Em(m) accumulates the block of m in harmonic order
    for (int idx = threadIdx.x; idx < 3 * G.M; idx += blockDim.x) {
      const int comp = idx / G.M, m = idx % G.M;
      const int h0 = m == 0 ? 0 : G.K + (m - 1) * (2 * G.K - 1);
      const int h1 = m == 0 ? G.K : h0 + 2 * G.K - 1;
      double s = 0.0;
      for (int h = h0; h < h1; ++h) s = s + db[comp * G.Hp + h];
      if (m < 256) Em[comp][m] = s;
    }
    if (threadIdx.x < 3) {
      double s = 0.0;  // Ek(0): harmonics with k = 0 in order, m = 0..M-1
      for (int m = 0; m < G.M; ++m) {
        const int h = m == 0 ? 0 : G.K + (m - 1) * (2 * G.K - 1) + (G.K - 1);
        s = s + db[threadIdx.x * G.Hp + h];
      }
      Ek0s[threadIdx.x] = s;
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (diag_now) {
    double E = 0.0, Ek0 = 0.0, Em0 = 0.0;
    for (int comp = 0; comp < 3; ++comp) {
      double e = 0.0;
      for (int m = 0; m < G.M; ++m) e = e + Em[comp][m];
      E = E + e;
      Ek0 = Ek0 + Ek0s[comp];
      Em0 = Em0 + Em[comp][0];
    }
    DiagRec* rec = B.diag + (size_t)slot * c->diag_cap + (c->ndiag % c->diag_cap);
    rec->E = E;
    rec->Ek0 = Ek0;
    rec->Em0 = Em0;
    rec->ncorr = -1.0;
    rec->dterr = 0.0;
  }
  if (c->mode == MODE_PRED) {
    // vel_predictor done: u_ := old current, N_ := this evaluation's raw N
    const int old_base = c->ubase;
    c->ubase = c->ucur;
    c->ucur = c->unext;
    c->unext = old_base;
    c->npred = c->ncur;
    c->ncur = 1 - c->ncur;
    c->it = 1;
    c->ncorr = 0;
    c->lasterr = 1e99;
    c->mode = MODE_CORR;
    if (diag_now) c->ndiag = c->ndiag + 1;
    return;
  }
  // corrector done
  const double dterr = sqrt(__longlong_as_double((long long)c->dterr_bits));
  c->dterr_bits = 0ull;
  c->ncorr = c->ncorr + 1;
  const int old_cur = c->ucur;
  c->ucur = c->unext;
  c->unext = old_cur;
  c->last_dterr = dterr;
  if (c->host_driven) return;   // the caller applies tim_check_cgce itself
  int it = c->it;
  if (it == 1) {
    c->corr_dt = prm.dt * sqrt(prm.dterr_tol / dterr);
    c->lasterr = 1e99;
  }
  int fail = OPF_FAIL_NONE;
  if (!(dterr == dterr)) {
    fail = OPF_FAIL_NAN;
    it = 0;
  } else if (prm.dt < 1e-9 && c->step > 30) {
    fail = OPF_FAIL_DT0;
    it = 0;
  } else if (it > 10) {
    fail = OPF_FAIL_ITS;
    it = 0;
  } else if (dterr > c->lasterr) {
    if (dterr > 2.0 * prm.dterr_tol) fail = OPF_FAIL_INCERR;
    else c->corr_dt = prm.dt / (10.0 * prm.courant);
    it = 0;
  } else if (dterr > prm.dterr_tol) {
    c->lasterr = dterr;
    it = it + 1;
  } else {
    it = 0;
  }
  if (c->corr_fixed > 0) it = c->ncorr >= c->corr_fixed ? 0 : (it == 0 ? 1 : it);
  if (fail != OPF_FAIL_NONE && c->corr_fixed <= 0) {
    c->status = fail;
    c->active = 0;
    return;
  }
  c->it = it;
  if (it == 0) {
    // step complete: tim_t += dt, tim_step += 1 (main.f90:43-44)
    if (c->diag_every > 0 && c->ndiag > 0) {
      DiagRec* rec = B.diag + (size_t)slot * c->diag_cap + ((c->ndiag - 1) % c->diag_cap);
      if ((int)rec->step == c->step) {
        rec->ncorr = c->ncorr;
        rec->dterr = dterr;
      }
    }
    c->last_ncorr = c->ncorr;
    c->t = c->t + prm.dt;
    c->step = c->step + 1;
    c->mode = MODE_PRED;
    if (c->target_steps >= 0 && c->step >= c->target_steps) c->active = 0;
  }
}

// tim_lumesh_invert (timestep.f90:102-123) of one coll field x[n][Hp] with
// the slot's factor `which`; (0,0) of the pressure matrix gives zero.
__global__ void k_lumesh_invert(Geo G, Bufs B, int slot, int which, double2* x) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= G.H) return;
  int k, m;
  harmonic_km(h, G.K, k, m);
  const int N = G.N, P = G.P;
  const int p = m * G.K + (k < 0 ? -k : k);
  if (which == MAT_LNP && k == 0 && m == 0) {
    for (int n = 0; n < N; ++n) x[(size_t)n * G.Hp + h] = make_double2(0.0, 0.0);
    return;
  }
  const double* F = B.LU + ((size_t)(slot * 4 + which) * N) * NLU * P + p;
  const unsigned char* pv = B.PV + ((size_t)(slot * 4 + which) * N) * P + p;
  lu_solve<double2>(F, pv, P, x + h, G.Hp, N);
}

}  // namespace opf
