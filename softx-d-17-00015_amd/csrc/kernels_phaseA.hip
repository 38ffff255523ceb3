// The fused per-harmonic step kernel: vel_predictor / vel_corrector
// (velocity.f90:400-432) with vel_addHPF (:338-361), tim_nlincorr
// (timestep.f90:174-184), vel_step (velocity.f90:367-394), tim_measurecorr
// (timestep.f90:190-207) and the curl of the new velocity for the next
// nonlinear evaluation (var_coll_curl, variables.f90:408-436).
//
// One workgroup owns a group of `gp` consecutive (|k|, m) pairs, i.e. the
// harmonics +k and -k of each pair (2*gp "harmonic slots").  Everything a
// harmonic needs between its input and output stays in LDS:
//   V[4][N][GH]  the three right-hand sides (u+, u-, uz) and the pressure
// Threads map to (radial point, harmonic) for the stages that are parallel in
// r (right-hand side, divergence, gradient, influence correction, curl).  The
// banded LU solves are serial in r; they run one lane per (harmonic, real or
// imaginary part, system) in wave 0 while the other waves stream the factor
// columns of the group's pairs into a double-buffered LDS chunk, so the
// solve lanes never wait on HBM.  Factors are stored once per pair and shared
// by the +k / -k harmonics and their real and imaginary parts.
//
// Arithmetic follows the reference operation by operation (see the cited
// routines); products and sums are rounded separately (no contraction).
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
  double2* UB;         // u_ staged for the right-hand-side stencil [3][N][GH]
  double* FC;          // [2][3][gp][FCP]
  double* E;           // energy terms [3][N][GH] (aliases FC; used before the solves)
  unsigned char* PVS;  // [4][gp][N]
  double* AC;          // [GH][8]
  double* RED;         // [8]
  int GH, FCP;
};

__device__ __forceinline__ double2& Vat(const Lds& L, int c, int n, int lh, int N) {
  return L.V[((size_t)c * N + n) * L.GH + lh];
}

// Load chunk `c` of the factors of matrices mats[0..nmat) for the group into
// LDS buffer `buf`, using threads [first, STEP_THREADS).
// The systems of a solve phase use consecutive factor matrices (LDp, LDm,
// LDz = 0, 1, 2 or LNp alone) acting on consecutive LDS components.
struct Mats {
  int m0;   // first factor matrix
  int c0;   // first LDS component
};

__device__ __forceinline__ double2 chunk_src(const Bufs& B, int slot, int m0, int np, int per, int P, int N, int j0,
                                             int p0, int i) {
  const int mi = i / (np * per);
  const int rem = i - mi * np * per;
  const int pl = rem / per;
  const int q = rem - pl * per;
  return ((const double2*)(lu_block(B, slot, m0 + mi, p0 + pl, P, N) + (size_t)j0 * LUS))[q];
}

__device__ __forceinline__ void chunk_dst(const Lds& L, int np, int per, int gp, int buf, int i, double2 v) {
  const int mi = i / (np * per);
  const int rem = i - mi * np * per;
  const int pl = rem / per;
  const int q = rem - pl * per;
  ((double2*)(L.FC + (((size_t)buf * 3 + mi) * gp + pl) * L.FCP))[q] = v;
}

// Load chunk `c` (columns [c*cs, c*cs+cs)) of the factors of the phase's
// matrices for the group into LDS buffer `buf`, using threads [first, 256).
// Every loader thread issues its (at most four) loads before its stores, so
// a chunk costs one memory latency.
__device__ __forceinline__ void load_chunk(const Lds& L, const Bufs& B, int slot, const Mats& ms, int nmat, int p0,
                                           int np, int gp, int P, int N, int cs, int c, int buf, int first) {
  const int j0 = c * cs;
  const int ncol = min(cs, N - j0);
  const int per = ncol * (LUS / 2);           // double2 per (matrix, pair)
  const int total = nmat * np * per;
  const int nl = STEP_THREADS - first;
  const int i0 = threadIdx.x - first;
  const int i1 = i0 + nl, i2 = i0 + 2 * nl, i3 = i0 + 3 * nl;
  double2 t0, t1, t2, t3;
  if (i0 < total) t0 = chunk_src(B, slot, ms.m0, np, per, P, N, j0, p0, i0);
  if (i1 < total) t1 = chunk_src(B, slot, ms.m0, np, per, P, N, j0, p0, i1);
  if (i2 < total) t2 = chunk_src(B, slot, ms.m0, np, per, P, N, j0, p0, i2);
  if (i3 < total) t3 = chunk_src(B, slot, ms.m0, np, per, P, N, j0, p0, i3);
  if (i0 < total) chunk_dst(L, np, per, gp, buf, i0, t0);
  if (i1 < total) chunk_dst(L, np, per, gp, buf, i1, t1);
  if (i2 < total) chunk_dst(L, np, per, gp, buf, i2, t2);
  if (i3 < total) chunk_dst(L, np, per, gp, buf, i3, t3);
}

// One solve phase: systems `nsys` (matrices mats[], LDS components comps[]),
// solver lane t < 2*GH*nsys -> (system, harmonic slot, part).
__device__ __forceinline__ void solve_phase(const Lds& L, const Bufs& B, int slot, const Mats& ms, int nsys, int p0,
                                            int np, int gp, int P, int K, int N, int cs, bool pressure) {
  const int tid = threadIdx.x;
  const int nsolv = 2 * L.GH * nsys;
  const int first_loader = nsolv <= 64 ? 64 : ((nsolv + 63) / 64) * 64;
  const bool solver = tid < nsolv;
  const bool loader = tid >= first_loader;
  const int sys = solver ? tid / (2 * L.GH) : 0;
  const int rem = tid - sys * 2 * L.GH;
  const int lh = rem >> 1, part = rem & 1;
  const Slot S = slot_of(lh, gp, np, p0, K);
  const bool act = solver && S.valid && !(pressure && S.h00);
  const int comp = ms.c0 + sys;
  const unsigned char* piv = L.PVS + ((size_t)(ms.m0 + sys) * gp + S.pl) * N;
  double* x = (double*)&Vat(L, comp, 0, lh, N) + part;
  const size_t xs = (size_t)2 * L.GH;   // doubles between radial rows
  const int nch = (N + cs - 1) / cs;

  double w[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) w[i] = 0.0;
  if (loader) load_chunk(L, B, slot, ms, nsys, p0, np, gp, P, N, cs, 0, 0, first_loader);
  if (act) {
#pragma unroll
    for (int i = 0; i < 5; ++i) w[i] = i < N ? x[i * xs] : 0.0;
  }
  __syncthreads();
  // forward: L solve with the row interchanges of DGBTRS.  The pivot and the
  // multipliers of column j+1 are read from LDS while column j is applied.
  // Multiplier slots below the last row hold exact zeros (dgbtf2 never
  // writes them and the band is zero there), so the updates need no guard.
  for (int c = 0; c < nch; ++c) {
    if (loader && c + 1 < nch) load_chunk(L, B, slot, ms, nsys, p0, np, gp, P, N, cs, c + 1, (c + 1) & 1, first_loader);
    const int jbeg = c * cs, jend = min((c + 1) * cs, N - 1);
    if (act && jbeg < jend) {
      const double* F = L.FC + (((size_t)(c & 1) * 3 + sys) * gp + S.pl) * L.FCP - (size_t)jbeg * LUS;
      int off = piv[jbeg];
      double l1 = F[jbeg * LUS + 9], l2 = F[jbeg * LUS + 10], l3 = F[jbeg * LUS + 11], l4 = F[jbeg * LUS + 12];
      for (int j = jbeg; j < jend; ++j) {
        const bool more = j + 1 < jend;
        const int offn = more ? piv[j + 1] : 0;
        const double* Fn = F + (size_t)(more ? j + 1 : j) * LUS;
        const double n1 = Fn[9], n2 = Fn[10], n3 = Fn[11], n4 = Fn[12];
        const double xn = j + 5 < N ? x[(j + 5) * xs] : 0.0;
        const double a0 = w[0];
        double sv = w[0];
        sv = off == 1 ? w[1] : sv;
        sv = off == 2 ? w[2] : sv;
        sv = off == 3 ? w[3] : sv;
        sv = off == 4 ? w[4] : sv;
        w[0] = sv;
        w[1] = off == 1 ? a0 : w[1];
        w[2] = off == 2 ? a0 : w[2];
        w[3] = off == 3 ? a0 : w[3];
        w[4] = off == 4 ? a0 : w[4];
        const double t = -w[0];
        w[1] = w[1] + l1 * t;
        w[2] = w[2] + l2 * t;
        w[3] = w[3] + l3 * t;
        w[4] = w[4] + l4 * t;
        x[j * xs] = w[0];
        w[0] = w[1];
        w[1] = w[2];
        w[2] = w[3];
        w[3] = w[4];
        w[4] = xn;
        off = offn;
        l1 = n1; l2 = n2; l3 = n3; l4 = n4;
      }
    }
    __syncthreads();
  }
  // backward: DTBSV with 8 superdiagonals; the last chunk is still resident.
  // Rows above the first hold exact zeros in the band, so the updates of the
  // window's padding entries are harmless and need no guard.
  double v[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) v[q] = 0.0;
  if (act) {
    x[(N - 1) * xs] = w[0];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int idx = N - 9 + q;
      v[q] = idx >= 0 ? x[idx * xs] : 0.0;
    }
    v[8] = w[0];
  }
  for (int c = nch - 1; c >= 0; --c) {
    // chunk c-1 is already in its buffer when c is the last chunk
    if (loader && c - 1 >= 0 && c != nch - 1)
      load_chunk(L, B, slot, ms, nsys, p0, np, gp, P, N, cs, c - 1, (c - 1) & 1, first_loader);
    const int jbeg = min((c + 1) * cs, N) - 1, jlow = c * cs;
    if (act) {
      const double* F = L.FC + (((size_t)(c & 1) * 3 + sys) * gp + S.pl) * L.FCP - (size_t)jlow * LUS;
      double u[9];
#pragma unroll
      for (int d = 0; d <= 8; ++d) u[d] = F[jbeg * LUS + 8 - d];
      for (int j = jbeg; j >= jlow; --j) {
        const double* Fn = F + (size_t)(j - 1 >= jlow ? j - 1 : j) * LUS;
        double un[9];
#pragma unroll
        for (int d = 0; d <= 8; ++d) un[d] = Fn[8 - d];
        const double xn = j - 9 >= 0 ? x[(j - 9) * xs] : 0.0;
        v[8] = v[8] / u[0];
        const double t = v[8];
        x[j * xs] = t;
#pragma unroll
        for (int d = 1; d <= 8; ++d) v[8 - d] = v[8 - d] - t * u[d];
#pragma unroll
        for (int q = 8; q >= 1; --q) v[q] = v[q - 1];
        v[0] = xn;
#pragma unroll
        for (int d = 0; d <= 8; ++d) u[d] = un[d];
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

}  // namespace

__global__ void __launch_bounds__(STEP_THREADS) k_step(Geo G, Bufs B, StepTiles ST, int slot0) {
  const int slot = slot0 + blockIdx.z;
  SlotCtl* ctl = &B.ctl[slot];
  if (!ctl->active) return;
  // XCD-aware group order: consecutive workgroup ids go to different XCDs,
  // so give each XCD a contiguous range of pairs (neighbouring harmonics
  // share cache lines of the spectral fields).
  const int per = (ST.ngroups + 7) / 8;
  const int grp = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (grp >= ST.ngroups) return;
  const int gp = ST.gp, cs = ST.cs;
  const int N = G.N, P = G.P, K = G.K;
  const int p0 = grp * gp;
  const int np = min(gp, P - p0);
  const int tid = threadIdx.x;
  const size_t hs = G.Hp, csz = G.coll;
  const MeshOff mo = MeshOff::make(N);
  const double* me = G.mesh;
  const double* dr1c = me + mo.dr1 + (KL + 1);
  const SlotParams prm = B.prm[slot];
  const bool corr = ctl->mode == MODE_CORR;
  const size_t s3 = (size_t)slot * 3 * csz;
  const double2* uc = B.U[ctl->ucur] + s3;                 // current iterate
  const double2* ub = B.U[corr ? ctl->ubase : ctl->ucur] + s3;   // u_ (vel_predictor copies u -> u_)
  double2* un = B.U[ctl->unext] + s3;
  const double2* Nc = B.Nraw[ctl->ncur] + s3;
  const double2* Np = B.Nraw[ctl->npred] + s3;
  double2* om = B.om + s3;
  const bool want_diag = !corr && ctl->diag_every > 0 && (ctl->step % ctl->diag_every) == 0;

  extern __shared__ double2 lds2[];
  Lds L;
  L.GH = 2 * gp;
  L.FCP = cs * LUS + 2;
  unsigned char* base = (unsigned char*)lds2;
  L.V = (double2*)base;
  base += (size_t)4 * N * L.GH * sizeof(double2);
  L.UB = (double2*)base;
  base += (size_t)3 * N * L.GH * sizeof(double2);
  L.FC = (double*)base;
  L.E = L.FC;
  const size_t fc_bytes = (size_t)2 * 3 * gp * L.FCP * sizeof(double);
  const size_t e_bytes = (size_t)3 * N * L.GH * sizeof(double);
  base += (fc_bytes > e_bytes ? fc_bytes : e_bytes);
  L.AC = (double*)base;
  base += (size_t)L.GH * 8 * sizeof(double);
  L.RED = (double*)base;
  base += 8 * sizeof(double);
  L.PVS = base;

  // pivots of the group's four matrices
  for (int i = tid; i < 4 * gp * N; i += STEP_THREADS) {
    const int mat = i / (gp * N), rem = i - mat * gp * N, pl = rem / N, n = rem - pl * N;
    L.PVS[i] = pl < np ? pv_block(B, slot, mat, p0 + pl, P, N)[n] : 0;
  }

  const double th = prm.implicit, th1 = 1.0 - prm.implicit;
  const double c1 = 1.0 / prm.dt;
  const double c2 = (1.0 - prm.implicit) * prm.nu / prm.Re;
  const double wnorm = 4.0 * D_PI * D_PI / prm.alpha;
  const int NI = N * L.GH;

  // ---- stage 0: u_ of the group into LDS (one memory latency) ------------
  for (int it = tid; it < NI; it += STEP_THREADS) {
    const int n = it / L.GH, lh = it - n * L.GH;
    const Slot S = slot_of(lh, gp, np, p0, K);
    double2 a0 = make_double2(0.0, 0.0), a1 = a0, a2 = a0;
    if (S.valid) {
      const size_t o = (size_t)n * hs + S.h;
      a0 = ub[o];
      a1 = ub[csz + o];
      a2 = ub[2 * csz + o];
    }
    L.UB[((size_t)0 * N + n) * L.GH + lh] = a0;
    L.UB[((size_t)1 * N + n) * L.GH + lh] = a1;
    L.UB[((size_t)2 * N + n) * L.GH + lh] = a2;
  }
  __syncthreads();

  // ---- stage 1: effective N and RHS = Lt u_ + N (tim_meshmult) -----------
  for (int it = tid; it < NI; it += STEP_THREADS) {
    const int n = it / L.GH, lh = it - n * L.GH;
    const Slot S = slot_of(lh, gp, np, p0, K);
    if (!S.valid) {
      for (int c = 0; c < 4; ++c) Vat(L, c, n, lh, N) = make_double2(0.0, 0.0);
      if (want_diag)
        for (int c = 0; c < 3; ++c) L.E[((size_t)c * N + n) * L.GH + lh] = 0.0;
      continue;
    }
    const size_t o = (size_t)n * hs + S.h;
    const double ak = prm.alpha * (double)S.k;
    const double a = ak * me[mo.Ubase + n];
    const double b = me[mo.twor + n];
    const double2 ucr = uc[o], uct = uc[csz + o], ucz = uc[2 * csz + o];
    double2 nr = Nc[o], nt = Nc[csz + o], nz = Nc[2 * csz + o];
    nr = make_double2(nr.x + a * ucr.y, nr.y - a * ucr.x);
    nt = make_double2(nt.x + a * uct.y, nt.y - a * uct.x);
    nz = make_double2((nz.x + a * ucz.y) + b * ucr.x, (nz.y - a * ucz.x) + b * ucr.y);
    double pr_c = 0.0;
    if (S.h00 && prm.const_flux) {
      pr_c = pr0_of(uc + 2 * csz, hs, N, dr1c);
      nz.x = nz.x + (4.0 * pr_c) / prm.Re;
    }
    if (corr) {
      const double2 bur = L.UB[((size_t)0 * N + n) * L.GH + lh];
      const double2 but = L.UB[((size_t)1 * N + n) * L.GH + lh];
      const double2 buz = L.UB[((size_t)2 * N + n) * L.GH + lh];
      double2 pr = Np[o], pt = Np[csz + o], pz = Np[2 * csz + o];
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
    const double r2i = me[mo.rinv2 + n];
#pragma unroll
    for (int dd = -KL; dd <= KL; ++dd) {
      const int j = n + dd;
      if (j < 0 || j > N - 1) continue;
      const double lap = me[mo.lapb + (KL - dd) * N + j];
      double ap = c2 * lap, am = c2 * lap, az = c2 * lap;
      if (dd == 0) {
        ap = (ap + c2 * diag_d(r2i, G.Mp, S.m, prm.alpha, S.k, 1)) + c1;
        am = (am + c2 * diag_d(r2i, G.Mp, S.m, prm.alpha, S.k, -1)) + c1;
        az = (az + c2 * diag_d(r2i, G.Mp, S.m, prm.alpha, S.k, 0)) + c1;
      }
      const double2 br = L.UB[((size_t)0 * N + j) * L.GH + lh];
      const double2 bt = L.UB[((size_t)1 * N + j) * L.GH + lh];
      const double2 bz = L.UB[((size_t)2 * N + j) * L.GH + lh];
      rp.x = rp.x + ap * (br.x - bt.y);
      rp.y = rp.y + ap * (br.y + bt.x);
      rm.x = rm.x + am * (br.x + bt.y);
      rm.y = rm.y + am * (br.y - bt.x);
      rz.x = rz.x + az * bz.x;
      rz.y = rz.y + az * bz.y;
    }
    Vat(L, 0, n, lh, N) = rp;
    Vat(L, 1, n, lh, N) = rm;
    Vat(L, 2, n, lh, N) = rz;
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
        rec->Ub = prm.const_flux ? pr0_of(uc + 2 * csz, hs, N, dr1c) : 0.0;
        rec->Uc = 1.0 + s;
        rec->Ufr = sqrt(fabs((f - 2.0) / prm.Re));
      }
    }
    if (S.h00 && n == 0) ctl->Pr0_hpf = pr_c;
  }
  __syncthreads();
  // energy partials (var_coll_norm, variables.f90:732-735), n ascending
  if (want_diag && tid < 3 * L.GH) {
    const int c = tid / L.GH, lh = tid - c * L.GH;
    const Slot S = slot_of(lh, gp, np, p0, K);
    if (S.valid) {
      double bsum = 0.0;
      for (int n = 0; n < N; ++n) bsum = bsum + L.E[((size_t)c * N + n) * L.GH + lh] * me[mo.intrdr + n];
      if (S.h00) bsum = 0.5 * bsum;
      B.diagb[(size_t)slot * 3 * hs + (size_t)c * hs + S.h] = bsum;
    }
  }

  // ---- stage 2: pressure RHS = div(pm2rt(R+, R-), Rz) (vel_adjPPE(1)) ----
  for (int it = tid; it < NI; it += STEP_THREADS) {
    const int n = it / L.GH, lh = it - n * L.GH;
    const Slot S = slot_of(lh, gp, np, p0, K);
    double2 dv = make_double2(0.0, 0.0);
    if (S.valid && !S.h00 && n != 0 && n != N - 1) {
      const double mr = (me[mo.rinv + n] * (double)G.Mp) * (double)S.m;
      const double ak = prm.alpha * (double)S.k;
      double2 d = make_double2(0.0, 0.0);
#pragma unroll
      for (int dd = -KL; dd <= KL; ++dd) {
        const int j = n + dd;
        if (j < 0 || j > N - 1) continue;
        const double2 a = Vat(L, 0, j, lh, N), bm = Vat(L, 1, j, lh, N);
        const double wgt = me[mo.dr1b + (KL - dd) * N + j];
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
  {
    const Mats ms{MAT_LNP, 3};
    solve_phase(L, B, slot, ms, 1, p0, np, gp, P, K, N, cs, true);
  }

  // ---- stage 4: subtract grad p (rt2pm), zero the BC rows ----------------
  for (int it = tid; it < NI; it += STEP_THREADS) {
    const int n = it / L.GH, lh = it - n * L.GH;
    const Slot S = slot_of(lh, gp, np, p0, K);
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
      const int j = n + dd;
      if (j < 0 || j > N - 1) continue;
      const double wgt = me[mo.dr1b + (KL - dd) * N + j];
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
  {
    const Mats ms{MAT_LDP, 0};
    solve_phase(L, B, slot, ms, 3, p0, np, gp, P, K, N, cs, false);
  }

  // ---- stage 6: influence-matrix correction (vel_adjPPE(2)) --------------
  if (tid < L.GH) {
    const int lh = tid;
    const Slot S = slot_of(lh, gp, np, p0, K);
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
      const double* Ai = B.Ainv + (size_t)slot * 16 * G.H + S.h;
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
    for (int i = 0; i < 4; ++i) {
      L.AC[lh * 8 + i] = aR[i];
      L.AC[lh * 8 + 4 + i] = aI[i];
    }
  }
  __syncthreads();
  for (int it = tid; it < NI; it += STEP_THREADS) {
    const int n = it / L.GH, lh = it - n * L.GH;
    const Slot S = slot_of(lh, gp, np, p0, K);
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
  __syncthreads();
  // vel_adjustFlux(1) (velocity.f90:170-175) on harmonic (0,0): n ascending
  const bool has00 = p0 == 0;
  if (has00 && tid == 0) {
    double d3 = 0.0, Pr0 = ctl->Pr0_hpf;
    if (prm.const_flux) {
      double s = 0.0;
      for (int n = 0; n < N; ++n) s = s + Vat(L, 2, n, 0, N).x * me[mo.intrdr + n];
      const double d2 = 2.0 * s;
      d3 = -d2 / ctl->flux_d1;
      Pr0 = Pr0 + (d3 * prm.Re) / 4.0;
    }
    L.RED[0] = d3;
    ctl->Pr0 = Pr0;
  }
  __syncthreads();
  double dmax = 0.0;
  for (int it = tid; it < NI; it += STEP_THREADS) {
    const int n = it / L.GH, lh = it - n * L.GH;
    const Slot S = slot_of(lh, gp, np, p0, K);
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
  for (int it = tid; it < NI; it += STEP_THREADS) {
    const int n = it / L.GH, lh = it - n * L.GH;
    const Slot S = slot_of(lh, gp, np, p0, K);
    if (!S.valid) continue;
    const double mr = (me[mo.rinv + n] * (double)G.Mp) * (double)S.m;
    const double ak = prm.alpha * (double)S.k;
    const double2 vr = Vat(L, 0, n, lh, N), vt = Vat(L, 1, n, lh, N), vz = Vat(L, 2, n, lh, N);
    double2 d1 = make_double2(0.0, 0.0), d2 = make_double2(0.0, 0.0);
#pragma unroll
    for (int dd = -KL; dd <= KL; ++dd) {
      const int j = n + dd;
      if (j < 0 || j > N - 1) continue;
      const double wgt = me[mo.dr1b + (KL - dd) * N + j];
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
    // orders like its value: an integer atomic max is exact
    for (int off = 32; off >= 1; off >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, off));
    if ((tid & 63) == 0) atomicMax(&ctl->dterr_bits, (unsigned long long)__double_as_longlong(dmax));
  }
}

}  // namespace opf
