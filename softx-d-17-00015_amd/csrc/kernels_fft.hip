// Dealiased spectral <-> physical transforms with the u x curl(u) products.
//
// The reference transforms each radial plane with FFTW (transform.fftw3.f90:
// 65-147): a complex length-3K transform along z for every azimuthal mode
// (exponent +1 to physical space, -1 back), and a real length-3M transform
// along theta for every z line, with the 3/2 zero padding and truncation done
// by copies.  Here the same DFTs run as LDS-staged Stockham FFTs (radix 4, 2,
// 3) in three kernels with intermediates laid out for coalesced access:
//
//   k_zfwd : 6 coll fields (u, curl u) -> Yf[slot][n][z][6][M]   (z transform)
//   k_theta: Yf -> theta transforms, products, scaling, inverse theta
//            transforms -> Yb[slot][n][z][3][M]
//   k_zbwd : Yb -> z transform, truncation -> Nraw coll fields
//
// The padding, the m = 0 Hermitian fill and the coll <-> spec layout change
// (variables.f90:184-294) are folded into the loads and stores.  Two real
// theta transforms share one complex FFT (a + ib), so the three c2r
// transforms per pair of fields and the r2c transforms run as 3 + 2 complex
// FFTs per line.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace opf {

__device__ __forceinline__ double2 cmul_tw(double2 a, double2 w, double sgn) {
  // a * (w.x + i sgn w.y)
  const double wy = sgn * w.y;
  return make_double2(a.x * w.x - a.y * wy, a.x * wy + a.y * w.x);
}

template <int R> __device__ __forceinline__ void dft(double2* v, double sgn);

template <> __device__ __forceinline__ void dft<2>(double2* v, double) {
  const double2 a = v[0], b = v[1];
  v[0] = make_double2(a.x + b.x, a.y + b.y);
  v[1] = make_double2(a.x - b.x, a.y - b.y);
}

template <> __device__ __forceinline__ void dft<3>(double2* v, double sgn) {
  const double s3 = sgn * 0.86602540378443864676;  // sgn * sin(2 pi / 3)
  const double2 a = v[0], b = v[1], c = v[2];
  const double2 t1 = make_double2(b.x + c.x, b.y + c.y);
  const double2 t2 = make_double2(a.x - 0.5 * t1.x, a.y - 0.5 * t1.y);
  const double2 t3 = make_double2(s3 * (b.x - c.x), s3 * (b.y - c.y));
  v[0] = make_double2(a.x + t1.x, a.y + t1.y);
  // X1 = t2 + i t3, X2 = t2 - i t3
  v[1] = make_double2(t2.x - t3.y, t2.y + t3.x);
  v[2] = make_double2(t2.x + t3.y, t2.y - t3.x);
}

template <> __device__ __forceinline__ void dft<4>(double2* v, double sgn) {
  const double2 a = v[0], b = v[1], c = v[2], d = v[3];
  const double2 s02 = make_double2(a.x + c.x, a.y + c.y);
  const double2 d02 = make_double2(a.x - c.x, a.y - c.y);
  const double2 s13 = make_double2(b.x + d.x, b.y + d.y);
  const double2 d13 = make_double2(b.x - d.x, b.y - d.y);
  // w = sgn * i ; w * z = sgn * (-z.y, z.x)
  const double2 wd = make_double2(-sgn * d13.y, sgn * d13.x);
  v[0] = make_double2(s02.x + s13.x, s02.y + s13.y);
  v[1] = make_double2(d02.x + wd.x, d02.y + wd.y);
  v[2] = make_double2(s02.x - s13.x, s02.y - s13.y);
  v[3] = make_double2(d02.x - wd.x, d02.y - wd.y);
}

// One Stockham pass of radix R over B transforms of length L stored at
// buf[b*L + i].  Ns = product of the radices of the previous passes.
template <int R>
__device__ void fft_pass(double2* buf, int L, int B, int Ns, double sgn, const double2* tw) {
  const int nb = L / R;
  const int total = nb * B;
  const int tid = threadIdx.x;
  const int stride_tw = L / (Ns * R);
  double2 v[FFT_MAXB][R];
#pragma unroll
  for (int i = 0; i < FFT_MAXB; ++i) {
    const int q = tid + i * FFT_THREADS;
    if (q < total) {
      const int b = q / nb, j = q - b * nb;
      const double2* x = buf + b * L;
      const int kk = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = x[j + r * nb];
#pragma unroll
      for (int r = 1; r < R; ++r) v[i][r] = cmul_tw(v[i][r], tw[kk * r * stride_tw], sgn);
      dft<R>(v[i], sgn);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < FFT_MAXB; ++i) {
    const int q = tid + i * FFT_THREADS;
    if (q < total) {
      const int b = q / nb, j = q - b * nb;
      double2* x = buf + b * L;
      const int kk = j % Ns;
      const int base = (j / Ns) * Ns * R + kk;
#pragma unroll
      for (int r = 0; r < R; ++r) x[base + r * Ns] = v[i][r];
    }
  }
  __syncthreads();
}

// Full transform: sgn = +1 (to physical space) or -1, unnormalised (FFTW).
__device__ void fft_lds(double2* buf, int B, const FftPlan& pl, double sgn, const double2* tw) {
  int Ns = 1;
  for (int p = 0; p < pl.nrad; ++p) {
    const int R = pl.rad[p];
    if (R == 4) fft_pass<4>(buf, pl.L, B, Ns, sgn, tw);
    else if (R == 2) fft_pass<2>(buf, pl.L, B, Ns, sgn, tw);
    else fft_pass<3>(buf, pl.L, B, Ns, sgn, tw);
    Ns *= R;
  }
}

extern __shared__ double2 smem[];

// spectral coefficient (k, m) at radial row `row` of a coll field, k in
// (-K, K); for m = 0 the k < 0 half is the conjugate of k > 0
// (transform.fftw3.f90:72-74).
__device__ __forceinline__ double2 coef(const double2* row, int K, int k, int m) {
  if (m == 0) {
    if (k >= 0) return row[k];
    const double2 c = row[-k];
    return make_double2(c.x, -c.y);
  }
  return row[K + (m - 1) * (2 * K - 1) + k + K - 1];
}

// Load the padded z column of (k, m) into buf[0..Z): X(j) = a(k=j) for j < K,
// X(Z+k) = a(k) for k < 0, zero in between (transform.fftw3.f90:72-84).
__device__ __forceinline__ void load_zcol(double2* buf, const double2* row, int K, int Z, int m, int j) {
  double2 v = make_double2(0.0, 0.0);
  if (j < K) v = coef(row, K, j, m);
  else if (j > 2 * K) v = coef(row, K, j - Z, m);
  buf[j] = v;
}

// ---------------------------------------------------------------------------
// z transform of u and curl u for all active slots.
// grid: x = m tiles, y = N * 6 (plane, field), z = slot.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(FFT_THREADS) k_zfwd(Geo G, Bufs B, FftTiles T, int slot0) {
  const int slot = slot0 + blockIdx.z;
  const SlotCtl* c = &B.ctl[slot];
  if (!c->active) return;
  const int n = blockIdx.y / 6, f = blockIdx.y % 6;
  const int m0 = blockIdx.x * T.tm;
  const int tm = min(T.tm, G.M - m0);
  const int Z = G.Z, K = G.K, M = G.M;
  const double2* src = (f < 3 ? B.U[c->ucur] + (size_t)slot * 3 * G.coll + (size_t)f * G.coll
                              : B.om + (size_t)slot * 3 * G.coll + (size_t)(f - 3) * G.coll) +
                       (size_t)n * G.Hp;
  double2* buf = smem;
  for (int i = threadIdx.x; i < tm * Z; i += FFT_THREADS) {
    const int b = i / Z, j = i - b * Z;
    load_zcol(buf + b * Z, src, K, Z, m0 + b, j);
  }
  __syncthreads();
  fft_lds(buf, tm, T.pz, 1.0, B.twZ);
  double2* Y = B.Yf + (((size_t)slot * G.N + n) * Z) * 6 * M;
  for (int i = threadIdx.x; i < tm * Z; i += FFT_THREADS) {
    const int z = i / tm, b = i - z * tm;
    Y[((size_t)z * 6 + f) * M + m0 + b] = buf[b * Z + z];
  }
}

// ---------------------------------------------------------------------------
// theta transforms and products.  grid: x = z tiles, y = N, z = slot.
// LDS: buf[(pair*tz + line)*T + theta] for the three packed pairs, so the two
// forward transforms of every line are the first 2*tz contiguous transforms.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(FFT_THREADS) k_theta(Geo G, Bufs B, FftTiles T, int slot0) {
  const int slot = slot0 + blockIdx.z;
  const SlotCtl* c = &B.ctl[slot];
  if (!c->active) return;
  const int n = blockIdx.y;
  const int z0 = blockIdx.x * T.tz;
  const int tz = min(T.tz, G.Z - z0);
  const int M = G.M, TT = G.T;
  const double2* Yf = B.Yf + (((size_t)slot * G.N + n) * G.Z + z0) * 6 * M;
  double2* buf = smem;
  // pack (u_r,u_t), (u_z,w_r), (w_t,w_z) as a + ib into full Hermitian spectra;
  // c2r reads only the real part of the m = 0 bin (transform.fftw3.f90:97-98)
  for (int i = threadIdx.x; i < tz * 3 * TT; i += FFT_THREADS) {
    const int mm = i % TT;
    const int pl = i / TT;
    const int pr = pl / tz, line = pl - pr * tz;
    const double2* src = Yf + (size_t)line * 6 * M;
    double2 v = make_double2(0.0, 0.0);
    if (mm == 0) {
      v = make_double2(src[(2 * pr) * M].x, src[(2 * pr + 1) * M].x);
    } else if (mm < M) {
      const double2 A = src[(2 * pr) * M + mm], Bv = src[(2 * pr + 1) * M + mm];
      v = make_double2(A.x - Bv.y, A.y + Bv.x);
    } else if (mm > TT - M) {
      const int q = TT - mm;
      const double2 A = src[(2 * pr) * M + q], Bv = src[(2 * pr + 1) * M + q];
      v = make_double2(A.x + Bv.y, -A.y + Bv.x);
    }
    buf[i] = v;
  }
  __syncthreads();
  fft_lds(buf, 3 * tz, T.pt, 1.0, B.twT);
  // products (velocity.f90:325-329), scaled as tra_phys2spec (:114-117)
  const double scale = 1.0 / (double)(G.Z * TT);
  for (int i = threadIdx.x; i < tz * TT; i += FFT_THREADS) {
    const int line = i / TT, th = i - line * TT;
    double2* L0 = buf + (size_t)line * TT + th;
    const double2 w0 = L0[0], w1 = L0[(size_t)tz * TT], w2 = L0[(size_t)2 * tz * TT];
    const double ur = w0.x, ut = w0.y, uz = w1.x, cr = w1.y, ct = w2.x, cz = w2.y;
    const double Nr = ut * cz - uz * ct;
    const double Nt = uz * cr - ur * cz;
    const double Nz = ur * ct - ut * cr;
    L0[0] = make_double2(scale * Nr, scale * Nt);
    L0[(size_t)tz * TT] = make_double2(scale * Nz, 0.0);
  }
  __syncthreads();
  fft_lds(buf, 2 * tz, T.pt, -1.0, B.twT);
  // unpack the real transforms: A(m) = (X(m) + conj X(-m))/2, B(m) = (X(m) - conj X(-m))/2i
  double2* Yb = B.Yb + (((size_t)slot * G.N + n) * G.Z + z0) * 3 * M;
  for (int i = threadIdx.x; i < tz * 3 * M; i += FFT_THREADS) {
    const int line = i / (3 * M), rem = i - line * 3 * M;
    const int f = rem / M, mm = rem - f * M;
    double2 v;
    if (f == 2) {
      v = buf[((size_t)tz + line) * TT + mm];
    } else {
      const double2* L = buf + (size_t)line * TT;
      const double2 X = L[mm], Xc = L[(TT - mm) % TT];
      if (f == 0) v = make_double2(0.5 * (X.x + Xc.x), 0.5 * (X.y - Xc.y));
      else v = make_double2(0.5 * (X.y + Xc.y), -0.5 * (X.x - Xc.x));
    }
    Yb[(size_t)line * 3 * M + f * M + mm] = v;
  }
}

// ---------------------------------------------------------------------------
// inverse z transform and truncation into the raw nonlinear terms.
// grid: x = m tiles, y = N * 3, z = slot.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void store_zcol(double2* row, const double2* buf, int K, int Z, int m, int tid, int nth) {
  if (m == 0) {
    for (int k = tid; k < K; k += nth) row[k] = buf[k];
  } else {
    double2* blk = row + K + (m - 1) * (2 * K - 1) + (K - 1);
    for (int k = tid - (K - 1); k < K; k += nth) blk[k] = buf[k >= 0 ? k : k + Z];
  }
}

__global__ void __launch_bounds__(FFT_THREADS) k_zbwd(Geo G, Bufs B, FftTiles T, int slot0) {
  const int slot = slot0 + blockIdx.z;
  const SlotCtl* c = &B.ctl[slot];
  if (!c->active) return;
  const int n = blockIdx.y / 3, f = blockIdx.y % 3;
  const int m0 = blockIdx.x * T.tm;
  const int tm = min(T.tm, G.M - m0);
  const int Z = G.Z, K = G.K, M = G.M;
  const double2* Y = B.Yb + (((size_t)slot * G.N + n) * Z) * 3 * M;
  double2* buf = smem;
  for (int i = threadIdx.x; i < tm * Z; i += FFT_THREADS) {
    const int z = i / tm, b = i - z * tm;
    buf[b * Z + z] = Y[((size_t)z * 3 + f) * M + m0 + b];
  }
  __syncthreads();
  fft_lds(buf, tm, T.pz, -1.0, B.twZ);
  double2* row = B.Nraw[c->ncur] + (size_t)slot * 3 * G.coll + (size_t)f * G.coll + (size_t)n * G.Hp;
  for (int b = 0; b < tm; ++b) store_zcol(row, buf + b * Z, K, Z, m0 + b, threadIdx.x, FFT_THREADS);
}

// ---------------------------------------------------------------------------
// Granular transforms of one coll field (tra_coll2phys / tra_phys2coll of the
// drop-in interface): spec -> Y[n][z][m] -> phys[n][z][th] and back.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(FFT_THREADS) k_zfwd_one(Geo G, Bufs B, FftTiles T, const double2* a, double2* Y) {
  const int n = blockIdx.y;
  const int m0 = blockIdx.x * T.tm;
  const int tm = min(T.tm, G.M - m0);
  const int Z = G.Z, K = G.K, M = G.M;
  const double2* src = a + (size_t)n * G.Hp;
  double2* buf = smem;
  for (int i = threadIdx.x; i < tm * Z; i += FFT_THREADS) {
    const int b = i / Z, j = i - b * Z;
    load_zcol(buf + b * Z, src, K, Z, m0 + b, j);
  }
  __syncthreads();
  fft_lds(buf, tm, T.pz, 1.0, B.twZ);
  for (int i = threadIdx.x; i < tm * Z; i += FFT_THREADS) {
    const int z = i / tm, b = i - z * tm;
    Y[((size_t)n * Z + z) * M + m0 + b] = buf[b * Z + z];
  }
}

// c2r along theta of one field, lines paired as a + ib.  grid: x = z pairs, y = N.
__global__ void __launch_bounds__(FFT_THREADS) k_theta_c2r(Geo G, Bufs B, FftTiles T, const double2* Y, double* phys) {
  const int n = blockIdx.y;
  const int z0 = blockIdx.x * 2;
  const int M = G.M, TT = G.T;
  const bool two = z0 + 1 < G.Z;
  double2* buf = smem;
  const double2* A = Y + ((size_t)n * G.Z + z0) * M;
  const double2* Bp = A + M;
  for (int mm = threadIdx.x; mm < TT; mm += FFT_THREADS) {
    double2 v = make_double2(0.0, 0.0);
    if (mm == 0) v = make_double2(A[0].x, two ? Bp[0].x : 0.0);
    else if (mm < M) {
      const double2 a = A[mm], b = two ? Bp[mm] : make_double2(0.0, 0.0);
      v = make_double2(a.x - b.y, a.y + b.x);
    } else if (mm > TT - M) {
      const int q = TT - mm;
      const double2 a = A[q], b = two ? Bp[q] : make_double2(0.0, 0.0);
      v = make_double2(a.x + b.y, -a.y + b.x);
    }
    buf[mm] = v;
  }
  __syncthreads();
  fft_lds(buf, 1, T.pt, 1.0, B.twT);
  // phys layout of the reference: Re(0:3K-1, 0:3M-1, N), z fastest
  double* out = phys + (size_t)n * TT * G.Z + z0;
  for (int th = threadIdx.x; th < TT; th += FFT_THREADS) {
    out[(size_t)th * G.Z] = buf[th].x;
    if (two) out[(size_t)th * G.Z + 1] = buf[th].y;
  }
}

// r2c along theta (scaled by 1/(3K 3M) first, transform.fftw3.f90:114-117).
__global__ void __launch_bounds__(FFT_THREADS) k_theta_r2c(Geo G, Bufs B, FftTiles T, const double* phys, double2* Y) {
  const int n = blockIdx.y;
  const int z0 = blockIdx.x * 2;
  const int M = G.M, TT = G.T;
  const bool two = z0 + 1 < G.Z;
  const double scale = 1.0 / (double)(G.Z * TT);
  double2* buf = smem;
  const double* in = phys + (size_t)n * TT * G.Z + z0;
  for (int th = threadIdx.x; th < TT; th += FFT_THREADS)
    buf[th] = make_double2(scale * in[(size_t)th * G.Z], two ? scale * in[(size_t)th * G.Z + 1] : 0.0);
  __syncthreads();
  fft_lds(buf, 1, T.pt, -1.0, B.twT);
  double2* A = Y + ((size_t)n * G.Z + z0) * M;
  for (int mm = threadIdx.x; mm < M; mm += FFT_THREADS) {
    const double2 X = buf[mm], Xc = buf[(TT - mm) % TT];
    A[mm] = make_double2(0.5 * (X.x + Xc.x), 0.5 * (X.y - Xc.y));
    if (two) A[M + mm] = make_double2(0.5 * (X.y + Xc.y), -0.5 * (X.x - Xc.x));
  }
}

__global__ void __launch_bounds__(FFT_THREADS) k_zbwd_one(Geo G, Bufs B, FftTiles T, const double2* Y, double2* a) {
  const int n = blockIdx.y;
  const int m0 = blockIdx.x * T.tm;
  const int tm = min(T.tm, G.M - m0);
  const int Z = G.Z, K = G.K, M = G.M;
  double2* buf = smem;
  for (int i = threadIdx.x; i < tm * Z; i += FFT_THREADS) {
    const int z = i / tm, b = i - z * tm;
    buf[b * Z + z] = Y[((size_t)n * Z + z) * M + m0 + b];
  }
  __syncthreads();
  fft_lds(buf, tm, T.pz, -1.0, B.twZ);
  double2* row = a + (size_t)n * G.Hp;
  for (int b = 0; b < tm; ++b) store_zcol(row, buf + b * Z, K, Z, m0 + b, threadIdx.x, FFT_THREADS);
}

}  // namespace opf
