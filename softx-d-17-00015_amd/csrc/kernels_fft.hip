// Dealiased spectral <-> physical transforms with the u x curl(u) products.
//
// The reference transforms each radial plane with FFTW (transform.fftw3.f90:
// 65-147): a complex length-3K transform along z for every azimuthal mode
// (exponent +1 to physical space, -1 back), and a real length-3M transform
// along theta for every z line, with the 3/2 zero padding and truncation done
// by copies.  Here the same DFTs run as LDS-resident Stockham FFTs in three
// kernels with intermediates laid out for coalesced access:
//
//   k_zfwd : 6 coll fields (u, curl u) -> Yf[slot][n][z][6][M]   (z transform)
//   k_theta: Yf -> theta transforms, products, scaling, inverse theta
//            transforms -> Yb[slot][n][z][3][M]
//   k_zbwd : Yb -> z transform, truncation -> Nraw coll fields
//
// For grids whose radial plane fits in LDS, k_plane runs the same three
// stages for one plane in one workgroup, with Yf and Yb kept in LDS.
//
// The padding, the m = 0 Hermitian fill and the coll <-> spec layout change
// (variables.f90:184-294) are folded into the loads and stores.  Two real
// theta transforms share one complex FFT (a + ib), so the three c2r
// transforms per pair of fields and the r2c transforms run as 3 + 2 complex
// FFTs per line.
//
// Every pass is one radix-R butterfly per thread (R up to 16, built from
// radix 2/3/4 DFTs in registers), so the lengths 3K, 3M of the usual grids
// take two or three passes.  The twiddle table of the length is copied into
// LDS with the data; global loads are issued in batches so that a kernel's
// input costs one memory latency.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace opf {

namespace {

// 1 - r^2 with the product rounded before the difference, as the
// reference's host arithmetic does (no fused multiply-add)
__device__ __forceinline__ double one_minus_sq(double r) {
#pragma clang fp contract(off)
  return 1.0 - r * r;
}

__device__ __forceinline__ double2 cmul_tw(double2 a, double2 w, double sgn) {
  // a * (w.x + i sgn w.y)
  const double wy = sgn * w.y;
  return make_double2(a.x * w.x - a.y * wy, a.x * wy + a.y * w.x);
}

template <int R> __device__ __forceinline__ void dft(double2* v, double sgn);

template <> __device__ __forceinline__ void dft<1>(double2*, double) {}

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

// Radix R = R1 * R2 (R1, R2 <= 4) as R2 DFTs of length R1, the internal
// twiddles W_R^(n2 k1) and R1 DFTs of length R2 (Cooley-Tukey): input n =
// R2 n1 + n2, output k = k1 + R1 k2.  W_R^x = tw[x L / R], tw[q] =
// exp(2 pi i q / L).  Computed in place: output k is left in
// v[perm<R>(k)], perm(k1 + R1 k2) = R2 k1 + k2.
template <int R> struct Split;
template <> struct Split<2> { static constexpr int R1 = 2, R2 = 1; };
template <> struct Split<3> { static constexpr int R1 = 3, R2 = 1; };
template <> struct Split<4> { static constexpr int R1 = 4, R2 = 1; };
template <> struct Split<6> { static constexpr int R1 = 3, R2 = 2; };
template <> struct Split<8> { static constexpr int R1 = 4, R2 = 2; };
template <> struct Split<9> { static constexpr int R1 = 3, R2 = 3; };
template <> struct Split<12> { static constexpr int R1 = 4, R2 = 3; };
template <> struct Split<16> { static constexpr int R1 = 4, R2 = 4; };

template <int R>
__host__ __device__ constexpr int perm(int k) {
  return Split<R>::R2 * (k % Split<R>::R1) + k / Split<R>::R1;
}

template <int R>
__device__ __forceinline__ void bfly(double2 (&v)[R], double sgn, const double2* twg, int L) {
  constexpr int R1 = Split<R>::R1, R2 = Split<R>::R2;
  if constexpr (R2 == 1) {
    dft<R1>(v, sgn);
  } else {
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      double2 a[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) a[n1] = v[R2 * n1 + n2];
      dft<R1>(a, sgn);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) v[R2 * k1 + n2] = a[k1];
    }
    const int step = L / R;
#pragma unroll
    for (int n2 = 1; n2 < R2; ++n2)
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1)
        v[R2 * k1 + n2] = cmul_tw(v[R2 * k1 + n2], twg[((n2 * k1) % R) * step], sgn);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      double2 a[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) a[n2] = v[R2 * k1 + n2];
      dft<R2>(a, sgn);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) v[R2 * k1 + k2] = a[k2];
    }
  }
}

// LDS layout of a batch of transforms of length L: a transform's input,
// element i of transform b, at buf[b*Lp + i] (linear), its output at
// buf[b*Lp + pidx(i)], pidx(i) = i + i/16, Lp = padded_len(L).  The pad
// spreads the stride-R writes of the first Stockham pass over the LDS
// banks; the first pass reads the linear input, whose stride-L/R reads
// fall on consecutive addresses across a wave, with constant offsets.
__host__ __device__ __forceinline__ int pidx(int i) { return i + (i >> 4); }

// One Stockham pass of radix R over B transforms of length L stored at
// buf[b*Lp + pidx(i)] (B*L <= CAP; LIN: the first pass, reading the linear
// input layout buf[b*Lp + i]).  Ns = product of the radices of the
// previous passes (the twiddles of the first pass are all 1 and skipped);
// tw: the length-L twiddle table in LDS, twg: the same table in global
// memory, read with uniform indices for the butterflies' internal twiddles.
// Outer twiddles are fetched four at a time to bound register use.
// LC, NSC: the length and Ns at compile time (0: the run-time L_, Ns_), so
// that the index divisions become multiplications by constants.
template <int R, int LC, int NSC, int CAP = FFT_TILE, bool LIN = false, int NT = FFT_THREADS>
__device__ __forceinline__ void fft_pass(double2* buf, int L_, int B, int Ns_, double sgn, const double2* tw,
                                         const double2* twg) {
  constexpr int MB = (CAP + R * NT - 1) / (R * NT);
  const int L = LC ? LC : L_;
  const int Ns = LC ? NSC : Ns_;
  const int Lp = padded_len(L);
  const int nb = L / R;
  const int total = nb * B;
  const int tid = threadIdx.x;
  const int stride_tw = L / (Ns * R);
  // Where a stride is a compile-time multiple of the 16-element pad period,
  // pidx(a + r*s) = pidx(a) + r*(s + s/16): one address per thread and
  // constant offsets per element instead of a padded index per element.
  // The same holds for the writes of a first pass whose radix divides 16
  // (each thread's R outputs lie inside one period).
  constexpr int NBC = LC ? LC / R : 0;
  constexpr bool RSEP = LC && NBC % 16 == 0;
  constexpr bool WSEP = LC && (NSC % 16 == 0 || (NSC == 1 && 16 % R == 0));
  double2 v[MB][R];
  int dst[MB];
#pragma unroll
  for (int i = 0; i < MB; ++i) {
    const int q = tid + i * NT;
    dst[i] = -1;
    if (q < total) {
      const int b = q / nb, j = q - b * nb;
      const double2* x = buf + b * Lp;
      const int kk = j % Ns;
      if constexpr (LIN) {
        const double2* xr = x + j;
#pragma unroll
        for (int r = 0; r < R; ++r) v[i][r] = xr[r * nb];
      } else if constexpr (RSEP) {
        const double2* xr = x + pidx(j);
#pragma unroll
        for (int r = 0; r < R; ++r) v[i][r] = xr[r * (NBC + NBC / 16)];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[i][r] = x[pidx(j + r * nb)];
      }
      if (Ns > 1) {
#pragma unroll
        for (int g = 1; g < R; g += 4) {
          __builtin_amdgcn_sched_barrier(0);
          double2 w[4];
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (g + u < R) w[u] = tw[kk * (g + u) * stride_tw];
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (g + u < R) v[i][g + u] = cmul_tw(v[i][g + u], w[u], sgn);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      bfly<R>(v[i], sgn, twg, L);
      dst[i] = b * Lp + (j / Ns) * Ns * R + kk;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < MB; ++i) {
    if (dst[i] >= 0) {
      const int q = tid + i * NT;
      const int b = q / nb;
      const int e = dst[i] - b * Lp;          // unpadded index of output 0
      double2* x = buf + b * Lp;
      if constexpr (WSEP) {
        double2* xw = x + pidx(e);
#pragma unroll
        for (int r = 0; r < R; ++r) xw[r * (NSC + NSC / 16)] = v[i][perm<R>(r)];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) x[pidx(e + r * Ns)] = v[i][perm<R>(r)];
      }
    }
  }
  __syncthreads();
}

// Full transform: sgn = +1 (to physical space) or -1, unnormalised (FFTW).
template <int CAP = FFT_TILE>
__device__ __forceinline__ void fft_lds(double2* buf, int B, const FftPlan& pl, double sgn, const double2* tw,
                                        const double2* twg) {
  auto pass = [&](auto lin, int R, int Ns) {
    constexpr bool LN = decltype(lin)::value;
    switch (R) {
      case 16: fft_pass<16, 0, 0, CAP, LN>(buf, pl.L, B, Ns, sgn, tw, twg); break;
      case 12: fft_pass<12, 0, 0, CAP, LN>(buf, pl.L, B, Ns, sgn, tw, twg); break;
      case 9: fft_pass<9, 0, 0, CAP, LN>(buf, pl.L, B, Ns, sgn, tw, twg); break;
      case 8: fft_pass<8, 0, 0, CAP, LN>(buf, pl.L, B, Ns, sgn, tw, twg); break;
      case 6: fft_pass<6, 0, 0, CAP, LN>(buf, pl.L, B, Ns, sgn, tw, twg); break;
      case 4: fft_pass<4, 0, 0, CAP, LN>(buf, pl.L, B, Ns, sgn, tw, twg); break;
      case 3: fft_pass<3, 0, 0, CAP, LN>(buf, pl.L, B, Ns, sgn, tw, twg); break;
      default: fft_pass<2, 0, 0, CAP, LN>(buf, pl.L, B, Ns, sgn, tw, twg); break;
    }
  };
  pass(std::true_type{}, pl.rad[0], 1);
  int Ns = pl.rad[0];
  for (int p = 1; p < pl.nrad; ++p) {
    pass(std::false_type{}, pl.rad[p], Ns);
    Ns *= pl.rad[p];
  }
}

// length of the radix sequence R0 R1 R2 (0: not fixed at compile time)
template <int R0, int R1, int R2>
constexpr int plan_len() {
  return R0 * (R1 > 0 ? R1 : 1) * (R2 > 0 ? R2 : 1);
}

// Transform with the radix sequence fixed at compile time (R0 = 0: the
// plan's own sequence).  Each pass then keeps only its own registers live,
// and lengths and strides are constants.  CAP: the most values (B * L) a
// call transforms; NT: the workgroup's threads (the generic path takes
// FFT_THREADS).
template <int R0, int R1, int R2, int CAP = FFT_TILE, int NT = FFT_THREADS>
__device__ __forceinline__ void run_fft(double2* buf, int B, const FftPlan& pl, double sgn, const double2* tw,
                                        const double2* twg) {
  if constexpr (R0 == 0) {
    static_assert(NT == FFT_THREADS, "the generic transform runs on FFT_THREADS threads");
    fft_lds<CAP>(buf, B, pl, sgn, tw, twg);
  } else {
    constexpr int L = plan_len<R0, R1, R2>();
    fft_pass<R0, L, 1, CAP, true, NT>(buf, L, B, 1, sgn, tw, twg);
    if constexpr (R1 > 1) fft_pass<R1, L, R0, CAP, false, NT>(buf, L, B, R0, sgn, tw, twg);
    if constexpr (R2 > 1) fft_pass<R2, L, R0 * R1, CAP, false, NT>(buf, L, B, R0 * R1, sgn, tw, twg);
  }
}

// twiddle table of length L into LDS
__device__ __forceinline__ void load_tw(double2* twl, const double2* tw, int L) {
  for (int i = threadIdx.x; i < L; i += FFT_THREADS) twl[i] = tw[i];
}

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

// Padded z columns of modes m0 .. m0+tm-1 into the tile: X(j) = a(k=j)
// for j < K, X(Z+k) = a(k) for k < 0, zero in between
// (transform.fftw3.f90:72-84).  A tile holds at most FFT_TILE values, so
// every thread issues its (at most ZLD) loads in one batch.
constexpr int ZLD = FFT_TILE / FFT_THREADS;
__device__ __forceinline__ void load_zcols(double2* buf, const double2* row, int K, int Z, int Zp, int m0, int tm) {
  const int nk = 2 * K - 1;
  const int total = tm * nk;
  double2 v[ZLD];
  int d[ZLD];
#pragma unroll
  for (int u = 0; u < ZLD; ++u) {
    const int i = threadIdx.x + u * FFT_THREADS;
    const bool ok = i < total;
    const int ic = ok ? i : 0;
    const int b = ic / nk, e = ic - b * nk;
    const int k = e - (K - 1);
    v[u] = coef(row, K, k, m0 + b);
    d[u] = ok ? b * Zp + (k >= 0 ? k : k + Z) : -1;
  }
#pragma unroll
  for (int u = 0; u < ZLD; ++u)
    if (d[u] >= 0) buf[d[u]] = v[u];
  const int nzero = K + 1;   // j = K .. 2K
  for (int i = threadIdx.x; i < tm * nzero; i += FFT_THREADS) {
    const int b = i / nzero, j = K + (i - b * nzero);
    buf[b * Zp + j] = make_double2(0.0, 0.0);
  }
}

// column b of the tile (linear input layout) = Y[z*ystride + m0 + b], one batch.
// TMC: the full tile width at compile time (0: use tm), items of columns
// b >= tm are skipped.
template <int TMC>
__device__ __forceinline__ void load_ycols(double2* buf, const double2* Y, size_t ystride, int Z, int Zp, int m0, int tm) {
  const int tw = TMC ? TMC : tm;
  const int total = tw * Z;
  double2 v[ZLD];
  int d[ZLD];
#pragma unroll
  for (int u = 0; u < ZLD; ++u) {
    const int i = threadIdx.x + u * FFT_THREADS;
    const int ic = i < total ? i : 0;
    const int z = ic / tw, b = ic - z * tw;
    const bool ok = i < total && b < tm;
    v[u] = Y[(size_t)z * ystride + m0 + (ok ? b : 0)];
    d[u] = ok ? b * Zp + z : -1;
  }
#pragma unroll
  for (int u = 0; u < ZLD; ++u)
    if (d[u] >= 0) buf[d[u]] = v[u];
}

// truncation of a z column to the 2K-1 retained k into a coll row
__device__ __forceinline__ void store_zcol(double2* row, const double2* buf, int K, int Z, int m, int tid, int nth) {
  if (m == 0) {
    for (int k = tid; k < K; k += nth) row[k] = buf[pidx(k)];
  } else {
    double2* blk = row + K + (m - 1) * (2 * K - 1) + (K - 1);
    for (int k = tid - (K - 1); k < K; k += nth) blk[k] = buf[pidx(k >= 0 ? k : k + Z)];
  }
}

}  // namespace

extern __shared__ double2 smem[];

// ---------------------------------------------------------------------------
// z transform of u and curl u for all active slots.
// grid: x = 8 * per >= (m tile, plane, field) tiles, z = slot.
// LDS: tm columns at stride Zp, then Z twiddles.
// ---------------------------------------------------------------------------
template <int R0, int R1, int R2>
__global__ void __launch_bounds__(FFT_THREADS) k_zfwd_t(Geo G, Bufs B, FftTiles T, int slot0) {
  const int slot = slot_of(B, slot0);
  const SlotCtl* c = &B.ctl[slot];
  if (!c->active) return;
  constexpr int LZ = plan_len<R0, R1, R2>();
  constexpr int TMC = LZ ? tile_m(LZ) : 0;
  // a new evaluation: clear the slot's per-plane velocity maxima (the theta
  // kernel that follows records them)
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < 3 * G.N; i += FFT_THREADS) B.cflmax[(size_t)slot * 3 * G.N + i] = 0ull;
  // XCD-aware tiles as in k_zbwd_t: runs of neighbouring m tiles, which
  // share cache lines of Y and of the coll rows, stay on one XCD's L2
  const int gx = (G.M + T.tm - 1) / T.tm, ntile = gx * 6 * G.N;
  const int per = (ntile + 7) / 8;
  const int tile = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (tile >= ntile) return;
  const int ty = tile / gx, tx = tile - ty * gx;
  const int n = ty / 6, f = ty % 6;
  const int m0 = tx * T.tm;
  const int tm = min(T.tm, G.M - m0);
  const int Z = LZ ? LZ : G.Z, K = LZ ? LZ / 3 : G.K, M = G.M;
  const double2* src = (f < 3 ? B.U[c->ucur] + (size_t)slot * 3 * G.coll + (size_t)f * G.coll
                              : B.om + (size_t)slot * 3 * G.coll + (size_t)(f - 3) * G.coll) +
                       (size_t)n * G.Hp;
  const int Zp = padded_len(Z);
  double2* buf = smem;
  double2* twl = smem + (size_t)T.tm * Zp;
  load_tw(twl, B.twZ, Z);
  load_zcols(buf, src, K, Z, Zp, m0, tm);
  __syncthreads();
  run_fft<R0, R1, R2>(buf, tm, T.pz, 1.0, twl, B.twZ);
  double2* Y = B.Yf + (((size_t)slot * G.N + n) * Z) * 6 * M;
  const int tw = TMC ? TMC : tm;
  for (int i = threadIdx.x; i < tw * Z; i += FFT_THREADS) {
    const int z = i / tw, b = i - z * tw;
    if (b < tm) Y[((size_t)z * 6 + f) * M + m0 + b] = buf[b * Zp + pidx(z)];
  }
}

// ---------------------------------------------------------------------------
// theta transforms and products.  grid: x = z tiles, y = N, z = slot.
// LDS: transform (pair*tz + line) of the three packed pairs at stride Tp, so
// the two forward transforms of every line are the first 2*tz transforms;
// then T twiddles.
// ---------------------------------------------------------------------------
constexpr int TLD = THETA_ITEMS / FFT_THREADS;

// products (velocity.f90:325-329) at one physical point of radial plane n,
// scaled as tra_phys2spec (:114-117), and the running maxima of |u_r|,
// |u_t|, |u_z + 1 - r^2| for vel_maxtstep (velocity.f90:464-475).  Inputs:
// the packed transforms' values (u_r, u_t), (u_z, w_r), (w_t, w_z); outputs
// (N_r, N_t) and N_z.  One definition for every kernel that forms them, so
// they round alike.
struct NlPlane {
  double scale, ubase, cor_d1, cor_s, cor_c;
  bool cor;
  __device__ static NlPlane make(const Geo& G, const Bufs& B, int slot, int n, int TT) {
    NlPlane p;
    p.scale = 1.0 / (double)(G.Z * TT);
    p.ubase = one_minus_sq(G.mesh[MeshOff::make(G.N).r + n]);
    p.cor = B.prm[slot].coriolis != 0;
    p.cor_d1 = B.prm[slot].cor_d1;
    p.cor_s = B.prm[slot].cor_sphi;
    p.cor_c = B.prm[slot].cor_cphi;
    return p;
  }
  __device__ __forceinline__ void point(double2 w0, double2 w1, double2 w2, const double2* twl, int th, double& mr,
                                        double& mt, double& mz, double2& nrt, double& nz) const {
    const double ur = w0.x, ut = w0.y, uz = w1.x, cr = w1.y, ct = w2.x, cz = w2.y;
    mr = fmax(mr, fabs(ur));
    mt = fmax(mt, fabs(ut));
    mz = fmax(mz, fabs(uz + ubase));
    double Nr = ut * cz - uz * ct;
    double Nt = uz * cr - ur * cz;
    double Nz = ur * ct - ut * cr;
    if (cor) {   // Coriolis.f90:62-68: N += -1/(4 Ro) Omega^ x (U + u)
      const double2 cs = twl[th];   // (cos, sin)(2 pi th / T)
      const double Or = cor_s * cs.x, Ot = -cor_s * cs.y;
      const double uzU = uz + ubase;
      Nr += cor_d1 * (Ot * uzU - cor_c * ut);
      Nt += cor_d1 * (cor_c * ur - Or * uzU);
      Nz += cor_d1 * (Or * ut - Ot * ur);
    }
    nrt = make_double2(scale * Nr, scale * Nt);
    nz = scale * Nz;
  }
};
template <int R0, int R1, int R2>
__global__ void __launch_bounds__(FFT_THREADS) k_theta_t(Geo G, Bufs B, FftTiles T, int slot0) {
  const int slot = slot_of(B, slot0);
  const SlotCtl* c = &B.ctl[slot];
  if (!c->active) return;
  constexpr int LT = plan_len<R0, R1, R2>();
  const int n = blockIdx.y;
  const int z0 = blockIdx.x * T.tz;
  const int tz = min(T.tz, G.Z - z0);
  const int TT = LT ? LT : G.T, M = LT ? LT / 3 : G.M;   // T = 3M
  const double2* Yf = B.Yf + (((size_t)slot * G.N + n) * G.Z + z0) * 6 * M;
  const int Tp = padded_len(TT);
  double2* buf = smem;
  double2* twl = smem + (size_t)3 * T.tz * Tp;
  load_tw(twl, B.twT, TT);
  // pack (u_r,u_t), (u_z,w_r), (w_t,w_z) as a + ib into full Hermitian
  // spectra: item (line, pair, m) fills bins m and T-m; c2r reads only the
  // real part of the m = 0 bin (transform.fftw3.f90:97-98)
  const int total = tz * 3 * M;   // = tz * T <= THETA_ITEMS = TLD * FFT_THREADS
  double2 A[TLD], Bv[TLD];
  int it[TLD];
#pragma unroll
  for (int u = 0; u < TLD; ++u) {
    const int i = threadIdx.x + u * FFT_THREADS;
    const bool ok = i < total;
    const int ic = ok ? i : 0;
    const int line = ic / (3 * M), rem = ic - line * 3 * M;
    const int pr = rem / M, mm = rem - pr * M;
    const double2* src = Yf + (size_t)line * 6 * M + (size_t)(2 * pr) * M + mm;
    A[u] = src[0];
    Bv[u] = src[M];
    it[u] = ok ? i : -1;
  }
#pragma unroll
  for (int u = 0; u < TLD; ++u) {
    if (it[u] < 0) continue;
    const int i = it[u];
    const int line = i / (3 * M), rem = i - line * 3 * M;
    const int pr = rem / M, mm = rem - pr * M;
    double2* L = buf + ((size_t)pr * tz + line) * Tp;
    if (mm == 0) {
      L[0] = make_double2(A[u].x, Bv[u].x);
    } else {
      L[mm] = make_double2(A[u].x - Bv[u].y, A[u].y + Bv[u].x);
      L[TT - mm] = make_double2(A[u].x + Bv[u].y, -A[u].y + Bv[u].x);
    }
  }
  const int nzero = TT - 2 * M + 1;   // bins M .. T-M
  for (int i = threadIdx.x; i < 3 * tz * nzero; i += FFT_THREADS) {
    const int pl = i / nzero, mm = M + (i - pl * nzero);
    const int pr = pl / tz, line = pl - pr * tz;
    buf[((size_t)pr * tz + line) * Tp + mm] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  run_fft<R0, R1, R2, 3 * THETA_ITEMS>(buf, 3 * tz, T.pt, 1.0, twl, B.twT);
  // products and the plane's velocity maxima
  const NlPlane pl = NlPlane::make(G, B, slot, n, TT);
  double mr = 0.0, mt = 0.0, mz = 0.0;
  // N of this thread's items, written after a barrier as the forward
  // transforms' linear input (it overlaps the inverse transforms' output)
  double2 nrt[TLD];
  double nzv[TLD];
#pragma unroll
  for (int u = 0; u < TLD; ++u) {
    const int i = threadIdx.x + u * FFT_THREADS;
    if (i >= tz * TT) continue;
    const int line = i / TT, th = i - line * TT;
    const double2* L0 = buf + (size_t)line * Tp + pidx(th);
    pl.point(L0[0], L0[(size_t)tz * Tp], L0[(size_t)2 * tz * Tp], twl, th, mr, mt, mz, nrt[u], nzv[u]);
  }
  // maxima are exact and order free: non-negative doubles order like their
  // bit patterns, so one integer atomic max per workgroup and component.
  // Recorded only where they are read: an adaptive slot's predictor
  // evaluation (the device loop) and host-driven evaluations.
  const bool want_cfl = (B.prm[slot].adaptive && c->mode == MODE_PRED) || c->host_driven;
  if (want_cfl) {
    for (int off = 32; off >= 1; off >>= 1) {
      mr = fmax(mr, __shfl_xor(mr, off));
      mt = fmax(mt, __shfl_xor(mt, off));
      mz = fmax(mz, __shfl_xor(mz, off));
    }
    __shared__ double wmax[3][FFT_THREADS / 64];
    if ((threadIdx.x & 63) == 0) {
      wmax[0][threadIdx.x >> 6] = mr;
      wmax[1][threadIdx.x >> 6] = mt;
      wmax[2][threadIdx.x >> 6] = mz;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      double v = 0.0;
      for (int w = 0; w < FFT_THREADS / 64; ++w) v = fmax(v, wmax[threadIdx.x][w]);
      atomicMax(B.cflmax + ((size_t)slot * 3 + threadIdx.x) * G.N + n, (unsigned long long)__double_as_longlong(v));
    }
  }
  __syncthreads();
  // (N_r, N_t) of each line as transform `line`; N_z of lines 2q and 2q + 1
  // as the real and imaginary parts of transform tz + q (two real
  // transforms in one complex one; an odd last line pairs with zero)
  const int nzt = (tz + 1) / 2;
#pragma unroll
  for (int u = 0; u < TLD; ++u) {
    const int i = threadIdx.x + u * FFT_THREADS;
    if (i >= tz * TT) continue;
    const int line = i / TT, th = i - line * TT;
    buf[(size_t)line * Tp + th] = nrt[u];
    double* d = (double*)(buf + ((size_t)tz + (line >> 1)) * Tp + th);
    d[line & 1] = nzv[u];
    if ((line & 1) == 0 && line + 1 == tz) d[1] = 0.0;
  }
  __syncthreads();
  run_fft<R0, R1, R2, 3 * THETA_ITEMS>(buf, tz + nzt, T.pt, -1.0, twl, B.twT);
  // unpack the real transforms: A(m) = (X(m) + conj X(-m))/2, B(m) = (X(m) - conj X(-m))/2i
  double2* Yb = B.Yb + (((size_t)slot * G.N + n) * G.Z + z0) * 3 * M;
  for (int i = threadIdx.x; i < tz * 3 * M; i += FFT_THREADS) {
    const int line = i / (3 * M), rem = i - line * 3 * M;
    const int f = rem / M, mm = rem - f * M;
    double2 v;
    // f = 0, 1: N_r, N_t of the line's own transform; f = 2: N_z, the real
    // (even line) or imaginary (odd line) part of transform tz + line/2
    const double2* L = buf + (f == 2 ? (size_t)tz + (line >> 1) : (size_t)line) * Tp;
    const double2 X = L[pidx(mm)], Xc = L[pidx((TT - mm) % TT)];
    const bool re = f == 0 || (f == 2 && (line & 1) == 0);
    if (re) v = make_double2(0.5 * (X.x + Xc.x), 0.5 * (X.y - Xc.y));
    else v = make_double2(0.5 * (X.y + Xc.y), -0.5 * (X.x - Xc.x));
    Yb[(size_t)line * 3 * M + f * M + mm] = v;
  }
}

// ---------------------------------------------------------------------------
// inverse z transform and truncation into the raw nonlinear terms.
// grid: x = m tiles, y = N * 3, z = slot.  LDS: tm columns at stride Zp, then Z twiddles.
// ---------------------------------------------------------------------------
template <int R0, int R1, int R2>
__global__ void __launch_bounds__(FFT_THREADS) k_zbwd_t(Geo G, Bufs B, FftTiles T, int slot0) {
  const int slot = slot_of(B, slot0);
  const SlotCtl* c = &B.ctl[slot];
  if (!c->active) return;
  constexpr int LZ = plan_len<R0, R1, R2>();
  // XCD-aware tiles: launch order b goes to XCD b % 8, so tile t = (b % 8)
  // per + b / 8 keeps runs of neighbouring tiles (which share cache lines of
  // Yb) on one XCD's L2.  grid x = 8 per >= tiles; spare blocks return.
  const int gx = (G.M + T.tm - 1) / T.tm, ntile = gx * 3 * G.N;
  const int per = (ntile + 7) / 8;
  const int tile = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (tile >= ntile) return;
  const int ty = tile / gx, tx = tile - ty * gx;
  const int n = ty / 3, f = ty % 3;
  const int m0 = tx * T.tm;
  const int tm = min(T.tm, G.M - m0);
  const int Z = LZ ? LZ : G.Z, K = LZ ? LZ / 3 : G.K, M = G.M;
  const double2* Y = B.Yb + (((size_t)slot * G.N + n) * Z) * 3 * M + (size_t)f * M;
  const int Zp = padded_len(Z);
  double2* buf = smem;
  double2* twl = smem + (size_t)T.tm * Zp;
  load_tw(twl, B.twZ, Z);
  load_ycols<LZ ? tile_m(LZ) : 0>(buf, Y, (size_t)3 * M, Z, Zp, m0, tm);
  __syncthreads();
  run_fft<R0, R1, R2>(buf, tm, T.pz, -1.0, twl, B.twZ);
  double2* row = B.Nraw[c->ncur] + (size_t)slot * 3 * G.coll + (size_t)f * G.coll + (size_t)n * G.Hp;
  for (int b = 0; b < tm; ++b) store_zcol(row, buf + b * Zp, K, Z, m0 + b, threadIdx.x, FFT_THREADS);
}

// ---------------------------------------------------------------------------
// A radial plane's whole evaluation in one workgroup, for grids whose six
// z-transformed fields fit in LDS: the z transforms of u and curl u, the
// theta transforms, the products and the inverse theta and z transforms, as
// k_zfwd_t, k_theta_t and k_zbwd_t do them in turn (the same passes and
// operations, so the same bits), with the intermediates Y kept in LDS
// instead of making a round trip through HBM.  grid: x = N, z = slot.
// LDS: the six fields' z columns (field f, mode m at (f M + m) Zp), then the
// theta transforms of a chunk of PL z-lines (as k_theta_t's tile with tz =
// PL), then the Z and T twiddles.  The nonlinear terms' z columns, in the
// linear input layout, overwrite fields 0..2 chunk by chunk: a chunk's
// entries at linear z lie at or below the padded entries of the chunks
// still to come (pidx(z) >= z), and its own were read before.
// ---------------------------------------------------------------------------
template <int ZR0, int ZR1, int ZR2, int TR0, int TR1, int TR2, int PL>
__global__ void __launch_bounds__(PLANE_THREADS) k_plane_t(Geo G, Bufs B, FftTiles T, int slot0) {
  constexpr int NT = PLANE_THREADS;
  constexpr int Z = plan_len<ZR0, ZR1, ZR2>(), TT = plan_len<TR0, TR1, TR2>();
  constexpr int K = Z / 3, M = TT / 3, NK = 2 * K - 1;
  constexpr int Zp = padded_len(Z), Tp = padded_len(TT);
  static_assert(Z % PL == 0 && PL % 2 == 0, "whole, even chunks of z-lines");
  const int slot = slot_of(B, slot0);
  const SlotCtl* c = &B.ctl[slot];
  if (!c->active) return;
  const int n = blockIdx.x;
  const int tid = threadIdx.x;
  double2* A = smem;
  double2* Lb = A + (size_t)6 * M * Zp;
  double2* twz = Lb + (size_t)3 * PL * Tp;
  double2* twt = twz + Z;
  for (int i = tid; i < Z; i += NT) twz[i] = B.twZ[i];
  for (int i = tid; i < TT; i += NT) twt[i] = B.twT[i];
  // z columns of the six fields (load_zcols for every mode at once)
  const double2* U = B.U[c->ucur] + (size_t)slot * 3 * G.coll + (size_t)n * G.Hp;
  const double2* W = B.om + (size_t)slot * 3 * G.coll + (size_t)n * G.Hp;
  {
    constexpr int TOT = 6 * M * NK, IT = (TOT + NT - 1) / NT;
    double2 v[IT];
    int d[IT];
#pragma unroll
    for (int u = 0; u < IT; ++u) {
      const int i = tid + u * NT;
      const bool ok = i < TOT;
      const int ic = ok ? i : 0;
      const int f = ic / (M * NK), rem = ic - f * (M * NK);
      const int b = rem / NK, k = rem - b * NK - (K - 1);
      const double2* row = f < 3 ? U + (size_t)f * G.coll : W + (size_t)(f - 3) * G.coll;
      v[u] = coef(row, K, k, b);
      d[u] = ok ? (f * M + b) * Zp + (k >= 0 ? k : k + Z) : -1;
    }
#pragma unroll
    for (int u = 0; u < IT; ++u)
      if (d[u] >= 0) A[d[u]] = v[u];
    constexpr int NZ = K + 1;   // j = K .. 2K
    for (int i = tid; i < 6 * M * NZ; i += NT) {
      const int col = i / NZ;
      A[col * Zp + K + (i - col * NZ)] = make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  run_fft<ZR0, ZR1, ZR2, 6 * M * Z, NT>(A, 6 * M, T.pz, 1.0, twz, B.twZ);
  const NlPlane pl = NlPlane::make(G, B, slot, n, TT);
  double mr = 0.0, mt = 0.0, mz = 0.0;
  for (int z0 = 0; z0 < Z; z0 += PL) {
    // pack the chunk's lines as k_theta_t does from Y
    {
      constexpr int TOT = PL * 3 * M, IT = (TOT + NT - 1) / NT;
      double2 a[IT], bv[IT];
      int it[IT];
#pragma unroll
      for (int u = 0; u < IT; ++u) {
        const int i = tid + u * NT;
        const bool ok = i < TOT;
        const int ic = ok ? i : 0;
        const int line = ic / (3 * M), rem = ic - line * 3 * M;
        const int pr = rem / M, mm = rem - pr * M;
        const double2* src = A + (size_t)(2 * pr * M + mm) * Zp + pidx(z0 + line);
        a[u] = src[0];
        bv[u] = src[(size_t)M * Zp];
        it[u] = ok ? i : -1;
      }
#pragma unroll
      for (int u = 0; u < IT; ++u) {
        if (it[u] < 0) continue;
        const int i = it[u];
        const int line = i / (3 * M), rem = i - line * 3 * M;
        const int pr = rem / M, mm = rem - pr * M;
        double2* L = Lb + (size_t)(pr * PL + line) * Tp;
        if (mm == 0) {
          L[0] = make_double2(a[u].x, bv[u].x);
        } else {
          L[mm] = make_double2(a[u].x - bv[u].y, a[u].y + bv[u].x);
          L[TT - mm] = make_double2(a[u].x + bv[u].y, -a[u].y + bv[u].x);
        }
      }
      constexpr int NZ = TT - 2 * M + 1;   // bins M .. T-M
      for (int i = tid; i < 3 * PL * NZ; i += NT) {
        const int t = i / NZ;
        Lb[(size_t)t * Tp + M + (i - t * NZ)] = make_double2(0.0, 0.0);
      }
    }
    __syncthreads();
    run_fft<TR0, TR1, TR2, 3 * PL * TT, NT>(Lb, 3 * PL, T.pt, 1.0, twt, B.twT);
    {
      constexpr int IT = (PL * TT + NT - 1) / NT;
      double2 nrt[IT];
      double nzv[IT];
#pragma unroll
      for (int u = 0; u < IT; ++u) {
        const int i = tid + u * NT;
        if (i >= PL * TT) continue;
        const int line = i / TT, th = i - line * TT;
        const double2* L0 = Lb + (size_t)line * Tp + pidx(th);
        pl.point(L0[0], L0[(size_t)PL * Tp], L0[(size_t)2 * PL * Tp], twt, th, mr, mt, mz, nrt[u], nzv[u]);
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < IT; ++u) {
        const int i = tid + u * NT;
        if (i >= PL * TT) continue;
        const int line = i / TT, th = i - line * TT;
        Lb[(size_t)line * Tp + th] = nrt[u];
        double* d = (double*)(Lb + ((size_t)PL + (line >> 1)) * Tp + th);
        d[line & 1] = nzv[u];
      }
    }
    __syncthreads();
    run_fft<TR0, TR1, TR2, 3 * PL * TT, NT>(Lb, PL + PL / 2, T.pt, -1.0, twt, B.twT);
    // unpack as k_theta_t does, into the z columns of the nonlinear terms
    for (int i = tid; i < PL * 3 * M; i += NT) {
      const int line = i / (3 * M), rem = i - line * 3 * M;
      const int f = rem / M, mm = rem - f * M;
      const double2* L = Lb + (f == 2 ? (size_t)PL + (line >> 1) : (size_t)line) * Tp;
      const double2 X = L[pidx(mm)], Xc = L[pidx((TT - mm) % TT)];
      const bool re = f == 0 || (f == 2 && (line & 1) == 0);
      double2 v;
      if (re) v = make_double2(0.5 * (X.x + Xc.x), 0.5 * (X.y - Xc.y));
      else v = make_double2(0.5 * (X.y + Xc.y), -0.5 * (X.x - Xc.x));
      A[(size_t)(f * M + mm) * Zp + z0 + line] = v;
    }
    __syncthreads();
  }
  // the plane's velocity maxima where they are read (zero otherwise, as the
  // z kernel leaves them)
  const bool want_cfl = (B.prm[slot].adaptive && c->mode == MODE_PRED) || c->host_driven;
  if (!want_cfl) {
    if (tid < 3) B.cflmax[((size_t)slot * 3 + tid) * G.N + n] = 0ull;
  } else {
    __shared__ double wmax[3][NT / 64];
    for (int off = 32; off >= 1; off >>= 1) {
      mr = fmax(mr, __shfl_xor(mr, off));
      mt = fmax(mt, __shfl_xor(mt, off));
      mz = fmax(mz, __shfl_xor(mz, off));
    }
    if ((tid & 63) == 0) {
      wmax[0][tid >> 6] = mr;
      wmax[1][tid >> 6] = mt;
      wmax[2][tid >> 6] = mz;
    }
    __syncthreads();
    if (tid < 3) {
      double v = 0.0;
      for (int w = 0; w < NT / 64; ++w) v = fmax(v, wmax[tid][w]);
      B.cflmax[((size_t)slot * 3 + tid) * G.N + n] = (unsigned long long)__double_as_longlong(v);
    }
  }
  run_fft<ZR0, ZR1, ZR2, 6 * M * Z, NT>(A, 3 * M, T.pz, -1.0, twz, B.twZ);
  // truncation to the retained k (store_zcol for every mode at once)
  constexpr int PER = K + (M - 1) * NK;   // values per field
  for (int i = tid; i < 3 * PER; i += NT) {
    const int f = i / PER, r = i - f * PER;
    double2* row = B.Nraw[c->ncur] + (size_t)slot * 3 * G.coll + (size_t)f * G.coll + (size_t)n * G.Hp;
    if (r < K) {
      row[r] = A[(size_t)(f * M) * Zp + pidx(r)];
    } else {
      const int m = 1 + (r - K) / NK, k = (r - K) - (m - 1) * NK - (K - 1);
      row[K + (m - 1) * NK + k + K - 1] = A[(size_t)(f * M + m) * Zp + pidx(k >= 0 ? k : k + Z)];
    }
  }
}

// ---------------------------------------------------------------------------
// Granular transforms of one coll field (tra_coll2phys / tra_phys2coll of the
// drop-in interface): spec -> Y[n][z][m] -> phys[n][z][th] and back.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(FFT_THREADS) k_zfwd_one(Geo G, Bufs B, FftTiles T, const double2* a, double2* Y,
                                                          size_t fa, size_t fy) {
  a += blockIdx.z * fa;   // field blockIdx.z of a batch at strides fa, fy
  Y += blockIdx.z * fy;
  const int n = blockIdx.y;
  const int m0 = blockIdx.x * T.tm;
  const int tm = min(T.tm, G.M - m0);
  const int Z = G.Z, K = G.K, M = G.M;
  const double2* src = a + (size_t)n * G.Hp;
  const int Zp = padded_len(Z);
  double2* buf = smem;
  double2* twl = smem + (size_t)T.tm * Zp;
  load_tw(twl, B.twZ, Z);
  load_zcols(buf, src, K, Z, Zp, m0, tm);
  __syncthreads();
  fft_lds(buf, tm, T.pz, 1.0, twl, B.twZ);
  for (int i = threadIdx.x; i < tm * Z; i += FFT_THREADS) {
    const int z = i / tm, b = i - z * tm;
    Y[((size_t)n * Z + z) * M + m0 + b] = buf[b * Zp + pidx(z)];
  }
}

// c2r along theta of one field, lines paired as a + ib.  grid: x = z pairs, y = N.
// LDS: one padded line, then T twiddles.
__global__ void __launch_bounds__(FFT_THREADS) k_theta_c2r(Geo G, Bufs B, FftTiles T, const double2* Y, double* phys) {
  const int n = blockIdx.y;
  const int z0 = blockIdx.x * 2;
  const int M = G.M, TT = G.T;
  const bool two = z0 + 1 < G.Z;
  double2* buf = smem;
  double2* twl = smem + padded_len(TT);
  load_tw(twl, B.twT, TT);
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
  fft_lds(buf, 1, T.pt, 1.0, twl, B.twT);
  // phys layout of the reference: Re(0:3K-1, 0:3M-1, N), z fastest
  double* out = phys + (size_t)n * TT * G.Z + z0;
  for (int th = threadIdx.x; th < TT; th += FFT_THREADS) {
    out[(size_t)th * G.Z] = buf[pidx(th)].x;
    if (two) out[(size_t)th * G.Z + 1] = buf[pidx(th)].y;
  }
}

// r2c along theta (scaled by 1/(3K 3M) first, transform.fftw3.f90:114-117).
// The theta transforms of the variable-viscosity plugins: the z-lines of one
// radial plane packed two per complex transform (as k_theta_c2r does), BT
// transforms per workgroup, on physical fields in the plugins' own layout
// [n][z][theta] (theta fastest, so every line is one contiguous run).
// grid: x = blocks of 2 BT z-lines, y = N.
__global__ void __launch_bounds__(FFT_THREADS) k_c2r_lines(Geo G, Bufs B, FftTiles T, int BT, const double2* Y,
                                                           double* phys, size_t fy, size_t fp) {
  Y += blockIdx.z * fy;   // field blockIdx.z of a batch at strides fy, fp
  phys += blockIdx.z * fp;
  const int n = blockIdx.y;
  const int M = G.M, TT = G.T, Z = G.Z;
  const int Tp = padded_len(TT);
  const int z0 = blockIdx.x * 2 * BT;
  const int nt = min(BT, (Z - z0 + 1) / 2);
  double2* buf = smem;
  double2* twl = smem + (size_t)BT * Tp;
  load_tw(twl, B.twT, TT);
  for (int i = threadIdx.x; i < nt * TT; i += FFT_THREADS) {
    const int b = i / TT, mm = i - b * TT;
    const int z = z0 + 2 * b;
    const bool two = z + 1 < Z;
    const double2* A = Y + ((size_t)n * Z + z) * M;
    const double2* Bp = A + M;
    double2 v = make_double2(0.0, 0.0);
    if (mm == 0) v = make_double2(A[0].x, two ? Bp[0].x : 0.0);
    else if (mm < M) {
      const double2 a = A[mm], c = two ? Bp[mm] : make_double2(0.0, 0.0);
      v = make_double2(a.x - c.y, a.y + c.x);
    } else if (mm > TT - M) {
      const int q = TT - mm;
      const double2 a = A[q], c = two ? Bp[q] : make_double2(0.0, 0.0);
      v = make_double2(a.x + c.y, -a.y + c.x);
    }
    buf[(size_t)b * Tp + mm] = v;
  }
  __syncthreads();
  fft_lds(buf, nt, T.pt, 1.0, twl, B.twT);
  for (int i = threadIdx.x; i < nt * TT; i += FFT_THREADS) {
    const int b = i / TT, th = i - b * TT;
    const int z = z0 + 2 * b;
    const double2 v = buf[(size_t)b * Tp + pidx(th)];
    double* out = phys + ((size_t)n * Z + z) * TT + th;
    out[0] = v.x;
    if (z + 1 < Z) out[TT] = v.y;
  }
}

// The inverse of k_c2r_lines with the scaling of tra_phys2spec (1/(3K 3M),
// transform.fftw3.f90:114), into the z-transform intermediates Y.
__global__ void __launch_bounds__(FFT_THREADS) k_r2c_lines(Geo G, Bufs B, FftTiles T, int BT, const double* phys,
                                                           double2* Y, size_t fp, size_t fy) {
  phys += blockIdx.z * fp;   // field blockIdx.z of a batch at strides fp, fy
  Y += blockIdx.z * fy;
  const int n = blockIdx.y;
  const int M = G.M, TT = G.T, Z = G.Z;
  const int Tp = padded_len(TT);
  const int z0 = blockIdx.x * 2 * BT;
  const int nt = min(BT, (Z - z0 + 1) / 2);
  const double scale = 1.0 / (double)(Z * TT);
  double2* buf = smem;
  double2* twl = smem + (size_t)BT * Tp;
  load_tw(twl, B.twT, TT);
  for (int i = threadIdx.x; i < nt * TT; i += FFT_THREADS) {
    const int b = i / TT, th = i - b * TT;
    const int z = z0 + 2 * b;
    const double* in = phys + ((size_t)n * Z + z) * TT + th;
    buf[(size_t)b * Tp + th] = make_double2(scale * in[0], z + 1 < Z ? scale * in[TT] : 0.0);
  }
  __syncthreads();
  fft_lds(buf, nt, T.pt, -1.0, twl, B.twT);
  for (int i = threadIdx.x; i < nt * M; i += FFT_THREADS) {
    const int b = i / M, mm = i - b * M;
    const int z = z0 + 2 * b;
    const double2* L = buf + (size_t)b * Tp;
    const double2 X = L[pidx(mm)], Xc = L[pidx((TT - mm) % TT)];
    double2* A = Y + ((size_t)n * Z + z) * M;
    A[mm] = make_double2(0.5 * (X.x + Xc.x), 0.5 * (X.y - Xc.y));
    if (z + 1 < Z) A[M + mm] = make_double2(0.5 * (X.y + Xc.y), -0.5 * (X.x - Xc.x));
  }
}

__global__ void __launch_bounds__(FFT_THREADS) k_theta_r2c(Geo G, Bufs B, FftTiles T, const double* phys, double2* Y) {
  const int n = blockIdx.y;
  const int z0 = blockIdx.x * 2;
  const int M = G.M, TT = G.T;
  const bool two = z0 + 1 < G.Z;
  const double scale = 1.0 / (double)(G.Z * TT);
  double2* buf = smem;
  double2* twl = smem + padded_len(TT);
  load_tw(twl, B.twT, TT);
  const double* in = phys + (size_t)n * TT * G.Z + z0;
  for (int th = threadIdx.x; th < TT; th += FFT_THREADS)
    buf[th] = make_double2(scale * in[(size_t)th * G.Z], two ? scale * in[(size_t)th * G.Z + 1] : 0.0);
  __syncthreads();
  fft_lds(buf, 1, T.pt, -1.0, twl, B.twT);
  double2* A = Y + ((size_t)n * G.Z + z0) * M;
  for (int mm = threadIdx.x; mm < M; mm += FFT_THREADS) {
    const double2 X = buf[pidx(mm)], Xc = buf[pidx((TT - mm) % TT)];
    A[mm] = make_double2(0.5 * (X.x + Xc.x), 0.5 * (X.y - Xc.y));
    if (two) A[M + mm] = make_double2(0.5 * (X.y + Xc.y), -0.5 * (X.x - Xc.x));
  }
}

__global__ void __launch_bounds__(FFT_THREADS) k_zbwd_one(Geo G, Bufs B, FftTiles T, const double2* Y, double2* a,
                                                          size_t fy, size_t fa) {
  Y += blockIdx.z * fy;   // field blockIdx.z of a batch at strides fy, fa
  a += blockIdx.z * fa;
  const int n = blockIdx.y;
  const int m0 = blockIdx.x * T.tm;
  const int tm = min(T.tm, G.M - m0);
  const int Z = G.Z, K = G.K, M = G.M;
  const int Zp = padded_len(Z);
  double2* buf = smem;
  double2* twl = smem + (size_t)T.tm * Zp;
  load_tw(twl, B.twZ, Z);
  load_ycols<0>(buf, Y + (size_t)n * Z * M, (size_t)M, Z, Zp, m0, tm);
  __syncthreads();
  fft_lds(buf, tm, T.pz, -1.0, twl, B.twZ);
  double2* row = a + (size_t)n * G.Hp;
  for (int b = 0; b < tm; ++b) store_zcol(row, buf + b * Zp, K, Z, m0 + b, threadIdx.x, FFT_THREADS);
}

// Radix sequences with specialised transform kernels (the transform lengths
// 3K, 3M of the usual grids); other lengths use the generic kernels.
namespace {
template <int R0, int R1, int R2>
FftKerns kerns_of() {
  return FftKerns{k_zfwd_t<R0, R1, R2>, k_theta_t<R0, R1, R2>, k_zbwd_t<R0, R1, R2>};
}
struct PlanKerns {
  int r[3];
  FftKerns k;
};
}  // namespace

// Plan pairs (z, theta) with a whole-plane kernel: the grids whose six
// z-transformed fields fit in LDS.
PlaneKern plane_kernel(const FftPlan& pz, const FftPlan& pt, size_t* smem) {
  struct Entry {
    int z[3], t[3], pl;
    PlaneKern k;
  };
  static const Entry table[] = {
      {{12, 6, 1}, {12, 3, 1}, 24, k_plane_t<12, 6, 1, 12, 3, 1, 24>},   // (K, M) = (24, 12): 132 KiB
  };
  auto same = [](const FftPlan& p, const int r[3]) {
    if (p.nrad < 1 || p.nrad > 3) return false;
    for (int i = 0; i < 3; ++i)
      if ((i < p.nrad ? p.rad[i] : 1) != r[i]) return false;
    return true;
  };
  for (const auto& e : table)
    if (same(pz, e.z) && same(pt, e.t)) {
      const int Z = pz.L, TT = pt.L, M = TT / 3;
      *smem = ((size_t)6 * M * padded_len(Z) + (size_t)3 * e.pl * padded_len(TT) + Z + TT) * sizeof(double2);
      return e.k;
    }
  *smem = 0;
  return nullptr;
}

FftKerns fft_kernels(const FftPlan& pz, const FftPlan& pt) {
  static const PlanKerns table[] = {
      {{9, 2, 1}, kerns_of<9, 2, 1>()},     {{12, 1, 1}, kerns_of<12, 1, 1>()},
      {{12, 2, 1}, kerns_of<12, 2, 1>()},   {{12, 3, 1}, kerns_of<12, 3, 1>()},
      {{12, 6, 1}, kerns_of<12, 6, 1>()},   {{16, 3, 1}, kerns_of<16, 3, 1>()},
      {{16, 9, 1}, kerns_of<16, 9, 1>()},   {{16, 12, 1}, kerns_of<16, 12, 1>()},
      {{16, 12, 2}, kerns_of<16, 12, 2>()}, {{16, 16, 3}, kerns_of<16, 16, 3>()},
  };
  auto find = [&](const FftPlan& p) -> const FftKerns* {
    if (p.nrad < 1 || p.nrad > 3) return nullptr;
    int r[3] = {1, 1, 1};
    for (int i = 0; i < p.nrad; ++i) r[i] = p.rad[i];
    for (const auto& e : table)
      if (e.r[0] == r[0] && e.r[1] == r[1] && e.r[2] == r[2]) return &e.k;
    return nullptr;
  };
  const FftKerns generic = kerns_of<0, 0, 0>();
  FftKerns k = generic;
  if (const FftKerns* z = find(pz)) {
    k.zfwd = z->zfwd;
    k.zbwd = z->zbwd;
  }
  if (const FftKerns* t = find(pt)) k.theta = t->theta;
  return k;
}

}  // namespace opf
