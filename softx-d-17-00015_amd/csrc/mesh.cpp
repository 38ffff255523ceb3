// Host precompute of the radial mesh and finite-difference weights.
//
// Restates mes_precompute / mes_rdom_init / mes_weights (meshs.f90:66-225)
// and the Chebyshev helpers CHEBY / EVAL2 / DERIV2 (cheby.f:2-108) in their
// order of floating-point operations; the dense inverse follows LAPACK's
// unblocked DGETF2 + DGETRI as mes_mat_invert (meshs.f90:231-243) calls them.
// Runs once per resolution class; the results are uploaded to the device.
#include <cmath>
#include <stdexcept>
#include <vector>

#include "opf_internal.hpp"

namespace opf {

// DGETF2 + DGETRI (unblocked) on a row-major n x n matrix, in place.
void mat_invert(std::vector<double>& A, int n) {
  auto at = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
  std::vector<int> ipiv(n);
  for (int j = 0; j < n; ++j) {
    int p = j;
    double best = std::fabs(at(j, j));
    for (int i = j + 1; i < n; ++i)
      if (std::fabs(at(i, j)) > best) { best = std::fabs(at(i, j)); p = i; }
    ipiv[j] = p;
    if (at(p, j) == 0.0) throw std::runtime_error("matrix inversion error 1");
    if (p != j)
      for (int c = 0; c < n; ++c) std::swap(at(p, c), at(j, c));
    if (j < n - 1) {
      const double rc = 1.0 / at(j, j);
      for (int i = j + 1; i < n; ++i) at(i, j) = at(i, j) * rc;
      for (int i = j + 1; i < n; ++i) {
        const double lij = at(i, j);
        for (int c = j + 1; c < n; ++c) at(i, c) = at(i, c) - lij * at(j, c);
      }
    }
  }
  // DTRTI2: inverse of U in place
  std::vector<double> x(n);
  for (int j = 0; j < n; ++j) {
    at(j, j) = 1.0 / at(j, j);
    const double ajj = -at(j, j);
    for (int i = 0; i < j; ++i) x[i] = at(i, j);
    for (int jj = 0; jj < j; ++jj) {
      if (x[jj] != 0.0) {
        const double temp = x[jj];
        for (int i = 0; i < jj; ++i) x[i] = x[i] + temp * at(i, jj);
        x[jj] = x[jj] * at(jj, jj);
      }
    }
    for (int i = 0; i < j; ++i) at(i, j) = x[i] * ajj;
  }
  // solve inv(A)*L = inv(U)
  std::vector<double> work(n);
  for (int j = n - 1; j >= 0; --j) {
    for (int i = j + 1; i < n; ++i) { work[i] = at(i, j); at(i, j) = 0.0; }
    for (int jj = j + 1; jj < n; ++jj) {
      const double temp = -work[jj];
      if (temp != 0.0)
        for (int i = 0; i < n; ++i) at(i, j) = at(i, j) + temp * at(i, jj);
    }
  }
  for (int j = n - 2; j >= 0; --j) {
    const int jp = ipiv[j];
    if (jp != j)
      for (int i = 0; i < n; ++i) std::swap(at(i, j), at(i, jp));
  }
}

namespace {

// DERIV2 (cheby.f:69-108): coefficients of df/dx from those of f.
std::vector<double> deriv2(const std::vector<double>& a) {
  const int m = (int)a.size();
  std::vector<double> A(m + 1), AX(m + 1, 0.0);
  for (int i = 0; i < m; ++i) A[i + 1] = a[i];
  AX[m] = 0.0;
  if (m == 1) {
    AX[1] = 0.0;
  } else if (m == 2) {
    AX[1] = 2.0 * A[2];
  } else if (m == 3) {
    AX[1] = 2.0 * A[2];
    AX[2] = 8.0 * A[3];
  } else if (m == 4) {
    AX[1] = 2.0 * A[2] + 6.0 * A[4];
    AX[2] = 8.0 * A[3];
    AX[3] = 12.0 * A[4];
  } else {
    AX[m - 1] = -((4.0 * (1.0 - m)) * A[m]);
    AX[m - 2] = -((4.0 * (2.0 - m)) * A[m - 1]);
    for (int j = m - 3; j >= 2; --j) AX[j] = (4.0 * j) * A[j + 1] + AX[j + 2];
    AX[1] = 2.0 * A[2] + AX[3] / 2.0;
  }
  return std::vector<double>(AX.begin() + 1, AX.end());
}

// EVAL2 (cheby.f:38-67): Clenshaw evaluation on the shifted interval [0,1].
double eval2(const std::vector<double>& a, double x) {
  const int m = (int)a.size();
  std::vector<double> B(m + 3, 0.0);
  const double xa = 2.0 * x - 1.0;
  const double xb = 2.0 * xa;
  for (int j = m - 1; j >= 0; --j) B[j + 1] = (xb * B[j + 2] - B[j + 3]) + a[j];
  return B[1] - xa * B[2];
}

// CHEBY (cheby.f:2-36): nd-th derivative of T_m at x.
double cheby(int m, int nd, double x) {
  if (m == 0) return nd == 0 ? 1.0 : 0.0;
  std::vector<double> a(m + 1, 0.0);
  a[m] = 1.0;
  for (int i = 0; i < nd; ++i) a = deriv2(a);
  return eval2(a, x);
}


// Integer power as the reference's compiler lowers x**n for integer n.
double powi(double x, int n) {
  unsigned b = (unsigned)(n < 0 ? -n : n);
  double r = 1.0, a = x;
  for (;;) {
    if (b & 1u) r *= a;
    b >>= 1;
    if (b == 0) break;
    a *= a;
  }
  return n < 0 ? 1.0 / r : r;
}

// mes_weights (meshs.f90:200-225): weights of the i-th derivative at x0.
std::vector<double> weights(int i, const double* x, int n, double x0) {
  double xl = x0, xr = x0;
  for (int j = 0; j < n; ++j) { xl = std::fmin(xl, x[j]); xr = std::fmax(xr, x[j]); }
  std::vector<double> x_(n), A((size_t)n * n), d(n);
  for (int j = 0; j < n; ++j) x_[j] = (x[j] - xl) / (xr - xl);
  const double x0_ = (x0 - xl) / (xr - xl);
  for (int j = 0; j < n; ++j) {
    for (int ii = 0; ii < n; ++ii) A[(size_t)ii * n + j] = cheby(j, 0, x_[ii]);
    d[j] = cheby(j, i, x0_);
  }
  mat_invert(A, n);
  const double den = powi(xr - xl, i);
  std::vector<double> w(n);
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int ii = 0; ii < n; ++ii) s = s + d[ii] * A[(size_t)ii * n + j];
    w[j] = s / den;
  }
  return w;
}

}  // namespace

void build_mesh(int N, const double* r_in, HostMesh& me) {
  if (N < 2 * KL + 1 || N > MAXN) throw std::runtime_error("radial resolution out of range");
  me.N = N;
  me.r.assign(N, 0.0);
  if (r_in) {
    for (int n = 0; n < N; ++n) me.r[n] = r_in[n];
    if (me.r[0] < 0.0) throw std::runtime_error("mes_precompute: mesh err 2");
    if (std::fabs(me.r[N - 1] - 1.0) > 1e-8) throw std::runtime_error("mes_precompute: mesh err 3");
  } else {
    // default mesh (meshs.f90:90-99)
    const int N_ = N + (int)std::sqrt((double)N);
    for (int n = N_ - N + 1; n <= N_; ++n)
      me.r[n - N_ + N - 1] = 0.5 * (1.0 + std::cos(D_PI * (N_ - n) / (double)N_));
    for (int it = 0; it < 10; ++it) {
      const double dr = 1.5 * me.r[0] - 0.5 * me.r[1];
      for (int n = 0; n < N; ++n) me.r[n] = me.r[n] * (1.0 + dr) - dr;
    }
  }
  me.rinv.resize(N);
  me.rinv2.resize(N);
  me.r2.resize(N);
  for (int n = 0; n < N; ++n) {
    me.rinv[n] = powi(me.r[n], -1);
    me.rinv2[n] = powi(me.r[n], -2);
    me.r2[n] = powi(me.r[n], 2);
  }
  std::vector<double> r_(N + 1);
  r_[0] = 0.0;
  for (int n = 0; n < N; ++n) r_[n + 1] = me.r[n];

  std::vector<double> dr[KL + 1];
  for (int i = 1; i <= KL; ++i) dr[i].assign((size_t)LDA * N, 0.0);
  auto B = [&](std::vector<double>& band, int row, int col) -> double& {
    return band[(size_t)row * N + col];
  };
  for (int n = 1; n <= N; ++n) {
    const int l = std::max(1, n - KL), rr = std::min(n + KL, N);
    for (int i = 1; i <= KL; ++i) {
      auto w = weights(i, &r_[l], rr - l + 1, r_[n]);
      for (int j = l; j <= rr; ++j) B(dr[i], KL + n - j, j - 1) = w[j - l];
    }
  }
  me.dr0.assign((KL + 1) * (KL + 1), 0.0);
  me.dr1.assign((KL + 1) * (KL + 1), 0.0);
  for (int i = 0; i <= KL; ++i) {
    auto w0 = weights(i, &r_[1], KL + 1, r_[0]);
    auto w1 = weights(i, &r_[N - KL], KL + 1, r_[N]);
    for (int q = 0; q <= KL; ++q) {
      me.dr0[q + (KL + 1) * i] = w0[q];
      me.dr1[q + (KL + 1) * i] = w1[q];
    }
  }
  me.lapb.assign((size_t)LDA * N, 0.0);
  for (int j = 1; j <= N; ++j) {
    const int l = std::max(1, j - KL), rr = std::min(j + KL, N);
    for (int n = l; n <= rr; ++n)
      B(me.lapb, KL + n - j, j - 1) = B(dr[2], KL + n - j, j - 1) + B(dr[1], KL + n - j, j - 1) / r_[n];
  }
  me.intrdr.assign(N, 0.0);
  for (int n = 1; n <= N; ++n) {
    const int l = std::max(1, n - KL), rr = std::min(n + KL, N), nn = rr - l + 1;
    std::vector<std::vector<double>> A(nn);
    for (int i = 1; i <= nn; ++i) A[i - 1] = weights(i - 1, &r_[l], nn, r_[n]);
    double c = 1.0, e = 1.0;
    for (int i = 1; i <= nn; ++i) {
      c = c * (r_[std::min(n + 1, N)] - r_[n]) / (double)i;
      e = e * (r_[std::max(0, n - 1)] - r_[n]) / (double)i;
      for (int jj = 0; jj < nn; ++jj)
        me.intrdr[l - 1 + jj] = me.intrdr[l - 1 + jj] + ((0.5 * (c - e)) * A[i - 1][jj]) * r_[n];
    }
  }
  me.dr1b = dr[1];
  me.dr2b = dr[2];
}

// Radix plan for L = 2^a 3^b: radix-4 passes first, then 2, then 3.
FftPlan make_plan(int L) {
  FftPlan p;
  p.L = L;
  int n = L;
  // largest supported radix first: lengths 3K and 3M of the usual grids need
  // two or three passes (144 = 16*9, 192 = 16*12, 768 = 16*16*3)
  static const int radices[] = {16, 12, 9, 8, 6, 4, 3, 2};
  while (n > 1) {
    int r = 0;
    for (int c : radices)
      if (n % c == 0) { r = c; break; }
    if (r == 0 || p.nrad == 16) throw std::runtime_error("FFT length must be 2^a 3^b");
    p.rad[p.nrad++] = r;
    n /= r;
  }
  return p;
}

}  // namespace opf
