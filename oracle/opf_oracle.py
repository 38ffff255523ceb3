"""CPU restatement (oracle) of the Openpipeflow 1.12 time-step hot path.

TEST INFRASTRUCTURE ONLY.  This module is imported by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` as
the checker; the product package never imports it and never falls back to it.

Every function restates the reference algorithm operation by operation, in
the reference's order of floating-point operations, and cites the reference
file:line it follows.  The arrays follow the reference's ``coll`` layout
(radial index fastest is irrelevant in numpy; here a coll field is a complex
array ``a[n, h]`` of shape (N, H), n = radial point 0..N-1, h = harmonic in
the order of parallel.h:47-56).

Third-party algorithms the reference calls:
  * FFTW3 (version not pinned, Makefile:26): the DFT definitions of
    transform.fftw3.f90:35-147 are restated with numpy.fft (pocketfft).
  * LAPACK dgbtrf/dgbtrs (version not pinned): restated as LAPACK 3.x's
    unblocked DGBTF2 (DGBTRF takes this path for KL=4 < NB) and DGBTRS with
    DGER / DTBSV, vectorised over the harmonics.
  * LAPACK dgetrf/dgetri (mesh weights, 4x4 influence matrix): restated as
    unblocked DGETF2 + DGETRI (DTRTI2 + column sweep + column swaps).

Parity pins (tests/test_oracle_*.py, tests/test_state_io.py): golden vectors
produced by the compiled reference (oracle/_ref, tests/golden/make_golden.py),
including steporbit's call sequence; FD weights exact on polynomials; DFT brute
force; banded LU vs dense solve; laminar closed forms (E = pi^2/(3 alpha)
normalisation, u_tau = sqrt(2/Re)); the axisymmetric swirl decay rate;
var_coll_shift as a translation of the grid by whole points; io_interp exact
on polynomials of the field's parity across the axis; the plugins' forces in
closed form on Hagen-Poiseuille flow (Coriolis also about the pipe axis).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

KL = 4                      # parameters.f90:45
LDA = 2 * KL + 1            # band width of 'mesh' type (meshs.f90:37-39)
LDAB = 3 * KL + 1           # 'lumesh' leading dim (meshs.f90:42-45)
D_PI = 3.1415926535897931   # parameters.f90:54


# ---------------------------------------------------------------------------
# cheby.f
# ---------------------------------------------------------------------------
def _deriv2(a):
    """DERIV2 (cheby.f:69-108): Chebyshev coefficients of df/dx."""
    m = len(a)
    ax = [0.0] * m
    if m <= 0:
        raise ValueError("M .LE. 0 in Sub DERIV2")
    A = [None] + list(a)            # 1-based
    AX = [None] + ax
    AX[m] = 0.0
    if m == 1:
        AX[1] = 0.0
    elif m == 2:
        AX[1] = 2.0 * A[2]
    elif m == 3:
        AX[1] = 2.0 * A[2]
        AX[2] = 8.0 * A[3]
    elif m == 4:
        AX[1] = 2.0 * A[2] + 6.0 * A[4]
        AX[2] = 8.0 * A[3]
        AX[3] = 12.0 * A[4]
    else:
        AX[m - 1] = -((4.0 * (1.0 - m)) * A[m])
        AX[m - 2] = -((4.0 * (2.0 - m)) * A[m - 1])
        for j in range(m - 3, 1, -1):
            AX[j] = (4.0 * j) * A[j + 1] + AX[j + 2]
        AX[1] = 2.0 * A[2] + AX[3] / 2.0
    return AX[1:]


def _eval2(a, x):
    """EVAL2 (cheby.f:38-67): Clenshaw sum of a Chebyshev series on [0,1]."""
    m = len(a)
    A = [None] + list(a)
    out = []
    for xx in x:
        xa = 2.0 * xx - 1.0
        xb = 2.0 * xa
        B = [0.0] * (m + 3)     # B(0:M+2)
        for j in range(m - 1, -1, -1):
            B[j + 1] = (xb * B[j + 2] - B[j + 3]) + A[j + 1]
        out.append(B[1] - xa * B[2])
    return out


def cheby(m, nd, x):
    """CHEBY (cheby.f:2-36): nd-th derivative of T_m at the points x."""
    if m == 0:
        return [1.0 if nd == 0 else 0.0 for _ in x]
    a = [0.0] * m + [1.0]
    for _ in range(nd):
        a = _deriv2(a)
    return _eval2(a, x)


# ---------------------------------------------------------------------------
# dense LAPACK restatements (small matrices, pure Python)
# ---------------------------------------------------------------------------
def dgetf2(A):
    """Unblocked LU with partial pivoting (LAPACK DGETF2).  A: list of rows,
    modified in place.  Returns 0-based pivot list."""
    n = len(A)
    ipiv = []
    for j in range(n):
        p = max(range(j, n), key=lambda i: (abs(A[i][j]), -i))
        ipiv.append(p)
        if A[p][j] == 0.0:
            raise ArithmeticError("matrix inversion error 1")
        if p != j:
            A[p], A[j] = A[j], A[p]
        if j < n - 1:
            r = 1.0 / A[j][j]
            for i in range(j + 1, n):
                A[i][j] = A[i][j] * r
            for i in range(j + 1, n):
                lij = A[i][j]
                for c in range(j + 1, n):
                    A[i][c] = A[i][c] - lij * A[j][c]
    return ipiv


def dgetri(A, ipiv):
    """Inverse from the LU factors (LAPACK DGETRI, unblocked path)."""
    n = len(A)
    # DTRTI2 upper, non-unit: inv(U) in place
    for j in range(n):
        A[j][j] = 1.0 / A[j][j]
        ajj = -A[j][j]
        # DTRMV('Upper','No transpose','Non-unit', j, A, x = A(0:j, j))
        x = [A[i][j] for i in range(j)]
        for jj in range(j):
            if x[jj] != 0.0:
                temp = x[jj]
                for i in range(jj):
                    x[i] = x[i] + temp * A[i][jj]
                x[jj] = x[jj] * A[jj][jj]
        for i in range(j):
            A[i][j] = x[i] * ajj
    # solve inv(A)*L = inv(U)
    for j in range(n - 1, -1, -1):
        work = [0.0] * n
        for i in range(j + 1, n):
            work[i] = A[i][j]
            A[i][j] = 0.0
        if j < n - 1:
            # DGEMV: A(:,j) = A(:,j) - A(:,j+1:n) * work(j+1:n)
            for jj in range(j + 1, n):
                temp = -work[jj]
                if temp != 0.0:
                    for i in range(n):
                        A[i][j] = A[i][j] + temp * A[i][jj]
    for j in range(n - 2, -1, -1):
        jp = ipiv[j]
        if jp != j:
            for i in range(n):
                A[i][j], A[i][jp] = A[i][jp], A[i][j]
    return A


def mat_invert(A):
    """mes_mat_invert (meshs.f90:231-243): dgetrf + dgetri."""
    A = [list(map(float, row)) for row in A]
    piv = dgetf2(A)
    return dgetri(A, piv)


# ---------------------------------------------------------------------------
# meshs.f90
# ---------------------------------------------------------------------------
def _powi(x, n):
    """Integer power as LLVM's powi lowering does it (x**n, n integer)."""
    b = abs(n)
    r = 1.0
    a = x
    while True:
        if b & 1:
            r *= a
        b >>= 1
        if b == 0:
            break
        a *= a
    return 1.0 / r if n < 0 else r


def mes_weights(i, x, x0):
    """mes_weights (meshs.f90:200-225): weights of the i-th derivative at x0
    from the points x, via a local Chebyshev interpolant."""
    n = len(x)
    xl = min(x0, min(x))
    xr = max(x0, max(x))
    x_ = [(xx - xl) / (xr - xl) for xx in x]
    x0_ = (x0 - xl) / (xr - xl)
    A = [[0.0] * n for _ in range(n)]
    d = [0.0] * n
    for j in range(n):
        col = cheby(j, 0, x_)
        for ii in range(n):
            A[ii][j] = col[ii]
        d[j] = cheby(j, i, [x0_])[0]
    Ai = mat_invert(A)
    den = _powi(xr - xl, i)
    w = []
    for j in range(n):
        s = 0.0
        for ii in range(n):
            s = s + d[ii] * Ai[ii][j]
        w.append(s / den)
    return w


def default_mesh(N):
    """Default radial points (meshs.f90:90-99)."""
    N_ = N + int(math.sqrt(float(N)))
    r = [0.0] * N
    for n in range(N_ - N + 1, N_ + 1):
        r[n - N_ + N - 1] = 0.5 * (1.0 + math.cos(D_PI * (N_ - n) / float(N_)))
    for _ in range(10):
        dr = 1.5 * r[0] - 0.5 * r[1]
        r = [rr * (1.0 + dr) - dr for rr in r]
    return r


@dataclass
class Mesh:
    """type rdom (meshs.f90:48-55) for a serial run."""
    N: int
    r: np.ndarray                 # r[n] = r_n (n = 0..N-1)
    rp: dict                      # rp[j] = r**j, j in -3..3
    dr: list                      # dr[i] band (LDA, N), i = 1..KL (index 0 unused)
    radLap: np.ndarray            # band (LDA, N)
    dr0: np.ndarray               # (KL+1, KL+1): dr0[:, i]
    dr1: np.ndarray               # (KL+1, KL+1): dr1[:, i]
    intrdr: np.ndarray            # (N,)


def build_mesh(N, r=None):
    """mes_precompute + mes_rdom_init (meshs.f90:66-194)."""
    r1 = list(default_mesh(N) if r is None else r)
    rp = {}
    for j in range(-3, 4):
        rp[j] = np.array(r1) if j == 1 else np.array([_powi(x, j) for x in r1])
    r_ = [0.0] + r1                 # r_(0:N)
    dr = [None] + [np.zeros((LDA, N)) for _ in range(KL)]
    for n in range(1, N + 1):
        l = max(1, n - KL)
        rr = min(n + KL, N)
        for i in range(1, KL + 1):
            w = mes_weights(i, r_[l:rr + 1], r_[n])
            for j in range(l, rr + 1):
                dr[i][KL + n - j, j - 1] = w[j - l]
    dr0 = np.zeros((KL + 1, KL + 1))
    dr1 = np.zeros((KL + 1, KL + 1))
    for i in range(KL + 1):
        dr0[:, i] = mes_weights(i, r_[1:KL + 2], r_[0])
        dr1[:, i] = mes_weights(i, r_[N - KL:N + 1], r_[N])
    radLap = np.zeros((LDA, N))
    for j in range(1, N + 1):
        l = max(1, j - KL)
        rr = min(j + KL, N)
        for n in range(l, rr + 1):
            radLap[KL + n - j, j - 1] = dr[2][KL + n - j, j - 1] + dr[1][KL + n - j, j - 1] / r_[n]
    intrdr = [0.0] * N
    for n in range(1, N + 1):
        l = max(1, n - KL)
        rr = min(n + KL, N)
        nn = rr - l + 1
        A = [mes_weights(i - 1, r_[l:rr + 1], r_[n]) for i in range(1, nn + 1)]
        c = 1.0
        e = 1.0
        for i in range(1, nn + 1):
            c = c * (r_[min(n + 1, N)] - r_[n]) / float(i)
            e = e * (r_[max(0, n - 1)] - r_[n]) / float(i)
            for jj in range(nn):
                intrdr[l - 1 + jj] = intrdr[l - 1 + jj] + ((0.5 * (c - e)) * A[i - 1][jj]) * r_[n]
    return Mesh(N=N, r=np.array(r1), rp=rp, dr=dr, radLap=radLap, dr0=dr0, dr1=dr1,
                intrdr=np.array(intrdr))


def band_to_dense(M):
    """Dense A from 'mesh' band storage M(KL+1+n-j, j) = A(n,j)."""
    N = M.shape[1]
    A = np.zeros((N, N))
    for j in range(N):
        for n in range(max(0, j - KL), min(j + KL, N - 1) + 1):
            A[n, j] = M[KL + n - j, j]
    return A


# ---------------------------------------------------------------------------
# harmonics (parallel.h:47-56, variables.f90:60-88)
# ---------------------------------------------------------------------------
@dataclass
class Harmonics:
    K: int
    M: int
    Mp: int
    k: np.ndarray = field(init=False)
    m: np.ndarray = field(init=False)

    def __post_init__(self):
        ks, ms = [], []
        for m in range(self.M):
            for k in range(-(self.K - 1), self.K):
                if k < 0 and m == 0:
                    continue
                ks.append(k)
                ms.append(m)
        self.k = np.array(ks, dtype=np.int64)
        self.m = np.array(ms, dtype=np.int64)

    @property
    def H(self):
        return len(self.k)


def n_harmonics(K, M):
    return (2 * K - 1) * M - (K - 1)


class Ctx:
    """Module state of the reference for one run: parameters, mesh and the
    derived tables of var_precompute (variables.f90:60-88)."""

    def __init__(self, N, K, M, Mp=1, Re=2400.0, alpha=1.25, dt=0.01,
                 implicit=0.51, const_flux=True, dterr_tol=1e-5, r=None):
        self.N, self.K, self.M, self.Mp = N, K, M, Mp
        self.Re, self.alpha, self.dt = float(Re), float(alpha), float(dt)
        self.implicit = implicit
        self.const_flux = const_flux
        self.dterr_tol = dterr_tol
        self.mesh = build_mesh(N, r)
        self.harm = Harmonics(K, M, Mp)
        me = self.mesh
        hk, hm = self.harm.k, self.harm.m
        # ad_m1r1(:,m) = r^-1 * Mp*m ; ad_k1a1(k) = alpha*k  (variables.f90:79-86)
        self.m1r1 = (me.rp[-1][:, None] * float(Mp)) * hm[None, :].astype(float)
        self.k1a1 = self.alpha * hk.astype(float)
        self.nu = 1.0                                   # vel_precompute
        self.U = 1.0 - me.rp[2]                         # velocity.f90:53
        self.Up = -(2.0 * me.r)                         # velocity.f90:54
        self.Pr0 = 0.0
        self.t = 0.0
        self.step = 0
        self.corr_dt = 0.0        # tim_corr_dt (timestep.f90:35)
        self.cfl_dt = 0.0         # tim_cfl_dt (timestep.f90:36)
        self.cfl_dir = 0
        self._ntc = 0             # the saved counter i of tim_new_tstep (timestep.f90:261)
        self.matrices()

    # -- shapes --------------------------------------------------------------
    @property
    def H(self):
        return self.harm.H

    def zeros(self):
        return np.zeros((self.N, self.H), dtype=np.complex128)

    # -- variables.f90 operators ---------------------------------------------
    def meshmult(self, Mb, x):
        """var_coll_meshmult (variables.f90:378-402): y = A.x, accumulated over
        j ascending starting from 0."""
        N = self.N
        y = np.zeros_like(x)
        for n in range(N):
            acc = np.zeros(x.shape[1:], dtype=x.dtype)
            for j in range(max(0, n - KL), min(n + KL, N - 1) + 1):
                acc = acc + Mb[KL + n - j, j] * x[j]
            y[n] = acc
        return y

    def curl(self, r, t, z):
        """var_coll_curl (variables.f90:408-436)."""
        mr, ak = self.m1r1, self.k1a1[None, :]
        orr = (-(z.imag * mr) + t.imag * ak) + 1j * (z.real * mr - t.real * ak)
        d = self.meshmult(self.mesh.dr[1], z)
        ot = (-(r.imag * ak) - d.real) + 1j * (r.real * ak - d.imag)
        d = self.meshmult(self.mesh.dr[1], t)
        rm1 = self.mesh.rp[-1][:, None]
        oz = ((t.real * rm1 + d.real) + r.imag * mr) + 1j * ((t.imag * rm1 + d.imag) - r.real * mr)
        return orr, ot, oz

    def grad(self, p):
        """var_coll_grad (variables.f90:442-459)."""
        mr, ak = self.m1r1, self.k1a1[None, :]
        gr = self.meshmult(self.mesh.dr[1], p)
        gt = -(p.imag * mr) + 1j * (p.real * mr)
        gz = -(p.imag * ak) + 1j * (p.real * ak)
        return gr, gt, gz

    def div(self, r, t, z):
        """var_coll_div (variables.f90:465-481)."""
        mr, ak = self.m1r1, self.k1a1[None, :]
        rm1 = self.mesh.rp[-1][:, None]
        d = self.meshmult(self.mesh.dr[1], r)
        re = ((r.real * rm1 + d.real) - t.imag * mr) - z.imag * ak
        im = ((r.imag * rm1 + d.imag) + t.real * mr) + z.real * ak
        return re + 1j * im

    @staticmethod
    def imposesym0(r, t, z):
        """var_imposesym(0) (variables.f90:611-616) on harmonic 0."""
        r[:, 0] = 0.0
        t[:, 0] = t[:, 0].real
        z[:, 0] = z[:, 0].real

    def norm(self, a):
        """var_coll_norm (variables.f90:718-751): E, Ek, Em."""
        w = 4.0 * D_PI * D_PI / self.alpha
        f = w * (a.real * a.real + a.imag * a.imag)
        intrdr = self.mesh.intrdr
        b = np.zeros(self.H)
        for n in range(self.N):              # dot_product order, n ascending
            b = b + f[n] * intrdr[n]
        b[0] = 0.5 * b[0]
        Ek = np.zeros(self.K)
        Em = np.zeros(self.M)
        for h in range(self.H):
            Em[self.harm.m[h]] += b[h]
            Ek[abs(self.harm.k[h])] += b[h]
        E = 0.0
        for v in Em:
            E = E + v
        return E, Ek, Em

    # -- transform.fftw3.f90 -------------------------------------------------
    def spec2phys(self, a):
        """var_coll2spec + tra_spec2phys (variables.f90:184-197,
        transform.fftw3.f90:65-102).  Returns phys p[n, z, th] (3K x 3M)."""
        N, K, M = self.N, self.K, self.M
        Z, T = 3 * K, 3 * M
        X = np.zeros((N, Z, M), dtype=np.complex128)
        X[:, 0:K, 0] = a[:, 0:K]
        X[:, 2 * K + 1:, 0] = np.conj(a[:, K - 1:0:-1])
        nh = 2 * K - 1
        for m in range(1, M):
            X[:, 0:K, m] = a[:, nh:nh + K]
            X[:, 2 * K + 1:, m] = a[:, nh - K + 1:nh]
            nh += 2 * K - 1
        Y = np.fft.ifft(X, axis=1) * Z          # exponent +1, unnormalised
        Xs = np.zeros((N, Z, T // 2 + 1), dtype=np.complex128)
        Xs[:, :, 0:M] = Y
        Xs[:, :, 0] = Xs[:, :, 0].real          # c2r reads only Re of the DC bin
        if T % 2 == 0:
            Xs[:, :, T // 2] = Xs[:, :, T // 2].real
        return np.fft.irfft(Xs, n=T, axis=2) * T

    def phys2spec(self, p):
        """tra_phys2spec + var_spec2coll (transform.fftw3.f90:108-147,
        variables.f90:281-294)."""
        N, K, M = self.N, self.K, self.M
        Z, T = 3 * K, 3 * M
        scale = 1.0 / float(Z * T)
        Xs = np.fft.rfft(scale * p, axis=2)
        Y = Xs[:, :, 0:M]
        X = np.fft.fft(Y, axis=1)               # exponent -1, unnormalised
        a = np.zeros((N, self.H), dtype=np.complex128)
        a[:, 0:K] = X[:, 0:K, 0]
        nh = 2 * K - 1
        for m in range(1, M):
            a[:, nh:nh + K] = X[:, 0:K, m]
            a[:, nh - K + 1:nh] = X[:, 2 * K + 1:, m]
            nh += 2 * K - 1
        return a

    # -- LAPACK band restatements (vectorised over systems) ------------------
    @staticmethod
    def dgbtf2(AB):
        """LAPACK DGBTF2 with KL=KU=4 on AB[s, 0:13, 0:N] (0-based rows =
        Fortran rows-1).  In place; returns ipiv[s, j] (1-based rows)."""
        S, _, N = AB.shape
        KU = KL
        KV = KU + KL
        for j in range(KU + 2, min(KV, N) + 1):
            for i in range(KV - j + 2, KL + 1):
                AB[:, i - 1, j - 1] = 0.0
        ipiv = np.zeros((S, N), dtype=np.int64)
        ju = np.ones(S, dtype=np.int64)
        sidx = np.arange(S)
        for j in range(1, N + 1):
            if j + KV <= N:
                AB[:, 0:KL, j + KV - 1] = 0.0
            km = min(KL, N - j)
            col = np.abs(AB[:, KV:KV + km + 1, j - 1])
            jp = np.argmax(col, axis=1) + 1          # first maximum, as IDAMAX
            ipiv[:, j - 1] = jp + j - 1
            piv = AB[sidx, KV + jp - 1, j - 1]
            if np.any(piv == 0.0):
                raise ArithmeticError("tim_lumesh_init")
            ju = np.maximum(ju, np.minimum(j + KU + jp - 1, N))
            span = ju - j                            # swap/update columns j..ju
            for c in range(0, KV + 1):
                sel = (jp != 1) & (c <= span)
                if not np.any(sel):
                    continue
                s = sidx[sel]
                ra = KV + jp[sel] - c - 1
                rb = KV - c
                tmp = AB[s, ra, j - 1 + c].copy()
                AB[s, ra, j - 1 + c] = AB[s, rb, j - 1 + c]
                AB[s, rb, j - 1 + c] = tmp
            if km > 0:
                rcp = 1.0 / AB[:, KV, j - 1]
                AB[:, KV + 1:KV + 1 + km, j - 1] = AB[:, KV + 1:KV + 1 + km, j - 1] * rcp[:, None]
                for c in range(1, KV + 1):
                    sel = c <= span
                    if not np.any(sel):
                        continue
                    y = AB[:, KV - c, j - 1 + c]       # AB(KV+1-cc, J+cc)
                    temp = np.where(sel, -y, 0.0)
                    for ii in range(1, km + 1):
                        r = KV + ii - c                 # AB(KV+1+ii-cc, J+cc)
                        upd = AB[:, r, j - 1 + c] + AB[:, KV + ii, j - 1] * temp
                        AB[:, r, j - 1 + c] = np.where(sel & (y != 0.0), upd, AB[:, r, j - 1 + c])
        return ipiv

    @staticmethod
    def dgbtrs(AB, ipiv, b):
        """LAPACK DGBTRS('N', N, KL=4, KU=4, 1 rhs) for each system s on the
        real vector b[:, s] (n, s).  Returns the solution."""
        N, S = b.shape
        x = b.copy()
        KD = 2 * KL + 1
        sidx = np.arange(S)
        for j in range(1, N):
            lm = min(KL, N - j)
            l = ipiv[:, j - 1] - 1
            bj = x[j - 1].copy()
            bl = x[l, sidx].copy()
            swap = l != (j - 1)
            x[j - 1] = np.where(swap, bl, bj)
            x[l, sidx] = np.where(swap, bj, bl)
            temp = -x[j - 1]
            for i in range(1, lm + 1):
                x[j - 1 + i] = x[j - 1 + i] + AB[:, KD + i - 1, j - 1] * temp
        for j in range(N, 0, -1):
            xj = x[j - 1] / AB[:, KD - 1, j - 1]
            nz = x[j - 1] != 0.0
            x[j - 1] = np.where(nz, xj, x[j - 1])
            temp = x[j - 1]
            for i in range(j - 1, max(1, j - 2 * KL) - 1, -1):
                upd = x[i - 1] - temp * AB[:, KD - 1 + i - j, j - 1]
                x[i - 1] = np.where(nz, upd, x[i - 1])
        return x

    # -- timestep.f90 ----------------------------------------------------------
    def _d(self, PM):
        """Diagonal term d(n, h) of tim_mesh_init/tim_lumesh_init
        (timestep.f90:54-55, 76-77)."""
        me = self.mesh
        m = self.harm.m.astype(float)[None, :]
        k = self.harm.k.astype(float)[None, :]
        r2 = me.rp[-2][:, None]
        Mp = float(self.Mp)
        d = -((((r2 * Mp) * m) * Mp) * m) - (((self.alpha * k) * self.alpha) * k)
        if PM != 0:
            d = (d - r2) - ((((2.0 * PM) * Mp) * m) * r2)
        return d

    def mesh_init(self, PM, c1, c2):
        """tim_mesh_init (timestep.f90:46-60): band (LDA, N, H)."""
        A = np.repeat((c2 * self.mesh.radLap)[:, :, None], self.H, axis=2)
        d = self._d(PM)
        A[KL] = (A[KL] + c2 * d) + c1
        return A

    def lumesh_init(self, PM, BC, c1, c2):
        """tim_lumesh_init (timestep.f90:67-96): returns (AB[h,13,N], ipiv)."""
        N, H = self.N, self.H
        me = self.mesh
        AB = np.zeros((H, LDAB, N))
        AB[:, KL:, :] = (c2 * me.radLap)[None, :, :]
        d = self._d(PM)
        AB[:, 2 * KL, :] = (AB[:, 2 * KL, :] + c2 * d.T) + c1
        S = (self.harm.m * self.Mp + abs(PM)) % 2
        for j in range(1, KL + 2):
            AB[:, 2 * KL + 1 - j, j - 1] = me.dr0[j - 1, 1 - S]
        for j in range(N - KL, N + 1):
            AB[:, 2 * KL + N - j, j - 1] = me.dr1[KL - N + j, BC]
        skip = np.zeros(H, dtype=bool)
        if BC == 1:
            skip = (self.harm.k == 0) & (self.harm.m == 0)
        ipiv = np.tile(np.arange(1, N + 1), (H, 1))
        live = ~skip
        AB_live = AB[live].copy()
        ipiv[live] = self.dgbtf2(AB_live)
        AB[live] = AB_live
        return AB, ipiv, skip

    def lumesh_invert(self, L, b):
        """tim_lumesh_invert (timestep.f90:102-123): Re then Im."""
        AB, ipiv, skip = L
        re = self.dgbtrs(AB, ipiv, b.real.copy())
        im = self.dgbtrs(AB, ipiv, b.imag.copy())
        x = re + 1j * im
        x[:, skip] = 0.0
        return x

    @staticmethod
    def meshmult_lt(A, b, c):
        """tim_meshmult (timestep.f90:130-154): d = A b + c with A per harmonic."""
        N = b.shape[0]
        d = c.copy()
        for n in range(N):
            acc = c[n].copy()
            for j in range(max(0, n - KL), min(n + KL, N - 1) + 1):
                acc = acc + A[KL + n - j, j] * b[j]
            d[n] = acc
        return d

    @staticmethod
    def zerobc(a):
        """tim_zerobc (timestep.f90:160-169)."""
        a[0] = 0.0
        a[-1] = 0.0

    # -- velocity.f90 ----------------------------------------------------------
    @staticmethod
    def rt2pm(r, t):
        """vel_rt2pm (velocity.f90:80-95)."""
        up = (r.real - t.imag) + 1j * (r.imag + t.real)
        um = (r.real + t.imag) + 1j * (r.imag - t.real)
        return up, um

    @staticmethod
    def pm2rt(up, um):
        """vel_pm2rt (velocity.f90:98-114)."""
        r = (0.5 * (up.real + um.real)) + 1j * (0.5 * (up.imag + um.imag))
        t = (0.5 * (up.imag - um.imag)) + 1j * (-0.5 * (up.real - um.real))
        return r, t

    def matrices(self):
        """vel_matrices (velocity.f90:120-146)."""
        if not hasattr(self, "LNp"):
            self.LNp = self.lumesh_init(0, 1, 0.0, 1.0)
        d1 = 1.0 / self.dt
        d2 = -self.implicit * self.nu / self.Re
        self.LDp = self.lumesh_init(1, 0, d1, d2)
        self.LDm = self.lumesh_init(-1, 0, d1, d2)
        self.LDz = self.lumesh_init(0, 0, d1, d2)
        d1 = 1.0 / self.dt
        d2 = (1.0 - self.implicit) * self.nu / self.Re
        self.Ltp = self.mesh_init(1, d1, d2)
        self.Ltm = self.mesh_init(-1, d1, d2)
        self.Ltz = self.mesh_init(0, d1, d2)
        self.adjust_flux_init()
        self.adjppe_init()

    def adjust_flux_init(self):
        """vel_adjustFlux(0) (velocity.f90:160-168)."""
        if not self.const_flux:
            return
        N = self.N
        Ui = np.ones(N)
        Ui[0] = 0.0
        Ui[-1] = 0.0
        AB, ipiv, _ = self.LDz
        self.Ui = self.dgbtrs(AB[0:1], ipiv[0:1], Ui[:, None])[:, 0]
        s = 0.0
        for n in range(N):
            s = s + self.Ui[n] * self.mesh.intrdr[n]
        self.flux_d1 = 2.0 * s
        if self.flux_d1 == 0.0:
            raise ArithmeticError("vel_adjustFlux: err 2")

    def adjust_flux(self, uz):
        """vel_adjustFlux(1) (velocity.f90:170-175)."""
        if not self.const_flux:
            return
        s = 0.0
        for n in range(self.N):
            s = s + uz[n, 0].real * self.mesh.intrdr[n]
        d2 = 2.0 * s
        d3 = -d2 / self.flux_d1
        uz[:, 0] = (uz[:, 0].real + d3 * self.Ui) + 1j * uz[:, 0].imag
        self.Pr0 = self.Pr0 + (d3 * self.Re) / 4.0

    def evalBC(self, up, um, uz):
        """vel_evalBC (velocity.f90:262-298): B[4, h] complex."""
        N = self.N
        n = N - 1
        dr1c = self.mesh.dr1[:, 1]
        urRe = 0.5 * (up[n].real + um[n].real)
        urIm = 0.5 * (up[n].imag + um[n].imag)
        utRe = 0.5 * (up[n].imag - um[n].imag)
        utIm = -0.5 * (up[n].real - um[n].real)
        uzRe = uz[n].real
        uzIm = uz[n].imag
        sre = up[N - KL - 1:N].real + um[N - KL - 1:N].real
        sim = up[N - KL - 1:N].imag + um[N - KL - 1:N].imag
        drRe = np.zeros(self.H)
        drIm = np.zeros(self.H)
        for i in range(KL + 1):
            drRe = drRe + dr1c[i] * sre[i]
            drIm = drIm + dr1c[i] * sim[i]
        drRe = 0.5 * drRe
        drIm = 0.5 * drIm
        BRe = np.zeros((4, self.H))
        BIm = np.zeros((4, self.H))
        BRe[0] = up[n].real
        BIm[0] = up[n].imag
        BRe[1] = um[n].real
        BIm[1] = um[n].imag
        BRe[2] = -uzIm
        BIm[2] = uzRe
        d = self.mesh.rp[-1][n]
        m_ = (self.harm.m * self.Mp).astype(float)
        a_ = self.k1a1
        BRe[3] = ((urRe * d + drRe) - (d * m_) * utIm) - a_ * uzIm
        BIm[3] = ((urIm * d + drIm) + (d * m_) * utRe) + a_ * uzRe
        h00 = (self.harm.k == 0) & (self.harm.m == 0)
        BRe[:, h00] = 0.0
        BIm[:, h00] = 0.0
        return BRe, BIm

    def adjppe_init(self):
        """vel_adjPPE(0) (velocity.f90:191-225): influence vectors U and the
        inverted 4x4 influence matrices."""
        N, H = self.N, self.H
        Uv = np.zeros((N, H, 6))
        A = np.zeros((H, 4, 4))
        for j in range(1, 5):
            c1 = self.zeros()
            c2 = self.zeros()
            c3 = self.zeros()
            if j == 1:
                c1[N - 1, :] = 1.0
                c1 = self.lumesh_invert(self.LDp, c1)
                Uv[:, :, 0] = c1.real
            elif j == 2:
                c2[N - 1, :] = 1.0
                c2 = self.lumesh_invert(self.LDm, c2)
                Uv[:, :, 1] = c2.real
            elif j == 3:
                c3[N - 1, :] = 1j
                c3 = self.lumesh_invert(self.LDz, c3)
                Uv[:, :, 2] = c3.imag
            else:
                c1[N - 1, :] = -1.0
                c1 = self.lumesh_invert(self.LNp, c1)
                c1, c2, c3 = self.grad(c1)
                c1, c2 = self.rt2pm(c1, c2)
                Uv[:, :, 3] = c1.real
                Uv[:, :, 4] = c2.real
                Uv[:, :, 5] = c3.imag
            BRe, _ = self.evalBC(c1, c2, c3)
            A[:, :, j - 1] = BRe.T
        Ainv = np.zeros_like(A)
        for h in range(H):
            if self.harm.k[h] == 0 and self.harm.m[h] == 0:
                Ainv[h] = A[h]
                continue
            Ainv[h] = np.array(mat_invert(A[h].tolist()))
        self.Uinf = Uv
        self.Ainf = Ainv

    def adjppe_project(self, ur, ut, uz):
        """vel_adjPPE(1) (velocity.f90:227-236) on the RHS held in (ur=u+,
        ut=u-, uz)."""
        c1, c2 = self.pm2rt(ur, ut)
        c1 = self.div(c1, c2, uz)
        self.zerobc(c1)
        c1 = self.lumesh_invert(self.LNp, c1)
        c1, c2, c3 = self.grad(c1)
        c1, c2 = self.rt2pm(c1, c2)
        return ur - c1, ut - c2, uz - c3

    def adjppe_correct(self, ur, ut, uz):
        """vel_adjPPE(2) (velocity.f90:238-256)."""
        BRe, BIm = self.evalBC(ur, ut, uz)
        A = self.Ainf
        aR = np.zeros((4, self.H))
        aI = np.zeros((4, self.H))
        for i in range(4):
            sR = np.zeros(self.H)
            sI = np.zeros(self.H)
            for j in range(4):
                sR = sR + A[:, i, j] * BRe[j]
                sI = sI + A[:, i, j] * BIm[j]
            aR[i] = -sR
            aI[i] = -sI
        h00 = (self.harm.k == 0) & (self.harm.m == 0)
        aR[:, h00] = 0.0
        aI[:, h00] = 0.0
        U = self.Uinf
        live = ~h00
        def upd(x, re_add, im_add):
            re = np.where(live, x.real + re_add, x.real)
            im = np.where(live, x.imag + im_add, x.imag)
            return re + 1j * im
        ur = upd(ur, aR[0] * U[:, :, 0], aI[0] * U[:, :, 0])
        ut = upd(ut, aR[1] * U[:, :, 1], aI[1] * U[:, :, 1])
        uz = upd(uz, -(aI[2] * U[:, :, 2]), aR[2] * U[:, :, 2])
        ur = upd(ur, aR[3] * U[:, :, 3], aI[3] * U[:, :, 3])
        ut = upd(ut, aR[3] * U[:, :, 4], aI[3] * U[:, :, 4])
        uz = upd(uz, -(aI[3] * U[:, :, 5]), aR[3] * U[:, :, 5])
        return ur, ut, uz

    # -- step-level routines ---------------------------------------------------
    def set_state(self, ur, ut, uz):
        self.ur, self.ut, self.uz = ur.copy(), ut.copy(), uz.copy()

    def imposesym(self):
        """vel_imposesym (velocity.f90:61-74): mode 0, then the symmetries
        in self.symm (1 mirror, 2 shift-reflect, 3 shift-rotate; the
        b_mirrorsym, b_shiftrefl, b_shiftrott flags of parameters.f90:26-28)."""
        self.imposesym0(self.ur, self.ut, self.uz)
        for i in sorted(getattr(self, "symm", ())):
            self.ur, self.ut, self.uz = self.imposesym_mode(i, self.ur, self.ut, self.uz)

    def imposesym_mode(self, i, r, t, z):
        """var_imposesym(i) (variables.f90:538-618) for i = 1, 2, 3; returns
        new fields.  1, 2: average with the image of the reflection, which
        maps harmonic (k, m) to (-k, m) (to itself for m = 0) with factors
        a = 1 (mirror) or (-1)^k (shift-reflect), b = -1 on the imaginary
        part for m > 0, and -1 on u_theta; 3: zero the harmonics with k + m
        odd.  Mode 0 (harmonic 0) is applied last."""
        hk, hm = self.harm.k, self.harm.m
        r, t, z = r.copy(), t.copy(), z.copy()
        if i in (1, 2):
            where = {(int(k), int(m)): h for h, (k, m) in enumerate(zip(hk, hm))}
            img = np.array([h if m == 0 else where[(-int(k), int(m))] for h, (k, m) in enumerate(zip(hk, hm))])
            a = np.ones(hk.size)
            if i == 2:
                a[(hk % 2) != 0] = -1.0
            b = np.where(hm == 0, 1.0, -1.0)

            def image(f, sign):
                g = f[:, img]
                return (g.real * (sign * a)) + 1j * (g.imag * ((sign * a) * b))

            c1, c2, c3 = image(r, 1.0), image(t, -1.0), image(z, 1.0)
            r = ((r.real + c1.real) / 2.0) + 1j * ((r.imag + c1.imag) / 2.0)
            t = ((t.real + c2.real) / 2.0) + 1j * ((t.imag + c2.imag) / 2.0)
            z = ((z.real + c3.real) / 2.0) + 1j * ((z.imag + c3.imag) / 2.0)
        if i == 3:
            odd = ((hk + hm) % 2) != 0
            r[:, odd] = 0.0
            t[:, odd] = 0.0
            z[:, odd] = 0.0
        self.imposesym0(r, t, z)
        return r, t, z

    def transform(self):
        """vel_transform (velocity.f90:304-317)."""
        self.p_r = self.spec2phys(self.ur)
        self.p_t = self.spec2phys(self.ut)
        self.p_z = self.spec2phys(self.uz)
        c1, c2, c3 = self.curl(self.ur, self.ut, self.uz)
        self.p_cr = self.spec2phys(c1)
        self.p_ct = self.spec2phys(c2)
        self.p_cz = self.spec2phys(c3)

    def nonlinear(self):
        """vel_nonlinear + vel_addHPF (velocity.f90:323-361), then the
        var_null(2) hook of main.f90:23,38 (the Coriolis plugin when
        self.coriolis = (Ek0, phi))."""
        p = self.p_t * self.p_cz - self.p_z * self.p_ct
        self.Nr = self.phys2spec(p)
        p = self.p_z * self.p_cr - self.p_r * self.p_cz
        self.Nt = self.phys2spec(p)
        p = self.p_r * self.p_ct - self.p_t * self.p_cr
        self.Nz = self.phys2spec(p)
        self.add_hpf()
        if getattr(self, "coriolis", None) is not None:
            self.add_coriolis(*self.coriolis)
        if getattr(self, "nonnewtonian", None) is not None:
            self.add_nonnewtonian(**self.nonnewtonian)
        if getattr(self, "les", None) is not None:
            self.add_les(**self.les)

    def _total_velocity_gradient(self):
        """get_nu's copies u, with vel_U on the (0, 0) mode of u_z, and the
        gradient tensor A(i, j) on the physical grid (nonnewtonian.f90:108-147,
        LES.f90:114-150)."""
        ur, ut, uz = self.ur.copy(), self.ut.copy(), self.uz.copy()
        uz[:, 0] = (uz[:, 0].real + self.U) + 1j * uz[:, 0].imag
        dr1 = self.mesh.dr[1]
        rows = [[self.meshmult(dr1, ur), self.meshmult(dr1, ut), self.meshmult(dr1, uz)]]
        ri = self.mesh.rp[-1][:, None]
        mm = (self.harm.m * self.Mp).astype(float)[None, :]
        rows.append([ri * (-(ur.imag * mm) - ut.real) + 1j * (ri * (ur.real * mm - ut.imag)),
                     ri * (-(ut.imag * mm) + ur.real) + 1j * (ri * (ut.real * mm + ur.imag)),
                     ri * (-(uz.imag * mm)) + 1j * (ri * (uz.real * mm))])
        d1 = self.alpha * self.harm.k.astype(float)[None, :]
        rows.append([-(a.imag * d1) + 1j * (a.real * d1) for a in (ur, ut, uz)])
        return (ur, ut, uz), [[self.spec2phys(rows[i][j]) for j in range(3)] for i in range(3)]

    def les_lengths(self, filter):
        """Rdth, dz, Delta_r and the Smagorinsky width Delta (LES.f90:81-89, 199-206)."""
        r = self.mesh.r
        N = self.N
        Rdth = (2.0 * D_PI / float(self.Mp * 3 * self.M)) * (3.0 / 2.0)
        dz = (2.0 * D_PI / (self.alpha * float(3 * self.K))) * (3.0 / 2.0)
        Dr = np.zeros(N)
        Dr[0] = (r[1] - 0.0) / 2.0
        for n in range(1, N - 1):
            Dr[n] = (r[n + 1] - r[n - 1]) / 2.0
        Dr[N - 1] = 1.0 - r[N - 2]
        if filter == 1:
            Delta = (Dr * Rdth * dz) ** (1.0 / 3.0)
        elif filter == 2:
            Delta = (Dr * r * Rdth * dz) ** (1.0 / 3.0)
        else:
            Delta = np.sqrt((Dr ** 2 + (r * Rdth) ** 2 + dz ** 2) / 3.0)
        return Rdth, dz, Dr, Delta

    def add_les(self, model, C_s, filter, A_plus):
        """util(2) of the LES plugin (LES.f90:100-320): the eddy viscosity of
        Smagorinsky (model 1) or Vreman (model 2), update_nu_r at the
        predictor evaluation (tim_it = 0: vel_nu and vel_matrices) and
        update_nonlin_terms."""
        (ur, ut, uz), A = self._total_velocity_gradient()
        S = [[A[i][j] for j in range(3)] for i in range(3)]
        for i, j in ((0, 1), (0, 2), (1, 2)):
            S[i][j] = 0.5 * (A[i][j] + A[j][i])
            S[j][i] = S[i][j]
        sr = S[0][0] ** 2 + S[1][1] ** 2 + S[2][2] ** 2 + 2.0 * (S[0][1] ** 2 + S[0][2] ** 2 + S[1][2] ** 2)
        sr = np.sqrt(2.0 * sr)
        N = self.N
        w = self.mesh.dr1[:, 1]                      # mes_D%dr1(:,1)
        d1 = 0.0
        for i in range(KL + 1):
            d1 = d1 + self.uz[N - 1 - KL + i, 0].real * w[i]
        u_tau = math.sqrt(abs((d1 - 2.0) / self.Re))
        y_plus = (self.Re * u_tau) * (1.0 - self.mesh.r)
        Rdth, dz, Dr, Delta = self.les_lengths(filter)
        with np.errstate(divide="ignore", invalid="ignore"):
            vanDr = 1.0 - np.exp(-((y_plus / A_plus) ** 2))
        if model == 1:
            nu = (vanDr * (C_s * Delta) ** 2)[:, None, None] * sr
        else:
            AA = np.zeros_like(sr)
            for i in range(3):
                for j in range(3):
                    AA = AA + A[i][j] ** 2
            dl = [Dr[:, None, None], (self.mesh.r * Rdth if filter == 2 else np.full(N, Rdth))[:, None, None], dz]
            Bm = [[None] * 3 for _ in range(3)]
            for i in range(3):
                for j in range(3):
                    b = np.zeros_like(sr)
                    for m in range(3):
                        b = b + (dl[m] ** 2) * A[m][i] * A[m][j]
                    Bm[i][j] = b
            P = (Bm[0][0] * Bm[1][1] - Bm[0][1] ** 2 + Bm[0][0] * Bm[2][2] - Bm[0][2] ** 2
                 + Bm[1][1] * Bm[2][2] - Bm[1][2] ** 2)
            P = np.where(AA < 1e-12, 0.0, P)
            P = np.maximum(P, 0.0)
            AA = np.maximum(AA, 1e-12)
            nu = C_s * np.sqrt(P / AA)
            if A_plus > 0.0:
                dd = 0.03 / C_s
                dd = dd + (1.0 - dd) * vanDr
                nu = dd[:, None, None] * nu
        nuc = self.phys2spec(nu)
        self.les_nu, self.les_nuc = nu, nuc
        if getattr(self, "it", 0) == 0:              # update_nu_r (LES.f90:264-273)
            d1 = 1.0 + self.Re * nu.max()
            d1 = 1.10 * d1
            if self.nu < 0.93 * d1 or self.nu > 1.07 * d1:
                self.nu = d1
                self.matrices()
        g = [self.spec2phys(a) for a in self.grad(nuc)]
        q = [self.spec2phys(a) for a in self.curl(*self.curl(ur, ut, uz))]
        nu_ = self.Re * nu + (1.0 - self.nu)
        out = []
        for i in range(3):
            qi = (1.0 / self.Re) * nu_ * (-q[i])
            if i == 2:
                qi = qi + nu * (-4.0)
            for j in range(3):
                qi = qi + g[j] * S[j][i] * 2.0
            out.append(self.phys2spec(qi))
        self.Nr = self.Nr + out[0]
        self.Nt = self.Nt + out[1]
        self.Nz = self.Nz + out[2]

    def add_nonnewtonian(self, rho, radius, mu_0, mu_inf, lam_CY, n_CY, a_CY):
        """util(2) of the non-Newtonian plugin with the Carreau-Yasuda model:
        get_nu (nonnewtonian.f90:93-193) and update_nonlin_terms (:203-244),
        vel_nu = 1.  The velocity includes the base flow U on the (0, 0) mode."""
        ur, ut, uz = self.ur.copy(), self.ut.copy(), self.uz.copy()
        uz[:, 0] = (uz[:, 0].real + self.U) + 1j * uz[:, 0].imag
        lam_ = self.Re * lam_CY * mu_0 / (rho * radius ** 2)
        dr1 = self.mesh.dr[1]
        rows = [[self.meshmult(dr1, ur), self.meshmult(dr1, ut), self.meshmult(dr1, uz)]]
        ri = self.mesh.rp[-1][:, None]
        mm = (self.harm.m * self.Mp).astype(float)[None, :]
        rows.append([ri * (-(ur.imag * mm) - ut.real) + 1j * (ri * (ur.real * mm - ut.imag)),
                     ri * (-(ut.imag * mm) + ur.real) + 1j * (ri * (ut.real * mm + ur.imag)),
                     ri * (-(uz.imag * mm)) + 1j * (ri * (uz.real * mm))])
        d1 = self.alpha * self.harm.k.astype(float)[None, :]
        rows.append([-(a.imag * d1) + 1j * (a.real * d1) for a in (ur, ut, uz)])
        S = [[self.spec2phys(rows[i][j]) for j in range(3)] for i in range(3)]
        for i, j in ((0, 1), (0, 2), (1, 2)):
            S[i][j] = 0.5 * (S[i][j] + S[j][i])
            S[j][i] = S[i][j]
        sr = S[0][0] ** 2 + S[1][1] ** 2 + S[2][2] ** 2 + 2.0 * (S[0][1] ** 2 + S[0][2] ** 2 + S[1][2] ** 2)
        sr = np.sqrt(2.0 * sr)
        d1 = mu_inf / mu_0
        d2 = -n_CY / a_CY
        nu = d1 + (1.0 - d1) * (1.0 + (lam_ * sr) ** a_CY) ** d2
        nuc = self.phys2spec(nu)
        self.nn_nu = nu
        self.nn_nuc = nuc
        g = [self.spec2phys(a) for a in self.grad(nuc)]
        q = [self.spec2phys(a) for a in self.curl(*self.curl(ur, ut, uz))]
        nu_ = nu - 1.0
        out = []
        for i in range(3):
            qi = nu_ * (-q[i])
            for j in range(3):
                qi = qi + g[j] * S[j][i] * 2.0
            if i == 2:
                qi = qi + (1.0 - 1.0) * (-4.0)
            out.append(self.phys2spec(qi / self.Re))
        self.Nr = self.Nr + out[0]
        self.Nt = self.Nt + out[1]
        self.Nz = self.Nz + out[2]

    def add_coriolis(self, Ek0, phi):
        """util(2) of Coriolis.f90:54-71: N += -1/(4 Ro) Omega^ x (U + u) on
        the physical grid, Ro = Re Ek0, Omega^ = (sin phi cos theta,
        -sin phi sin theta, cos phi); each component transformed and added."""
        if self.Mp != 1:
            raise ValueError("add_coriolis: i_Mp/=1")
        T = 3 * self.M
        d1 = 2.0 * D_PI / float(T)
        th = d1 * np.arange(T, dtype=np.float64)
        Or = math.sin(phi) * np.cos(th)
        Ot = -math.sin(phi) * np.sin(th)
        Oz = math.cos(phi)
        Ro = self.Re * Ek0
        d1 = -1.0 / (4.0 * Ro)
        U = self.U[:, None, None]
        uz = self.p_z + U
        Nr = d1 * (Ot[None, None, :] * uz - Oz * self.p_t)
        Nt = d1 * (Oz * self.p_r - Or[None, None, :] * uz)
        Nz = d1 * (Or[None, None, :] * self.p_t - Ot[None, None, :] * self.p_r)
        self.Nr = self.Nr + self.phys2spec(Nr)
        self.Nt = self.Nt + self.phys2spec(Nt)
        self.Nz = self.Nz + self.phys2spec(Nz)

    def add_hpf(self):
        """vel_addHPF (velocity.f90:338-361)."""
        b = -self.Up
        a = (self.alpha * self.harm.k.astype(float))[None, :] * self.U[:, None]
        ur, ut, uz = self.ur, self.ut, self.uz
        self.Nr = (self.Nr.real + a * ur.imag) + 1j * (self.Nr.imag - a * ur.real)
        self.Nt = (self.Nt.real + a * ut.imag) + 1j * (self.Nt.imag - a * ut.real)
        self.Nz = ((self.Nz.real + a * uz.imag) + b[:, None] * ur.real) + \
            1j * ((self.Nz.imag - a * uz.real) + b[:, None] * ur.imag)
        if self.const_flux:
            s = 0.0
            for i in range(KL + 1):
                s = s + uz[self.N - KL - 1 + i, 0].real * self.mesh.dr1[i, 1]
            self.Pr0 = s / (-2.0)
            self.Nz[:, 0] = (self.Nz[:, 0].real + (4.0 * self.Pr0) / self.Re) + 1j * self.Nz[:, 0].imag

    def vstep(self):
        """vel_step (velocity.f90:367-394)."""
        c1, c2 = self.rt2pm(self.Nr, self.Nt)
        up, um = self.rt2pm(self.ur_, self.ut_)
        up = self.meshmult_lt(self.Ltp, up, c1)
        um = self.meshmult_lt(self.Ltm, um, c2)
        uz = self.meshmult_lt(self.Ltz, self.uz_, self.Nz)
        up, um, uz = self.adjppe_project(up, um, uz)
        self.zerobc(up)
        self.zerobc(um)
        self.zerobc(uz)
        up = self.lumesh_invert(self.LDp, up)
        um = self.lumesh_invert(self.LDm, um)
        uz = self.lumesh_invert(self.LDz, uz)
        up, um, uz = self.adjppe_correct(up, um, uz)
        ur, ut = self.pm2rt(up, um)
        self.adjust_flux(uz)
        ur[:, 0] = 0.0
        ut[:, 0] = ut[:, 0].real
        uz[:, 0] = uz[:, 0].real
        self.ur, self.ut, self.uz = ur, ut, uz

    def predictor(self):
        """vel_predictor (velocity.f90:400-410)."""
        self.ur_, self.ut_, self.uz_ = self.ur.copy(), self.ut.copy(), self.uz.copy()
        self.Nr_, self.Nt_, self.Nz_ = self.Nr.copy(), self.Nt.copy(), self.Nz.copy()
        self.vstep()

    def corrector(self):
        """vel_corrector (velocity.f90:416-432) with tim_nlincorr and
        tim_measurecorr (timestep.f90:174-207)."""
        r, t, z = self.ur.copy(), self.ut.copy(), self.uz.copy()
        d1 = self.implicit
        d2 = 1.0 - self.implicit
        self.Nr = (d1 * self.Nr.real + d2 * self.Nr_.real) + 1j * (d1 * self.Nr.imag + d2 * self.Nr_.imag)
        self.Nt = (d1 * self.Nt.real + d2 * self.Nt_.real) + 1j * (d1 * self.Nt.imag + d2 * self.Nt_.imag)
        self.Nz = (d1 * self.Nz.real + d2 * self.Nz_.real) + 1j * (d1 * self.Nz.imag + d2 * self.Nz_.imag)
        self.vstep()
        dterr = 0.0
        for a, b in ((self.ur, r), (self.ut, t), (self.uz, z)):
            d = a - b
            v = d.real * d.real + d.imag * d.imag
            dterr = max(dterr, float(v.max()))
        self.dterr = max(getattr(self, "dterr", 0.0), math.sqrt(dterr))

    def check_cgce(self):
        """tim_check_cgce (timestep.f90:214-253) with d_timestep > 0 and
        tim_pcn = -1.  Returns 'ok' or an error string."""
        status = "ok"
        if self.it == 1:
            self.corr_dt = self.dt * math.sqrt(self.dterr_tol / self.dterr) if self.dterr > 0 else math.inf
            self.lasterr = 1e99
        if self.dt < 1e-9 and self.step > 30:
            status, self.it = "dt->0", 0
        elif self.it > 10:
            status, self.it = "too many its", 0
        elif self.dterr > self.lasterr:
            if self.dterr > 2.0 * self.dterr_tol:
                status = "increasing error"
            else:
                self.corr_dt = self.dt / (10.0 * 0.5)     # d_courant = 0.5
            self.it = 0
        elif self.dterr > self.dterr_tol:
            self.lasterr = self.dterr
            self.it += 1
        else:
            self.it = 0
        self.dterr = 0.0
        return status

    def maxtstep(self):
        """vel_maxtstep (velocity.f90:438-481) on one rank: the CFL limits of
        u_r (radial spacing), u_theta (2 pi / (3M Mp)) and u_z plus the base
        flow 1 - r^2 (2 pi / (alpha 3K)), from the physical fields of the last
        transform; tim_cfl_dir is the last direction whose limit equals the
        minimum."""
        N, r = self.N, self.mesh.r
        dts = [1e99, 1e99, 1e99]
        for n in range(N):
            if n == 0:
                d = r[1] - r[0]
            elif n == N - 1:
                d = r[N - 1] - r[N - 2]
            else:
                d = min(r[n] - r[n - 1], r[n + 1] - r[n])
            mx = float(np.abs(self.p_r[n]).max())
            if mx != 0.0:
                dts[0] = min(dts[0], d / mx)
            d = 2.0 * math.pi / float(3 * self.M * self.Mp)
            mx = float(np.abs(self.p_t[n]).max())
            if mx != 0.0:
                dts[1] = min(dts[1], d / mx)
            d = 2.0 * math.pi / (self.alpha * float(3 * self.K))
            mx = float(np.abs(self.p_z[n] + (1.0 - r[n] * r[n])).max())
            if mx != 0.0:
                dts[2] = min(dts[2], d / mx)
        self.cfl_dt = min(dts)
        for i in range(3):
            if self.cfl_dt == dts[i]:
                self.cfl_dir = i + 1
        return self.cfl_dt, self.cfl_dir

    def new_tstep(self, maxdt=1e99, courant=0.5):
        """tim_new_tstep (timestep.f90:259-278): grow by at most 11% (not
        again within 99 steps of a growth), cap by the CFL and corrector
        limits, change only beyond 5% / 10%.  Returns tim_new_dt."""
        dt = min(self.dt * 1.11, maxdt)
        if self.step == 0 and self.corr_dt == 0.0:
            dt = min(dt, self.cfl_dt * 0.1)
        if self.cfl_dt > 0.0:
            dt = min(dt, self.cfl_dt * courant)
        if self.corr_dt > 0.0:
            dt = min(dt, self.corr_dt * 0.95)
        self._ntc -= 1
        new = (dt < self.dt * 0.95 or (dt > self.dt * 1.10 and self._ntc < 0)
               or (dt == maxdt and self.dt < dt and self._ntc < 0))
        if new:
            if dt >= self.dt:
                self._ntc = 99
            self.dt = dt
        return new

    def diagnostics(self):
        """io_write_energy (io.f90:501-524) and io_write_friction
        (io.f90:627-639)."""
        E, Ek, Em = self.norm(self.ur)
        Ek0, Em0 = Ek[0], Em[0]
        E_, Ek, Em = self.norm(self.ut)
        E, Ek0, Em0 = E + E_, Ek0 + Ek[0], Em0 + Em[0]
        E_, Ek, Em = self.norm(self.uz)
        E, Ek0, Em0 = E + E_, Ek0 + Ek[0], Em0 + Em[0]
        uz0 = self.uz[:, 0].real
        s = 0.0
        for i in range(KL + 1):
            s = s + uz0[i] * self.mesh.dr0[i, 0]
        Uc = 1.0 + s
        s = 0.0
        for i in range(KL + 1):
            s = s + uz0[self.N - KL - 1 + i] * self.mesh.dr1[i, 1]
        Ufr = math.sqrt(abs((s - 2.0) / self.Re))
        return dict(t=self.t, E=E, Ek0=Ek0, Em0=Em0, Pr0=self.Pr0, Uc=Uc, Ufr=Ufr)

    def coll_shift(self, a, dth, dz):
        """var_coll_shift (variables.f90:487-528): harmonic (k, m) times
        e^{-i alpha k dz}, then e^{-i m Mp dth}, i.e. the field u(theta - dth,
        z - dz); each rotation as Re c + Im s, Im c - Re s."""
        b = a.copy()
        for d, idx, f in ((dz, self.harm.k, self.alpha), (dth, self.harm.m, float(self.Mp))):
            if d == 0.0:
                continue
            arg = (d * idx.astype(float)) * f
            c, sn = np.cos(arg), np.sin(arg)
            re, im = b.real.copy(), b.imag.copy()
            b = (re * c + im * sn) + 1j * (im * c - re * sn)
        return b

    def steporbit(self, ndts, T, sth, sz):
        """steporbit (newton.f90:721-775) on the current state: for ndts /= 1,
        tim_dt = T/ndts and vel_matrices; ndts steps with one corrector each
        (imposesym, predictor, corrector, imposesym); for ndts /= 1 the result
        shifted back, var_coll_shift(-sth, -sz)."""
        if ndts != 1:
            self.dt = T / float(ndts)
            self.matrices()
        self.step = 0
        self.run(ndts, ncorr_fixed=1)
        self.imposesym()
        if ndts != 1:
            self.ur, self.ut, self.uz = (self.coll_shift(f, -sth, -sz) for f in (self.ur, self.ut, self.uz))

    def run(self, nsteps, ncorr_fixed=None, record=None, adaptive=False, maxdt=1e99, courant=0.5):
        """The loop of main.f90:18-46 (no I/O).  With ncorr_fixed the corrector
        runs exactly that many times per step (the steporbit schedule of
        newton.f90:743-787 uses one).  adaptive: the d_timestep < 0 block of
        main.f90:24-29 (vel_maxtstep, tim_new_tstep, vel_matrices on a
        change).  Returns the per-step series; adaptive runs add the
        timestep quantities of io_write_timestep (io.f90:645-657)."""
        series = []
        while self.step < nsteps:
            self.it = 0               # tim_it at the predictor's evaluation
            self.imposesym()
            self.transform()
            self.nonlinear()
            tq = None
            if adaptive:
                self.maxtstep()
                new = self.new_tstep(maxdt, courant)
                if new:
                    self.matrices()
                tq = dict(dt=self.dt, cfl_dt=self.cfl_dt, cfl_dir=self.cfl_dir, corr_dt=self.corr_dt,
                          new_dt=int(new))
            diag = self.diagnostics()
            if tq is not None:
                diag.update(tq)
            self.predictor()
            self.it = 1
            ncorr = 0
            self.dterr = 0.0
            while self.it != 0:
                self.transform()
                self.nonlinear()
                self.corrector()
                ncorr += 1
                last = self.dterr
                status = self.check_cgce()
                if ncorr_fixed is not None:
                    self.it = 0 if ncorr >= ncorr_fixed else self.it
                if status != "ok":
                    raise RuntimeError("tim_check_cgce: " + status)
            diag.update(step=self.step, ncorr=ncorr, dterr=last)
            series.append(diag)
            self.t = self.t + self.dt
            self.step += 1
            if record is not None:
                record(self)
        return series


# ---------------------------------------------------------------------------
# state exchange with the compiled reference (raw coll layout)
# ---------------------------------------------------------------------------
def read_coll3(path, N, H):
    """Read ur, ut, uz written as Re(1:N,0:H-1), Im(1:N,0:H-1) (Fortran order)."""
    a = np.fromfile(path, dtype="<f8")
    if a.size != 6 * N * H:
        raise ValueError(f"{path}: {a.size} doubles, expected {6 * N * H}")
    a = a.reshape(6, H, N)            # Fortran (N,H) column-major == C (H,N)
    out = []
    for c in range(3):
        out.append((a[2 * c] + 1j * a[2 * c + 1]).T.copy())
    return out


def write_coll3(path, ur, ut, uz):
    parts = []
    for f in (ur, ut, uz):
        parts.append(np.ascontiguousarray(f.real.T))
        parts.append(np.ascontiguousarray(f.imag.T))
    np.concatenate([p.ravel() for p in parts]).astype("<f8").tofile(path)


# ---------------------------------------------------------------------------
# io.f90: loading a state onto another grid
# ---------------------------------------------------------------------------
def io_interp_wts(xi, xo):
    """io_interp_wts (io.f90:238-264): for each output point, the first index
    l of its stencil and the first row of the inverted local Taylor matrix
    (the weights of the interpolating polynomial's value at xo)."""
    xi = [float(v) for v in xi]
    ni = len(xi)
    out = []
    for x in xo:
        j = 0
        while xi[j] < x - 1e-8 and j < ni - 1:
            j += 1
        lo = max(0, j - KL)
        hi = min(j + KL, ni - 1)
        nn = hi - lo + 1
        A = [[1.0] * nn for _ in range(nn)]
        for c in range(1, nn):
            for i in range(nn):
                A[i][c] = A[i][c - 1] * (xi[lo + i] - float(x)) / float(c)
        Ainv = mat_invert(A)
        out.append((lo, list(Ainv[0])))
    return out


def io_interp(wts, fi):
    """io_interp (io.f90:270-289): fo(n) = dot_product(A(1,1:nn,n), fi(l:r))."""
    fo = np.zeros(len(wts))
    for n, (lo, w) in enumerate(wts):
        s = 0.0
        for i, wi in enumerate(w):
            s = s + wi * float(fi[lo + i])
        fo[n] = s
    return fo


def io_load_coll(a_file, Kf, Mf, r_file, K, M, Mp, r, S):
    """io_load_coll (io.f90:171-232): the file field a_file (Nf, Hf) on
    harmonics (Kf, Mf) and radial points r_file mapped onto (K, M) and r.
    Harmonics are matched by (k, m); the file is interpolated when its radial
    points differ, continued across the axis as fn(1-i) = (-1)^(m Mp + S) fn(i).
    The file index follows the file's own harmonic order (io.f90:205 also
    counts harmonics that neither grid holds, which differs only for a file
    larger in one of K, M and smaller in the other)."""
    r_file = np.asarray(r_file, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    Nf = r_file.size
    interp = Nf != r.size or np.max(np.abs(r - r_file)) > 1e-8
    xi = np.concatenate([-r_file[KL - 1::-1], r_file])
    wts = io_interp_wts(xi, r) if interp else None
    hf = Harmonics(Kf, Mf, Mp)
    index_f = {(int(k), int(m)): i for i, (k, m) in enumerate(zip(hf.k, hf.m))}
    hg = Harmonics(K, M, Mp)
    out = np.zeros((r.size, hg.H), dtype=np.complex128)
    for h, (k, m) in enumerate(zip(hg.k, hg.m)):
        i = index_f.get((int(k), int(m)))
        if i is None:
            continue
        col = a_file[:, i]
        if not interp:
            out[:, h] = col
            continue
        sgn = -1.0 if (m * Mp + S) % 2 == 1 else 1.0
        parts = []
        for part in (col.real, col.imag):
            fn = np.concatenate([sgn * part[KL - 1::-1], part])
            parts.append(io_interp(wts, fn))
        out[:, h] = parts[0] + 1j * parts[1]
    return out
