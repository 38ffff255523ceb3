"""Newton-Krylov-hookstep search for (relative) periodic orbits over the
device time step: the client of newton.f90 with the solver of
NewtonHook.f90 and GMRESm.f90.

The solver (``NewtonHook``) is generic: it finds a root of f(x) = 0 with
GMRES(m) on the Jacobian action and Viswanath's hookstep in a trust region,
exactly as newtonhook / gmresm / hookstep do (NewtonHook.f90:48-179,
GMRESm.f90:27-153).  ``OrbitNewton`` is newton.f90's problem: the Newton
vector holds, per shooting segment, the theta shift, z shift and period
(scaled by scaleT) and the scaled velocity field (rtz2vec), f(x) = F(x) - x
with F = steporbit (newton.f90:374-393, 721-775), and the Jacobian action by
finite differences plus the constraints of no update along the shift
directions and the trajectory (multJ, newton.f90:399-456).

The time integrations run on the device through ``opf_steporbit``: the
segments of multiple shooting advance together in one context (one slot
each), so getrhs costs one batched integration rather than ms sequential
ones.  Any object with the ``Stepper`` interface can stand in for the device
(the tests use the oracle's CPU time step to check the device path).
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg

from .ic import energy, harmonic_km

D_PI = 3.1415926535897931


# ---------------------------------------------------------------------------
# NewtonHook.f90 / GMRESm.f90
# ---------------------------------------------------------------------------
def hookstep(h, j, beta, delta):
    """hookstep (GMRESm.f90:121-153): the y minimising |beta e1 - H y| with
    |y| <= delta (Viswanath 2008), from the SVD of the (j+1) x j Hessenberg
    matrix; returns y and the predicted residual |beta e1 - H y|."""
    a = h[:j + 1, :j]
    u, s, vt = np.linalg.svd(a, full_matrices=True)
    p = beta * u[0, :j]
    mu = max(s[j - 1] * s[j - 1] * 1e-6, 1e-99)
    qn = 1e99
    q = np.zeros(j)
    while qn > delta:
        mu = mu * 1.1
        q = p * s / (mu + s * s)
        qn = math.sqrt(float(np.dot(q, q)))
    y = q @ vt
    r = -(a @ y)
    r[0] = r[0] + beta
    return y, math.sqrt(float(np.dot(r, r)))


class Gmres:
    """gmresm (GMRESm.f90:27-115): GMRES(m) with modified Gram-Schmidt and
    least squares by LAPACK dgelsy on the Hessenberg matrix; keeps the Krylov
    basis, H, j and beta between calls so that info = 2 recomputes the
    hookstep for a new trust region without new products."""

    def __init__(self, m, n):
        self.m, self.n = m, n
        self.h = np.zeros((m + 1, m))
        self.v = np.zeros((m + 1, n))     # row i = Krylov vector i+1
        self.j = 0
        self.beta = 0.0

    def _combine(self, y, j):
        z = np.zeros(self.n)
        for i in range(j):
            z = z + y[i] * self.v[i]
        return z

    def solve(self, x, b, matvec, psolve, dotprd, res, delta, its, info):
        """Returns (x, res, delta, its, info) as gmresm leaves its inout
        arguments: info 0 converged, 1 breakdown (residual grew), 2 max its."""
        m = self.m
        if info == 2:
            y, delta = hookstep(self.h, self.j, self.beta, delta)
            z = psolve(self._combine(y, self.j))
            return z, res, delta, its, 0
        tol = res
        imx = its
        its = 0
        self.v[:] = 0.0
        while True:   # restart
            res_ = 1e99
            stgn = 1.0 - 1e-14
            beta = math.sqrt(dotprd(x, x))
            w = np.zeros(self.n) if beta == 0.0 else matvec(x)
            w = b - w
            beta = math.sqrt(dotprd(w, w))
            self.beta = beta
            self.v[0] = w / beta
            self.h[:] = 0.0
            for j in range(1, m + 1):
                self.j = j
                its += 1
                z = psolve(self.v[j - 1].copy())
                w = matvec(z)
                for i in range(j):
                    self.h[i, j - 1] = dotprd(w, self.v[i])
                    w = w - self.h[i, j - 1] * self.v[i]
                self.h[j, j - 1] = math.sqrt(dotprd(w, w))
                self.v[j] = w / self.h[j, j - 1]
                p = np.zeros(j + 1)
                p[0] = beta
                y = scipy.linalg.lstsq(self.h[:j + 1, :j], p, lapack_driver="gelsy")[0]
                p = -(self.h[:j + 1, :j] @ y)
                p[0] = p[0] + beta
                res = math.sqrt(float(np.dot(p, p)))
                done = res <= tol or its == imx or res > res_
                if done or j == m:
                    if delta > 0.0:
                        y, delta = hookstep(self.h, j, beta, delta)
                    z = psolve(self._combine(y, j))
                    x = x + z
                    if its == imx:
                        info = 2
                    if res > res_:
                        info = 1
                    if res <= tol:
                        info = 0
                    if done:
                        return x, res, delta, its, info
                    break   # j == m: restart from the new x
                res_ = res * stgn


class NewtonHook:
    """newtonhook (NewtonHook.f90:48-179) with the state of its module
    (new_x, new_fx, new_tol, new_del, new_nits, new_gits)."""

    def __init__(self, x0):
        self.new_x = np.array(x0, dtype=np.float64)
        self.new_fx = np.zeros_like(self.new_x)
        self.new_tol = 0.0
        self.new_del = 0.0
        self.new_nits = 0
        self.new_gits = 0
        self.log = []

    def _say(self, msg):
        self.log.append(msg)

    def solve(self, f, df, dfp, sub, dp, m, gtol, tol, delta, mndl, mxdl, nits, info=0):
        """f(x) -> f; df(x) -> Jacobian action about new_x; dfp(x) -> x
        (preconditioner); sub() after every accepted iteration; dp(a, b).
        Returns info: 0 converged, 2 max iterations, 3 trust region too small."""
        n = self.new_x.size
        gm = Gmres(m, n)
        self.new_nits = 0
        self.new_gits = 0
        self.new_del = delta
        mxdl_ = mxdl
        ginfo = info
        self.new_fx = f(self.new_x)
        self.new_tol = math.sqrt(dp(self.new_fx, self.new_fx))
        if delta < 0.0:
            self.new_del = self.new_tol / 10.0
            mxdl_ = 1e99
        self._say(f"newton: nits={self.new_nits} res={self.new_tol:.6e}")
        sub()
        x_, fx_, tol_ = self.new_x.copy(), self.new_fx.copy(), self.new_tol
        tol__ = 1e99
        x__ = fx__ = None
        del__ = 0.0
        if self.new_tol < tol:
            self._say("newton: input already converged")
            return 0
        while True:
            if self.new_del < mndl:
                self._say("newton: trust region too small")
                return 3
            s = np.zeros(n)
            gres = gtol * self.new_tol
            gdel = self.new_del
            if ginfo != 2:
                self.new_gits = m
            s, gres, gdel, self.new_gits, ginfo = gm.solve(s, fx_, df, dfp, dp, gres, gdel, self.new_gits,
                                                           ginfo)
            ginfo = info
            self.new_x = x_ - s
            self.new_fx = f(self.new_x)
            self.new_tol = math.sqrt(dp(self.new_fx, self.new_fx))
            snrm = math.sqrt(dp(s, s))
            ared = tol_ - self.new_tol
            pred = tol_ - gdel
            with np.errstate(divide="ignore", invalid="ignore"):
                ratio = float(np.float64(ared) / np.float64(pred))   # IEEE, as the Fortran division
            self._say(f"newton: nits={self.new_nits} res={self.new_tol:.6e} gits={self.new_gits} "
                      f"del={self.new_del:.6e} |s|={snrm:.6e} ared/pred={ratio:.4f}")
            if self.new_tol > tol__:
                self._say("newton: accepting previous step")
                self.new_x, self.new_fx, self.new_tol, self.new_del = x__, fx__, tol__, del__
            elif ared < 0.0:
                self._say("newton: norm increased, try smaller step")
                self.new_del = snrm * 0.5
                ginfo = 2
            elif ratio < 0.75:
                self._say("newton: step ok, trying smaller step")
                x__, fx__, tol__ = self.new_x, self.new_fx, self.new_tol
                del__ = snrm if ratio > 0.1 else snrm * 0.5
                self.new_del = snrm * 0.7
                ginfo = 2
            elif snrm < self.new_del * 0.9:
                self._say("newton: step good, took full newton step")
                self.new_del = min(mxdl_, snrm * 2.0)
            elif self.new_del < mxdl_ * 0.9:
                self._say("newton: step good, trying larger step")
                x__, fx__, tol__, del__ = self.new_x, self.new_fx, self.new_tol, self.new_del
                self.new_del = min(mxdl_, snrm * 2.0)
                ginfo = 2
            if ginfo == 2:
                continue
            self.new_nits += 1
            sub()
            x_, fx_, tol_ = self.new_x.copy(), self.new_fx.copy(), self.new_tol
            tol__ = 1e99
            if self.new_tol < tol:
                self._say("newton: converged")
                return 0
            if self.new_nits == nits:
                self._say("newton: reached max its")
                return 2


# ---------------------------------------------------------------------------
# arnoldi.f
# ---------------------------------------------------------------------------
class Arnoldi:
    """arnold (arnoldi.f:22-246) in reverse communication: call ``step(sv)``
    with sv = A q_k (the start vector on the first call); it returns
    (ifail, sv) with ifail 0: apply A to the returned sv and call again;
    -1: the ncgd eigenvalues of largest magnitude converged (consecutive
    Hessenberg estimates closer than tol), ``wr2``/``wi2`` hold them and
    ``b`` the eigenvectors (a real column per real eigenvalue, the real and
    imaginary parts for a complex pair); 1: kmax reached; 2: invariant
    subspace.  Gram-Schmidt with one reorthogonalisation, as arnold does."""

    SMALL = 1e-80

    def __init__(self, n, kmax, ncgd, dp, tol):
        self.n, self.kmax, self.ncgd, self.dp, self.tol = n, kmax, ncgd, dp, tol
        self.k = 0
        self.q = np.zeros((kmax, n))
        self.h = np.zeros((kmax, kmax))
        self.wr2 = np.zeros(ncgd)
        self.wi2 = np.zeros(ncgd)
        self.b = np.zeros((2 * ncgd, n))

    def step(self, sv):
        k, q, h = self.k, self.q, self.h
        if k >= self.kmax:
            return 1, sv
        q[k] = sv
        for j in range(k):                  # modified Gram-Schmidt
            h[j, k - 1] = self.dp(q[k], q[j])
            q[k] = q[k] - h[j, k - 1] * q[j]
        for j in range(k):                  # reorthogonalise
            nm = self.dp(q[k], q[j])
            q[k] = q[k] - nm * q[j]
            h[j, k - 1] = h[j, k - 1] + nm
        norm = math.sqrt(self.dp(q[k], q[k]))
        if norm < self.SMALL:
            return 2, sv
        q[k] = q[k] / norm
        sv = q[k].copy()
        ifail = 0
        if k == 0:
            h[:] = 0.0
        else:
            h[k, k - 1] = norm
            if k >= self.ncgd:
                ifail = self._eigen(k)
        self.k = k + 1
        return ifail, sv

    def _eigen(self, k):
        """Eigenvalues of the k x k Hessenberg matrix (LAPACK's balanced QR),
        the ncgd largest in magnitude compared with the last estimates."""
        w, vr = scipy.linalg.eig(self.h[:k, :k])
        wr, wi = w.real, w.imag
        sel = np.zeros(k, dtype=bool)
        p = 0
        picked = []
        for i in range(self.ncgd):
            best, m = -1.0, -1
            for j in range(k):
                if not sel[j]:
                    nm = wr[j] * wr[j] + wi[j] * wi[j]
                    if nm > best:
                        best, m = nm, j
            sel[m] = True
            if abs(wi[m]) > self.SMALL:     # the conjugate partner
                partner = m + 1 if m + 1 < k and wi[m + 1] == -wi[m] else int(np.argmin(np.abs(w - np.conj(w[m]))))
                sel[partner] = True
            if math.sqrt((wr[m] - self.wr2[i]) ** 2 + (wi[m] - self.wi2[i]) ** 2) < self.tol:
                p += 1
            self.wr2[i], self.wi2[i] = wr[m], wi[m]
            picked.append(m)
        if p < self.ncgd:
            return 0
        cols = []
        Qk = self.q[:k]
        for m in picked:
            v = vr[:, m]
            if abs(wi[m]) > self.SMALL:
                if wi[m] < 0:
                    v = np.conj(v)
                cols.append(v.real @ Qk)
                cols.append(v.imag @ Qk)
            else:
                cols.append(v.real @ Qk)
        self.b = np.array(cols)
        return -1


def logeig(wr, wi, t):
    """logeig (arnoldi.f:256-302): Floquet multipliers wr + i wi = e^{sigma t}
    to exponents sigma; the angle lies in [-pi, pi) (a negative real
    multiplier gives -pi) as in arnoldi.f."""
    sr, si = np.zeros(len(wr)), np.zeros(len(wr))
    for i, (re, im) in enumerate(zip(wr, wi)):
        if abs(re) < Arnoldi.SMALL:
            if abs(im) < Arnoldi.SMALL:
                raise ValueError("logeig: zero multiplier")
            th = 0.5 * D_PI if im > 0 else -0.5 * D_PI
        elif re > 0:
            th = math.atan(im / re)
        else:
            th = D_PI + math.atan(im / re) if im > 0 else -D_PI + math.atan(im / re)
        sr[i] = math.log(math.sqrt(re * re + im * im)) / t
        si[i] = th / t
    return sr, si


# ---------------------------------------------------------------------------
# newton.f90: the orbit problem
# ---------------------------------------------------------------------------
class DeviceStepper:
    """steporbit on the device: one slot per segment of a context whose slots
    are set up with the run's Re, alpha (and symmetries)."""

    def __init__(self, ctx, slots):
        self.ctx = ctx
        self.slots = list(slots)

    def steporbit(self, fields, ndts, T, sth, sz, dt1):
        """fields[i] = (ur, ut, uz) of segment i; ndts[i] == 1 entries take one
        step of size dt1 (the tim_dt of the last integration).  Returns the
        integrated and shifted fields."""
        ctx = self.ctx
        sl = self.slots[:len(fields)]
        for s, nd, f in zip(sl, ndts, fields):
            if nd == 1 and ctx.timestep(s)["dt"] != dt1:
                ctx.vel_matrices(s, dt1)
            ctx.put_state(s, *f)
        ctx.steporbit(sl, ndts, T, sth, sz)
        return [ctx.get_state(s)[:3] for s in sl]

    def imposesym(self, fields):
        """vel_imposesym (velocity.f90:61-74) of host fields with the first
        slot's symmetries."""
        s = self.slots[0]
        self.ctx.put_state(s, *fields)
        self.ctx.imposesym(s)
        return self.ctx.get_state(s)[:3]


class BatchedStepper:
    """Several independent searches (one thread each) on one context: every
    steporbit request waits until each live search has posted one, then all
    of them advance in one opf_steporbit launch sequence over their slots.
    The searches' own arithmetic is untouched, so each follows exactly the
    iterates it would follow alone; only the device work is batched.

    ``stepper(i)`` is the Stepper of search i (slots ``slots[i]``).  Run each
    search's ``solve`` in its own thread (``run_searches`` does this)."""

    def __init__(self, ctx, slots):
        import threading
        self.ctx = ctx
        self.slots = [list(s) for s in slots]
        self.cv = threading.Condition()
        self.live = len(self.slots)
        self.pending = {}
        self.results = {}
        self.error = None
        self.launches = 0

    def stepper(self, i):
        outer = self

        class _One:
            def steporbit(self, fields, ndts, T, sth, sz, dt1):
                return outer._request(i, (fields, ndts, T, sth, sz, dt1))

            def imposesym(self, fields):
                with outer.cv:   # outside the batched phase: done alone, serialised
                    s = outer.slots[i][0]
                    outer.ctx.put_state(s, *fields)
                    outer.ctx.imposesym(s)
                    return outer.ctx.get_state(s)[:3]
        return _One()

    def finished(self, i):
        with self.cv:
            self.live -= 1
            self._maybe_flush()
            self.cv.notify_all()

    def _request(self, i, req):
        with self.cv:
            self.pending[i] = req
            self._maybe_flush()
            while i not in self.results and self.error is None:
                self.cv.wait()
            if self.error is not None:
                raise RuntimeError("batched steporbit failed") from self.error
            return self.results.pop(i)

    def _maybe_flush(self):
        if not self.pending or len(self.pending) < self.live:
            return
        ctx = self.ctx
        try:
            sl, nd, Ts, sths, szs, owners = [], [], [], [], [], []
            for i, (fields, ndts, T, sth, sz, dt1) in sorted(self.pending.items()):
                for j, f in enumerate(fields):
                    s = self.slots[i][j]
                    if ndts[j] == 1 and ctx.timestep(s)["dt"] != dt1:
                        ctx.vel_matrices(s, dt1)
                    ctx.put_state(s, *f)
                    sl.append(s)
                    nd.append(ndts[j])
                    Ts.append(T[j])
                    sths.append(sth[j])
                    szs.append(sz[j])
                    owners.append(i)
            ctx.steporbit(sl, nd, Ts, sths, szs)
            self.launches += 1
            out = {i: [] for i in self.pending}
            for s, i in zip(sl, owners):
                out[i].append(ctx.get_state(s)[:3])
            self.results.update(out)
        except Exception as e:   # every waiting search sees the failure
            self.error = e
        self.pending = {}
        self.cv.notify_all()


def run_searches(batch, searches, **solve_args):
    """Run OrbitNewton searches (each built on ``batch.stepper(i)``) in
    threads with their device integrations batched; returns each search's
    (info, NewtonHook)."""
    import threading
    res = [None] * len(searches)

    def work(i):
        try:
            res[i] = searches[i].solve(**solve_args)
        except Exception as e:   # noqa: BLE001  (reported to the caller below)
            res[i] = e
        finally:
            batch.finished(i)

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(searches))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for r in res:
        if isinstance(r, Exception):
            raise r
    return res


class OrbitNewton:
    """newton.f90's search for a relative periodic orbit on ms shooting
    segments.  x0 fields: (ur, ut, uz) per segment (or one state, segments
    after the first are then produced by integration as load_ms_states
    does); guesses sth, sz, T per segment (shifts.in); ndts steps per
    segment.  symm: b_mirrorsym / b_shiftrefl flags (they freeze the theta
    shift, newton.f90:448-450); T_fixed freezes the period."""

    def __init__(self, stepper, N, K, M, Mp, alpha, intrdr, x0, sth, sz, T, ndts, T_fixed=False,
                 freeze_theta_shift=False, r=None):
        self.stepper = stepper
        self.N, self.K, self.M, self.Mp, self.alpha = N, K, M, Mp, float(alpha)
        self.H = (2 * K - 1) * M - (K - 1)
        self.n = 8 + 6 * N * self.H
        self.intrdr = np.asarray(intrdr, dtype=np.float64)
        self.ndts = [int(v) for v in np.atleast_1d(ndts)]
        self.ms = len(self.ndts)
        self.T_fixed = bool(T_fixed)
        self.freeze_theta = bool(freeze_theta_shift)
        self.tim_dt = 0.0
        ks, ms_ = harmonic_km(K, M)
        self._k, self._m = ks, ms_
        # rtz2vec weights (newton.f90:572-580): d = sqrt|intrdr|, w(nh)
        self._d = np.sqrt(np.abs(self.intrdr))
        if r is None:
            from .engine import host_mesh
            r = host_mesh(N)["r"]
        r = np.asarray(r, dtype=np.float64)
        self._r2 = r * r                                  # mes_D%r(:,2)
        self._w = 4.0 / (4.0 + (self.alpha * np.abs(ks)) + (float(Mp) * np.abs(ms_)))
        states = [x0] if isinstance(x0[0], np.ndarray) and np.ndim(x0[0]) == 2 else list(x0)
        self.get_scales(*states[0])
        x = np.zeros(self.n * self.ms)
        sth, sz, T = (np.broadcast_to(np.asarray(v, dtype=np.float64), (self.ms,)) for v in (sth, sz, T))
        for p in range(self.ms):
            x[p * self.n:p * self.n + 3] = (sth[p], sz[p], T[p])
        x[8:self.n] = self.rtz2vec(*states[0])[8:]
        d = self.dotprod(-1, x[:self.n], x[:self.n])
        self.scaleT = x[2] / math.sqrt(d)            # get_scales (newton.f90:266-272)
        for p in range(self.ms):
            x[p * self.n:p * self.n + 3] /= self.scaleT
        for p in range(1, self.ms):                  # load_ms_states without state files
            if p < len(states):
                x[p * self.n + 8:(p + 1) * self.n] = self.rtz2vec(*states[p])[8:]
            else:
                y = self.steporbit_many([x[(p - 1) * self.n:p * self.n]], [self.ndts[p - 1]])[0]
                x[p * self.n + 8:(p + 1) * self.n] = y[8:]
        self.x0 = x
        self.history = []

    # -- scaled vector <-> fields (rtz2vec / vec2rtz, newton.f90:560-652) --------
    def get_scales(self, ur, ut, uz):
        """get_scales (newton.f90:242-275): scaleU from the energies of the
        components and of the mean axial flow."""
        e = lambda a: energy([a], self.intrdr, self.K, self.M, self.alpha)  # noqa: E731
        Er, Et, Ez = e(ur), e(ut), e(uz)
        c1 = uz.copy()
        c1[:, 0] = 0.0 + 1j * c1[:, 0].imag
        Ez00 = Ez - e(c1)
        Ez = Ez - Ez00
        self.scaleU = np.array([math.sqrt(0.2 * Ez00 / Er), math.sqrt(0.2 * Ez00 / Et),
                                math.sqrt(0.5 * Ez00 / Ez)])

    def _scales(self):
        s = (self._d[:, None, None] * self._w[None, :, None]) * self.scaleU[None, None, :]   # (N, H, 3)
        s[:, 0, 2] = self._d
        return s

    def rtz2vec(self, ur, ut, uz):
        s = self._scales()
        v = np.zeros(self.n)
        blk = np.empty((self.H, self.N, 6))
        for c, a in enumerate((ur, ut, uz)):
            blk[:, :, 2 * c] = (a.real * s[:, :, c]).T
            blk[:, :, 2 * c + 1] = (a.imag * s[:, :, c]).T
        v[8:] = blk.ravel()
        return v

    def vec2rtz(self, v):
        s = self._scales()
        blk = v[8:self.n].reshape(self.H, self.N, 6)
        return tuple((blk[:, :, 2 * c].T / s[:, :, c]) + 1j * (blk[:, :, 2 * c + 1].T / s[:, :, c])
                     for c in range(3))

    # -- dot products (newton.f90:658-697) --------------------------------------
    def dotprod(self, n_, a, b):
        n1 = 8 if n_ == -1 else 0
        return float(np.dot(a[n1:self.n], b[n1:self.n]))

    def dotprod_ms(self, n_, a, b):
        d = 0.0
        for p in range(self.ms):
            d = d + self.dotprod(n_, a[p * self.n:(p + 1) * self.n], b[p * self.n:(p + 1) * self.n])
        return d

    def normE(self, v):
        """normE (newton.f90:703-715)."""
        E = 0.0
        for p in range(self.ms):
            for a in self.vec2rtz(v[p * self.n:(p + 1) * self.n]):
                E = E + energy([a], self.intrdr, self.K, self.M, self.alpha)
        return math.sqrt(E)

    # -- time integration -------------------------------------------------------
    def steporbit_many(self, xs, ndts):
        """steporbit (newton.f90:721-775) of several segment vectors at once."""
        fields = [self.vec2rtz(x) for x in xs]
        P = [x[:3] * self.scaleT for x in xs]
        T = [p[2] for p in P]
        for nd, t in zip(ndts, T):
            if nd != 1:
                self.tim_dt = t / float(nd)
        ys = self.stepper.steporbit(fields, ndts, T, [p[0] for p in P], [p[1] for p in P], self.tim_dt)
        out = []
        for x, f in zip(xs, ys):
            y = self.rtz2vec(*f)
            y[:8] = x[:8]
            out.append(y)
        return out

    def getshiftdir(self, i, x):
        """getshiftdir (newton.f90:535-556): d x / d(theta shift) or d(z shift)."""
        ds = (float(self.Mp) * self._m) if i == 1 else (self.alpha * self._k)
        f = [(-a.imag * ds[None, :]) + 1j * (a.real * ds[None, :]) for a in self.vec2rtz(x)]
        return self.rtz2vec(*f)

    def getrhs(self, x):
        """getrhs (newton.f90:374-393): F(x) - x over the segments, constraint rows 0."""
        n, ms = self.n, self.ms
        ys = self.steporbit_many([x[p * n:(p + 1) * n] for p in range(ms)], self.ndts)
        y_ = np.zeros(n * ms)
        for p in range(ms):
            p1 = (p + 1) % ms
            y_[p1 * n:(p1 + 1) * n] = ys[p]
        y = y_ - x
        for p in range(ms):
            y[p * n:p * n + 8] = 0.0
        return y

    def multJ(self, nh, dx):
        """multJ (newton.f90:399-456): the Jacobian action by finite
        differences and the constraint rows."""
        n, ms = self.n, self.ms
        eps = math.sqrt(self.dotprod_ms(1, dx, dx))
        if eps == 0.0:
            raise RuntimeError("multJ: eps=0 (1)")
        eps = self.epsJ * math.sqrt(self.dotprod_ms(1, nh.new_x, nh.new_x)) / eps
        if eps == 0.0:
            raise RuntimeError("multJ: eps=0 (2)")
        y = nh.new_x + eps * dx
        s = self.getrhs(y)
        y = (s - nh.new_fx) / eps
        xs = [nh.new_x[p * n:(p + 1) * n] for p in range(ms)]
        traj = self.steporbit_many(xs, [1] * ms)
        for p in range(ms):
            xp, dxp = xs[p], dx[p * n:(p + 1) * n]
            y[p * n + 0] = self.dotprod(-1, self.getshiftdir(1, xp), dxp)
            y[p * n + 1] = self.dotprod(-1, self.getshiftdir(2, xp), dxp)
            st = (traj[p] - xp) / self.tim_dt
            y[p * n + 2] = self.dotprod(-1, st, dxp)
            if self.freeze_theta:
                y[p * n + 0] = dxp[0]
            if self.T_fixed:
                y[p * n + 2] = dxp[2]
        return y

    # -- the search (newton.f90:46-121) -----------------------------------------
    def solve(self, m=100, nits=20, rel_err=1e-8, delta=-1.0, mndl=1e-12, mxdl=1e3, gtol=1e-3, epsJ=1e-7,
              on_iteration=None):
        """newtonhook on this problem with the tolerances of params.in scaled
        by |x0| as newton.f90:91-95 does.  Returns (info, NewtonHook)."""
        self.epsJ = float(epsJ)
        d = self.dotprod_ms(-1, self.x0, self.x0)
        tol = rel_err * math.sqrt(d)
        scale = math.sqrt(d)
        nh = NewtonHook(self.x0)

        def sub():   # saveorbit (newton.f90:462-529): the newton.dat quantities
            norm_x = math.sqrt(self.dotprod_ms(-1, nh.new_x, nh.new_x))
            rec = dict(nits=nh.new_nits, gits=nh.new_gits, tol=nh.new_tol, delta=nh.new_del,
                       rel=nh.new_tol / norm_x, norm_x=norm_x,
                       normE_ratio=self.normE(nh.new_fx) / self.normE(nh.new_x),
                       shifts=[tuple(nh.new_x[p * self.n:p * self.n + 3] * self.scaleT)
                               for p in range(self.ms)])
            self.history.append(rec)
            if on_iteration is not None:
                on_iteration(self, nh)

        info = nh.solve(self.getrhs, lambda v: self.multJ(nh, v), lambda v: v, sub,
                        lambda a, b: self.dotprod_ms(1, a, b), m, gtol, tol,
                        delta * scale if delta > 0 else delta, mndl * scale, mxdl * scale, nits)
        self.result = nh
        return info, nh

    # -- Floquet analysis (getEigen, newton.f90:850-968) -------------------------
    def multA(self, x, y0, epsA):
        """multA (newton.f90:814-842): the linearised map over all segments
        about new_x by finite differences."""
        n = self.n
        y = x.copy()
        for p in range(self.ms):
            xp = self.xn[p * n:(p + 1) * n]
            eps = math.sqrt(self.dotprod(-1, y, y))
            if eps == 0.0:
                raise RuntimeError("multA: eps=0 (1)")
            eps = epsA * math.sqrt(self.dotprod(-1, xp, xp)) / eps
            if eps == 0.0:
                raise RuntimeError("multA: eps=0 (2)")
            y = xp + eps * y
            y[:8] = xp[:8]
            y = self.steporbit_many([y], [self.ndts[p]])[0]
            y = (y - y0[p * n:(p + 1) * n]) / eps
        return y

    def get_eigen(self, x=None, ncgd=4, kmax=60, epsA=1e-7, rel_err=1e-8, gtol=1e-3):
        """getEigen (newton.f90:850-968): the ncgd leading Floquet multipliers
        of the orbit x (default: the last Newton iterate) by Arnoldi on multA.
        Returns dict(k, ifail, multipliers (complex), exponents (complex),
        vectors: (ur, ut, uz) per eigenvector column scaled to the orbit's
        energy as the saved states are)."""
        n = self.n
        self.xn = np.array(self.result.new_x if x is None else x, dtype=np.float64)
        rerr = min(rel_err, gtol)
        t = float(np.sum(self.xn[2::n][:self.ms])) * self.scaleT
        y0 = np.concatenate(self.steporbit_many([self.xn[p * n:(p + 1) * n] for p in range(self.ms)],
                                                self.ndts))
        # initial input: the orbit's state shifted by 0.3 of each period, with
        # (1 - r^2) r^2 put into harmonics (1, 0) and (0, 1) of u_r and u_theta
        ur, ut, uz = self.vec2rtz(self.xn[:n])
        st, sz = 0.3 * (2.0 * D_PI), 0.3 * (2.0 * D_PI / self.alpha)
        ph = np.exp(-1j * ((sz * self._k.astype(float)) * self.alpha)) * \
            np.exp(-1j * ((st * self._m.astype(float)) * float(self.Mp)))
        ur, ut, uz = (a * ph[None, :] for a in (ur, ut, uz))
        r2 = self._r2
        prof = (1.0 - r2) * r2
        for h in (1, 2 * self.K - 1):    # var_harmonic(1, 0), var_harmonic(0, 1)
            ur[:, h] = prof + 1j * ur[:, h].imag
            ut[:, h] = prof + 1j * ut[:, h].imag
        if hasattr(self.stepper, "imposesym"):
            ur, ut, uz = self.stepper.imposesym((ur, ut, uz))
        sv = self.rtz2vec(ur, ut, uz)
        ar = Arnoldi(n, kmax, ncgd, lambda a, b: self.dotprod(-1, a, b), rerr)
        while True:
            sv[:8] = 0.0
            ifail, sv = ar.step(sv)
            if ifail != 0:
                break
            sv = self.multA(sv, y0, epsA)
        mult = ar.wr2 + 1j * ar.wi2
        # converged (-1) or at kmax (1): the last estimates, as getEigen writes them
        sr, si = logeig(ar.wr2, ar.wi2, t) if ifail in (-1, 1) else (np.full(ncgd, np.nan), np.full(ncgd, np.nan))
        E0 = sum(energy([a], self.intrdr, self.K, self.M, self.alpha) for a in self.vec2rtz(self.xn[:n]))
        vecs = []
        if ifail == -1:
            j = 0
            for i in range(ncgd):
                cols = ar.b[j:j + (2 if ar.wi2[i] != 0.0 else 1)]
                E = sum(energy([a], self.intrdr, self.K, self.M, self.alpha)
                        for c in cols for a in self.vec2rtz(c))
                for c in cols:
                    v = c * math.sqrt(E0 / E) if E > 0.0 else c
                    vecs.append(self.vec2rtz(v))
                j += len(cols)
        return dict(k=ar.k, ifail=ifail, t=t, multipliers=mult, exponents=sr + 1j * si, vectors=vecs,
                    arnoldi=ar)

    def fields(self, x, p=0):
        """The velocity fields of segment p of a Newton vector."""
        return self.vec2rtz(x[p * self.n:(p + 1) * self.n])

    def shifts(self, x, p=0):
        """(sth, sz, T) of segment p, unscaled."""
        return tuple(x[p * self.n:p * self.n + 3] * self.scaleT)
