"""Newton-Krylov-hookstep search for (relative) periodic orbits over the
device time step: the client of newton.f90.

``HookstepNewton`` is a generic root finder for f(x) = 0: restarted GMRES
(``Gmres``, Saad & Schultz 1986) on a Jacobian action, Viswanath's hookstep
(``hookstep``) inside a trust region, and Dennis & Schnabel's region
updates.  Its thresholds and the order of its decisions are those of the
reference's solver (NewtonHook.f90:48-179, GMRESm.f90:27-153), so a search
follows the reference's iterates; the parameters are in ``HookParams`` and
``TrustRegionParams``.

``OrbitNewton`` is newton.f90's problem: the Newton vector holds, per
shooting segment, the theta shift, z shift and period (scaled by scaleT) and
the scaled velocity field (rtz2vec), f(x) = F(x) - x with F the integration
over one period followed by the shift back (newton.f90:374-393, 721-775),
and the Jacobian action by finite differences with constraint rows that
forbid updates along the shift directions and the trajectory
(newton.f90:399-456).

The time integrations run on the device through ``opf_steporbit``: the
segments of multiple shooting advance together in one context (one slot
each), so a residual costs one batched integration rather than ms
sequential ones.  Any object with the ``Stepper`` interface can stand in for
the device (the tests use the oracle's CPU time step to check the device
path).
"""
from __future__ import annotations

import math
import time

import numpy as np
import scipy.linalg

from .ic import energy, harmonic_km

D_PI = 3.1415926535897931


# ---------------------------------------------------------------------------
# Krylov solver, hookstep and trust-region Newton iteration
# ---------------------------------------------------------------------------
class HookParams:
    """Parameters of the hookstep's search for the Levenberg-Marquardt
    multiplier (NewtonHook.f90 / GMRESm.f90:121-153 use these values): the
    search starts at max(sigma_min^2 * start_rel, start_abs) and grows the
    multiplier by the factor `growth` until the step fits the region."""
    start_rel = 1e-6
    start_abs = 1e-99
    growth = 1.1


class TrustRegionParams:
    """The trust-region rules of newtonhook (NewtonHook.f90:120-160), in
    Dennis & Schnabel's terms: rho = actual / predicted reduction of |f|.
    rho < accept_ratio: keep the trial as a fallback and shrink the region
    to shrink_ok * |s|; the fallback's own region is |s| when
    rho > poor_ratio, else halve * |s|.  |f| grew: shrink to halve * |s|.
    A good step shorter than full_fraction of the region doubles the region
    (expand * |s|, capped) and is taken; a good step that filled the region
    is tried again with the region expanded while the region is below
    full_fraction of its cap."""
    accept_ratio = 0.75
    poor_ratio = 0.1
    shrink_ok = 0.7
    halve = 0.5
    full_fraction = 0.9
    expand = 2.0


class Status:
    """Exit codes shared with newtonhook's info and gmresm's info."""
    CONVERGED = 0
    KRYLOV_STALLED = 1      # GMRES: the residual stopped decreasing
    MAX_ITERATIONS = 2
    REGION_COLLAPSED = 3    # Newton: trust region below its minimum


def hookstep(h, j, beta, delta):
    """Viswanath's hookstep (Viswanath 2007, J. Fluid Mech. 580; also
    arXiv:0809.1498): among the y with |y| <= delta, the one minimising
    |beta e1 - Hbar y| for the (j+1) x j Hessenberg matrix Hbar = h[:j+1, :j].
    With Hbar = U diag(sigma) V^T, y(mu) = V c with
    c_i = g_i sigma_i / (sigma_i^2 + mu) and g = beta U^T e1; mu grows
    geometrically from a small start until |c| <= delta.  Returns y and the
    residual |beta e1 - Hbar y| it leaves (the predicted |f| after the step)."""
    hb = h[:j + 1, :j]
    u, sigma, vt = np.linalg.svd(hb, full_matrices=True)
    g = beta * u[0, :j]
    mu = max(sigma[j - 1] * sigma[j - 1] * HookParams.start_rel, HookParams.start_abs)
    while True:
        mu = mu * HookParams.growth
        c = g * sigma / (mu + sigma * sigma)
        if math.sqrt(float(np.dot(c, c))) <= delta:
            break
    y = c @ vt
    resid = -(hb @ y)
    resid[0] = resid[0] + beta
    return y, math.sqrt(float(np.dot(resid, resid)))


class Gmres:
    """Restarted GMRES(m) (Saad & Schultz 1986, SIAM J. Sci. Stat. Comput.
    7:856) with modified Gram-Schmidt and right preconditioning, solving
    A x = b for the Newton step.  After each Arnoldi column the small least-
    squares problem min |beta e1 - Hbar y| is solved by a rank-revealing QR
    (LAPACK dgelsy, as gmresm does) to monitor the residual.  At the end of a
    cycle the step is either that solution or, with a trust region
    (radius > 0), the hookstep.  The Arnoldi basis of the last cycle is kept
    so that `rescale` returns the hookstep for another radius without new
    products of A (newtonhook's retries, GMRESm.f90:50-55).

    Stopping (GMRESm.f90:76-104): |r| <= tol, the iteration budget spent, or
    |r| no longer below (1 - stall) of the previous column's value."""

    stall = 1e-14

    def __init__(self, m, n):
        self.m, self.n = m, n
        self.hess = np.zeros((m + 1, m))
        self.basis = np.zeros((m + 1, n))     # row i: Krylov vector i
        self.cols = 0
        self.beta = 0.0

    def _expand(self, y, cols, precond):
        z = np.zeros(self.n)
        for i in range(cols):
            z = z + y[i] * self.basis[i]
        return precond(z)

    def rescale(self, radius, precond):
        """Hookstep of the last cycle's basis for a new radius: (step,
        predicted residual)."""
        y, pred = hookstep(self.hess, self.cols, self.beta, radius)
        return self._expand(y, self.cols, precond), pred

    def _residual_of_lsq(self, cols, beta):
        hb = self.hess[:cols + 1, :cols]
        e1 = np.zeros(cols + 1)
        e1[0] = beta
        y = scipy.linalg.lstsq(hb, e1, lapack_driver="gelsy")[0]
        r = -(hb @ y)
        r[0] = r[0] + beta
        return y, math.sqrt(float(np.dot(r, r)))

    def solve(self, b, apply, precond, dot, tol, budget, radius, x=None):
        """Returns (x, |r|, predicted residual, iterations, status).  The
        predicted residual is the hookstep's when radius > 0, else 0."""
        x = np.zeros(self.n) if x is None else x
        its = 0
        pred = 0.0
        self.basis[:] = 0.0
        while True:                                      # one cycle of m columns
            r = b if math.sqrt(dot(x, x)) == 0.0 else b - apply(x)
            beta = math.sqrt(dot(r, r))
            self.beta = beta
            self.basis[0] = r / beta
            self.hess[:] = 0.0
            last = math.inf
            for cols in range(1, self.m + 1):
                self.cols = cols
                its += 1
                w = apply(precond(self.basis[cols - 1].copy()))
                for i in range(cols):                    # modified Gram-Schmidt
                    self.hess[i, cols - 1] = dot(w, self.basis[i])
                    w = w - self.hess[i, cols - 1] * self.basis[i]
                self.hess[cols, cols - 1] = math.sqrt(dot(w, w))
                with np.errstate(divide="ignore", invalid="ignore"):
                    self.basis[cols] = w / self.hess[cols, cols - 1]
                y, res = self._residual_of_lsq(cols, beta)
                converged, spent, stalled = res <= tol, its == budget, res > last
                if not (converged or spent or stalled or cols == self.m):
                    last = res * (1.0 - self.stall)
                    continue
                if radius > 0.0:
                    y, pred = hookstep(self.hess, cols, beta, radius)
                x = x + self._expand(y, cols, precond)
                if converged:
                    return x, res, pred, its, Status.CONVERGED
                if stalled:
                    return x, res, pred, its, Status.KRYLOV_STALLED
                if spent:
                    return x, res, pred, its, Status.MAX_ITERATIONS
                break                                    # full cycle: restart from x


class _Trial:
    """A point of the Newton iteration: x, f(x), |f(x)| and, for a kept
    fallback, the trust radius to resume with."""
    __slots__ = ("x", "fx", "res", "radius")

    def __init__(self, x, fx, res, radius=0.0):
        self.x, self.fx, self.res, self.radius = x, fx, res, radius


class HookstepNewton:
    """Newton's method with a GMRES-hookstep trust region (Viswanath 2007;
    the region management of Dennis & Schnabel 1996, ch. 6) for f(x) = 0,
    the solver newton.f90 calls (newtonhook, NewtonHook.f90:48-179).

    Each iteration solves J s = f(x) for the step by GMRES(m) to a residual
    krylov_rel * |f(x)|, constrained to |s| <= radius by the hookstep, and
    tries x - s.  The region then shrinks or grows by TrustRegionParams;
    shrinking re-solves the hookstep on the same Krylov basis.  Rejected
    trials that still reduced |f| are kept as a fallback, taken when a later
    trial from the same point is worse.

    Public state after `solve` (and inside the callbacks): x, fx, res = |f(x)|,
    radius, newton_its, krylov_its (GMRES iterations of the last solve), and
    `log`, one line per trial and per decision."""

    def __init__(self, x0):
        self.x = np.array(x0, dtype=np.float64)
        self.fx = np.zeros_like(self.x)
        self.res = 0.0
        self.radius = 0.0
        self.newton_its = 0
        self.krylov_its = 0
        self.log = []

    def _note(self, msg):
        self.log.append(msg)

    def _take(self, p):
        self.x, self.fx, self.res = p.x, p.fx, p.res

    def solve(self, func, jac, precond, on_accept, dot, m, krylov_rel, tol, radius, radius_min, radius_max,
              max_its):
        """func(x) -> f(x); jac(v) -> J v about self.x; precond(v) -> v;
        on_accept() after the start point and every accepted iterate;
        dot(a, b).  radius = 0 disables the trust region; radius < 0 starts
        it at |f(x0)|/10 with no cap.  Returns a Status code."""
        tr = TrustRegionParams
        norm = lambda v: math.sqrt(dot(v, v))  # noqa: E731
        krylov = Gmres(m, self.x.size)
        self.newton_its = 0
        self.krylov_its = 0
        self.radius = radius
        cap = radius_max
        self.fx = func(self.x)
        self.res = norm(self.fx)
        if radius < 0.0:
            self.radius = self.res / 10.0
            cap = 1e99
        self._note(f"start: |f| = {self.res:.6e}")
        on_accept()
        base = _Trial(self.x.copy(), self.fx.copy(), self.res)
        fallback = None
        if self.res < tol:
            self._note("start point within tolerance")
            return Status.CONVERGED
        reuse_basis = False
        while True:
            if self.radius < radius_min:
                self._note(f"stop: trust radius {self.radius:.3e} below {radius_min:.3e}")
                return Status.REGION_COLLAPSED
            if reuse_basis:
                step, pred_res = krylov.rescale(self.radius, precond)
            else:
                step, _, pred_res, self.krylov_its, _ = krylov.solve(
                    base.fx, jac, precond, dot, krylov_rel * base.res, m, self.radius)
            trial = _Trial(base.x - step, None, 0.0)
            trial.fx = func(trial.x)
            trial.res = norm(trial.fx)
            self._take(trial)
            snorm = norm(step)
            actual = base.res - trial.res
            predicted = base.res - pred_res
            with np.errstate(divide="ignore", invalid="ignore"):
                rho = float(np.float64(actual) / np.float64(predicted))
            self._note(f"trial {self.newton_its}: |f| = {trial.res:.6e}, gmres its {self.krylov_its}, "
                       f"radius {self.radius:.6e}, |s| = {snorm:.6e}, rho = {rho:.4f}")
            reuse_basis = True
            if fallback is not None and trial.res > fallback.res:
                self._note("  worse than the kept fallback: take the fallback")
                self._take(fallback)
                self.radius = fallback.radius
                reuse_basis = False
            elif actual < 0.0:
                self._note("  |f| grew: halve the region")
                self.radius = snorm * tr.halve
            elif rho < tr.accept_ratio:
                self._note("  weak reduction: keep as fallback, shrink the region")
                fallback = _Trial(trial.x, trial.fx, trial.res,
                                  snorm if rho > tr.poor_ratio else snorm * tr.halve)
                self.radius = snorm * tr.shrink_ok
            elif snorm < self.radius * tr.full_fraction:
                self._note("  good step inside the region: take it, expand the region")
                self.radius = min(cap, snorm * tr.expand)
                reuse_basis = False
            elif self.radius < cap * tr.full_fraction:
                self._note("  good step at the region's edge: keep as fallback, try a larger region")
                fallback = _Trial(trial.x, trial.fx, trial.res, self.radius)
                self.radius = min(cap, snorm * tr.expand)
            else:
                reuse_basis = False
            if reuse_basis:
                continue
            self.newton_its += 1
            on_accept()
            base = _Trial(self.x.copy(), self.fx.copy(), self.res)
            fallback = None
            if self.res < tol:
                self._note(f"converged after {self.newton_its} iterations")
                return Status.CONVERGED
            if self.newton_its == max_its:
                self._note(f"stop: {max_its} iterations without convergence")
                return Status.MAX_ITERATIONS


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


def coll_shift(field, K, M, Mp, alpha, dth, dz):
    """var_coll_shift (variables.f90:487-528) of a host field (N, H): harmonic
    (k, m) times e^{-i alpha k dz}, then e^{-i m Mp dth}, each factor rounded
    as the reference applies it (Re' = Re c + Im s, Im' = Im c - Re s), with
    the C library's cos and sin."""
    ks, ms = harmonic_km(K, M)
    out = np.array(field, dtype=np.complex128)
    re, im = out.real.copy(), out.imag.copy()

    def rot(c, s_):
        nonlocal re, im
        r0 = re
        re = r0 * c[None, :] + im * s_[None, :]
        im = im * c[None, :] - r0 * s_[None, :]

    if dz != 0.0:
        a = [dz * float(k) * alpha for k in ks]
        rot(np.array([math.cos(v) for v in a]), np.array([math.sin(v) for v in a]))
    if dth != 0.0:
        a = [dth * float(m) * float(Mp) for m in ms]
        rot(np.array([math.cos(v) for v in a]), np.array([math.sin(v) for v in a]))
    return re + 1j * im


class ServerStepper:
    """steporbit (newton.f90:721-775) through the request server: every
    segment is one integration request (ndts steps of dt = T/ndts, or one
    step of the last dt for multJ's trajectory direction; one corrector per
    step), so the searches' integrations are continuously batched with
    whatever else the server runs.  The end of steporbit, vel_imposesym and
    the shift back by the orbit's shifts, is applied to the returned states
    on the host with the device's arithmetic, so results equal opf_steporbit's
    bit for bit.  Runs without a symmetric subspace (the server's slots
    impose none)."""

    def __init__(self, server, cls, Re, alpha, Mp=1):
        self.srv, self.cls = server, int(cls)
        self.Re, self.alpha, self.Mp = float(Re), float(alpha), int(Mp)
        _, self.K, self.M = server.classes[self.cls]
        self.requests = 0

    def steporbit(self, fields, ndts, T, sth, sz, dt1):
        ids = []
        for f, nd, t in zip(fields, ndts, T):
            dt = t / float(nd) if nd != 1 else dt1
            ids.append(self.srv.submit(self.cls, self.Re, self.alpha, dt, int(nd), f, corr_fixed=1))
        self.requests += len(ids)
        out = []
        for rid, nd, a, b in zip(ids, ndts, sth, sz):
            while not self.srv.poll(rid)["done"]:
                time.sleep(0.0005)
            st = self.srv.poll(rid)["status"]
            if st != 0:
                raise RuntimeError(f"steporbit request failed with status {st}")
            u = [np.array(x) for x in self.srv.result(rid, self.cls)[:3]]
            u[0][:, 0] = 0.0                          # vel_imposesym: var_imposesym(0) on harmonic 0
            u[1][:, 0] = u[1][:, 0].real
            u[2][:, 0] = u[2][:, 0].real
            if nd != 1:
                u = [coll_shift(x, self.K, self.M, self.Mp, self.alpha, -a, -b) for x in u]
            out.append(tuple(u))
        return out


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
    (status, HookstepNewton)."""
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

    def _segments(self, x):
        return [x[p * self.n:(p + 1) * self.n] for p in range(self.ms)]

    def residual(self, x):
        """f(x) = F(x) - x of multiple shooting (getrhs, newton.f90:374-393):
        segment p is integrated over its period and compared with the start
        of segment p + 1 (cyclically); the 8 leading entries of every segment
        (shifts, period and spares) carry the constraints, zero here."""
        n = self.n
        images = self.steporbit_many(self._segments(x), self.ndts)
        f = np.empty(n * self.ms)
        for p, img in enumerate(images):
            q = (p + 1) % self.ms
            f[q * n:(q + 1) * n] = img - x[q * n:(q + 1) * n]
        f.reshape(self.ms, n)[:, :8] = 0.0
        return f

    def _fd_increment(self, x, dx):
        """The finite-difference increment h of the Jacobian action:
        h |dx| = epsJ |x| (newton.f90:411-418)."""
        ndx = math.sqrt(self.dotprod_ms(1, dx, dx))
        if ndx == 0.0:
            raise ValueError("Jacobian action requested for a zero vector")
        h = self.epsJ * math.sqrt(self.dotprod_ms(1, x, x)) / ndx
        if h == 0.0:
            raise ValueError("Jacobian action about x = 0 has no finite-difference increment")
        return h

    def jacobian_action(self, solver, dx):
        """J dx about the solver's current point (multJ, newton.f90:399-456):
        the forward difference (f(x + h dx) - f(x)) / h in the velocity
        entries; per segment, the constraint rows are the components of dx
        along the theta-shift, z-shift and time directions of x (so that the
        update does not move along the orbit's symmetry group), or dx's own
        shift / period entry where that one is frozen."""
        x, fx = solver.x, solver.fx
        h = self._fd_increment(x, dx)
        jv = (self.residual(x + h * dx) - fx) / h
        xs = self._segments(x)
        # the time direction: one time step of each segment
        forward = self.steporbit_many(xs, [1] * self.ms)
        rows = jv.reshape(self.ms, self.n)
        for p, (xp, ahead) in enumerate(zip(xs, forward)):
            dxp = dx[p * self.n:(p + 1) * self.n]
            along_t = (ahead - xp) / self.tim_dt
            rows[p, 0] = dxp[0] if self.freeze_theta else self.dotprod(-1, self.getshiftdir(1, xp), dxp)
            rows[p, 1] = self.dotprod(-1, self.getshiftdir(2, xp), dxp)
            rows[p, 2] = dxp[2] if self.T_fixed else self.dotprod(-1, along_t, dxp)
        return jv

    # -- the search (newton.f90:46-121) -----------------------------------------
    def solve(self, m=100, nits=20, rel_err=1e-8, delta=-1.0, mndl=1e-12, mxdl=1e3, gtol=1e-3, epsJ=1e-7,
              on_iteration=None):
        """The hookstep Newton search on this problem, with the tolerances of
        newton_params.in relative to |x0| as newton.f90:91-95 takes them.
        Returns (status, HookstepNewton)."""
        self.epsJ = float(epsJ)
        scale = math.sqrt(self.dotprod_ms(-1, self.x0, self.x0))
        solver = HookstepNewton(self.x0)

        def record():   # the quantities saveorbit writes to newton.dat (newton.f90:462-529)
            norm_x = math.sqrt(self.dotprod_ms(-1, solver.x, solver.x))
            self.history.append(dict(
                nits=solver.newton_its, gits=solver.krylov_its, tol=solver.res, delta=solver.radius,
                rel=solver.res / norm_x, norm_x=norm_x,
                normE_ratio=self.normE(solver.fx) / self.normE(solver.x),
                shifts=[self.shifts(solver.x, p) for p in range(self.ms)]))
            if on_iteration is not None:
                on_iteration(self, solver)

        status = solver.solve(self.residual, lambda v: self.jacobian_action(solver, v), lambda v: v, record,
                              lambda a, b: self.dotprod_ms(1, a, b), m, gtol, rel_err * scale,
                              delta * scale if delta > 0 else delta, mndl * scale, mxdl * scale, nits)
        self.result = solver
        return status, solver

    # -- Floquet analysis (getEigen, newton.f90:850-968) -------------------------
    def multA(self, x, y0, epsA):
        """multA (newton.f90:814-842): the linearised map over all segments
        about the orbit self.xn by finite differences."""
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
        self.xn = np.array(self.result.x if x is None else x, dtype=np.float64)
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
