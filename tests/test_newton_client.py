"""The Newton-Krylov-hookstep solver of the orbit client (opf/newton.py)
against the reference's own newtonhook / gmresm / hookstep (NewtonHook.f90,
GMRESm.f90), compiled from its sources by oracle/build_newton_ref.sh and run
on a small algebraic system by oracle/newton_driver.f90
(tests/golden/golden_newton_alg.npz): every accepted iterate (Newton and
GMRES counts, residual, trust region, x) and the exit code.  Plus the pieces
of the orbit problem that need no time step: the scaled vector round trip,
the shift directions against a finite shift, and the hookstep's constraint.
No GPU is used."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import load_golden


def _system(n, amp):
    """oracle/newton_driver.f90's f_i(x) = x_i + 0.4 sin(2 x_{i+1}) + 0.2 x_i^3 - b_i."""
    i = np.arange(1, n + 1, dtype=np.float64)
    xs = np.cos(i)

    def f0(x):
        return x + 0.4 * np.sin(2.0 * np.roll(x, -1)) + 0.2 * x ** 3

    b = f0(xs)
    return (lambda x: f0(x) - b), xs + amp * np.sin(3.0 * i)


@pytest.mark.parametrize("case", range(4))
def test_newtonhook_matches_reference(case):
    from opf.newton import HookstepNewton
    g = load_golden("newton_alg")
    n, m, nits, gtol, tol, dl, mndl, mxdl, epsJ, amp = g["cases"][case]
    n, m, nits = int(n), int(m), int(nits)
    f, x0 = _system(n, amp)
    nh = HookstepNewton(x0)
    rows = []

    def df(v):   # multJ's finite difference (newton.f90:411-418)
        eps = math.sqrt(float(v @ v))
        eps = epsJ * math.sqrt(float(nh.x @ nh.x)) / eps
        return (f(nh.x + eps * v) - nh.fx) / eps

    def sub():
        rows.append([nh.newton_its, nh.krylov_its, nh.res, nh.radius, *nh.x])

    info = nh.solve(f, df, lambda v: v, sub, lambda a, b: float(a @ b), m, gtol, tol, dl, mndl, mxdl, nits)
    ref = g[f"iters_{case}"]
    got = np.array(rows)
    assert info == int(g[f"info_{case}"])
    assert got.shape == ref.shape
    assert np.array_equal(got[:, :2], ref[:, :2])                  # Newton and GMRES iteration counts
    big = ref[:, 2] > 1e-6
    assert np.allclose(got[big, 2], ref[big, 2], rtol=1e-6, atol=0)   # residuals
    assert np.allclose(got[:, 3], ref[:, 3], rtol=1e-6, atol=1e-12)   # trust regions
    assert np.abs(got[:, 4:] - ref[:, 4:]).max() <= 1e-7              # iterates


def test_hookstep_respects_trust_region():
    """hookstep: |y| <= delta, and a delta above the unconstrained least-squares
    step returns that step (mu -> 0 limit within its 10% growth)."""
    from opf.newton import hookstep
    rng = np.random.default_rng(2)
    j = 6
    h = np.zeros((j + 1, j))
    for c in range(j):
        h[:c + 2, c] = rng.standard_normal(c + 2)
    beta = 1.7
    for delta in (1e-3, 0.1, 1.0):
        y, pred = hookstep(h, j, beta, delta)
        assert np.linalg.norm(y) <= delta * (1 + 1e-12)
        r = -(h[:j + 1, :j] @ y)
        r[0] += beta
        assert abs(pred - np.linalg.norm(r)) <= 1e-12 * max(1.0, pred)
    e1 = np.zeros(j + 1)
    e1[0] = beta
    yls = np.linalg.lstsq(h[:j + 1, :j], e1, rcond=None)[0]
    y, _ = hookstep(h, j, beta, 10 * np.linalg.norm(yls))
    assert np.linalg.norm(y - yls) <= 1e-5 * np.linalg.norm(yls)


def test_gmres_solves_linear_system():
    """GMRES without a trust region on a well-conditioned system: the
    residual reached is the one reported and below the tolerance."""
    from opf.newton import Gmres
    rng = np.random.default_rng(4)
    n = 30
    A = np.eye(n) * 3 + rng.standard_normal((n, n)) * 0.3
    b = rng.standard_normal(n)
    gm = Gmres(40, n)
    x, res, _, its, status = gm.solve(b, lambda v: A @ v, lambda v: v, lambda a, c: float(a @ c),
                                      1e-10 * np.linalg.norm(b), 40, 0.0)
    assert status == 0 and its <= 40
    assert np.linalg.norm(A @ x - b) <= 1e-9 * np.linalg.norm(b)


def _orbit_problem(ms=1):
    import opf_oracle as O
    from opf import ic
    from opf.newton import OrbitNewton
    N, K, M, alpha = 16, 6, 4, 1.25
    me = O.build_mesh(N)
    u = ic.seeded_perturbation(me.r, me.intrdr, K, M, alpha, 3, 1e-3)

    class NoStep:
        def steporbit(self, fields, ndts, T, sth, sz, dt1):
            return fields

    return OrbitNewton(NoStep(), N, K, M, 1, alpha, me.intrdr, [u] * ms, 0.1, 0.2, 2.0, [10] * ms), u, O


def test_orbit_vector_round_trip_and_scales():
    """rtz2vec / vec2rtz invert each other; the period entry is scaled to the
    norm of the velocity part (get_scales, newton.f90:266-272)."""
    on, u, _ = _orbit_problem()
    v = on.rtz2vec(*u)
    back = on.vec2rtz(v)
    for a, b in zip(back, u):
        assert np.abs(a - b).max() <= 1e-15 * np.abs(b).max()
    d = on.dotprod(-1, on.x0, on.x0)
    assert abs(on.x0[2] - math.sqrt(d)) <= 1e-14 * math.sqrt(d)
    assert on.shifts(on.x0) == pytest.approx((0.1, 0.2, 2.0), rel=1e-15)


def test_shift_directions_match_finite_shift():
    """getshiftdir(i, x) is the derivative of x under a theta / z shift:
    against the oracle's var_coll_shift by a small angle (central
    difference; the sign follows the field u(theta - s, z - s))."""
    on, u, O = _orbit_problem()
    c = O.Ctx(on.N, on.K, on.M, alpha=on.alpha)
    x = on.rtz2vec(*u)
    h = 1e-6
    for i, (dth, dz) in ((1, (h, 0.0)), (2, (0.0, h))):
        fp = [c.coll_shift(a, -dth, -dz) for a in u]
        fm = [c.coll_shift(a, dth, dz) for a in u]
        fd = (on.rtz2vec(*fp) - on.rtz2vec(*fm)) / (2 * h)
        sd = on.getshiftdir(i, x)
        assert np.abs(fd[8:] - sd[8:]).max() <= 1e-6 * np.abs(sd[8:]).max()


def test_arnoldi_matches_reference():
    """arnold + logeig (arnoldi.f) of the reference on a fixed matrix with a
    known dominant spectrum (golden_arnoldi_mat): the same iteration count
    and exit code, the eigenvalues and exponents to 1e-12, and eigenvectors
    that span the same directions (each compared after removing its complex
    scale, which inverse iteration and the QR algorithm fix differently)."""
    from opf.newton import Arnoldi, logeig
    g = load_golden("arnoldi_mat")
    A, n, kmax, ncgd, tol = g["A"], int(g["n"]), int(g["kmax"]), int(g["ncgd"]), float(g["tol"])
    ar = Arnoldi(n, kmax, ncgd, lambda a, b: float(a @ b), tol)
    sv = 1.0 + np.sin(np.arange(1, n + 1, dtype=np.float64))
    while True:
        ifail, sv = ar.step(sv)
        if ifail != 0:
            break
        sv = A @ sv
    assert (ar.k, ifail) == (int(g["k"]), int(g["ifail"]))
    eig = g["eig"]
    assert np.abs(ar.wr2 - eig[:, 0]).max() <= 1e-12
    assert np.abs(ar.wi2 - eig[:, 1]).max() <= 1e-12
    sr, si = logeig(ar.wr2, ar.wi2, 1.0)
    assert np.abs(sr - eig[:, 2]).max() <= 1e-12 and np.abs(si - eig[:, 3]).max() <= 1e-12
    ref = g["vecs"]
    col = 0
    for i in range(ncgd):
        if abs(eig[i, 1]) > 0:
            v, w = ar.b[col] + 1j * ar.b[col + 1], ref[col] + 1j * ref[col + 1]
            col += 2
        else:
            v, w = ar.b[col] + 0j, ref[col] + 0j
            col += 1
        c = np.vdot(v, w) / np.vdot(v, v)
        assert np.abs(c * v - w).max() <= 1e-8 * np.abs(w).max(), i
        lam = eig[i, 0] + 1j * eig[i, 1]
        assert np.abs(A @ v - lam * v).max() <= 1e-8 * np.abs(v).max(), i


def test_logeig_quadrants():
    """logeig's angle equals atan2 in all four quadrants and on the axes,
    except on the negative real axis, where arnoldi.f:296-300 gives -pi."""
    from opf.newton import logeig
    wr = np.array([0.5, -0.5, -0.5, 0.5, 0.0, 0.0, -0.7])
    wi = np.array([0.3, 0.3, -0.3, -0.3, 0.2, -0.2, 0.0])
    sr, si = logeig(wr, wi, 2.0)
    want = np.arctan2(wi, wr)
    want[-1] = -np.pi
    assert np.allclose(si, want / 2.0, atol=1e-15)
    assert np.allclose(sr, np.log(np.hypot(wr, wi)) / 2.0, atol=1e-15)


def test_residual_wires_shooting_segments():
    """The residual (getrhs, newton.f90:374-393) with ms = 2 and an identity time step:
    segment p's image lands in segment p+1 (cyclically), the constraint rows
    are zero, so f = (x_1 - x_0, x_0 - x_1) in the velocity entries."""
    on, u, _ = _orbit_problem(ms=2)
    n = on.n
    x = on.x0.copy()
    x[n + 8:] = 0.5 * x[n + 8:]                 # make the two segments differ
    f = on.residual(x)
    assert np.all(f[:8] == 0.0) and np.all(f[n:n + 8] == 0.0)
    assert np.array_equal(f[n + 8:], x[8:n] - x[n + 8:])
    assert np.array_equal(f[8:n], x[n + 8:] - x[8:n])


def test_multj_constraint_rows_for_fixed_period_and_frozen_shift():
    """The Jacobian action's special cases (multJ, newton.f90:448-451): with T_fixed and a frozen
    theta shift (mirror or shift-reflect symmetry) the rows 1 and 3 of each
    segment return the update's own entries."""
    from opf.newton import HookstepNewton
    on, u, _ = _orbit_problem()
    on.T_fixed, on.freeze_theta, on.epsJ = True, True, 1e-7
    on.tim_dt = 0.2
    nh = HookstepNewton(on.x0)
    nh.fx = on.residual(nh.x)
    dx = np.random.default_rng(1).standard_normal(on.n)
    y = on.jacobian_action(nh, dx)
    assert y[0] == dx[0] and y[2] == dx[2]
    assert y[1] == on.dotprod(-1, on.getshiftdir(2, on.x0), dx)


def test_host_coll_shift_equals_oracle():
    """The host var_coll_shift the server-backed stepper applies equals the
    oracle's var_coll_shift (variables.f90:487-528) bit for bit, for a z
    shift, a theta shift and both."""
    from opf.newton import coll_shift
    on, u, O = _orbit_problem()
    c = O.Ctx(on.N, on.K, on.M, alpha=on.alpha)
    for dth, dz in ((0.0, 0.37), (0.21, 0.0), (-0.3, 1.1)):
        for a in u:
            assert np.array_equal(coll_shift(a, on.K, on.M, 1, on.alpha, dth, dz), c.coll_shift(a, dth, dz))
