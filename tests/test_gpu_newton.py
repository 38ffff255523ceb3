"""The Newton-Krylov client's device path through the C ABI:
opf_steporbit against the reference's own steporbit call sequence
(golden_orbit_*: oracle/ref_driver.f90 with OPF_ORBIT), batched segments
against the same segments run alone, opf_vel_shift against the oracle's
var_coll_shift, and the orbit search (opf/newton.py) with the device time
step against the same search with the oracle's CPU time step."""
from __future__ import annotations

import numpy as np
import pytest

import opf_oracle as O
from conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu


def _ctx(opf, g, capacity=1, symm=False):
    ctx = opf.Context(int(g["N"]), int(g["K"]), int(g["M"]), capacity=capacity)
    for s in range(capacity):
        if symm:
            ctx.symmetry(s, shiftrefl=True)
        ctx.setup(s, float(g["Re"]), float(g["alpha"]), float(g["dt"]))
    return ctx


@pytest.mark.parametrize("name,symm", [("orbit_mini", False), ("orbit_sr_mini", True)])
def test_steporbit_matches_reference(name, symm):
    import opf
    g = load_golden(name)
    ctx = _ctx(opf, g, symm=symm)
    for i, (ndts, T, sth, sz) in enumerate(g["cases"]):
        ctx.setup(0, float(g["Re"]), float(g["alpha"]), float(g["dt"]))   # tim_dt of the workload
        ctx.put_state(0, *g["ic"])
        print("steporbit", name, i, int(ndts), T, sth, sz, flush=True)
        ctx.steporbit([0], int(ndts), T, sth, sz)
        u = ctx.get_state(0)
        ref = g[f"orbit_{i}"]
        for c in range(3):
            assert rel_err(u[c], ref[c]) <= 1e-11, (i, c)
        # tim_t and tim_step advance only for ndts /= 1 (newton.f90:759-762)
        st = ctx.status(0)
        if int(ndts) == 1:
            assert st["step"] == 0 and st["t"] == 0.0, st
        else:
            assert st["step"] == int(ndts) and abs(st["t"] - T) <= 1e-12 * T, st


def test_steporbit_batched_equals_alone():
    """Three segments with different step counts, periods and shifts in one
    call give the same bits as each run alone."""
    import opf
    g = load_golden("orbit_mini")
    cases = [(20, 0.2, 0.3, 0.7), (25, 0.2, -0.1, 0.4), (12, 0.15, 0.0, -0.2)]
    rng = np.random.default_rng(9)
    states = [[a * (1.0 + 0.1 * rng.standard_normal()) for a in g["ic"]] for _ in cases]
    many = _ctx(opf, g, capacity=3)
    for s, st in enumerate(states):
        many.put_state(s, *st)
    many.steporbit([0, 1, 2], [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases],
                   [c[3] for c in cases])
    for s, (st, c) in enumerate(zip(states, cases)):
        one = _ctx(opf, g)
        one.put_state(0, *st)
        one.steporbit([0], *c)
        a, b = many.get_state(s), one.get_state(0)
        for k in range(3):
            assert np.array_equal(a[k], b[k]), (s, k)


def test_vel_shift_matches_oracle():
    import opf
    g = load_golden("orbit_mini")
    ctx = _ctx(opf, g)
    ctx.put_state(0, *g["ic"])
    ctx.shift(0, 0.37, -1.1)
    u = ctx.get_state(0)
    c = O.Ctx(int(g["N"]), int(g["K"]), int(g["M"]), alpha=float(g["alpha"]))
    for k in range(3):
        want = c.coll_shift(g["ic"][k], 0.37, -1.1)
        assert np.abs(u[k] - want).max() <= 1e-15 * np.abs(want).max(), k
    ctx.shift(0, -0.37, 1.1)
    u = ctx.get_state(0)
    for k in range(3):
        assert rel_err(u[k], g["ic"][k]) <= 1e-14


class OracleStepper:
    """steporbit of the oracle (CPU) behind the client's Stepper interface."""

    def __init__(self, N, K, M, Re, alpha):
        self.args = (N, K, M)
        self.Re, self.alpha = Re, alpha

    def steporbit(self, fields, ndts, T, sth, sz, dt1):
        out = []
        for f, nd, t, a, b in zip(fields, ndts, T, sth, sz):
            c = O.Ctx(*self.args, Re=self.Re, alpha=self.alpha, dt=dt1 if nd == 1 else t / nd)
            c.set_state(*f)
            c.steporbit(nd, t, a, b)
            out.append((c.ur, c.ut, c.uz))
        return out


def test_orbit_search_device_matches_oracle_stepper():
    """Two Newton iterations of the orbit search at (16, 6, 4), T = 0.1 in 10
    steps: the device-stepped search and the oracle-stepped search take the
    same decisions (GMRES counts, trust regions) and reach the same iterate.
    The finite-difference Jacobian amplifies rounding differences by about
    1/epsJ, hence the 1e-5 tolerance."""
    import opf
    from opf.newton import DeviceStepper, OrbitNewton
    g = load_golden("orbit_mini")
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    Re, alpha = float(g["Re"]), float(g["alpha"])
    intrdr = O.build_mesh(N).intrdr
    ctx = _ctx(opf, g)
    runs = []
    for stepper in (DeviceStepper(ctx, [0]), OracleStepper(N, K, M, Re, alpha)):
        on = OrbitNewton(stepper, N, K, M, 1, alpha, intrdr, g["ic"], 0.0, 0.0, 0.1, [10])
        print("orbit search with", type(stepper).__name__, flush=True)
        info, nh = on.solve(m=8, nits=2, rel_err=1e-10, delta=-1.0, epsJ=1e-7)
        runs.append((info, on.history, nh.x))
    (i0, h0, x0), (i1, h1, x1) = runs
    assert i0 == i1 == 2
    assert [r["gits"] for r in h0] == [r["gits"] for r in h1]
    for a, b in zip(h0, h1):
        assert abs(a["tol"] - b["tol"]) <= 1e-5 * b["tol"]
        assert abs(a["delta"] - b["delta"]) <= 1e-5 * b["delta"]
    assert h0[-1]["tol"] < h0[0]["tol"]
    assert np.abs(x0 - x1).max() <= 1e-5 * np.abs(x1).max()


def test_multiple_shooting_segments_share_one_launch():
    """ms = 2: the residual integrates both segments in one batched call; the
    result equals the segments stepped one at a time."""
    import opf
    from opf.newton import DeviceStepper, OrbitNewton
    g = load_golden("orbit_mini")
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    alpha = float(g["alpha"])
    intrdr = O.build_mesh(N).intrdr
    both = _ctx(opf, g, capacity=2)
    on2 = OrbitNewton(DeviceStepper(both, [0, 1]), N, K, M, 1, alpha, intrdr, g["ic"], 0.1, 0.2, 0.1, [10, 10])
    f2 = on2.residual(on2.x0)
    single = _ctx(opf, g)
    n = on2.n
    for p in range(2):
        on1 = OrbitNewton(DeviceStepper(single, [0]), N, K, M, 1, alpha, intrdr, g["ic"], 0.1, 0.2, 0.1, [10])
        on1.scaleU, on1.scaleT = on2.scaleU, on2.scaleT
        yp = on1.steporbit_many([on2.x0[p * n:(p + 1) * n]], [10])[0]
        p1 = (p + 1) % 2
        want = yp - on2.x0[p1 * n:(p1 + 1) * n]
        assert np.array_equal(f2[p1 * n + 8:(p1 + 1) * n], want[8:]), p


def test_floquet_arnoldi_device_matches_oracle_stepper():
    """getEigen's Arnoldi on multA (newton.f90:814-905) about the seeded state
    for T = 0.1 in 10 steps, 16 iterations with the device time step and with
    the oracle's: the same Hessenberg matrix and multiplier estimates (finite
    differences amplify rounding by about 1/epsA: 1e-5)."""
    import opf
    from opf.newton import DeviceStepper, OrbitNewton
    g = load_golden("orbit_mini")
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    Re, alpha = float(g["Re"]), float(g["alpha"])
    me = O.build_mesh(N)
    ctx = _ctx(opf, g)
    res = []
    for stepper in (DeviceStepper(ctx, [0]), OracleStepper(N, K, M, Re, alpha)):
        on = OrbitNewton(stepper, N, K, M, 1, alpha, me.intrdr, g["ic"], 0.0, 0.0, 0.1, [10], r=me.r)
        print("getEigen with", type(stepper).__name__, flush=True)
        res.append(on.get_eigen(on.x0, ncgd=2, kmax=16, epsA=1e-7, rel_err=1e-7))
    a, b = res
    assert a["ifail"] == b["ifail"] == 1 and a["k"] == b["k"] == 16
    ha, hb = a["arnoldi"].h, b["arnoldi"].h
    assert np.abs(ha - hb).max() <= 1e-5 * np.abs(hb).max()
    assert np.abs(a["multipliers"] - b["multipliers"]).max() <= 1e-5


def test_floquet_eigenpairs_satisfy_the_device_map():
    """A converged getEigen on the device about a near-laminar state (energy
    1e-10 at Re = 100, so the map is the linear evolution about Hagen-
    Poiseuille flow; T = 1 in 100 steps): every leading multiplier lies
    inside the unit circle (the flow is linearly stable), and each leading
    multiplier mu and its eigenvector v satisfy multA v = mu v, checked with
    one more device integration.  The finite-difference map carries noise of
    about 1e-7 relative, so the eigenvalue tolerance is 1e-6 and the
    eigenpair residual is checked to 1e-3."""
    import opf
    from opf import ic
    from opf.newton import DeviceStepper, OrbitNewton
    N, K, M, Re, alpha = 16, 6, 4, 100.0, 1.25
    me = O.build_mesh(N)
    u = ic.seeded_perturbation(me.r, me.intrdr, K, M, alpha, 4, 1e-10)
    ctx = opf.Context(N, K, M)
    ctx.setup(0, Re, alpha, 0.01)
    on = OrbitNewton(DeviceStepper(ctx, [0]), N, K, M, 1, alpha, me.intrdr, u, 0.0, 0.0, 1.0, [100], r=me.r)
    r = on.get_eigen(on.x0, ncgd=2, kmax=100, epsA=1e-7, rel_err=1e-6)
    print("getEigen:", r["k"], r["ifail"], r["multipliers"], flush=True)
    assert r["ifail"] == -1
    assert np.all(np.abs(r["multipliers"]) < 1.0)
    ar = r["arnoldi"]
    y0 = on.steporbit_many([on.x0], [100])[0]
    j = 0
    for i in range(2):
        mu = ar.wr2[i] + 1j * ar.wi2[i]
        if ar.wi2[i] != 0.0:
            vr, vi = ar.b[j], ar.b[j + 1]
            j += 2
            Avr, Avi = on.multA(vr, y0, 1e-7), on.multA(vi, y0, 1e-7)
            res = (Avr + 1j * Avi) - mu * (vr + 1j * vi)
            scale = np.linalg.norm((vr + 1j * vi)[8:])
        else:
            v = ar.b[j]
            j += 1
            res = on.multA(v, y0, 1e-7) - mu * v
            scale = np.linalg.norm(v[8:])
        print("eigenpair", i, mu, np.linalg.norm(res[8:]) / scale, flush=True)
        assert np.linalg.norm(res[8:]) <= 1e-3 * scale, i
    assert np.allclose(r["exponents"].real, np.log(np.abs(r["multipliers"])) / 1.0, rtol=1e-12)
    # the least damped Stokes mode of the pipe: sigma = -j_{1,1}^2 / Re (the
    # axisymmetric swirl, and the k = 0, m = 1 mode of the same rate), to the
    # discretisation error at N = 16, dt = 0.01
    j11 = 3.8317059702075123
    assert np.abs(r["exponents"].real + j11 * j11 / Re).max() <= 1e-4 * j11 * j11 / Re


def test_independent_searches_share_launches():
    """Three orbit searches from different states run as threads over one
    context (BatchedStepper): each reaches bit for bit the iterates it reaches
    alone, while their integrations advance together: one launch sequence per
    round of requests instead of one per request."""
    import opf
    from opf.newton import BatchedStepper, DeviceStepper, OrbitNewton, run_searches
    g = load_golden("orbit_mini")
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    alpha = float(g["alpha"])
    me = O.build_mesh(N)
    starts = [[a * s for a in g["ic"]] for s in (1.0, 0.7, 1.3)]
    args = dict(m=6, nits=2, rel_err=1e-10, delta=-1.0, epsJ=1e-7)
    many = _ctx(opf, g, capacity=3)
    batch = BatchedStepper(many, [[0], [1], [2]])
    searches = [OrbitNewton(batch.stepper(i), N, K, M, 1, alpha, me.intrdr, st, 0.0, 0.0, 0.1, [10], r=me.r)
                for i, st in enumerate(starts)]
    print("batched searches", flush=True)
    res = run_searches(batch, searches, **args)
    calls = []
    for i, st in enumerate(starts):
        one = _ctx(opf, g)
        counter = DeviceStepper(one, [0])
        n_calls = [0]
        orig = counter.steporbit

        def counted(*a, _o=orig, _n=n_calls):
            _n[0] += 1
            return _o(*a)

        counter.steporbit = counted
        alone = OrbitNewton(counter, N, K, M, 1, alpha, me.intrdr, st, 0.0, 0.0, 0.1, [10], r=me.r)
        info, nh = alone.solve(**args)
        assert info == res[i][0]
        assert np.array_equal(nh.x, res[i][1].x), i
        assert [r["tol"] for r in alone.history] == [r["tol"] for r in searches[i].history]
        calls.append(n_calls[0])
    # lock step: one launch sequence per round of requests, as many rounds as
    # the busiest search needs
    assert batch.launches == max(calls) < sum(calls)
    print("launches batched", batch.launches, "alone", sum(calls), flush=True)


def test_searches_through_the_request_server():
    """Three orbit searches run as threads whose integrations go through one
    request server (ServerStepper): their requests are batched with each
    other's on the server, and each search reaches bit for bit the iterates
    it reaches alone on a plain context (DeviceStepper)."""
    import threading
    import opf
    from opf.newton import DeviceStepper, OrbitNewton, ServerStepper
    g = load_golden("orbit_mini")
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    Re, alpha = float(g["Re"]), float(g["alpha"])
    me = O.build_mesh(N)
    starts = [[a * s for a in g["ic"]] for s in (1.0, 0.7, 1.3)]
    args = dict(m=6, nits=2, rel_err=1e-10, delta=-1.0, epsJ=1e-7)
    srv = opf.Server([(N, K, M)], block_slots=[4], pool_bytes=opf.bytes_needed(N, K, M, 4))
    steppers = [ServerStepper(srv, 0, Re, alpha) for _ in starts]
    searches = [OrbitNewton(st_, N, K, M, 1, alpha, me.intrdr, st, 0.0, 0.0, 0.1, [10], r=me.r)
                for st_, st in zip(steppers, starts)]
    res = [None] * len(searches)

    def work(i):
        res[i] = searches[i].solve(**args)

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(searches))]
    print("searches through the server", flush=True)
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.close()
    for i, st in enumerate(starts):
        one = _ctx(opf, g)
        alone = OrbitNewton(DeviceStepper(one, [0]), N, K, M, 1, alpha, me.intrdr, st, 0.0, 0.0, 0.1, [10], r=me.r)
        info, nh = alone.solve(**args)
        assert res[i] is not None and info == res[i][0], i
        assert np.array_equal(nh.x, res[i][1].x), i
        assert [r["tol"] for r in alone.history] == [r["tol"] for r in searches[i].history]
    assert all(s.requests > 0 for s in steppers)
