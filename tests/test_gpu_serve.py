"""Request server: integrations of mixed resolution admitted into free slots
from one device memory pool, between iterations of whatever else is
running, retired when done.  Every request's result is bit-identical to the
same integration run alone in a plain context, whatever it was served with,
and the reference's mini case served with one corrector per step (the
reference's run needed exactly one every step) reproduces the reference's
state after 100 steps."""
from __future__ import annotations

import time

import numpy as np
import pytest

from conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu

CLASSES = [(16, 6, 4), (32, 24, 12)]
# (class, Re, alpha, dt, nsteps, corr_fixed, E0, seed)
REQS = [(0, 2400.0, 1.25, 0.01, 100, 1, None, None),      # the reference's mini case
        (0, 1800.0, 1.0, 0.01, 23, 1, 1e-4, 21), (1, 3000.0, 1.25, 0.01, 40, 1, 1e-3, 22),
        (0, 2700.0, 1.7, 0.005, 9, 0, 1e-3, 23), (1, 2200.0, 1.0, 0.01, 17, 1, 1e-5, 24),
        (0, 2500.0, 1.25, 0.008, 31, 1, 3e-3, 25), (1, 2900.0, 1.7, 0.01, 5, 0, 2e-3, 26),
        (0, 2000.0, 1.0, 0.01, 12, 1, 5e-4, 27)]


def _ic(N, K, M, alpha, E0, seed):
    import opf
    from opf import ic
    me = opf.host_mesh(N)
    return ic.seeded_perturbation(me["r"], me["intrdr"], K, M, alpha, seed, E0)


def test_served_equals_alone_and_reference():
    import opf
    g = load_golden("mini")
    states = []
    for (c, Re, al, dt, n, cf, E0, seed) in REQS:
        N, K, M = CLASSES[c]
        states.append(tuple(g["ic"]) if E0 is None else _ic(N, K, M, al, E0, seed))
    # a pool of one block per class: 5 + 3 requests wait for 3 + 2 slots
    pool = opf.bytes_needed(*CLASSES[0], 3) + opf.bytes_needed(*CLASSES[1], 2)
    srv = opf.Server(CLASSES, block_slots=[3, 2], pool_bytes=pool)
    ids = [srv.submit(c, Re, al, dt, n, st, corr_fixed=cf)
           for (c, Re, al, dt, n, cf, _, _), st in zip(REQS, states)]
    print("submitted", len(ids), "requests", flush=True)
    deadline = time.time() + 60
    while not all(srv.poll(i)["done"] for i in ids):
        assert time.time() < deadline, srv.stats()
        time.sleep(0.01)
    assert srv.stats()["state_steps"] == sum(r[4] for r in REQS)
    for rid, (c, Re, al, dt, n, cf, _, _), st in zip(ids, REQS, states):
        p = srv.poll(rid)
        assert p["status"] == 0 and p["t_admit"] >= p["t_submit"] and p["t_done"] >= p["t_admit"]
        u = srv.result(rid, c)
        N, K, M = CLASSES[c]
        ctx = opf.Context(N, K, M)
        ctx.setup(0, Re, al, dt)
        ctx.put_state(0, *st[:3])
        ctx.run(n, corr_fixed=cf)
        v = ctx.get_state(0)
        for k in range(3):
            assert np.array_equal(u[k], v[k]), (rid, k)
        assert abs(u[3] - n * dt) < 1e-9 * max(1.0, n * dt)
    srv.close()
    # the reference's case
    srv = opf.Server(CLASSES[:1], block_slots=[1], pool_bytes=opf.bytes_needed(*CLASSES[0], 1))
    rid = srv.submit(0, 2400.0, 1.25, 0.01, 100, tuple(g["ic"]), corr_fixed=1)
    while not srv.poll(rid)["done"]:
        time.sleep(0.01)
    u = srv.result(rid, 0)
    for k in range(3):
        assert rel_err(u[k], g["state_100"][k]) <= 1e-10, k
    srv.close()


def test_pool_shared_between_classes():
    """A pool that holds one block of either class but not both: requests of
    the second class wait until the first class's block is idle and released,
    then run; the pool never holds more than its limit; results equal the
    integrations run alone."""
    import opf
    b0, b1 = opf.bytes_needed(*CLASSES[0], 2), opf.bytes_needed(*CLASSES[1], 2)
    pool = max(b0, b1) + min(b0, b1) // 2
    srv = opf.Server(CLASSES, block_slots=[2, 2], pool_bytes=pool)
    reqs = [(0, 2400.0, 1.25, 0.01, 30, 1, 1e-3, 31), (1, 2600.0, 1.0, 0.01, 20, 1, 1e-3, 32),
            (0, 2100.0, 1.7, 0.01, 15, 1, 1e-4, 33), (1, 2300.0, 1.25, 0.01, 10, 0, 2e-3, 34)]
    states = [_ic(*CLASSES[c], al, E0, seed) for (c, _, al, _, _, _, E0, seed) in reqs]
    ids = [srv.submit(c, Re, al, dt, n, st, corr_fixed=cf) for (c, Re, al, dt, n, cf, _, _), st in zip(reqs, states)]
    peak = 0
    deadline = time.time() + 60
    while not all(srv.poll(i)["done"] for i in ids):
        assert time.time() < deadline, srv.stats()
        pl = srv.pool()
        peak = max(peak, pl["used"])
        assert pl["used"] <= pool and pl["blocks"] <= 1, pl
        time.sleep(0.002)
    assert peak > 0
    for rid, (c, Re, al, dt, n, cf, _, _), st in zip(ids, reqs, states):
        assert srv.poll(rid)["status"] == 0
        u = srv.result(rid, c)
        ctx = opf.Context(*CLASSES[c])
        ctx.setup(0, Re, al, dt)
        ctx.put_state(0, *st)
        ctx.run(n, corr_fixed=cf)
        v = ctx.get_state(0)
        ctx.close()
        for k in range(3):
            assert np.array_equal(u[k], v[k]), (rid, k)
    srv.close()


def test_server_rejects_bad_requests():
    import opf
    with pytest.raises(opf.serve.ServerError):            # a pool smaller than one block
        opf.Server(CLASSES[:1], block_slots=[1], pool_bytes=opf.bytes_needed(*CLASSES[0], 1) - 1)
    srv = opf.Server(CLASSES[:1], block_slots=[1], pool_bytes=opf.bytes_needed(*CLASSES[0], 1))
    N, K, M = CLASSES[0]
    st = _ic(N, K, M, 1.25, 1e-3, 5)
    with pytest.raises(opf.serve.ServerError):
        srv.submit(3, 2400.0, 1.25, 0.01, 10, st)
    with pytest.raises(opf.serve.ServerError):
        srv.submit(0, 2400.0, 1.25, -0.01, 10, st)
    srv.close()


# The serving mix: sizes (32,24,12), (48,48,24), (64,64,48) in ratio 5:3:2,
# Poisson arrivals, Re ~ U[2000, 5300], alpha in {0.5, 1.25}, one corrector
# per step.  Horizons are cut from T ~ U[5, 50] time units to U[0.1, 1.0]
# at dt = 0.01 (10-100 steps) so that every request can also be run alone
# within the test's time.
MIX_CLASSES = [(32, 24, 12), (48, 48, 24), (64, 64, 48)]


def _mix_requests(n, seed=77):
    rng = np.random.default_rng(seed)
    ratio = np.array([5.0, 3.0, 2.0])
    out, t = [], 0.0
    for i in range(n):
        cls = int(rng.choice(3, p=ratio / ratio.sum()))
        Re = float(rng.uniform(2000.0, 5300.0))
        alpha = float(rng.choice([0.5, 1.25]))
        nsteps = int(round(float(rng.uniform(0.1, 1.0)) / 0.01))
        E0 = float(10.0 ** rng.uniform(-5.0, -3.0))
        t += float(rng.exponential(1.0 / 40.0))
        out.append(dict(cls=cls, Re=Re, alpha=alpha, dt=0.01, nsteps=nsteps, cf=1, E0=E0, seed=500 + i, at=t))
    return out


def test_serving_mix_alone_equals_in_mix_and_reference():
    """Requests of the three serving sizes arrive as a Poisson stream while
    others run.  Every result is bit for bit the same integration run alone
    in a plain context.  Two requests carry the reference's own cases and
    are integrated one step with the corrector to convergence: the
    (48,48,24) one meets golden_s48 and the (64,64,48) one golden_c1 after
    step 1 (sampled coefficients <= 1e-12 normwise relative)."""
    import opf
    reqs = _mix_requests(24)
    meshes = [opf.host_mesh(N) for (N, _, _) in MIX_CLASSES]
    from opf import ic
    states = []
    for r in reqs:
        N, K, M = MIX_CLASSES[r["cls"]]
        me = meshes[r["cls"]]
        states.append(ic.seeded_perturbation(me["r"], me["intrdr"], K, M, r["alpha"], r["seed"], r["E0"]))
    goldens = {1: load_golden("s48"), 2: load_golden("c1")}
    for k, (cls, g) in enumerate(goldens.items()):
        j = 3 + 7 * k
        N, K, M = MIX_CLASSES[cls]
        me = meshes[cls]
        reqs[j].update(cls=cls, Re=float(g["Re"]), alpha=float(g["alpha"]), dt=float(g["dt"]), nsteps=1, cf=0,
                       golden=cls)
        states[j] = ic.seeded_perturbation(me["r"], me["intrdr"], K, M, float(g["alpha"]), int(g["seed"]),
                                           float(g["E0"]))
    pool = sum(opf.bytes_needed(N, K, M, b) for (N, K, M), b in zip(MIX_CLASSES, [6, 4, 2]))
    srv = opf.Server(MIX_CLASSES, block_slots=[6, 4, 2], pool_bytes=pool)
    t0 = time.time()
    ids = []
    for r, st in zip(reqs, states):
        wait = r["at"] - (time.time() - t0)
        if wait > 0:
            time.sleep(wait)
        ids.append(srv.submit(r["cls"], r["Re"], r["alpha"], r["dt"], r["nsteps"], st, corr_fixed=r["cf"]))
    print("serving mix: submitted", len(ids), "requests", flush=True)
    deadline = time.time() + 120
    while not all(srv.poll(i)["done"] for i in ids):
        assert time.time() < deadline, srv.stats()
        time.sleep(0.01)
    results = []
    for rid, r in zip(ids, reqs):
        p = srv.poll(rid)
        assert p["status"] == 0, (rid, p)
        results.append(srv.result(rid, r["cls"]))
    srv.close()
    print("serving mix: checking each request alone", flush=True)
    for r, st, u in zip(reqs, states, results):
        N, K, M = MIX_CLASSES[r["cls"]]
        ctx = opf.Context(N, K, M)
        ctx.setup(0, r["Re"], r["alpha"], r["dt"])
        ctx.put_state(0, *st[:3])
        ctx.run(r["nsteps"], corr_fixed=r["cf"])
        v = ctx.get_state(0)
        ctx.close()
        for c in range(3):
            assert np.array_equal(u[c], v[c]), (r, c)
        if "golden" in r:
            g = goldens[r["golden"]]
            sel = g["sel"]
            for c in range(3):
                assert np.abs(u[c][:, sel] - g["state_1"][c]).max() / g["state_1_norm"][c] <= 1e-12, (r, c)
