"""Request server: integrations of mixed resolution admitted into free slots,
batched with whatever else is running, retired when done.  Every request's
result is bit-identical to the same integration run alone in a plain
context, whatever it was served with, and the reference's mini case served
with one corrector per step (the reference's run needed exactly one every
step) reproduces the reference's state after 100 steps."""
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
    srv = opf.Server(CLASSES, capacity=[3, 2], chunk=7)
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
    srv = opf.Server(CLASSES[:1], capacity=[1], chunk=50)
    rid = srv.submit(0, 2400.0, 1.25, 0.01, 100, tuple(g["ic"]), corr_fixed=1)
    while not srv.poll(rid)["done"]:
        time.sleep(0.01)
    u = srv.result(rid, 0)
    for k in range(3):
        assert rel_err(u[k], g["state_100"][k]) <= 1e-10, k
    srv.close()


def test_server_rejects_bad_requests():
    import opf
    srv = opf.Server(CLASSES[:1], capacity=[1], chunk=10)
    N, K, M = CLASSES[0]
    st = _ic(N, K, M, 1.25, 1e-3, 5)
    with pytest.raises(opf.serve.ServerError):
        srv.submit(3, 2400.0, 1.25, 0.01, 10, st)
    with pytest.raises(opf.serve.ServerError):
        srv.submit(0, 2400.0, 1.25, -0.01, 10, st)
    srv.close()
