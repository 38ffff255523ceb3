"""Batched mode: many independent states (own Re, alpha, dt, initial
energy) advanced together in one context, every kernel launch covering all
slots.  A state advanced inside a mixed batch gives bit-identical states and
diagnostics to the same state advanced alone, and a batch slot holding the
reference's mini configuration follows the reference's energy series."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu

N, K, M = 16, 6, 4
# (Re, alpha, dt, E0, seed); the first is the reference's mini configuration
MIX = [(2400.0, 1.25, 0.01, 1e-3, 1), (1800.0, 1.0, 0.01, 1e-5, 11), (3000.0, 1.7, 0.005, 1e-2, 12),
       (2100.0, 1.25, 0.01, 3e-4, 13), (2700.0, 1.0, 0.008, 5e-3, 14), (2400.0, 1.7, 0.01, 2e-4, 15)]
NSTEPS = 30


def _ic(ctx, alpha, E0, seed):
    from opf import ic
    me = ctx.mesh()
    return ic.seeded_perturbation(me["r"], me["intrdr"], K, M, alpha, seed, E0)


def test_batch_equals_alone_bitwise():
    import opf
    batch = opf.Context(N, K, M, capacity=len(MIX))
    for s, (Re, al, dt, E0, seed) in enumerate(MIX):
        batch.setup(s, Re, al, dt)
        batch.put_state(s, *_ic(batch, al, E0, seed))
    print("batch of", len(MIX), "slots,", NSTEPS, "steps", flush=True)
    batch.run(NSTEPS, diag_every=1)
    for s, (Re, al, dt, E0, seed) in enumerate(MIX):
        alone = opf.Context(N, K, M)
        alone.setup(0, Re, al, dt)
        alone.put_state(0, *_ic(alone, al, E0, seed))
        alone.run(NSTEPS, diag_every=1)
        ua, ub = alone.get_state(0), batch.get_state(s)
        for c in range(3):
            assert np.array_equal(ua[c], ub[c]), (s, c)
        assert np.array_equal(alone.diag(0), batch.diag(s)), s
        assert batch.status(s)["step"] == NSTEPS


def test_batch_slot_follows_reference_series():
    import opf
    g = load_golden("mini")
    batch = opf.Context(N, K, M, capacity=3)
    for s, (Re, al, dt, E0, seed) in enumerate(MIX[:3]):
        batch.setup(s, Re, al, dt)
        if s == 0:
            ic = g["ic"]
            batch.put_state(s, ic[0], ic[1], ic[2])
        else:
            batch.put_state(s, *_ic(batch, al, E0, seed))
    batch.run(100, diag_every=1)
    d = batch.diag(0)
    rs = g["series"]
    assert np.max(np.abs(d[:, 2] - rs[:, 2]) / rs[:, 2]) <= 1e-10
    assert np.max(np.abs(d[:, 7] - rs[:, 7]) / rs[:, 7]) <= 1e-10
    assert np.array_equal(d[:, 8].astype(int), rs[:, 8].astype(int))
    u = batch.get_state(0)
    for c in range(3):
        assert rel_err(u[c], g["state_100"][c]) <= 1e-10, c


# The 512-state ensemble at (32,24,12): Re ~ U[1800, 3000], alpha in
# {1.0, 1.25, 1.7}, E0 log-uniform in [1e-5, 1e-2], one seed per state,
# dt = 0.01, 200 steps, corrector to convergence.  Slot 0 holds the
# reference's (32,24,12) case (Re 2400, alpha 1.25, E0 1e-3, seed 1).
ENS_N, ENS_K, ENS_M, ENS_DT, ENS_STEPS = 32, 24, 12, 0.01, 200


def ensemble_params(n, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        Re = float(rng.uniform(1800.0, 3000.0))
        alpha = float(rng.choice([1.0, 1.25, 1.7]))
        E0 = float(10.0 ** rng.uniform(-5.0, -2.0))
        out.append((Re, alpha, E0, i + 1))
    out[0] = (2400.0, 1.25, 1e-3, 1)
    return out


def test_ensemble_512_states():
    """512 states in one context, every launch covering all of them.  Slot 0
    meets the reference's own (32,24,12) run: energy and friction velocity
    of steps 0..99 <= 1e-10 relative with equal corrector counts, and the
    state after 100 steps <= 1e-10 normwise.  Eight sampled slots end the
    200 steps bit for bit equal to the same states advanced alone, with equal
    diagnostics records."""
    import opf
    from opf import ic
    g = load_golden("c0")
    params = ensemble_params(512)
    ctx = opf.Context(ENS_N, ENS_K, ENS_M, capacity=len(params))
    me = ctx.mesh()
    states = {}
    for s, (Re, al, E0, seed) in enumerate(params):
        ctx.setup(s, Re, al, ENS_DT)
        f = tuple(g["ic"]) if s == 0 else ic.seeded_perturbation(me["r"], me["intrdr"], ENS_K, ENS_M, al,
                                                                  seed, E0)
        states[s] = f
        ctx.put_state(s, *f)
    print("ensemble: 512 slots set up; 100 steps", flush=True)
    ctx.run(100, diag_every=1)
    d = ctx.diag(0)
    rs = g["series"]
    assert d.shape[0] == 100
    assert np.max(np.abs(d[:, 2] - rs[:, 2]) / rs[:, 2]) <= 1e-10
    assert np.max(np.abs(d[:, 7] - rs[:, 7]) / rs[:, 7]) <= 1e-10
    assert np.array_equal(d[:, 8].astype(int), rs[:, 8].astype(int))
    u = ctx.get_state(0)
    for c in range(3):
        assert rel_err(u[c], g["state_100"][c]) <= 1e-10, c
    print("ensemble: 100 more steps", flush=True)
    ctx.run(ENS_STEPS - 100, diag_every=1)
    for s in range(len(params)):
        st = ctx.status(s)
        assert st["status"] == 0 and st["step"] == ENS_STEPS, (s, st)
    sample = [1, 2, 97, 200, 255, 256, 400, 511]
    for s in sample:
        Re, al, E0, seed = params[s]
        alone = opf.Context(ENS_N, ENS_K, ENS_M)
        alone.setup(0, Re, al, ENS_DT)
        alone.put_state(0, *states[s])
        alone.run(100, diag_every=1)
        alone.run(ENS_STEPS - 100, diag_every=1)
        ua, ub = alone.get_state(0), ctx.get_state(s)
        for c in range(3):
            assert np.array_equal(ua[c], ub[c]), (s, c)
        assert np.array_equal(alone.diag(0), ctx.diag(s)), s
        alone.close()
    ctx.close()
