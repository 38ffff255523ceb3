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
