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
    return ic.seeded_perturbation(me["r"], me["intrdr"], ctx.K, ctx.M, alpha, seed, E0)


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


def _mix_context(opf, graph, compact, slow_dt=0.08, streams=1, grid=(N, K, M)):
    """The MIX slots plus one slot at a large time step, which takes several
    correctors per step and so iterates on alone at the end of every run."""
    ctx = opf.Context(*grid, capacity=len(MIX) + 1)
    ctx.launch_mode(graph=graph, compact=compact, streams=streams)
    for s, (Re, al, dt, E0, seed) in enumerate(MIX + [(2400.0, 1.25, slow_dt, 1e-3, 21)]):
        ctx.setup(s, Re, al, dt)
        ctx.put_state(s, *_ic(ctx, al, E0, seed))
    return ctx


@pytest.mark.parametrize("grid,slow_dt", [((N, K, M), 0.08), ((32, 24, 12), 0.03)])
def test_graph_replay_and_compaction_equal_eager_launches(grid, slow_dt):
    """The time loop replayed as a captured HIP graph (4 iterations per
    graph, the remainder launched eagerly), and split over two streams, with the run's tail
    launched over the list of slots still iterating, against eager launches
    over the whole range: every slot's state and diagnostics are bit for bit
    equal after each of three runs, with host reads of the states and an
    eager launch (vel_imposesym) between the runs.  At (32,24,12) the
    evaluation is the whole-plane kernel, so it is captured, compacted and
    split too."""
    import opf
    eager = _mix_context(opf, 0, 0, slow_dt, grid=grid)
    fast = _mix_context(opf, 4, 1, slow_dt, grid=grid)
    split = _mix_context(opf, 0, 1, slow_dt, streams=2, grid=grid)
    for r in range(3):
        print(f"run {r}: 10 steps, eager and graph", flush=True)
        eager.run(10, diag_every=1)
        fast.run(10, diag_every=1)
        split.run(10, diag_every=1)
        for s in range(len(MIX) + 1):
            ue = eager.get_state(s)
            for other in (fast, split):
                uf = other.get_state(s)
                for c in range(3):
                    assert np.array_equal(ue[c], uf[c]), (r, s, c)
                assert np.array_equal(eager.diag(s), other.diag(s)), (r, s)
                assert eager.status(s) == other.status(s), (r, s)
        for ctx in (eager, fast, split):
            ctx.imposesym(0)
    st = fast.status(len(MIX))
    assert st["last_ncorr"] >= 2, st            # the slow slot iterated on alone
    counts = fast.launch_mode()
    assert counts["graph_launches"] > 0 and counts["compact_iterations"] > 0, counts
    assert eager.launch_mode()["graph_launches"] == 0


def test_launch_shapes_do_not_change_results(tmp_path, monkeypatch):
    """Every launch shape the autotuner may choose (step pairs per
    workgroup, z-transform tile widths, even theta tile widths) gives bit for
    bit the same states after 5 steps at (32,24,12) and (64,64,48); an odd
    theta tile is refused; a tuning table row is taken by opf_ctx_create for
    its (N, K, M) and capacity."""
    import opf
    from opf import ic
    for (n, k, m, Re, al, dt) in ((32, 24, 12, 2400.0, 1.25, 0.01), (64, 64, 48, 5300.0, 0.5, 0.005)):
        base = opf.Context(n, k, m)
        p0, tm0, tz0 = base.tile_config()
        base.close()
        me = opf.host_mesh(n)
        f = ic.seeded_perturbation(me["r"], me["intrdr"], k, m, al, 5, 1e-3)
        ref = None
        tz_half = tz0 if tz0 % 2 else max(2, 2 * (tz0 // 4))
        for shape in ((p0, tm0, tz0), (1, tm0, tz0), (2, max(1, tm0 // 2), tz0), (4, tm0, tz_half),
                      (2, 1, 2 if tz0 % 2 == 0 else tz0)):
            print(f"({n},{k},{m}) shape {shape}", flush=True)
            ctx = opf.Context(n, k, m, pairs=shape[0], tm=shape[1], tz=shape[2])
            assert ctx.tile_config() == shape
            ctx.setup(0, Re, al, dt)
            ctx.put_state(0, *f)
            ctx.run(5)
            u = ctx.get_state(0)
            ctx.close()
            if ref is None:
                ref = u
            for c in range(3):
                assert np.array_equal(u[c], ref[c]), (shape, c)
    with pytest.raises(opf.OpfError):
        opf.Context(32, 24, 12, tz=9)
    table = tmp_path / "tuning.tsv"
    table.write_text("# test table\n32 24 12 1 1 3 2 0 0 0\n32 24 12 4 2 4 6 16 0 0\n")
    monkeypatch.setenv("OPF_TUNING", str(table))
    assert opf.Context(32, 24, 12).tile_config() == (1, 3, 2)
    c6 = opf.Context(32, 24, 12, capacity=6)
    assert c6.tile_config() == (2, 4, 6)
    assert c6.launch_mode()["graph"] == 16
    assert opf.Context(32, 24, 12).launch_mode()["graph"] == 0


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
