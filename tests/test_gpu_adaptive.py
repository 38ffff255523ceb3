"""The adaptive time step (d_timestep < 0, main.f90:24-29) through the C ABI:
vel_maxtstep (velocity.f90:438-481) and tim_new_tstep (timestep.f90:259-278)
on the device, vel_matrices where tim_dt changes.  Against the reference's
own adaptive runs (golden_adapt_*: dt changes at steps 0 and 100 for
(16,6,4), at 0, 1 and 101 for (32,24,12)): the time series t (so the dt of
every step), E and u_tau to 1e-10 relative, equal corrector counts, the final
state to 1e-10 and the final tim_dt / tim_cfl_dt to 1e-12."""
from __future__ import annotations

import numpy as np
import pytest

import opf_oracle as O
from conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu


def _adaptive_ctx(opf, g, capacity=1, slot=0):
    ctx = opf.Context(int(g["N"]), int(g["K"]), int(g["M"]), capacity=capacity)
    ctx.adaptive(slot)
    ctx.setup(slot, float(g["Re"]), float(g["alpha"]), abs(float(g["dt"])))
    ic = g["ic"]
    ctx.put_state(slot, ic[0], ic[1], ic[2])
    return ctx


def _check_against(g, d, u, ts):
    rs = g["series"]
    n = int(g["nsteps"])
    assert d.shape[0] == n
    assert np.array_equal(d[:, 0], rs[:, 0])
    assert np.max(np.abs(d[1:, 1] - rs[1:, 1]) / rs[1:, 1]) <= 1e-12           # t: every step's dt
    assert np.max(np.abs(d[:, 2] - rs[:, 2]) / rs[:, 2]) <= 1e-10                # E
    assert np.max(np.abs(d[:, 7] - rs[:, 7]) / rs[:, 7]) <= 1e-10                # u_tau
    assert np.array_equal(d[:, 8].astype(int), rs[:, 8].astype(int))            # ncorr
    ref = g[f"state_{n}"]
    for c in range(3):
        assert rel_err(u[c], ref[c]) <= 1e-10, c
    last = g["dtser"][-1]       # step tim_dt tim_cfl_dt tim_cfl_dir tim_corr_dt tim_new_dt
    assert abs(ts["dt"] - last[1]) <= 1e-12 * last[1]
    assert ts["cfl_dir"] == int(last[3])


@pytest.mark.parametrize("name", ["adapt_mini", "adapt_c0"])
def test_adaptive_run_matches_reference(name):
    import opf
    g = load_golden(name)
    ctx = _adaptive_ctx(opf, g)
    ctx.run(int(g["nsteps"]), diag_every=1)
    ts = ctx.timestep(0)
    _check_against(g, ctx.diag(0), ctx.get_state(0), ts)
    # tim_cfl_dt recorded before the last step's predictor
    last = g["dtser"][-1]
    assert abs(ts["cfl_dt"] - last[2]) <= 1e-12 * last[2]


def test_adaptive_and_fixed_slots_share_a_context():
    """Slot 0 adaptive, slot 1 on the fixed tim_dt of golden_mini: the gated
    vel_matrices touch slot 0 only, each slot meets its own reference run."""
    import opf
    ga, gf = load_golden("adapt_mini"), load_golden("mini")
    ctx = _adaptive_ctx(opf, ga, capacity=2, slot=0)
    ctx.setup(1, float(gf["Re"]), float(gf["alpha"]), float(gf["dt"]))
    ic = gf["ic"]
    ctx.put_state(1, ic[0], ic[1], ic[2])
    ctx.run(int(ga["nsteps"]), diag_every=1)
    _check_against(ga, ctx.diag(0), ctx.get_state(0), ctx.timestep(0))
    d1 = ctx.diag(1)[:100]
    rs = gf["series"]
    assert np.max(np.abs(d1[:, 2] - rs[:, 2]) / rs[:, 2]) <= 1e-10
    assert ctx.timestep(1)["dt"] == float(gf["dt"])


def test_vel_maxtstep_host_driven_matches_oracle():
    import opf
    g = load_golden("c0")
    ctx = opf.Context(int(g["N"]), int(g["K"]), int(g["M"]))
    ctx.setup(0, float(g["Re"]), float(g["alpha"]), float(g["dt"]))
    ic = g["ic"]
    ctx.put_state(0, ic[0], ic[1], ic[2])
    ctx.transform(0)
    ctx.nonlinear(0)
    cfl, d = ctx.vel_maxtstep(0)
    oc = O.Ctx(int(g["N"]), int(g["K"]), int(g["M"]), Re=float(g["Re"]), alpha=float(g["alpha"]),
               dt=float(g["dt"]))
    oc.set_state(ic[0], ic[1], ic[2])
    oc.imposesym()
    oc.transform()
    ocfl, od = oc.maxtstep()
    assert d == od
    assert abs(cfl - ocfl) <= 1e-12 * ocfl


def test_vel_matrices_equals_fresh_setup():
    import opf
    g = load_golden("mini")
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    ic = g["ic"]
    a = opf.Context(N, K, M)
    a.setup(0, float(g["Re"]), float(g["alpha"]), 0.01)
    a.vel_matrices(0, 0.0071)
    b = opf.Context(N, K, M)
    b.setup(0, float(g["Re"]), float(g["alpha"]), 0.0071)
    for ctx in (a, b):
        ctx.put_state(0, ic[0], ic[1], ic[2])
        ctx.run(3)
    ua, ub = a.get_state(0), b.get_state(0)
    for c in range(3):
        assert np.array_equal(ua[c], ub[c]), c
    assert a.timestep(0)["dt"] == 0.0071
