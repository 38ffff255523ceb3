"""Symmetric subspaces (b_mirrorsym, b_shiftrefl, b_shiftrott,
parameters.f90:26-28): var_imposesym modes 1-3 (variables.f90:538-618) on
the device, through the C ABI.  The projection of one state against the
reference's own var_imposesym output bit for bit, and the loop with the
shift-reflect symmetry imposed every step against the reference's run built
with b_shiftrefl (energy series 1e-10, state after 1 step 1e-12 and after 20
steps 1e-10)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu


def _ctx(g):
    import opf
    ctx = opf.Context(int(g["N"]), int(g["K"]), int(g["M"]))
    ctx.setup(0, float(g["Re"]), float(g["alpha"]), float(g["dt"]))
    return ctx


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_imposesym_matches_reference_bitwise(mode):
    g = load_golden("sym_mini")
    ctx = _ctx(g)
    ctx.symmetry(0, mirror=mode == 1, shiftrefl=mode == 2, shiftrott=mode == 3)
    ic = g["ic"]
    ctx.put_state(0, ic[0], ic[1], ic[2])
    ctx.imposesym(0)
    u = ctx.get_state(0)
    for c in range(3):
        assert np.array_equal(u[c], g[f"sym_{mode}"][c]), c


def test_shift_reflect_loop_matches_reference():
    g = load_golden("sr_mini")
    ctx = _ctx(g)
    ctx.symmetry(0, shiftrefl=True)
    ic = g["ic"]
    ctx.put_state(0, ic[0], ic[1], ic[2])
    ctx.run(1, diag_every=1)
    u1 = ctx.get_state(0)
    for c in range(3):
        assert rel_err(u1[c], g["state_1"][c]) <= 1e-12, c
    ctx.run(int(g["nsteps"]) - 1, diag_every=1)
    d = ctx.diag(0)
    rs = g["series"]
    assert d.shape[0] == rs.shape[0]
    assert np.max(np.abs(d[:, 2] - rs[:, 2]) / rs[:, 2]) <= 1e-10
    assert np.array_equal(d[:, 8].astype(int), rs[:, 8].astype(int))
    u = ctx.get_state(0)
    for c in range(3):
        assert rel_err(u[c], g["state_20"][c]) <= 1e-10, c


def test_symmetry_off_leaves_the_loop_unchanged():
    """A slot without symmetries in a context where another slot has one:
    bitwise the same as alone."""
    import opf
    g = load_golden("mini")
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    ic = g["ic"]
    a = opf.Context(N, K, M, capacity=2)
    b = opf.Context(N, K, M)
    for ctx, slots in ((a, (0, 1)), (b, (0,))):
        for s in slots:
            ctx.setup(s, float(g["Re"]), float(g["alpha"]), float(g["dt"]))
            ctx.put_state(s, ic[0], ic[1], ic[2])
    a.symmetry(1, shiftrefl=True)
    a.run(5)
    b.run(5)
    ua, ub = a.get_state(0), b.get_state(0)
    for c in range(3):
        assert np.array_equal(ua[c], ub[c]), c
