"""io_save_state / io_load_state (io.f90:98-392) of device-resident slots
through the C ABI: a saved state continues bit for bit in a fresh context,
an adaptive slot takes tim_dt, tim_corr_dt and tim_cfl_dt from the file,
and a state loaded onto another resolution equals the host reading of the
same file and the oracle's io_load_coll (interpolation and resizing)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _ctx(opf, g, slot=0, adaptive=False):
    ctx = opf.Context(int(g["N"]), int(g["K"]), int(g["M"]))
    if adaptive:
        ctx.adaptive(slot)
    ctx.setup(slot, float(g["Re"]), float(g["alpha"]), abs(float(g["dt"])))
    return ctx


def test_saved_state_continues_bitwise(tmp_path):
    import opf
    g = load_golden("mini")
    a = _ctx(opf, g)
    a.put_state(0, *g["ic"], t=0.0)
    a.run(7)
    p = tmp_path / "state0001.cdf.dat"
    a.save_state(0, p)
    info = opf.state_file_info(p)
    assert (info["N"], info["K"], info["M"]) == (int(g["N"]), int(g["K"]), int(g["M"]))
    assert info["dt"] == float(g["dt"]) and info["Re"] == float(g["Re"])
    b = _ctx(opf, g)
    b.load_state(0, p)                      # t from the file
    ua, ub = a.get_state(0), b.get_state(0)
    for c in range(3):
        assert np.array_equal(ua[c], ub[c])
    assert ua[3] == ub[3] == b.status(0)["t"]
    a.run(5)
    b.run(5)
    ua, ub = a.get_state(0), b.get_state(0)
    for c in range(3):
        assert np.array_equal(ua[c], ub[c])
    b.load_state(0, p, t=100.0)             # t given by the caller
    assert b.get_state(0)[3] == 100.0


def test_adaptive_slot_takes_time_step_from_file(tmp_path):
    import opf
    g = load_golden("adapt_mini")
    a = _ctx(opf, g, adaptive=True)
    a.put_state(0, *g["ic"])
    a.run(3)
    ts = a.timestep(0)
    assert ts["dt"] != abs(float(g["dt"]))   # tim_new_tstep changed it
    p = tmp_path / "adapt.cdf.dat"
    a.save_state(0, p)
    b = _ctx(opf, g, adaptive=True)
    b.load_state(0, p)
    tb = b.timestep(0)
    assert tb["dt"] == ts["dt"]
    assert tb["corr_dt"] == ts["corr_dt"] and tb["cfl_dt"] == ts["cfl_dt"]
    c = _ctx(opf, g)                          # fixed tim_dt: d_timestep kept
    c.load_state(0, p)
    assert c.timestep(0)["dt"] == abs(float(g["dt"]))


def test_load_onto_other_resolution_matches_host_reading_and_oracle(tmp_path):
    import opf
    g = load_golden("mini")
    a = _ctx(opf, g)
    a.put_state(0, *g["ic"])
    a.run(4)
    p = tmp_path / "s.cdf.dat"
    a.save_state(0, p)
    N, K, M = int(g["N"]) + 7, int(g["K"]) + 2, int(g["M"]) - 1
    b = opf.Context(N, K, M)
    b.setup(0, float(g["Re"]), float(g["alpha"]), float(g["dt"]))
    b.load_state(0, p)
    host = opf.read_state_file(p, N, K, M)
    ub = b.get_state(0)
    for c, nm in enumerate(("ur", "ut", "uz")):
        assert np.array_equal(ub[c], host[nm]), nm
    # against the oracle's io_load_coll of the saved fields (io.f90:171-232,
    # axis parity S = 1, 1, 0), then vel_imposesym's harmonic-0 projection
    import opf_oracle as O
    ua = a.get_state(0)
    rf, r = opf.host_mesh(int(g["N"]))["r"], b.mesh()["r"]
    want = [O.io_load_coll(ua[c], int(g["K"]), int(g["M"]), rf, K, M, 1, r, S) for c, S in enumerate((1, 1, 0))]
    O.Ctx.imposesym0(*want)
    for c in range(3):
        assert np.abs(ub[c] - want[c]).max() <= 1e-12 * np.abs(want[c]).max(), c
    b.run(2)
    assert b.status(0)["status"] == 0
    # the same radial points, other (K, M): harmonics matched by (k, m), bit for bit
    d = opf.Context(int(g["N"]), K, M)
    d.setup(0, float(g["Re"]), float(g["alpha"]), float(g["dt"]))
    d.load_state(0, p)
    ud = d.get_state(0)
    want = [O.io_load_coll(ua[c], int(g["K"]), int(g["M"]), rf, K, M, 1, rf, S) for c, S in enumerate((1, 1, 0))]
    O.Ctx.imposesym0(*want)
    for c in range(3):
        assert np.array_equal(ud[c], want[c]), c
