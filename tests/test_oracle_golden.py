"""The oracle against the compiled reference's golden vectors.

The fixtures were written by the reference Fortran itself (oracle/_ref, via
tests/golden/make_golden.py) from the same seeded initial states.  Bars:
one-step spectral coefficients <= 1e-12 normwise relative per field; energy
and friction velocity series <= 1e-10 relative; identical corrector counts.
"""
import numpy as np
import pytest

import opf_oracle as O
from conftest import load_golden, rel_err

TOL_STEP = 1e-12
TOL_DIAG = 1e-10


def _ctx(g):
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    return O.Ctx(N, K, M, Re=float(g["Re"]), alpha=float(g["alpha"]), dt=float(g["dt"]))


MESH_TOL = 5e-13   # per stencil: max|difference| / max|stencil entry|


def _stencils(a, key, N):
    """Rows that are one stencil each: a band column, a row of dr0/dr1."""
    shape = {"intrdr": (1, N), "dr0": (5, 5), "dr1": (5, 5), "dr1band": (N, 9), "dr2band": (N, 9),
             "radlap": (N, 9)}[key]
    return np.asarray(a).reshape(shape)


@pytest.mark.parametrize("N", [16, 32, 48, 64, 96])
def test_mesh_matches_reference_mesh(N):
    """The oracle's and the library's mes_D against the reference's own
    mes_precompute (dumped by oracle/ref_driver.f90 from the compiled
    reference, meshs.f90:66-225).  The points are equal bit for bit.  The
    weights come from a dense inverse (mes_mat_invert, meshs.f90:231-243),
    which the reference takes from MKL's dgetrf/dgetri and the restatement
    from unblocked LAPACK; per stencil they agree to a few 1e-14 of the
    stencil's largest weight (the Vandermonde-like systems lose digits in
    the small weights, so the bar is normwise per stencil)."""
    from opf import host_mesh
    g = load_golden("mesh")
    lib = host_mesh(N)
    orc = O.build_mesh(N)
    assert np.array_equal(lib["r"], g[f"r_{N}"])
    assert np.array_equal(orc.r, g[f"r_{N}"])
    # the library's arrays in the dump's column-major order
    mine = {"intrdr": lib["intrdr"], "dr0": lib["dr0"].T.ravel(), "dr1": lib["dr1"].T.ravel(),
            "dr1band": lib["dr1band"].T.ravel(), "radlap": lib["radlap"].T.ravel()}
    for key, a in mine.items():
        ref = _stencils(g[f"{key}_{N}"], key, N)
        a = _stencils(a, key, N)
        scale = np.abs(ref).max(axis=1)
        assert np.all(np.abs(a - ref).max(axis=1) <= MESH_TOL * scale), (N, key)
    # the oracle's intrdr and bands (dr(1), dr(2), radLap) against the same dump
    theirs = {"intrdr": orc.intrdr, "dr0": orc.dr0.T.ravel(), "dr1": orc.dr1.T.ravel(),
              "dr1band": orc.dr[1].T.ravel(), "dr2band": orc.dr[2].T.ravel(), "radlap": orc.radLap.T.ravel()}
    for key, a in theirs.items():
        ref = _stencils(g[f"{key}_{N}"], key, N)
        a = _stencils(a, key, N)
        assert np.all(np.abs(a - ref).max(axis=1) <= MESH_TOL * np.abs(ref).max(axis=1)), (N, key)


@pytest.mark.parametrize("name", ["mini", "c0"])
def test_one_step_parity(name):
    g = load_golden(name)
    ctx = _ctx(g)
    ic = g["ic"]
    ctx.set_state(ic[0], ic[1], ic[2])
    ctx.run(1)
    ref = g["state_1"]
    for c, u in enumerate((ctx.ur, ctx.ut, ctx.uz)):
        assert rel_err(u, ref[c]) <= TOL_STEP, (name, c)


def test_stage_parity_mini():
    """Step-1 intermediates: N after each nonlinear evaluation, u after the
    predictor and the corrector (localises a mismatch to a stage)."""
    g = load_golden("mini")
    ctx = _ctx(g)
    ic = g["ic"]
    ctx.set_state(ic[0], ic[1], ic[2])
    stages = {}

    ctx.imposesym()
    ctx.transform()
    ctx.nonlinear()
    stages["01N"] = (ctx.Nr, ctx.Nt, ctx.Nz)
    ctx.predictor()
    stages["02u"] = (ctx.ur, ctx.ut, ctx.uz)
    ctx.transform()
    ctx.nonlinear()
    stages["03N"] = (ctx.Nr, ctx.Nt, ctx.Nz)
    ctx.dterr = 0.0
    ctx.corrector()
    stages["04u"] = (ctx.ur, ctx.ut, ctx.uz)
    for tag, fields in stages.items():
        ref = g[f"stage_{tag}"]
        for c in range(3):
            assert rel_err(fields[c], ref[c]) <= TOL_STEP, (tag, c)


def test_multi_step_mini():
    g = load_golden("mini")
    ctx = _ctx(g)
    ic = g["ic"]
    ctx.set_state(ic[0], ic[1], ic[2])
    states = {}
    ser = ctx.run(100, record=lambda c: states.__setitem__(c.step, (c.ur.copy(), c.ut.copy(), c.uz.copy())))
    for s in (1, 2, 10, 100):
        ref = g[f"state_{s}"]
        for c in range(3):
            assert rel_err(states[s][c], ref[c]) <= 1e-11, (s, c)
    rs = g["series"]
    E = np.array([d["E"] for d in ser])
    Ufr = np.array([d["Ufr"] for d in ser])
    nc = np.array([d["ncorr"] for d in ser])
    assert np.max(np.abs(E - rs[:, 2]) / rs[:, 2]) <= TOL_DIAG
    assert np.max(np.abs(Ufr - rs[:, 7]) / rs[:, 7]) <= TOL_DIAG
    assert np.array_equal(nc, rs[:, 8].astype(int))


def test_series_c0_first_steps():
    g = load_golden("c0")
    ctx = _ctx(g)
    ic = g["ic"]
    ctx.set_state(ic[0], ic[1], ic[2])
    ser = ctx.run(10)
    rs = g["series"][:10]
    for i, d in enumerate(ser):
        assert abs(d["E"] - rs[i, 2]) <= TOL_DIAG * rs[i, 2]
        assert abs(d["Ufr"] - rs[i, 7]) <= TOL_DIAG * rs[i, 7]
        assert abs(d["Pr0"] - rs[i, 5]) <= 1e-10 * max(abs(rs[i, 5]), 1e-6)
        assert d["ncorr"] == int(rs[i, 8])


def test_one_step_s48_sampled():
    g = load_golden("s48")
    ctx = _ctx(g)
    # The fixture keeps every 37th harmonic; regenerate the full initial state
    # with the package's generator and check it reproduces the stored sample.
    from opf import ic as icgen
    ur, ut, uz = icgen.seeded_perturbation(ctx.mesh.r, ctx.mesh.intrdr, ctx.K, ctx.M,
                                           ctx.alpha, int(g["seed"]), float(g["E0"]))
    sel = g["sel"]
    assert rel_err(np.stack([ur, ut, uz])[:, :, sel], g["ic_sel"]) <= 1e-15
    ctx.set_state(ur, ut, uz)
    ctx.run(1)
    ref = g["state_1"]
    for c, u in enumerate((ctx.ur, ctx.ut, ctx.uz)):
        den = g["state_1_norm"][c]
        assert np.abs(u[:, sel] - ref[c]).max() / den <= TOL_STEP


def test_adaptive_timestep_mini():
    """The d_timestep < 0 loop (main.f90:24-29): vel_maxtstep and
    tim_new_tstep of the oracle against the reference's own per-step
    timestep quantities over 120 steps (dt changes at steps 0 and 100: the
    first-step CFL cap and a growth after the 99-step hold), the energy
    series and the final state."""
    g = load_golden("adapt_mini")
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    ctx = O.Ctx(N, K, M, Re=float(g["Re"]), alpha=float(g["alpha"]), dt=abs(float(g["dt"])))
    ic = g["ic"]
    ctx.set_state(ic[0], ic[1], ic[2])
    ser = ctx.run(int(g["nsteps"]), adaptive=True)
    ref = g["dtser"]            # step tim_dt tim_cfl_dt tim_cfl_dir tim_corr_dt tim_new_dt
    assert len(ser) == ref.shape[0]
    new = np.array([s["new_dt"] for s in ser])
    assert np.array_equal(new, ref[:, 5].astype(int))
    assert np.nonzero(new)[0].tolist() == [0, 100]
    assert np.array_equal([s["cfl_dir"] for s in ser], ref[:, 3].astype(int))
    for key, col in (("dt", 1), ("cfl_dt", 2), ("corr_dt", 4)):
        got = np.array([s[key] for s in ser])
        assert np.max(np.abs(got - ref[:, col]) / np.maximum(np.abs(ref[:, col]), 1e-300)) <= 1e-12, key
    E = np.array([s["E"] for s in ser])
    assert np.max(np.abs(E - g["series"][:, 2]) / g["series"][:, 2]) <= TOL_DIAG
    for c, u in enumerate((ctx.ur, ctx.ut, ctx.uz)):
        assert rel_err(u, g["state_120"][c]) <= 1e-10, c


def test_imposesym_modes_match_reference():
    """var_imposesym(1..3) (variables.f90:538-618) of the oracle against the
    reference's own output on the same state, bit for bit; each mode is an
    idempotent projection (its reflection is exact, so P(P u) == P u)."""
    g = load_golden("sym_mini")
    ctx = O.Ctx(int(g["N"]), int(g["K"]), int(g["M"]))
    ic = g["ic"]
    for i in (1, 2, 3):
        got = ctx.imposesym_mode(i, ic[0], ic[1], ic[2])
        for c in range(3):
            assert np.array_equal(got[c], g[f"sym_{i}"][c]), (i, c)
        again = ctx.imposesym_mode(i, *got)
        for c in range(3):
            assert np.array_equal(again[c], got[c]), (i, c)


def test_shift_reflect_run_mini():
    """b_shiftrefl (parameters.f90:27): vel_imposesym projects every step;
    the reference's 20-step run compiled with the flag set."""
    g = load_golden("sr_mini")
    ctx = _ctx(g)
    ctx.symm = (2,)
    ic = g["ic"]
    ctx.set_state(ic[0], ic[1], ic[2])
    ser = ctx.run(int(g["nsteps"]))
    E = np.array([s["E"] for s in ser])
    assert np.max(np.abs(E - g["series"][:, 2]) / g["series"][:, 2]) <= TOL_DIAG
    for c, u in enumerate((ctx.ur, ctx.ut, ctx.uz)):
        assert rel_err(u, g["state_20"][c]) <= 1e-10, c


@pytest.mark.parametrize("name,symm", [("orbit_mini", ()), ("orbit_sr_mini", (2,))])
def test_steporbit_matches_reference(name, symm):
    """steporbit (newton.f90:721-775): the reference's own call sequence
    (oracle/ref_driver.f90, OPF_ORBIT) for new and unchanged tim_dt, with and
    without the final shift, and in the shift-reflect subspace."""
    g = load_golden(name)
    for i, (ndts, T, sth, sz) in enumerate(g["cases"]):
        c = O.Ctx(int(g["N"]), int(g["K"]), int(g["M"]), Re=float(g["Re"]), alpha=float(g["alpha"]),
                  dt=float(g["dt"]))
        c.symm = symm
        c.set_state(*g["ic"])
        c.steporbit(int(ndts), T, sth, sz)
        ref = g[f"orbit_{i}"]
        for k, u in enumerate((c.ur, c.ut, c.uz)):
            assert rel_err(u, ref[k]) <= 1e-12, (i, k)


PLUGIN_RUNS = ["nn_mini", "les_smag_mini", "les_smag3_mini", "les_vreman_mini"]


def plugin_params(g):
    """The plugin and its parameters recorded in a plugin fixture."""
    prm = {k: float(v) for k, v in g["plugin"]}
    if "model" in prm:
        prm["model"], prm["filter"] = int(prm["model"]), int(prm["filter"])
        return "les", prm
    return "nonnewtonian", prm


@pytest.mark.parametrize("name", PLUGIN_RUNS)
def test_plugin_run_matches_reference(name):
    """The oracle's variable-viscosity forcing (add_nonnewtonian, add_les)
    against the reference's own plugin procedures (get_nu, Smagorinsky /
    Vreman, update_nu_r, update_nonlin_terms of nonnewtonian.f90 and LES.f90,
    compiled from those files by oracle/build_ref.sh) on a fully
    three-dimensional state, where every strain component and every
    component of grad(nu) is nonzero: the state after 1 step <= 1e-12 and
    after 10 steps <= 1e-10, energy series <= 1e-10, one corrector per
    step."""
    g = load_golden(name)
    kind, prm = plugin_params(g)
    ctx = _ctx(g)
    setattr(ctx, kind, prm)
    ic = g["ic"]
    ctx.set_state(ic[0], ic[1], ic[2])
    states = {}
    ser = ctx.run(10, ncorr_fixed=1,
                  record=lambda c: states.__setitem__(c.step, (c.ur.copy(), c.ut.copy(), c.uz.copy())))
    for s, tol in ((1, TOL_STEP), (10, TOL_DIAG)):
        ref = g[f"state_{s}"]
        for c in range(3):
            assert rel_err(states[s][c], ref[c]) <= tol, (s, c)
    rs = g["series"]
    E = np.array([d["E"] for d in ser])
    assert np.max(np.abs(E - rs[:, 2]) / rs[:, 2]) <= TOL_DIAG
    assert np.all(rs[:, 8] == 1)
    if kind == "les":
        assert ctx.nu > 1.0          # update_nu_r moved the reference viscosity, as the reference's did
