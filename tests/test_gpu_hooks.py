"""Entry points a Fortran host uses around the step: var_coll_norm and the
friction quantities (io_write_energy / io_write_friction), the physical
fields of vel_transform and forcing added to vel_N* for var_null(2) hooks."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu


def _ctx(g, const_flux=True):
    import opf
    ctx = opf.Context(int(g["N"]), int(g["K"]), int(g["M"]))
    ctx.setup(0, float(g["Re"]), float(g["alpha"]), float(g["dt"]), const_flux=const_flux)
    ic = g["ic"]
    ctx.put_state(0, ic[0], ic[1], ic[2])
    return ctx


@pytest.mark.parametrize("name", ["mini", "c0"])
def test_norm_and_friction_match_reference(name):
    g = load_golden(name)
    ctx = _ctx(g)
    E, Ek0, Em0 = 0.0, 0.0, 0.0
    for comp in range(3):
        e, ek, em = ctx.var_norm(0, comp)
        assert ek.shape == (int(g["K"]),) and em.shape == (int(g["M"]),)
        assert abs(e - em.sum()) <= 1e-14 * abs(e)
        E, Ek0, Em0 = E + e, Ek0 + ek[0], Em0 + em[0]
    ref = g["series"][0]          # step 0: the reference's io quantities of the initial state
    assert abs(E - ref[2]) <= 1e-12 * ref[2]
    assert abs(Ek0 - ref[3]) <= 1e-12 * abs(ref[3])
    assert abs(Em0 - ref[4]) <= 1e-12 * abs(ref[4])
    Ub, Uc, Ufr = ctx.friction(0)
    assert abs(Ub - ref[5]) <= 1e-12 * abs(ref[5])
    assert abs(Uc - ref[6]) <= 1e-13 * abs(ref[6])
    assert abs(Ufr - ref[7]) <= 1e-12 * abs(ref[7])


@pytest.mark.parametrize("name", ["mini", "c0"])
def test_coll_norm_of_any_field(name):
    """var_coll_norm on host fields (opf_var_coll_norm): the three
    components of the reference's initial state give the reference's step-0
    E, Ek(0), Em(0) (1e-12) and equal the norms of the slot's own state; a
    field that is no slot's state (u_r + 2 u_z) equals the oracle's
    var_coll_norm to 1e-14 (the same operations in the same order)."""
    import opf_oracle as O
    g = load_golden(name)
    ctx = _ctx(g)
    E, Ek0, Em0 = 0.0, 0.0, 0.0
    for comp in range(3):
        e, ek, em = ctx.coll_norm(g["ic"][comp], 0)
        e2, ek2, em2 = ctx.var_norm(0, comp)
        assert abs(e - e2) <= 1e-14 * abs(e2)
        assert np.allclose(ek, ek2, rtol=1e-14, atol=0) and np.allclose(em, em2, rtol=1e-14, atol=0)
        E, Ek0, Em0 = E + e, Ek0 + ek[0], Em0 + em[0]
    ref = g["series"][0]
    assert abs(E - ref[2]) <= 1e-12 * ref[2]
    assert abs(Ek0 - ref[3]) <= 1e-12 * abs(ref[3])
    assert abs(Em0 - ref[4]) <= 1e-12 * abs(ref[4])
    f = g["ic"][0] + 2.0 * g["ic"][2]
    oc = O.Ctx(int(g["N"]), int(g["K"]), int(g["M"]), Re=float(g["Re"]), alpha=float(g["alpha"]),
               dt=float(g["dt"]))
    eo, eko, emo = oc.norm(f)
    e, ek, em = ctx.coll_norm(f, 0)
    assert abs(e - eo) <= 1e-14 * eo
    assert np.abs(ek - eko).max() <= 1e-14 * eo and np.abs(em - emo).max() <= 1e-14 * eo


def test_friction_without_constant_flux_and_diag_record():
    g = load_golden("mini")
    ctx = _ctx(g, const_flux=False)
    Ub, _, _ = ctx.friction(0)
    me = ctx.mesh()
    uz0 = g["ic"][2][:, 0].real
    ub_ref = 0.5 + 2.0 * float(np.dot(uz0, me["intrdr"]))      # io.f90:630
    assert abs(Ub - ub_ref) <= 1e-13
    ctx.run(1, diag_every=1)
    rec = ctx.diag(0)[0]
    assert abs(rec[5] - ub_ref) <= 1e-13


def test_phys_fields_and_curl():
    import opf_oracle as O
    g = load_golden("mini")
    ctx = _ctx(g)
    u = ctx.get_state(0)
    for c in range(3):
        assert np.array_equal(ctx.get_phys(0, c), ctx.coll2phys(u[c]))
    oc = O.Ctx(int(g["N"]), int(g["K"]), int(g["M"]), Re=float(g["Re"]), alpha=float(g["alpha"]),
               dt=float(g["dt"]))
    oc.set_state(u[0], u[1], u[2])
    curl = oc.curl(oc.ur, oc.ut, oc.uz)
    for c in range(3):
        p = ctx.get_phys(0, 3 + c)
        assert rel_err(p, oc.spec2phys(curl[c]).transpose(0, 2, 1)) <= 1e-12, c   # oracle p[n, z, th]


def test_add_nonlinear_is_added_to_the_evaluation():
    g = load_golden("mini")
    ctx = _ctx(g)
    ctx.transform(0)
    ctx.nonlinear(0)
    n1 = ctx.get_nonlinear(0)
    rng = np.random.default_rng(3)
    d = [1e-3 * (rng.standard_normal(n1[0].shape) + 1j * rng.standard_normal(n1[0].shape)) for _ in range(3)]
    ctx.add_nonlinear(0, *d)
    n2 = ctx.get_nonlinear(0)
    for c in range(3):
        assert np.max(np.abs(n2[c] - (n1[c] + d[c]))) <= 1e-15 * max(1.0, np.abs(n1[c]).max())


@pytest.mark.parametrize("phi", [0.0, 0.6])
def test_coriolis_plugin_matches_oracle(phi):
    """The Coriolis plugin (Coriolis.f90:54-71) fused into the device's
    products against the oracle's util(2) after vel_nonlinear: 1 step to
    1e-12 and 10 steps to 1e-10 at (16, 6, 4), Ek0 = 0.01 (a force of
    order 1/(4 Re Ek0) ~ 0.01 against N of order 1e-3)."""
    import opf_oracle as O
    g = load_golden("mini")
    ctx = _ctx(g)
    ctx.coriolis(0, Ek0=0.01, phi=phi)
    ctx.put_state(0, *g["ic"])
    o = O.Ctx(int(g["N"]), int(g["K"]), int(g["M"]), Re=float(g["Re"]), alpha=float(g["alpha"]),
              dt=float(g["dt"]))
    o.coriolis = (0.01, phi)
    o.set_state(*g["ic"])
    for n, tol in ((1, 1e-12), (10, 1e-10)):
        print("coriolis", phi, n, flush=True)
        ctx.run(n - o.step)
        o.run(n)
        u = ctx.get_state(0)
        for c, ref in enumerate((o.ur, o.ut, o.uz)):
            assert rel_err(u[c], ref) <= tol, (n, c)
    # the plugin changes the flow: against the run without it
    plain = _ctx(g)
    plain.run(10)
    assert rel_err(plain.get_state(0)[0], o.ur) > 1e-6
    ctx.coriolis(0, None)                      # removed again
    with pytest.raises(Exception):
        import opf
        bad = opf.Context(16, 6, 4, Mp=2)
        bad.coriolis(0, Ek0=0.01)              # io: 'add_coriolis: i_Mp/=1'


def test_nonnewtonian_plugin_matches_oracle():
    """The non-Newtonian plugin (nonnewtonian.f90:93-244, Carreau-Yasuda)
    on the device against the oracle's util(2): 1 step to 1e-12 and 10
    steps to 1e-10 at (16, 6, 4), both with one corrector per step
    (tim_pcn = 1); the viscosity falls to about a third of mu_0 at the wall."""
    import opf_oracle as O
    g = load_golden("mini")
    prm = dict(rho=1000.0, radius=0.01, mu_0=0.056, mu_inf=0.0056, lam_CY=3e-3, n_CY=0.6, a_CY=2.0)
    ctx = _ctx(g)
    ctx.nonnewtonian(0, **prm)
    ctx.put_state(0, *g["ic"])
    o = O.Ctx(int(g["N"]), int(g["K"]), int(g["M"]), Re=float(g["Re"]), alpha=float(g["alpha"]),
              dt=float(g["dt"]))
    o.nonnewtonian = prm
    o.set_state(*g["ic"])
    for n, tol in ((1, 1e-12), (10, 1e-10)):
        print("nonnewtonian", n, flush=True)
        ctx.run(n - o.step)
        o.run(n, ncorr_fixed=1)
        u = ctx.get_state(0)
        for c, ref in enumerate((o.ur, o.ut, o.uz)):
            assert rel_err(u[c], ref) <= tol, (n, c)
    assert ctx.status(0)["last_ncorr"] == 1
    assert 0.2 < o.nn_nu.min() < 0.5


@pytest.mark.parametrize("model,filt,Ap", [(1, 1, 25.0), (1, 3, 25.0), (2, 2, 25.0), (2, 1, 0.0)])
def test_les_plugin_matches_oracle(model, filt, Ap):
    """The LES plugin (LES.f90:100-320) on the device against the oracle's
    util(2): Smagorinsky and Vreman eddy viscosity, update_nu_r at every
    predictor with the matrices refactored on the device when vel_nu moves,
    one corrector per step; 1 step to 1e-12 and 10 steps to 1e-10 at
    (16, 6, 4)."""
    import opf_oracle as O
    g = load_golden("mini")
    prm = dict(model=model, C_s=0.1, filter=filt, A_plus=Ap)
    ctx = _ctx(g)
    ctx.les(0, **prm)
    ctx.put_state(0, *g["ic"])
    o = O.Ctx(int(g["N"]), int(g["K"]), int(g["M"]), Re=float(g["Re"]), alpha=float(g["alpha"]),
              dt=float(g["dt"]))
    o.les = prm
    o.set_state(*g["ic"])
    for n, tol in ((1, 1e-12), (10, 1e-10)):
        print("les", model, filt, n, flush=True)
        ctx.run(n - o.step)
        o.run(n, ncorr_fixed=1)
        u = ctx.get_state(0)
        for c, ref in enumerate((o.ur, o.ut, o.uz)):
            assert rel_err(u[c], ref) <= tol, (n, c)
    assert o.nu > 1.0                   # vel_nu moved off 1 and the matrices followed
    ctx.les(0, None)


def _seeded(N, K, M, Mp, alpha, seed, E0):
    import opf_oracle as O
    from opf import ic
    me = O.build_mesh(N)
    return ic.seeded_perturbation(me.r, me.intrdr, K, M, alpha, seed, E0, Mp=Mp)


@pytest.mark.parametrize("plugin", ["nonnewtonian", "les"])
def test_viscosity_plugins_with_mp2_match_oracle(plugin):
    """The variable-viscosity plugins at Mp = 2 (the azimuthal wavenumbers
    m Mp in the gradient tensor, Vreman's and Smagorinsky's R dtheta):
    device against oracle, 1 step to 1e-12 and 8 steps to 1e-10."""
    import opf
    import opf_oracle as O
    N, K, M, Mp, Re, alpha, dt = 16, 4, 4, 2, 2000.0, 1.1, 0.01
    u = _seeded(N, K, M, Mp, alpha, 7, 1e-3)
    prm = (dict(rho=1000.0, radius=0.01, mu_0=0.056, mu_inf=0.0056, lam_CY=3e-3, n_CY=0.6, a_CY=2.0)
           if plugin == "nonnewtonian" else dict(model=2, C_s=0.1, filter=2, A_plus=25.0))
    ctx = opf.Context(N, K, M, Mp=Mp)
    ctx.setup(0, Re, alpha, dt)
    getattr(ctx, plugin)(0, **prm)
    ctx.put_state(0, *u)
    o = O.Ctx(N, K, M, Mp=Mp, Re=Re, alpha=alpha, dt=dt)
    setattr(o, plugin, prm)
    o.set_state(*u)
    for n, tol in ((1, 1e-12), (8, 1e-10)):
        print(plugin, "Mp=2", n, flush=True)
        ctx.run(n - o.step)
        o.run(n, ncorr_fixed=1)
        got = ctx.get_state(0)
        for c, ref in enumerate((o.ur, o.ut, o.uz)):
            assert rel_err(got[c], ref) <= tol, (n, c)


def test_les_with_adaptive_time_step_matches_oracle():
    """LES and the adaptive time step on one slot: before each predictor the
    device first refactors for a new vel_nu (update_nu_r inside var_null(2))
    and then, if tim_new_tstep changes tim_dt, again (main.f90:23-29); 12
    steps against the oracle to 1e-10, with the same time series."""
    import opf
    import opf_oracle as O
    g = load_golden("mini")
    prm = dict(model=1, C_s=0.1, filter=1, A_plus=25.0)
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    ctx = opf.Context(N, K, M)
    ctx.adaptive(0)
    ctx.setup(0, float(g["Re"]), float(g["alpha"]), 0.01)
    ctx.les(0, **prm)
    ctx.put_state(0, *g["ic"])
    ctx.run(12, diag_every=1)
    o = O.Ctx(N, K, M, Re=float(g["Re"]), alpha=float(g["alpha"]), dt=0.01)
    o.les = prm
    o.set_state(*g["ic"])
    ser = o.run(12, ncorr_fixed=1, adaptive=True)
    u = ctx.get_state(0)
    for c, ref in enumerate((o.ur, o.ut, o.uz)):
        assert rel_err(u[c], ref) <= 1e-10, c
    d = ctx.diag(0)
    t_ref = np.array([r["t"] for r in ser]) if "t" in ser[0] else None
    assert abs(ctx.timestep(0)["dt"] - o.dt) <= 1e-12 * o.dt
    assert o.dt != 0.01 and o.nu > 1.0
    if t_ref is not None:
        assert np.max(np.abs(d[:, 1] - t_ref)) <= 1e-12


@pytest.mark.parametrize("name", ["nn_mini", "les_smag_mini", "les_smag3_mini", "les_vreman_mini"])
def test_viscosity_plugins_match_reference_plugin_runs(name):
    """The device's variable-viscosity plugins against the reference's own
    plugin procedures (nonnewtonian.f90 / LES.f90 compiled by
    oracle/build_ref.sh, run by the reference's time loop with the plugin at
    the var_null(2) points) from a three-dimensional state: the state after
    1 step <= 1e-12 and after 10 steps <= 1e-10 normwise relative, energy
    series <= 1e-10, one corrector per step."""
    from test_oracle_golden import plugin_params
    g = load_golden(name)
    kind, prm = plugin_params(g)
    ctx = _ctx(g)
    getattr(ctx, kind)(0, **prm)
    ctx.put_state(0, *g["ic"])
    for n, tol in ((1, 1e-12), (10, 1e-10)):
        print(name, n, flush=True)
        ctx.run(n - ctx.status(0)["step"], diag_every=1)
        u = ctx.get_state(0)
        for c in range(3):
            assert rel_err(u[c], g[f"state_{n}"][c]) <= tol, (n, c)
    d = ctx.diag(0)
    rs = g["series"]
    assert np.max(np.abs(d[:, 2] - rs[:, 2]) / rs[:, 2]) <= 1e-10
    assert np.all(d[:, 8] == 1)
