"""HIP path against the oracle and the reference's golden vectors.

Bars (the target tolerances, stated here): one-step spectral coefficients
<= 1e-12 normwise relative per field; energy and friction velocity series
<= 1e-10 relative; identical corrector counts; factors and transforms
<= 1e-13 relative.  Every call goes through the C ABI (libopf.so).
"""
import numpy as np
import pytest

import opf_oracle as O
from conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12
TOL_DIAG = 1e-10


@pytest.fixture(scope="module")
def opf():
    import opf as pkg
    return pkg


def _setup(opf, g, capacity=1):
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    ctx = opf.Context(N, K, M, capacity=capacity)
    for s in range(capacity):
        ctx.setup(s, float(g["Re"]), float(g["alpha"]), float(g["dt"]))
    return ctx


def _oracle(g):
    return O.Ctx(int(g["N"]), int(g["K"]), int(g["M"]), Re=float(g["Re"]),
                 alpha=float(g["alpha"]), dt=float(g["dt"]))


@pytest.mark.parametrize("name", ["mini", "c0"])
def test_lumesh_factors_match_oracle(opf, name):
    g = load_golden(name)
    ctx = _setup(opf, g)
    oc = _oracle(g)
    for which, L in enumerate((oc.LDp, oc.LDm, oc.LDz, oc.LNp)):
        AB, ipiv, skip = L
        ab, piv = ctx.lumesh_get(0, which)
        live = ~skip
        assert np.array_equal(piv[live], ipiv[live]), which
        scale = np.abs(AB[live]).max()
        assert np.abs(ab[live] - AB[live]).max() <= 1e-13 * scale, which


@pytest.mark.parametrize("name", ["mini", "c0"])
def test_lumesh_init_of_any_operator(opf, name):
    """tim_lumesh_init(PM, BC, c1, c2, A) into a slot's factor set
    (opf_tim_lumesh_init): with vel_matrices' own coefficients it rebuilds
    the set bit for bit; with others (u- with c1 = 3, c2 = -0.2, BC = 1,
    and the plain Laplacian with c1 = 0.5, c2 = 1, BC = 0) its factors and
    pivots match the oracle's tim_lumesh_init and tim_lumesh_invert(BC, A, b)
    its solve, including the zero (0,0) harmonic for BC = 1."""
    g = load_golden(name)
    ctx = _setup(opf, g)
    oc = _oracle(g)
    before = ctx.lumesh_get(0, 0)
    ctx.lumesh_init(0, 0, 1, 0, 1.0 / float(g["dt"]), -0.51 / float(g["Re"]))
    after = ctx.lumesh_get(0, 0)
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    rng = np.random.default_rng(11)
    b = rng.standard_normal((oc.N, oc.H)) + 1j * rng.standard_normal((oc.N, oc.H))
    for which, (PM, BC, c1, c2) in ((1, (-1, 1, 3.0, -0.2)), (2, (0, 0, 0.5, 1.0))):
        ctx.lumesh_init(0, which, PM, BC, c1, c2)
        AB, ipiv, skip = oc.lumesh_init(PM, BC, c1, c2)
        ab, piv = ctx.lumesh_get(0, which)
        live = ~skip
        assert np.array_equal(piv[live], ipiv[live]), which
        assert np.abs(ab[live] - AB[live]).max() <= 1e-13 * np.abs(AB[live]).max(), which
        x = ctx.lumesh_invert(0, which, b, BC=BC)
        ref = oc.lumesh_invert((AB, ipiv, skip), b)
        assert rel_err(x, ref) <= 1e-12, which
        if BC == 1:
            assert np.all(x[:, 0] == 0.0)


def test_lumesh_init_of_a_singular_operator_is_reported(opf):
    """tim_lumesh_init with c1 = c2 = 0 leaves only the axis and boundary rows,
    so DGBTF2 meets a zero pivot in the second column of every pair; the
    reference stops there (timestep.f90:92, dgbtrf info /= 0), the call
    returns OPF_ERR_SINGULAR.  The slot's status is cleared, and the regular
    coefficients rebuild the factor set bit for bit."""
    g = load_golden("c0")
    ctx = _setup(opf, g)
    before = ctx.lumesh_get(0, 2)
    with pytest.raises(opf.OpfError) as err:
        ctx.lumesh_init(0, 2, 0, 0, 0.0, 0.0)
    assert err.value.code == 3 and "tim_lumesh_init" in str(err.value)
    ctx.lumesh_init(0, 2, 0, 0, 1.0 / float(g["dt"]), -0.51 / float(g["Re"]))
    after = ctx.lumesh_get(0, 2)
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])


@pytest.mark.parametrize("name", ["mini", "c0"])
def test_lumesh_invert_matches_oracle(opf, name):
    g = load_golden(name)
    ctx = _setup(opf, g)
    oc = _oracle(g)
    rng = np.random.default_rng(7)
    b = rng.standard_normal((oc.N, oc.H)) + 1j * rng.standard_normal((oc.N, oc.H))
    for which, L in enumerate((oc.LDp, oc.LDm, oc.LDz, oc.LNp)):
        x = ctx.lumesh_invert(0, which, b)
        ref = oc.lumesh_invert(L, b)
        # the Neumann pressure matrix is nearly singular for the lowest modes
        # and amplifies rounding; the Dirichlet ones are well conditioned
        assert rel_err(x, ref) <= (1e-13 if which < 3 else 1e-11), which
        # residual of the banded system, harmonic by harmonic
        AB, ipiv, skip = L
        for h in range(0, oc.H, max(1, oc.H // 16)):
            if skip[h]:
                continue
            A = _dense_from_lu_input(oc, which, h)
            r = A @ x[:, h] - _zero_bc_rows(b[:, h], which)
            assert np.abs(r).max() <= 1e-12 * np.abs(A).max() * np.abs(x[:, h]).max(), (which, h)


def _dense_from_lu_input(oc, which, h):
    """Dense matrix assembled as tim_lumesh_init does before dgbtrf."""
    PM, BC, c1, c2 = [(1, 0, 1 / oc.dt, -oc.implicit / oc.Re), (-1, 0, 1 / oc.dt, -oc.implicit / oc.Re),
                      (0, 0, 1 / oc.dt, -oc.implicit / oc.Re), (0, 1, 0.0, 1.0)][which]
    N = oc.N
    A = O.band_to_dense(oc.mesh.radLap) * c2
    A[np.arange(N), np.arange(N)] += c2 * oc._d(PM)[:, h] + c1
    S = (int(oc.harm.m[h]) * oc.Mp + abs(PM)) % 2
    A[0, :] = 0.0
    A[0, :5] = oc.mesh.dr0[:, 1 - S]
    A[N - 1, :] = 0.0
    A[N - 1, N - 5:] = oc.mesh.dr1[:, BC]
    return A


def _zero_bc_rows(v, which):
    return v          # the solve treats rows 1 and N as given right-hand sides


@pytest.mark.parametrize("shape", [(16, 6, 4), (32, 24, 12), (24, 8, 6)])
def test_transforms_match_oracle(opf, shape):
    N, K, M = shape
    ctx = opf.Context(N, K, M)
    oc = O.Ctx(N, K, M)
    rng = np.random.default_rng(N + K + M)
    a = rng.standard_normal((N, oc.H)) + 1j * rng.standard_normal((N, oc.H))
    p = ctx.coll2phys(a)
    pref = oc.spec2phys(a).transpose(0, 2, 1)          # oracle p[n, z, th]
    assert rel_err(p, pref) <= 1e-13
    q = rng.standard_normal(p.shape)
    c = ctx.phys2coll(q)
    cref = oc.phys2spec(q.transpose(0, 2, 1))
    assert rel_err(c, cref) <= 1e-13
    # round trip on band-limited data
    a[:, 0] = a[:, 0].real
    assert rel_err(ctx.phys2coll(ctx.coll2phys(a)), a) <= 1e-13


def test_unsupported_transform_length_rejected(opf):
    """3K and 3M must be 2^a 3^b; other lengths fail loudly."""
    from opf._lib import OpfError
    with pytest.raises(OpfError):
        opf.Context(16, 6, 5)


def test_nonlinear_and_stages_mini(opf):
    """vel_nonlinear, vel_predictor, vel_corrector one call each (the Fortran
    drop-in path) against the reference's step-1 stage dumps."""
    g = load_golden("mini")
    ctx = _setup(opf, g)
    ic = g["ic"]
    ctx.put_state(0, ic[0], ic[1], ic[2])
    ctx.transform(0)
    ctx.nonlinear(0)
    for c, f in enumerate(ctx.get_nonlinear(0)):
        assert rel_err(f, g["stage_01N"][c]) <= TOL_STEP, ("01N", c)
    ctx.predictor(0)
    u = ctx.get_state(0)
    for c in range(3):
        assert rel_err(u[c], g["stage_02u"][c]) <= TOL_STEP, ("02u", c)
    ctx.transform(0)
    ctx.nonlinear(0)
    for c, f in enumerate(ctx.get_nonlinear(0)):
        assert rel_err(f, g["stage_03N"][c]) <= TOL_STEP, ("03N", c)
    d = ctx.corrector(0)
    u = ctx.get_state(0)
    for c in range(3):
        assert rel_err(u[c], g["stage_04u"][c]) <= TOL_STEP, ("04u", c)
    assert abs(d - g["series"][0, 9]) <= 1e-8 * g["series"][0, 9]


@pytest.mark.parametrize("K", [6, 3])
def test_nonlinear_long_theta_lines(opf, K):
    """vel_nonlinear with 3M = 384 theta points, where the theta kernel
    takes two lines per workgroup at its largest register tile (the theta
    length of the (96,256,128) size), against the oracle's products and
    transforms on the same state.  K = 3 gives an odd number of z lines, so
    the last workgroup holds one line."""
    from opf import ic as icgen
    N, M, Re, alpha = 12, 128, 3000.0, 1.25
    ctx = opf.Context(N, K, M)
    ctx.setup(0, Re, alpha, 0.01)
    me = ctx.mesh()
    u = icgen.seeded_perturbation(me["r"], me["intrdr"], K, M, alpha, 7, 1e-2)
    ctx.put_state(0, *u)
    ctx.transform(0)
    ctx.nonlinear(0)
    got = ctx.get_nonlinear(0)
    oc = O.Ctx(N, K, M, Re=Re, alpha=alpha, dt=0.01)
    oc.set_state(*u)
    oc.transform()
    oc.nonlinear()
    for c, ref in enumerate((oc.Nr, oc.Nt, oc.Nz)):
        assert rel_err(got[c], ref) <= TOL_STEP, c


def test_nonlinear_long_z_lines(opf):
    """vel_nonlinear with 3K = 768 z points (the z length of the
    (96,256,128) size), where a z-transform workgroup holds two m-columns,
    against the oracle."""
    from opf import ic as icgen
    N, K, M, Re, alpha = 12, 256, 4, 3000.0, 0.4
    ctx = opf.Context(N, K, M)
    ctx.setup(0, Re, alpha, 0.01)
    me = ctx.mesh()
    u = icgen.seeded_perturbation(me["r"], me["intrdr"], K, M, alpha, 9, 1e-2)
    ctx.put_state(0, *u)
    ctx.transform(0)
    ctx.nonlinear(0)
    got = ctx.get_nonlinear(0)
    oc = O.Ctx(N, K, M, Re=Re, alpha=alpha, dt=0.01)
    oc.set_state(*u)
    oc.transform()
    oc.nonlinear()
    for c, ref in enumerate((oc.Nr, oc.Nt, oc.Nz)):
        assert rel_err(got[c], ref) <= TOL_STEP, c


@pytest.mark.parametrize("name", ["mini", "c0"])
def test_one_step_golden(opf, name):
    g = load_golden(name)
    ctx = _setup(opf, g)
    ic = g["ic"]
    ctx.put_state(0, ic[0], ic[1], ic[2])
    ctx.run(1, diag_every=1)
    u = ctx.get_state(0)
    for c in range(3):
        assert rel_err(u[c], g["state_1"][c]) <= TOL_STEP, c
    st = ctx.status(0)
    assert st["step"] == 1 and st["status"] == 0
    assert abs(st["t"] - float(g["dt"])) < 1e-15


@pytest.mark.parametrize("name", ["mini", "c0"])
def test_hundred_steps_golden(opf, name):
    g = load_golden(name)
    ctx = _setup(opf, g)
    ic = g["ic"]
    ctx.put_state(0, ic[0], ic[1], ic[2])
    ctx.run(100, diag_every=1)
    d = ctx.diag(0)
    rs = g["series"]
    assert d.shape[0] == 100
    assert np.array_equal(d[:, 0], rs[:, 0])
    assert np.max(np.abs(d[:, 2] - rs[:, 2]) / rs[:, 2]) <= TOL_DIAG     # E
    assert np.max(np.abs(d[:, 7] - rs[:, 7]) / rs[:, 7]) <= TOL_DIAG     # u_tau
    assert np.max(np.abs(d[:, 5] - rs[:, 5]) / np.abs(rs[:, 5]).max()) <= 1e-9   # Pr0
    assert np.array_equal(d[:, 8].astype(int), rs[:, 8].astype(int))     # ncorr
    u = ctx.get_state(0)
    for c in range(3):
        assert rel_err(u[c], g["state_100"][c]) <= 1e-10, c


def test_one_step_s48_sampled(opf):
    g = load_golden("s48")
    ctx = _setup(opf, g)
    from opf import ic as icgen
    me = ctx.mesh()
    ur, ut, uz = icgen.seeded_perturbation(me["r"], me["intrdr"], ctx.K, ctx.M, float(g["alpha"]),
                                           int(g["seed"]), float(g["E0"]))
    sel = g["sel"]
    assert rel_err(np.stack([ur, ut, uz])[:, :, sel], g["ic_sel"]) <= 1e-13
    ctx.put_state(0, ur, ut, uz)
    ctx.run(1)
    u = ctx.get_state(0)
    for c in range(3):
        assert np.abs(u[c][:, sel] - g["state_1"][c]).max() / g["state_1_norm"][c] <= TOL_STEP


def test_c1_series(opf):
    """(64,64,48): the sampled state after step 1, then the reference's whole
    1000-step run: energy and friction velocity of every step <= 1e-10
    relative, corrector counts equal."""
    g = load_golden("c1")
    ctx = _setup(opf, g)
    from opf import ic as icgen
    me = ctx.mesh()
    ur, ut, uz = icgen.seeded_perturbation(me["r"], me["intrdr"], ctx.K, ctx.M, float(g["alpha"]),
                                           int(g["seed"]), float(g["E0"]))
    ctx.put_state(0, ur, ut, uz)
    ctx.run(1, diag_every=1)
    u = ctx.get_state(0)
    sel = g["sel"]
    for c in range(3):
        assert np.abs(u[c][:, sel] - g["state_1"][c]).max() / g["state_1_norm"][c] <= TOL_STEP
    rs = g["series"]
    assert rs.shape[0] == 1000
    print("c1: 999 more steps", flush=True)
    recs = {int(r[0]): r for r in ctx.diag(0)}
    done = 1
    while done < rs.shape[0]:             # the records ring holds 256 steps: read every 200
        n = min(200, rs.shape[0] - done)
        ctx.run(n, diag_every=1)
        done += n
        recs.update({int(r[0]): r for r in ctx.diag(0)})
    d = np.array([recs[k] for k in sorted(recs)])
    assert d.shape[0] == rs.shape[0]
    assert np.array_equal(d[:, 0].astype(int), rs[:, 0].astype(int))
    assert np.max(np.abs(d[:, 2] - rs[:, 2]) / rs[:, 2]) <= TOL_DIAG
    assert np.max(np.abs(d[:, 7] - rs[:, 7]) / rs[:, 7]) <= TOL_DIAG
    assert np.array_equal(d[:, 8].astype(int), rs[:, 8].astype(int))


def test_c2_reference_run(opf):
    """(96,256,128), alpha 0.4, Re 5300, dt 0.0025: the reference's own run
    (oracle/_ref built at this size) from the same seeded state.  Sampled
    spectral coefficients after step 1 <= 1e-12 and after step 20 <= 1e-10
    normwise relative per field; energy and friction velocity of all 20
    steps <= 1e-10 relative; equal corrector counts."""
    g = load_golden("c2")
    ctx = _setup(opf, g)
    from opf import ic as icgen
    me = ctx.mesh()
    ur, ut, uz = icgen.seeded_perturbation(me["r"], me["intrdr"], ctx.K, ctx.M, float(g["alpha"]),
                                           int(g["seed"]), float(g["E0"]))
    sel = g["sel"]
    for c, f in enumerate((ur, ut, uz)):
        assert np.array_equal(f[:, sel], g["ic_sel"][c])
    ctx.put_state(0, ur, ut, uz)
    print("c2: step 1", flush=True)
    ctx.run(1, diag_every=1)
    u = ctx.get_state(0)
    for c in range(3):
        assert np.abs(u[c][:, sel] - g["state_1"][c]).max() / g["state_1_norm"][c] <= TOL_STEP, c
    rs = g["series"]
    print(f"c2: {rs.shape[0] - 1} more steps", flush=True)
    ctx.run(rs.shape[0] - 1, diag_every=1)
    u = ctx.get_state(0)
    last = int(g["nsteps"])
    for c in range(3):
        assert np.abs(u[c][:, sel] - g[f"state_{last}"][c]).max() / g[f"state_{last}_norm"][c] <= TOL_DIAG, c
    d = ctx.diag(0)
    assert d.shape[0] == rs.shape[0]
    assert np.max(np.abs(d[:, 2] - rs[:, 2]) / rs[:, 2]) <= TOL_DIAG
    assert np.max(np.abs(d[:, 7] - rs[:, 7]) / rs[:, 7]) <= TOL_DIAG
    assert np.array_equal(d[:, 8].astype(int), rs[:, 8].astype(int))


def test_large_size_properties(opf):
    """The largest size, (N, K, M) = (96, 256, 128): no oracle run at this size;
    checks the properties the method fixes: the influence correction makes
    u = 0 at the wall to machine precision, the (0,0) mode keeps zero net
    flux, harmonic 0 keeps u_r = 0 and real u_t, u_z, and the transforms
    round-trip."""
    N, K, M = 96, 256, 128
    ctx = opf.Context(N, K, M)
    ctx.setup(0, 5300.0, 0.4, 0.0025)
    me = ctx.mesh()
    from opf import ic as icgen
    ur, ut, uz = icgen.seeded_perturbation(me["r"], me["intrdr"], K, M, 0.4, 3, 1e-2)
    rng = np.random.default_rng(1)
    a = ur.copy()
    a[:, 0] = a[:, 0].real
    assert rel_err(ctx.phys2coll(ctx.coll2phys(a)), a) <= 1e-13
    ctx.put_state(0, ur, ut, uz)
    ctx.run(2)
    u = ctx.get_state(0)
    scale = max(np.abs(x).max() for x in u[:3])
    for c in range(3):
        assert np.abs(u[c][-1]).max() <= 1e-13 * scale, c
    flux = np.dot(u[2][:, 0].real, me["intrdr"])
    assert abs(flux) <= 1e-14
    assert np.all(u[0][:, 0] == 0) and np.all(u[1][:, 0].imag == 0) and np.all(u[2][:, 0].imag == 0)
    st = ctx.status(0)
    assert st["status"] == 0 and st["step"] == 2
    del rng


@pytest.mark.parametrize("name", ["mini", "c0"])
def test_pairs_per_workgroup_agree(opf, name, monkeypatch):
    """The step kernel's grouping (1, 2 or 4 (|k|, m) pairs per workgroup)
    changes only which lanes and band widths a harmonic shares, never its
    arithmetic: the states agree bit for bit, and four pairs per workgroup
    meets the golden series."""
    g = load_golden(name)
    ic = g["ic"]
    states = {}
    for gp in (1, 2, 4):
        monkeypatch.setenv("OPF_STEP_GP", str(gp))
        ctx = _setup(opf, g)
        assert ctx.step_pairs() == gp
        ctx.put_state(0, ic[0], ic[1], ic[2])
        ctx.run(100, diag_every=1)
        states[gp] = (ctx.get_state(0), ctx.diag(0))
    monkeypatch.delenv("OPF_STEP_GP")
    for gp in (1, 4):
        for c in range(3):
            assert np.array_equal(states[gp][0][c], states[2][0][c]), (gp, c)
        assert np.array_equal(states[gp][1], states[2][1]), gp
    d = states[4][1]
    rs = g["series"]
    assert np.max(np.abs(d[:, 2] - rs[:, 2]) / rs[:, 2]) <= TOL_DIAG
    assert np.array_equal(d[:, 8].astype(int), rs[:, 8].astype(int))


@pytest.mark.parametrize("N", [80, 96])
def test_four_pairs_beyond_64_points(opf, N, monkeypatch):
    """Four pairs per workgroup for 64 < N <= 96 (at N = 96 the workgroup
    asks for more than 64 KiB of LDS): the states and diagnostics after 20
    steps equal those of one and two pairs per workgroup bit for bit."""
    from opf import ic as icgen
    K, M, alpha = 6, 4, 1.25
    res = {}
    for gp in (1, 2, 4):
        monkeypatch.setenv("OPF_STEP_GP", str(gp))
        print(f"N={N} pairs per workgroup={gp}", flush=True)
        ctx = opf.Context(N, K, M)
        assert ctx.step_pairs() == gp
        ctx.setup(0, 2400.0, alpha, 0.01)
        me = ctx.mesh()
        ctx.put_state(0, *icgen.seeded_perturbation(me["r"], me["intrdr"], K, M, alpha, 5, 1e-3))
        ctx.run(20, diag_every=1)
        res[gp] = (ctx.get_state(0), ctx.diag(0))
        ctx.close()
    monkeypatch.delenv("OPF_STEP_GP")
    assert len(res[2][1]) == 20
    for gp in (1, 4):
        for c in range(3):
            assert np.array_equal(res[gp][0][c], res[2][0][c]), (gp, c)
        assert np.array_equal(res[gp][1], res[2][1]), gp


@pytest.mark.parametrize("case", ["plain", "adaptive", "coriolis+mirror", "les+nonnewtonian"])
def test_plane_kernel_equals_transform_kernels(opf, case, monkeypatch):
    """At (K, M) = (24, 12) one workgroup per radial plane runs the whole
    evaluation with its intermediates in LDS.  It makes the same passes and
    operations as the three transform kernels, so two slots of different
    states agree with the three-kernel run bit for bit, diagnostics
    included: plain steps; an adaptive slot (the plane's velocity maxima
    feed the time step); a slot with the Coriolis terms (which read the
    theta twiddles) beside one with the mirror symmetry imposed; the LES
    and non-Newtonian plugins, which add their terms to the N the
    evaluation leaves."""
    g = load_golden("c0")
    ic = g["ic"]
    res = {}
    for plane in ("1", "0"):
        monkeypatch.setenv("OPF_PLANE", plane)
        print(f"whole-plane kernel {plane}, case {case}", flush=True)
        ctx = _setup(opf, g, capacity=2)
        ctx.put_state(0, ic[0], ic[1], ic[2])
        ctx.put_state(1, 1.5 * ic[0], 0.5 * ic[1], ic[2])
        if case == "adaptive":
            ctx.adaptive(1, True, maxdt=float(g["dt"]), courant=0.5)
        elif case == "coriolis+mirror":
            ctx.coriolis(0, Ek0=0.01, phi=0.3)
            ctx.symmetry(1, mirror=True)
        elif case == "les+nonnewtonian":
            ctx.les(0, model=2, C_s=0.1, filter=2, A_plus=25.0)
            ctx.nonnewtonian(1, rho=1000.0, radius=0.01, mu_0=0.056, mu_inf=0.0056, lam_CY=3e-3, n_CY=0.6,
                             a_CY=2.0)
        ctx.kernel_timing(True)
        ctx.run(30, diag_every=1)
        used = ctx.kernel_time("plane")[1]
        ctx.kernel_timing(False)
        assert (used > 0) == (plane == "1"), used
        res[plane] = [(ctx.get_state(s), ctx.diag(s), ctx.timestep(s)) for s in range(2)]
        ctx.close()
    monkeypatch.delenv("OPF_PLANE")
    for s in range(2):
        for c in range(3):
            assert np.array_equal(res["1"][s][0][c], res["0"][s][0][c]), (s, c)
        assert len(res["1"][s][1]) == 30
        assert np.array_equal(res["1"][s][1], res["0"][s][1]), s
        assert res["1"][s][2] == res["0"][s][2], s


@pytest.mark.parametrize("dt,expect", [(0.1, None), (0.2, 2), (0.4, 3)])
def test_corrector_decisions_at_large_dt(dt, expect):
    """tim_check_cgce's branches in the device loop at (16, 6, 4): dt = 0.1
    needs 4-5 correctors a step (states and corrector counts against the
    oracle over 20 steps); dt = 0.2 fails at step 0 with 'too many its'
    (status 2) and dt = 0.4 with 'increasing error' (status 3), as the
    oracle's tim_check_cgce does."""
    import opf
    import opf_oracle as O
    from opf._lib import OpfError
    g = load_golden("mini")
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    ctx = opf.Context(N, K, M)
    ctx.setup(0, float(g["Re"]), float(g["alpha"]), dt)
    ctx.put_state(0, *g["ic"])
    o = O.Ctx(N, K, M, Re=float(g["Re"]), alpha=float(g["alpha"]), dt=dt)
    o.set_state(*g["ic"])
    print("large dt", dt, flush=True)
    if expect is None:
        ctx.run(20, diag_every=1)
        ser = o.run(20)
        d = ctx.diag(0)
        assert np.array_equal(d[:, 8].astype(int), np.array([s["ncorr"] for s in ser]))
        assert max(s["ncorr"] for s in ser) >= 4
        u = ctx.get_state(0)
        for c, ref in enumerate((o.ur, o.ut, o.uz)):
            assert rel_err(u[c], ref) <= 1e-10, c
    else:
        with pytest.raises(OpfError):
            ctx.run(5)
        st = ctx.status(0)
        assert (st["status"], st["step"]) == (expect, 0)
        with pytest.raises(RuntimeError, match="too many its" if expect == 2 else "increasing error"):
            o.run(5)
        assert o.step == 0
