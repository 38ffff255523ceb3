"""Pins of the oracle that do not need the reference: closed forms, exactness
of the finite-difference weights, DFT by brute force, the banded LU against a
dense solve and against scipy's LAPACK, and the linear decay of an
axisymmetric swirl mode.  Each check is chosen so that a plausible mistake
(dropped term, wrong sign, transposed band index) fails it."""
import math

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import opf_oracle as O

N = 24


@pytest.fixture(scope="module")
def mesh():
    return O.build_mesh(N)


def test_default_mesh_shape(mesh):
    r = mesh.r
    assert abs(r[-1] - 1.0) < 1e-14
    assert r[0] > 0.0
    assert np.all(np.diff(r) > 0)
    # clustering at the wall: last spacing smaller than the middle spacing
    assert r[-1] - r[-2] < r[N // 2] - r[N // 2 - 1]


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4])
def test_fd_first_derivative_exact_low_degree(mesh, p):
    r = mesh.r
    d1 = O.band_to_dense(mesh.dr[1])
    d2 = O.band_to_dense(mesh.dr[2])
    f = r ** p
    exact1 = p * r ** (p - 1) if p > 0 else 0 * r
    exact2 = p * (p - 1) * r ** (p - 2) if p > 1 else 0 * r
    assert np.allclose(d1 @ f, exact1, atol=1e-8 * max(1, p * p), rtol=1e-10)
    assert np.allclose(d2 @ f, exact2, atol=1e-6 * max(1, p ** 3), rtol=1e-9)


def test_fd_interior_exact_degree_8(mesh):
    r = mesh.r
    d1 = O.band_to_dense(mesh.dr[1])
    f = r ** 8
    interior = slice(4, N - 4)
    assert np.allclose((d1 @ f)[interior], (8 * r ** 7)[interior], rtol=1e-8, atol=1e-9)


def test_radlap_is_d2_plus_d1_over_r(mesh):
    r = mesh.r
    L = O.band_to_dense(mesh.radLap)
    f = r ** 4          # (1/r) d/dr + d2/dr2 of r^4 = 16 r^2
    assert np.allclose(L @ f, 16 * r ** 2, rtol=1e-9, atol=1e-8)


def test_axis_and_wall_weights(mesh):
    r = mesh.r
    x = np.concatenate([[0.0], r])
    for p in range(0, 5):
        f = x ** p
        # dr0(:,i): i-th derivative at r=0 from r_1..r_5
        v0 = mesh.dr0[:, 0] @ f[1:6]
        assert abs(v0 - (1.0 if p == 0 else 0.0)) < 1e-10
        v1 = mesh.dr0[:, 1] @ f[1:6]
        assert abs(v1 - (1.0 if p == 1 else 0.0)) < 1e-9
        # dr1(:,1): derivative at r=1
        w1 = mesh.dr1[:, 1] @ f[N - 4:N + 1]
        assert abs(w1 - p) < 1e-9 * max(1, p)
    # dr1(:,0): value at the wall point itself is the identity weight
    assert np.allclose(mesh.dr1[:, 0], [0, 0, 0, 0, 1], atol=1e-13)


def test_intrdr_quadrature_order():
    """intrdr (meshs.f90:176-192) weights a Taylor integral of f over the half
    cells around r_n by r_n: exact for f = 1 (the rule is then the trapezoid
    rule for r, which is linear) and second order otherwise."""
    m1, m2 = O.build_mesh(24), O.build_mesh(48)
    assert abs(np.dot(m1.intrdr, np.ones(24)) - 0.5) < 1e-14
    for p in range(1, 7):
        e1 = np.dot(m1.intrdr, m1.r ** p) - 1.0 / (p + 2)
        e2 = np.dot(m2.intrdr, m2.r ** p) - 1.0 / (p + 2)
        assert abs(e1) < 5e-4
        assert 3.5 < e1 / e2 < 4.5


def test_laminar_energy_normalisation():
    """E of U = 1 - r^2 stored in harmonic 0 is pi^2/(3 alpha) (the laminar
    energy io.f90:548 normalises by), up to intrdr's second-order error."""
    E = []
    for n in (32, 64):
        ctx = O.Ctx(n, 2, 2, alpha=1.25)
        uz = ctx.zeros()
        uz[:, 0] = 1.0 - ctx.mesh.r ** 2
        E.append(ctx.norm(uz)[0])
    exact = math.pi ** 2 / (3 * 1.25)
    e1, e2 = E[0] - exact, E[1] - exact
    assert abs(e1) / exact < 2e-3
    assert 3.5 < e1 / e2 < 4.5


def test_friction_velocity_laminar():
    ctx = O.Ctx(16, 3, 2, Re=2400.0)
    ctx.set_state(ctx.zeros(), ctx.zeros(), ctx.zeros())
    d = ctx.diagnostics()
    assert abs(d["Ufr"] - math.sqrt(2.0 / 2400.0)) < 1e-15


def _brute_spec2phys(ctx, a):
    """Direct double sum of the Fourier series on the 3/2 grid."""
    K, M = ctx.K, ctx.M
    Z, T = 3 * K, 3 * M
    zj = np.arange(Z)[:, None]
    tj = np.arange(T)[None, :]
    out = np.zeros((a.shape[0], Z, T))
    for h in range(ctx.H):
        k, m = int(ctx.harm.k[h]), int(ctx.harm.m[h])
        ph = np.exp(2j * np.pi * (k * zj / Z + m * tj / T))
        term = a[:, h][:, None, None] * ph[None]
        if m == 0 and k == 0:
            out += term.real
        elif m == 0:
            out += 2 * term.real        # k<0 partner is the conjugate
        else:
            out += 2 * term.real        # m<0 partner is the conjugate
    return out


def test_spec2phys_brute_force_and_roundtrip():
    ctx = O.Ctx(6, 3, 3)
    rng = np.random.default_rng(3)
    a = rng.standard_normal((6, ctx.H)) + 1j * rng.standard_normal((6, ctx.H))
    a[:, 0] = a[:, 0].real
    p = ctx.spec2phys(a)
    assert np.allclose(p, _brute_spec2phys(ctx, a), atol=1e-12)
    back = ctx.phys2spec(p)
    assert np.allclose(back, a, atol=1e-13)


def test_single_mode_grid_values():
    ctx = O.Ctx(6, 4, 3)
    a = ctx.zeros()
    h = int(np.nonzero((ctx.harm.k == -2) & (ctx.harm.m == 1))[0][0])
    a[:, h] = 0.5
    p = ctx.spec2phys(a)
    Z, T = 12, 9
    zz, tt = np.meshgrid(np.arange(Z), np.arange(T), indexing="ij")
    expect = np.cos(2 * np.pi * (-2 * zz / Z + tt / T))
    assert np.allclose(p[0], expect, atol=1e-13)


def test_dealiased_product_is_exact_convolution():
    """The 3/2 grid makes the quadratic product alias-free: phys2spec of a
    product equals the truncated spectral convolution."""
    ctx = O.Ctx(6, 3, 2)
    rng = np.random.default_rng(5)
    a = rng.standard_normal((6, ctx.H)) + 1j * rng.standard_normal((6, ctx.H))
    b = rng.standard_normal((6, ctx.H)) + 1j * rng.standard_normal((6, ctx.H))
    a[:, 0] = a[:, 0].real
    b[:, 0] = b[:, 0].real
    c = ctx.phys2spec(ctx.spec2phys(a) * ctx.spec2phys(b))

    def full(x):
        d = {}
        for h in range(ctx.H):
            k, m = int(ctx.harm.k[h]), int(ctx.harm.m[h])
            d[(k, m)] = x[:, h]
            d[(-k, -m)] = np.conj(x[:, h])
        return d
    fa, fb = full(a), full(b)
    expect = np.zeros_like(c)
    for h in range(ctx.H):
        k, m = int(ctx.harm.k[h]), int(ctx.harm.m[h])
        s = 0
        for (k1, m1), v in fa.items():
            w = fb.get((k - k1, m - m1))
            if w is not None:
                s = s + v * w
        expect[:, h] = s
    assert np.allclose(c, expect, atol=1e-12)


def _random_band(rng, S, N):
    AB = np.zeros((S, 13, N))
    AB[:, 4:, :] = rng.standard_normal((S, 9, N))
    return AB


def test_band_lu_matches_lapack_and_dense():
    rng = np.random.default_rng(11)
    S, n = 7, 20
    AB = _random_band(rng, S, n)
    b = rng.standard_normal((n, S))
    ABo = AB.copy()
    ipiv = O.Ctx.dgbtf2(ABo)
    x = O.Ctx.dgbtrs(ABo, ipiv, b)
    for s in range(S):
        lu, piv, info = lapack.dgbtrf(AB[s].copy(), 4, 4)
        assert info == 0
        assert np.array_equal(piv + 1, ipiv[s])
        assert np.allclose(lu, ABo[s], rtol=0, atol=1e-12 * np.abs(lu).max())
        xs, info = lapack.dgbtrs(lu, 4, 4, b[:, s], piv)
        assert np.allclose(x[:, s], xs, rtol=1e-13, atol=1e-13 * np.abs(xs).max())
        dense = np.zeros((n, n))
        for j in range(n):
            for i in range(max(0, j - 4), min(n, j + 5)):
                dense[i, j] = AB[s, 8 + i - j, j]
        assert np.allclose(dense @ x[:, s], b[:, s], atol=1e-10)


def test_mat_invert_4x4():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((4, 4))
    Ai = np.array(O.mat_invert(A.tolist()))
    assert np.allclose(Ai @ A, np.eye(4), atol=1e-12)


def test_swirl_mode_decays_at_stokes_rate():
    """u_theta = J1(j11 r) at k = m = 0 is an exact solution decaying like
    exp(-j11^2 t / Re); the nonlinear centrifugal term only feeds u_r, which
    the step discards.  The theta-scheme factor per step is
    (1/dt - (1-th) lam/Re) / (1/dt + th lam/Re), lam = j11^2."""
    from scipy.special import j1
    j11 = 3.8317059702075125
    Re, dt = 100.0, 0.01
    ctx = O.Ctx(32, 2, 2, Re=Re, dt=dt)
    ut = ctx.zeros()
    ut[:, 0] = j1(j11 * ctx.mesh.r)
    ctx.set_state(ctx.zeros(), ut, ctx.zeros())
    ctx.run(30)
    a = ctx.ut[:, 0].real
    ctx.run(40)
    b = ctx.ut[:, 0].real
    ratio = np.dot(b, a) / np.dot(a, a)
    lam = j11 * j11
    g = (1 / dt - (1 - 0.51) * lam / Re) / (1 / dt + 0.51 * lam / Re)
    assert abs(ratio - g ** 10) < 1e-6
    assert np.abs(ctx.ur).max() == 0.0


def test_maxtstep_laminar_and_single_mode():
    """vel_maxtstep: for the laminar state only the axial limit is finite,
    2 pi / (alpha 3K (1 - r_1^2)) at the innermost point (the base flow peaks
    there); a single radial mode u_r = a(r) cos(alpha z) (real amplitude a)
    reaches |a(r_n)| on the grid, so the radial limit is min_n d_n / |a(r_n)|
    with the one-sided spacing at both ends."""
    N, K, M, alpha = 12, 4, 3, 1.25
    c = O.Ctx(N, K, M, alpha=alpha)
    z = np.zeros((N, c.harm.k.size), dtype=complex)
    c.set_state(z, z.copy(), z.copy())
    c.transform()
    cfl, d = c.maxtstep()
    r = c.mesh.r
    assert d == 3
    assert abs(cfl - 2.0 * math.pi / (alpha * 3 * K * (1.0 - r[0] ** 2))) <= 1e-15 * cfl
    # u_r of harmonic (k=1, m=0): physical value 2 a(r) cos(alpha z)
    a = 1e-3 * r * (1.0 - r ** 2)
    ur = z.copy()
    h = int(np.nonzero((c.harm.k == 1) & (c.harm.m == 0))[0][0])
    ur[:, h] = a
    c.ur = ur
    c.transform()
    c.maxtstep()
    gaps = np.diff(r)
    dr = np.concatenate([[gaps[0]], np.minimum(gaps[:-1], gaps[1:]), [gaps[-1]]])
    live = a != 0.0             # the wall point r = 1 carries no velocity (mx = 0 is skipped)
    want = np.min(dr[live] / (2.0 * np.abs(a[live])))
    assert abs(c.cfl_dt - min(want, cfl)) <= 1e-14 * want
    assert c.cfl_dir == (1 if want < cfl else 3)


def test_coll_shift_is_a_grid_translation():
    """var_coll_shift(dth, dz) gives the field u(theta - dth, z - dz): a shift
    by one grid spacing of the 3/2 grid (2 pi/(alpha 3K) in z, 2 pi/(Mp 3M) in
    theta) rolls the physical field by one point; shifting back restores it."""
    ctx = O.Ctx(6, 4, 3, Mp=2, alpha=1.3)
    rng = np.random.default_rng(8)
    a = rng.standard_normal((6, ctx.H)) + 1j * rng.standard_normal((6, ctx.H))
    a[:, 0] = a[:, 0].real
    p = ctx.spec2phys(a)
    Z, T = 3 * ctx.K, 3 * ctx.M
    dz = 2 * np.pi / (ctx.alpha * Z)
    dth = 2 * np.pi / (ctx.Mp * T)
    assert np.allclose(ctx.spec2phys(ctx.coll_shift(a, 0.0, dz)), np.roll(p, 1, axis=1), atol=1e-12)
    assert np.allclose(ctx.spec2phys(ctx.coll_shift(a, dth, 0.0)), np.roll(p, 1, axis=2), atol=1e-12)
    both = ctx.coll_shift(a, 2 * dth, -3 * dz)
    assert np.allclose(ctx.spec2phys(both), np.roll(np.roll(p, 2, axis=2), -3, axis=1), atol=1e-12)
    assert np.allclose(ctx.coll_shift(both, -2 * dth, 3 * dz), a, atol=1e-14)


def _coriolis_delta(ctx, ur, ut, uz, Ek0, phi):
    """The term util(2) adds, alone (the N it is added to set to zero)."""
    ctx.set_state(ur, ut, uz)
    ctx.transform()
    ctx.Nr, ctx.Nt, ctx.Nz = ctx.zeros(), ctx.zeros(), ctx.zeros()
    ctx.add_coriolis(Ek0, phi)
    return ctx.Nr, ctx.Nt, ctx.Nz


def test_coriolis_about_the_pipe_axis_is_a_rotation_of_u():
    """phi = 0 (rotation about the pipe axis): Omega^ x (U + u) = (-u_t, u_r, 0),
    so the added N is -1/(4 Ro) (-u_t, u_r, 0) harmonic by harmonic."""
    ctx = O.Ctx(8, 4, 3, Re=1500.0, alpha=1.1)
    rng = np.random.default_rng(12)
    u = [rng.standard_normal((8, ctx.H)) + 1j * rng.standard_normal((8, ctx.H)) for _ in range(3)]
    O.Ctx.imposesym0(*u)
    Ek0 = 0.02
    d1 = -1.0 / (4.0 * 1500.0 * Ek0)
    dr, dt_, dz = _coriolis_delta(ctx, *u, Ek0, 0.0)
    scale = abs(d1) * max(np.abs(u[0]).max(), np.abs(u[1]).max())
    assert np.abs(dr - (-d1) * u[1]).max() <= 1e-13 * scale
    assert np.abs(dt_ - d1 * u[0]).max() <= 1e-13 * scale
    assert np.abs(dz).max() <= 1e-13 * scale


def test_coriolis_on_laminar_flow_is_a_single_harmonic():
    """u = 0, tilted axis: the force d1 (Ot U, -Or U, 0) with Or = sin(phi)
    cos(theta), Ot = -sin(phi) sin(theta) lives in harmonic (k, m) = (0, 1):
    N_r = i d1 sin(phi) U / 2, N_t = -d1 sin(phi) U / 2, N_z = 0."""
    ctx = O.Ctx(10, 3, 4, Re=2000.0, alpha=1.25)
    z = ctx.zeros()
    Ek0, phi = 0.05, 0.7
    d1 = -1.0 / (4.0 * 2000.0 * Ek0)
    dr, dt_, dz = _coriolis_delta(ctx, z, z.copy(), z.copy(), Ek0, phi)
    h = int(np.nonzero((ctx.harm.k == 0) & (ctx.harm.m == 1))[0][0])
    U = ctx.U
    s = abs(d1) * math.sin(phi)
    assert np.abs(dr[:, h] - 1j * d1 * math.sin(phi) * U / 2).max() <= 1e-14 * s
    assert np.abs(dt_[:, h] - (-d1 * math.sin(phi) * U / 2)).max() <= 1e-14 * s
    others = np.ones(ctx.H, dtype=bool)
    others[h] = False
    assert np.abs(dr[:, others]).max() <= 1e-14 * s and np.abs(dt_[:, others]).max() <= 1e-14 * s
    assert np.abs(dz).max() <= 1e-14 * s


NN_PARAMS = dict(rho=1000.0, radius=0.01, mu_0=0.056, mu_inf=0.0056, lam_CY=3e-3, n_CY=0.6, a_CY=2.0)


def _nn_term(ctx, u, **prm):
    """The term the non-Newtonian util(2) adds, alone."""
    ctx.set_state(*u)
    ctx.Nr, ctx.Nt, ctx.Nz = ctx.zeros(), ctx.zeros(), ctx.zeros()
    ctx.add_nonnewtonian(**prm)
    return ctx.Nr, ctx.Nt, ctx.Nz


def test_nonnewtonian_on_laminar_flow_closed_form():
    """On Hagen-Poiseuille flow U = 1 - r^2 the strain rate is 2r, nu(r) =
    d1 + (1 - d1)(1 + (2 lam r)^a)^(-n/a), and the added force is axial and
    axisymmetric: N_z = (1/r) d/dr(r (nu - 1) U') / Re = (-4 (nu - 1) - 2 r
    nu') / Re.  nu' comes from finite differences of the transformed nu, so
    the comparison holds to the derivative's discretisation error, which
    falls from 2e-4 at N = 32 to 2.5e-6 at 64 and 6e-8 at 96."""
    N = 96
    ctx = O.Ctx(N, 3, 3, Re=2400.0, alpha=1.25)
    z = ctx.zeros()
    nr, nt, nz = _nn_term(ctx, (z, z.copy(), z.copy()), **NN_PARAMS)
    p = NN_PARAMS
    lam = 2400.0 * p["lam_CY"] * p["mu_0"] / (p["rho"] * p["radius"] ** 2)
    d1, d2, a = p["mu_inf"] / p["mu_0"], -p["n_CY"] / p["a_CY"], p["a_CY"]
    r = ctx.mesh.r
    x = 2.0 * lam * r
    nu = d1 + (1 - d1) * (1 + x ** a) ** d2
    dnu = (1 - d1) * d2 * (1 + x ** a) ** (d2 - 1) * a * x ** (a - 1) * 2.0 * lam
    want = (-4.0 * (nu - 1.0) - 2.0 * r * dnu) / 2400.0
    assert np.abs(ctx.nn_nuc[:, 0].real - nu).max() <= 2e-13
    scale = np.abs(want).max()
    assert np.abs(nz[:, 0].real - want).max() <= 2e-7 * scale
    assert np.abs(nz[:, 0].imag).max() <= 1e-14 * scale
    assert np.abs(nz[:, 1:]).max() <= 1e-14 * scale
    assert np.abs(nr).max() <= 1e-14 * scale and np.abs(nt).max() <= 1e-14 * scale


def test_nonnewtonian_with_unit_viscosity_adds_nothing():
    """mu_inf = mu_0 gives nu = 1 everywhere: the plugin's force vanishes
    (to the rounding of differentiating a constant)."""
    ctx = O.Ctx(16, 4, 3, Re=2400.0, alpha=1.25)
    rng = np.random.default_rng(6)
    u = [1e-2 * (rng.standard_normal((16, ctx.H)) + 1j * rng.standard_normal((16, ctx.H))) for _ in range(3)]
    O.Ctx.imposesym0(*u)
    prm = dict(NN_PARAMS, mu_inf=NN_PARAMS["mu_0"])
    out = _nn_term(ctx, u, **prm)
    assert max(np.abs(a).max() for a in out) <= 1e-13


def _les_term(ctx, u, **prm):
    """The LES util(2) term alone, at a predictor evaluation (tim_it = 0)."""
    ctx.set_state(*u)
    ctx.it = 0
    ctx.Nr, ctx.Nt, ctx.Nz = ctx.zeros(), ctx.zeros(), ctx.zeros()
    ctx.add_les(**prm)
    return ctx.Nr, ctx.Nt, ctx.Nz


def test_les_vreman_vanishes_on_laminar_flow():
    """Vreman's model is zero for a one-directional shear: on Hagen-Poiseuille
    flow only A_13 = dU/dr is non-zero, so P = 0 and nu_t = 0; update_nu_r
    then sets vel_nu = 1.1 (1 + 0) and the force is the reference-viscosity
    correction (1/Re)(1 - vel_nu)(-Lap U) = 0.4/Re on u_z of the (0, 0) mode."""
    ctx = O.Ctx(24, 3, 3, Re=2400.0, alpha=1.25)
    z = ctx.zeros()
    nr, nt, nz = _les_term(ctx, (z, z.copy(), z.copy()), model=2, C_s=0.03, filter=1, A_plus=0.0)
    assert np.abs(ctx.les_nu).max() == 0.0
    assert ctx.nu == 1.1
    want = 0.4 / 2400.0
    assert np.abs(nz[:, 0].real - want).max() <= 1e-12 * want
    assert np.abs(nz[:, 1:]).max() <= 1e-14 * want and np.abs(nz[:, 0].imag).max() <= 1e-14 * want
    assert np.abs(nr).max() <= 1e-14 * want and np.abs(nt).max() <= 1e-14 * want


@pytest.mark.parametrize("filt", [1, 2, 3])
def test_les_smagorinsky_viscosity_on_laminar_flow(filt):
    """Smagorinsky on Hagen-Poiseuille flow: strain rate 2r, u_tau = sqrt(2/Re)
    from the wall shear, y+ = Re u_tau (1 - r), so nu_t(r) = (1 -
    exp(-(y+/A+)^2)) (C_s Delta(r))^2 2r with the filter widths of
    LES.f90:199-206 (Delta = (Delta_r R dtheta dz)^(1/3), with r, or the rms);
    the force is axisymmetric and axial."""
    Re, Cs, Ap = 2400.0, 0.1, 25.0
    ctx = O.Ctx(24, 3, 3, Re=Re, alpha=1.25)
    z = ctx.zeros()
    nr, nt, nz = _les_term(ctx, (z, z.copy(), z.copy()), model=1, C_s=Cs, filter=filt, A_plus=Ap)
    r = ctx.mesh.r
    Rdth = 2 * np.pi / (3 * 3) * 1.5
    dz = 2 * np.pi / (1.25 * 9) * 1.5
    Dr = np.concatenate([[r[1] / 2], (r[2:] - r[:-2]) / 2, [1 - r[-2]]])
    Delta = {1: np.cbrt(Dr * Rdth * dz), 2: np.cbrt(Dr * r * Rdth * dz),
             3: np.sqrt((Dr ** 2 + (r * Rdth) ** 2 + dz ** 2) / 3)}[filt]
    yp = Re * np.sqrt(2 / Re) * (1 - r)
    nu = (1 - np.exp(-(yp / Ap) ** 2)) * (Cs * Delta) ** 2 * 2 * r
    assert np.abs(ctx.les_nuc[:, 0].real - nu).max() <= 1e-13 * nu.max()
    assert np.abs(ctx.les_nu.max() - nu.max()) <= 1e-14 * nu.max()
    assert abs(ctx.nu - 1.1 * (1 + Re * ctx.les_nu.max())) <= 1e-15 * ctx.nu
    s = np.abs(nz[:, 0]).max()
    assert np.abs(nz[:, 1:]).max() <= 1e-13 * s and np.abs(nr).max() <= 1e-13 * s and np.abs(nt).max() <= 1e-13 * s


def _cgce(ctx, seq, dt=0.01, step=0):
    """Feed tim_check_cgce a sequence of corrector errors from it = 1."""
    ctx.dt, ctx.step, ctx.it = dt, step, 1
    out = []
    for e in seq:
        ctx.dterr = e
        st = ctx.check_cgce()
        out.append((st, ctx.it))
        if ctx.it == 0:
            break
    return out


def test_check_cgce_decisions():
    """tim_check_cgce (timestep.f90:214-253) on synthetic corrector errors,
    d_dterr = 1e-5: convergence sets tim_corr_dt = dt sqrt(tol/err1); a
    decreasing error above tol iterates; eleven iterations fail ('too many
    its'); a growing error fails above 2 tol and otherwise stops with
    tim_corr_dt = dt/(10 d_courant); dt < 1e-9 after step 30 fails."""
    c = O.Ctx(9, 2, 1)
    assert _cgce(c, [1e-6]) == [("ok", 0)]
    assert abs(c.corr_dt - 0.01 * math.sqrt(10.0)) <= 1e-15
    assert _cgce(c, [1e-4, 5e-5, 1e-6]) == [("ok", 2), ("ok", 3), ("ok", 0)]
    assert abs(c.corr_dt - 0.01 * math.sqrt(0.1)) <= 1e-15          # from the first error only
    seq = [1e-3 / (1 + i) for i in range(12)]
    out = _cgce(c, seq)
    assert out[-1] == ("too many its", 0) and len(out) == 11
    assert _cgce(c, [1e-4, 2e-4]) == [("ok", 2), ("increasing error", 0)]
    assert _cgce(c, [1.5e-5, 1.8e-5]) == [("ok", 2), ("ok", 0)]
    assert abs(c.corr_dt - 0.01 / 5.0) <= 1e-17
    assert _cgce(c, [1e-6], dt=1e-10, step=31) == [("dt->0", 0)]
    assert _cgce(c, [1e-6], dt=1e-10, step=30) == [("ok", 0)]
