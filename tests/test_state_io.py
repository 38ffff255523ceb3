"""State files (io.f90) on the host: the netCDF classic codec and the mapping
of io_load_state / io_load_coll onto another grid (io.f90:98-232).

The codec is checked against an independent implementation of the format
(scipy.io.netcdf_file): files the library writes are read by scipy and files
scipy writes in the reference's layout are read by the library.  The radial
interpolation of the oracle (io_interp_wts / io_interp, io.f90:238-289) is
pinned by exactness on polynomials of the right parity across the axis;
the library's reading is then compared with the oracle on random fields.
No GPU is used."""
from __future__ import annotations

import numpy as np
import pytest
from scipy.io import netcdf_file

import opf_oracle as O

KL = O.KL


def _harmonics(K, M):
    h = O.Harmonics(K, M, 1)
    return list(zip(h.k.tolist(), h.m.tolist()))


def _rand_fields(rng, N, H):
    return [rng.standard_normal((N, H)) + 1j * rng.standard_normal((N, H)) for _ in range(3)]


def _write_reference_layout(path, r, K, M, Mp, fields, t, Re, alpha, dt, dtcor, dtcfl, version=1):
    """A state file as io_save_state defines it (io.f90:295-392), written by
    scipy: Fortran dims (r, H, ReIm) appear reversed in the file."""
    N = len(r)
    H = (2 * K - 1) * M - (K - 1)
    with netcdf_file(path, "w", version=version) as f:
        f.t = np.float64(t)
        f.Re = np.float64(Re)
        f.alpha = np.float64(alpha)
        f.createDimension("r", N)
        f.createDimension("H", H)
        f.createDimension("ReIm", 2)
        vr = f.createVariable("r", "d", ("r",))
        vr[:] = r
        for nm, v in (("dt", dt), ("dtcor", dtcor), ("dtcfl", dtcfl)):
            f.createVariable(nm, "d", ())
            f.variables[nm].data[...] = v
        for nm, a in zip(("Ur", "Ut", "Uz"), fields):
            v = f.createVariable(nm, "d", ("ReIm", "H", "r"))
            v.K = np.int32(K)
            v.M = np.int32(M)
            v.Mp = np.int32(Mp)
            v[0] = a.real.T
            v[1] = a.imag.T


# --------------------------------------------------------------------------
# the oracle's interpolation, pinned by polynomial exactness
# --------------------------------------------------------------------------
def _parity_poly(r, odd, c):
    f = sum(cj * r ** (2 * j) for j, cj in enumerate(c))
    return r * f if odd else f


@pytest.mark.parametrize("odd", [False, True])
def test_oracle_interp_exact_on_parity_polynomials(odd):
    """Across the axis the file's points are continued as -r(KL:1:-1) with
    the field's parity (-1)^(m Mp + S).  A local polynomial through the
    stencil is exact on polynomials it can represent: degree <= 4 everywhere
    (the wall stencil keeps 5 points), degree <= 8 where the full 9-point
    stencil fits.  A wrong sign across the axis fails the points near r = 0."""
    rng = np.random.default_rng(3 + odd)
    rf = O.build_mesh(40).r
    ro = O.build_mesh(57).r
    xi = np.concatenate([-rf[KL - 1::-1], rf])
    wts = O.io_interp_wts(xi, ro)
    full = np.array([len(w) == 2 * KL + 1 for _, w in wts])
    assert full.sum() > 40 and not full[-1]
    sgn = -1.0 if odd else 1.0
    for nterms, mask in ((2 if odd else 3, np.ones_like(full)), (4 if odd else 5, full)):
        c = rng.standard_normal(nterms)       # degree 2*nterms - 1 (odd) or 2*nterms - 2 (even)
        f = _parity_poly(rf, odd, c)
        want = _parity_poly(ro, odd, c)
        got = O.io_interp(wts, np.concatenate([sgn * f[KL - 1::-1], f]))
        err = np.abs(got - want)[mask].max() / np.abs(want).max()
        assert err <= 1e-12, (nterms, err)
        bad = O.io_interp(wts, np.concatenate([-sgn * f[KL - 1::-1], f]))
        assert np.abs(bad - want).max() > 1e-6 * np.abs(want).max()


def test_oracle_interp_identity_on_same_points():
    """Output points on the file's points take the file's values (the
    Taylor matrix row of the expansion point is e_1)."""
    rf = O.build_mesh(32).r
    xi = np.concatenate([-rf[KL - 1::-1], rf])
    f = np.cos(3 * rf) + rf ** 3
    fn = np.concatenate([f[KL - 1::-1], f])
    got = O.io_interp(O.io_interp_wts(xi, rf), fn)
    assert np.abs(got - f).max() <= 1e-13


def test_oracle_load_coll_matches_harmonics_by_km():
    rng = np.random.default_rng(5)
    r = O.build_mesh(20).r
    Kf, Mf, K, M = 3, 4, 4, 2
    a = _rand_fields(rng, 20, O.n_harmonics(Kf, Mf))[0]
    out = O.io_load_coll(a, Kf, Mf, r, K, M, 1, r, 0)
    hf = _harmonics(Kf, Mf)
    for h, km in enumerate(_harmonics(K, M)):
        if km in hf:
            assert np.array_equal(out[:, h], a[:, hf.index(km)])
        else:
            assert not out[:, h].any()


# --------------------------------------------------------------------------
# the library's codec and reading
# --------------------------------------------------------------------------
def test_library_file_read_by_scipy(tmp_path):
    import opf
    rng = np.random.default_rng(11)
    N, K, M, Mp = 24, 3, 4, 2
    H = O.n_harmonics(K, M)
    r = O.build_mesh(N).r
    fields = _rand_fields(rng, N, H)
    p = tmp_path / "state0000.cdf.dat"
    opf.write_state_file(p, r, K, M, Mp, *fields, t=12.5, Re=2400.0, alpha=1.25, dt=0.01, dtcor=0.02,
                         dtcfl=0.03)
    with open(p, "rb") as fh:
        assert fh.read(4) == b"CDF\x01"      # classic format, as nf90_create(nf90_clobber)
    with netcdf_file(p, "r", mmap=False) as f:
        assert f.t == 12.5 and f.Re == 2400.0 and f.alpha == 1.25
        assert f.dimensions == {"r": N, "H": H, "ReIm": 2}
        assert np.array_equal(f.variables["r"][:], r)
        assert f.variables["dt"].getValue() == 0.01
        assert f.variables["dtcor"].getValue() == 0.02
        assert f.variables["dtcfl"].getValue() == 0.03
        for nm, a in zip(("Ur", "Ut", "Uz"), fields):
            v = f.variables[nm]
            assert v.dimensions == ("ReIm", "H", "r")
            assert (int(v.K), int(v.M), int(v.Mp)) == (K, M, Mp)
            assert np.array_equal(v[0], a.real.T) and np.array_equal(v[1], a.imag.T)


@pytest.mark.parametrize("version", [1, 2])
def test_reference_layout_read_same_grid_bitwise(tmp_path, version):
    import opf
    rng = np.random.default_rng(12)
    N, K, M, Mp = 20, 4, 3, 1
    r = O.build_mesh(N).r
    fields = _rand_fields(rng, N, O.n_harmonics(K, M))
    p = tmp_path / "s.cdf.dat"
    _write_reference_layout(p, r, K, M, Mp, fields, 3.25, 1800.0, 0.75, 0.005, 0.004, 0.006, version=version)
    info = opf.state_file_info(p)
    assert (info["N"], info["K"], info["M"], info["Mp"]) == (N, K, M, Mp)
    assert (info["t"], info["Re"], info["alpha"], info["dt"]) == (3.25, 1800.0, 0.75, 0.005)
    s = opf.read_state_file(p, N, K, M, Mp)
    assert (s["t"], s["dt"], s["dtcor"], s["dtcfl"]) == (3.25, 0.005, 0.004, 0.006)
    for nm, a in zip(("ur", "ut", "uz"), fields):
        assert np.array_equal(s[nm], a)


def test_round_trip_bitwise(tmp_path):
    import opf
    rng = np.random.default_rng(13)
    N, K, M, Mp = 33, 5, 6, 3
    r = O.build_mesh(N).r
    fields = _rand_fields(rng, N, O.n_harmonics(K, M))
    p = tmp_path / "rt.cdf.dat"
    opf.write_state_file(p, r, K, M, Mp, *fields, t=1.0, Re=5000.0, alpha=1.0, dt=1e-3)
    s = opf.read_state_file(p, N, K, M, Mp, r=r)
    for nm, a in zip(("ur", "ut", "uz"), fields):
        assert np.array_equal(s[nm], a)


@pytest.mark.parametrize("Kf,Mf,K,M", [(3, 4, 2, 6), (2, 2, 4, 5), (5, 3, 3, 5), (3, 5, 5, 3)])
def test_resize_in_k_and_m_matches_oracle(tmp_path, Kf, Mf, K, M):
    """Harmonics carried by (k, m); the last two cases are larger in one of
    K, M and smaller in the other."""
    import opf
    rng = np.random.default_rng(20 + Kf * 10 + Mf)
    N, Mp = 18, 1
    r = O.build_mesh(N).r
    fields = _rand_fields(rng, N, O.n_harmonics(Kf, Mf))
    p = tmp_path / "rs.cdf.dat"
    _write_reference_layout(p, r, Kf, Mf, Mp, fields, 0.0, 2400.0, 1.25, 0.01, 0.0, 0.0)
    s = opf.read_state_file(p, N, K, M, Mp)
    hf = _harmonics(Kf, Mf)
    for nm, a, S in zip(("ur", "ut", "uz"), fields, (1, 1, 0)):
        want = O.io_load_coll(a, Kf, Mf, r, K, M, Mp, r, S)
        assert np.array_equal(s[nm], want)
        kept = [km for km in _harmonics(K, M) if km in hf]
        assert len(kept) == min(Kf, K) * 2 * min(Mf, M) - min(Mf, M) - (min(Kf, K) - 1)


@pytest.mark.parametrize("Nf,N,Mp", [(40, 57, 1), (64, 36, 2), (30, 30, 1)])
def test_radial_interpolation_matches_oracle(tmp_path, Nf, N, Mp):
    """A file on other radial points (or the same count on a perturbed mesh)
    read onto the default mesh: the library's result equals the oracle's
    io_load_coll to rounding on random (non-polynomial) fields."""
    import opf
    rng = np.random.default_rng(Nf + N)
    K, M = 3, 3
    rf = O.build_mesh(Nf).r
    if Nf == N:
        rf = rf.copy()
        rf[:-1] = rf[:-1] * (1.0 - 1e-3 * rf[:-1] * (1 - rf[:-1]))
    fields = _rand_fields(rng, Nf, O.n_harmonics(K, M))
    p = tmp_path / "ri.cdf.dat"
    _write_reference_layout(p, rf, K, M, Mp, fields, 0.0, 2400.0, 1.25, 0.01, 0.0, 0.0)
    r = O.build_mesh(N).r
    s = opf.read_state_file(p, N, K, M, Mp)
    for nm, a, S in zip(("ur", "ut", "uz"), fields, (1, 1, 0)):
        want = O.io_load_coll(a, K, M, rf, K, M, Mp, r, S)
        assert np.abs(s[nm] - want).max() <= 1e-12 * np.abs(want).max(), nm


def test_read_errors(tmp_path):
    import opf
    from opf._lib import OpfError
    with pytest.raises(OpfError, match="not found"):
        opf.read_state_file(tmp_path / "missing.cdf.dat", 16, 2, 2)
    junk = tmp_path / "junk.dat"
    junk.write_bytes(b"HDF\x01" + bytes(60))
    with pytest.raises(OpfError, match="netCDF"):
        opf.read_state_file(junk, 16, 2, 2)
    r = O.build_mesh(16).r
    p = tmp_path / "mp.cdf.dat"
    fields = _rand_fields(np.random.default_rng(1), 16, O.n_harmonics(2, 4))
    _write_reference_layout(p, r, 2, 4, 2, fields, 0.0, 1.0, 1.0, 0.01, 0.0, 0.0)
    with pytest.raises(OpfError, match="Mp"):          # io.f90:193
        opf.read_state_file(p, 16, 2, 4, 1)
    trunc = tmp_path / "trunc.cdf.dat"
    trunc.write_bytes(p.read_bytes()[:-64])
    with pytest.raises(OpfError, match="truncated"):
        opf.read_state_file(trunc, 16, 2, 4, 2)
