"""Seeded initial states: Hagen-Poiseuille plus a random perturbation.

The stored state is the perturbation only; the base flow U(r) = 1 - r^2 is
implicit in the time step (velocity.f90:53-54).  The perturbation is built
from a toroidal potential psi and a poloidal potential phi per harmonic,

    u = curl(psi z^) + curl curl(phi z^),

evaluated analytically (exact polynomial derivatives) at the radial points:

    u_r = (i m/r) psi + i a k phi'
    u_t = -psi' - (m a k / r) phi
    u_z = -(phi'' + phi'/r - m^2 phi / r^2)

with psi = c_psi r^m (1-r^2)^2 and phi = c_phi r^m (1-r^2)^3.  The field is
therefore exactly divergence-free, regular on the axis and satisfies no-slip
at r = 1 (psi(1) = psi'(1) = 0, phi(1) = phi'(1) = phi''(1) = 0); at k = m = 0
it carries zero net axial flux, matching the constant-flux runs.  The complex
coefficients are standard Gaussians from numpy's PCG64 generator, drawn per
harmonic in the harmonic order of parallel.h:47-56, scaled by
exp(-(k^2 + m^2)/16).  The state is finally rescaled so that its energy, as
io_write_energy sums var_coll_norm over the three components
(io.f90:501-524, variables.f90:718-751), equals the requested value.
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial import polynomial as P

D_PI = 3.1415926535897931


def harmonic_km(K, M):
    """(k, m) per harmonic index, parallel.h:47-56."""
    ks, ms = [], []
    for m in range(M):
        for k in range(-(K - 1), K):
            if k < 0 and m == 0:
                continue
            ks.append(k)
            ms.append(m)
    return np.array(ks), np.array(ms)


def energy(fields, intrdr, K, M, alpha):
    """Sum over components of var_coll_norm's E (variables.f90:718-751)."""
    ks, ms = harmonic_km(K, M)
    w = 4.0 * D_PI * D_PI / alpha
    E = 0.0
    for a in fields:
        f = w * (a.real * a.real + a.imag * a.imag)
        b = np.zeros(a.shape[1])
        for n in range(a.shape[0]):
            b = b + f[n] * intrdr[n]
        b[0] = 0.5 * b[0]
        Em = np.zeros(M)
        for h in range(a.shape[1]):
            Em[ms[h]] += b[h]
        s = 0.0
        for v in Em:
            s = s + v
        E = E + s
    return E


def _poly_rm(m, p):
    """Coefficients of r^m (1 - r^2)^p."""
    base = P.polypow([1.0, 0.0, -1.0], p)
    return np.concatenate([np.zeros(m), base])


def _div_r(c, times=1):
    """Exact division of a polynomial by r**times (low coefficients must vanish)."""
    c = np.asarray(c, dtype=float)
    if times <= 0:
        return c
    if np.any(np.abs(c[:times]) > 1e-12 * max(1.0, np.abs(c).max())):
        raise ValueError("polynomial not divisible by r")
    return c[times:] if len(c) > times else np.zeros(1)


def _radial_profiles(r, mm):
    """Values at r of the potentials' derivative combinations for azimuthal
    index mm: psi/r, psi', phi/r, phi', and the Laplacian term of u_z."""
    N = r.size
    psi = _poly_rm(mm, 2)
    phi = _poly_rm(mm, 3)
    dpsi = P.polyder(psi)
    dphi = P.polyder(phi)
    ddphi = P.polyder(dphi)
    # (r phi' - m^2 phi) / r^2
    q = P.polysub(P.polymulx(dphi), mm * mm * phi)
    lap_rest = _div_r(q, 2)
    v_dpsi = P.polyval(r, dpsi)
    v_dphi = P.polyval(r, dphi)
    v_lap = P.polyval(r, ddphi) + P.polyval(r, lap_rest)
    if mm > 0:
        v_psi_r = P.polyval(r, _div_r(psi, 1))
        v_phi_r = P.polyval(r, _div_r(phi, 1))
    else:
        v_psi_r = np.zeros(N)
        v_phi_r = np.zeros(N)
    return v_psi_r, v_dpsi, v_phi_r, v_dphi, v_lap


def seeded_perturbation(r, intrdr, K, M, alpha, seed, target_energy, Mp=1):
    """Return (ur, ut, uz), complex arrays of shape (N, H), with the given
    energy.  Deterministic in (r, K, M, Mp, alpha, seed, target_energy)."""
    r = np.asarray(r, dtype=float)
    N = r.size
    ks, ms = harmonic_km(K, M)
    H = ks.size
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((H, 4))      # the per-harmonic draws of 4, in harmonic order
    env = np.exp(-(ks * ks + ms * ms) / 16.0)
    cpsi = (g[:, 0] + 1j * g[:, 1]) * env
    cphi = (g[:, 2] + 1j * g[:, 3]) * env
    cpsi[0] = cpsi[0].real
    cphi[0] = cphi[0].real
    ur = np.zeros((N, H), np.complex128)
    ut = np.zeros((N, H), np.complex128)
    uz = np.zeros((N, H), np.complex128)
    for m in range(M):
        sel = np.nonzero(ms == m)[0]
        mm = m * Mp
        v_psi_r, v_dpsi, v_phi_r, v_dphi, v_lap = _radial_profiles(r, mm)
        ak = alpha * ks[sel]
        cp, cf = cpsi[sel], cphi[sel]
        ur[:, sel] = (1j * mm * v_psi_r)[:, None] * cp[None, :] + (1j * v_dphi[:, None] * ak[None, :]) * cf[None, :]
        ut[:, sel] = (-v_dpsi)[:, None] * cp[None, :] - (v_phi_r[:, None] * (mm * ak)[None, :]) * cf[None, :]
        uz[:, sel] = (-v_lap)[:, None] * cf[None, :]
    ur[:, 0] = 0.0
    ut[:, 0] = ut[:, 0].real
    uz[:, 0] = uz[:, 0].real
    E = energy((ur, ut, uz), intrdr, K, M, alpha)
    s = np.sqrt(target_energy / E)
    return ur * s, ut * s, uz * s
