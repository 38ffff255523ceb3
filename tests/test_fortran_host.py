"""The Fortran drop-in (softx-d-17-00015_amd/fortran): modules named and
shaped as the reference's (mpif, meshs, variables, transform, timestep,
velocity, io) whose procedures call the C ABI.

CPU: the modules build against a parameters module and export the
reference's procedure names.  GPU: the reference's own main.f90, compiled
unmodified with its own parameters.f90 against these modules
(fortran/build_dropin.sh, from __graft_entry__.build), run from a state file
of the reference's initial state: its vel_energy.dat, vel_friction.dat,
tim_step.dat and final state file match the reference's own run."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import PKG, load_golden, rel_err

FDIR = os.path.join(PKG, "fortran")
DROPIN = os.path.join(FDIR, "_dropin")


@pytest.mark.skipif(shutil.which("amdflang") is None, reason="amdflang not available")
def test_dropin_modules_build():
    """The modules compile against a parameters module and provide the
    reference's procedures (timestep.f90, velocity.f90, io.f90 names)."""
    with tempfile.TemporaryDirectory() as td:
        subprocess.run(["make", "-s", "-C", FDIR, f"BUILD={td}"], check=True)
        lib = os.path.join(td, "libopf_dropin.a")
        assert os.path.exists(lib)
        syms = subprocess.run(["nm", lib], check=True, capture_output=True, text=True).stdout
        for mod, proc in (("velocity", "vel_predictor"), ("velocity", "vel_corrector"), ("velocity", "vel_matrices"),
                          ("velocity", "vel_precompute"), ("timestep", "tim_lumesh_init"),
                          ("timestep", "tim_lumesh_invert"), ("timestep", "tim_check_cgce"),
                          ("transform", "tra_spec2phys"), ("transform", "tra_phys2spec"),
                          ("variables", "var_coll_norm"), ("variables", "var_coll2spec"),
                          ("io", "io_load_state"), ("io", "io_write2files")):
            assert f"M{mod}P{proc}" in syms, (mod, proc)
        for sym in ("opf_tim_lumesh_init", "opf_vel_predictor", "opf_var_coll_norm", "opf_state_file_read"):
            assert sym in syms, sym


def _run_dropin(name, td, state_in, timeout=300):
    exe = os.path.join(DROPIN, f"opf_dropin_{name}")
    if not os.access(exe, os.X_OK):
        pytest.skip(f"{exe} not built (needs the reference's main.f90 at build time)")
    shutil.copy(state_in, os.path.join(td, "state.cdf.in"))
    print("drop-in main:", name, flush=True)
    r = subprocess.run([exe], cwd=td, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "...done!" in r.stdout, r.stdout
    return r.stdout


def _state_file(opf, path, g, t=0.0, dt=None):
    N, K, M = int(g["N"]), int(g["K"]), int(g["M"])
    r = opf.host_mesh(N)["r"]
    ic = g["ic"]
    opf.write_state_file(path, r, K, M, 1, ic[0], ic[1], ic[2], t=t, Re=float(g["Re"]),
                         alpha=float(g["alpha"]), dt=abs(float(g["dt"])) if dt is None else dt)


@pytest.mark.gpu
def test_reference_main_c0_matches_reference_run():
    """main.f90 + parameters.f90 (32,24,12), Re 2400, alpha 1.25, dt 0.01,
    100 steps (i_maxtstep), energy and friction every 10 steps
    (i_save_rate2): every line of vel_energy.dat and vel_friction.dat equals
    the reference's step to 1e-10 (the files print 12 digits), the state
    saved at the end (state0001.cdf.dat) equals the reference's after 100
    steps to 1e-10, with t = 1."""
    import opf
    g = load_golden("c0")
    with tempfile.TemporaryDirectory() as td:
        _state_file(opf, os.path.join(td, "in.cdf"), g)
        _run_dropin("c0", td, os.path.join(td, "in.cdf"))
        en = np.loadtxt(os.path.join(td, "vel_energy.dat"), ndmin=2)
        fr = np.loadtxt(os.path.join(td, "vel_friction.dat"), ndmin=2)
        end = opf.read_state_file(os.path.join(td, "state0001.cdf.dat"), 32, 24, 12)
        start = opf.read_state_file(os.path.join(td, "state0000.cdf.dat"), 32, 24, 12)
    rs = g["series"]
    steps = np.arange(0, 100, 10)
    assert en.shape == (10, 4) and fr.shape == (10, 4)
    assert np.allclose(en[:, 0], rs[steps, 1], rtol=0, atol=1e-10)
    for col, ref in ((1, 2), (2, 3), (3, 4)):
        assert np.max(np.abs(en[:, col] - rs[steps, ref]) / np.abs(rs[steps, ref])) <= 1e-10, col
    for col, ref in ((1, 5), (2, 6), (3, 7)):
        assert np.max(np.abs(fr[:, col] - rs[steps, ref]) / np.abs(rs[steps, ref])) <= 1e-10, col
    for c, nm in enumerate(("ur", "ut", "uz")):
        assert rel_err(end[nm], g["state_100"][c]) <= 1e-10, nm
    assert abs(end["t"] - 1.0) <= 1e-12 and start["t"] == 0.0


@pytest.mark.gpu
def test_reference_main_mini_every_step_and_restart():
    """main.f90 at (16,6,4): the energy of every step of a 10-step run; a
    4-step run whose final state file restarts a run with d_time < 0 (tim_t
    from the file) for 6 more steps ends bit for bit where the 10-step run
    ends, and on the reference's state after 10 steps (1e-10)."""
    import opf
    g = load_golden("mini")
    with tempfile.TemporaryDirectory() as td:
        a, b, c = (os.path.join(td, x) for x in ("a", "b", "c"))
        for d in (a, b, c):
            os.mkdir(d)
        _state_file(opf, os.path.join(td, "in.cdf"), g)
        _run_dropin("mini", a, os.path.join(td, "in.cdf"))
        en = np.loadtxt(os.path.join(a, "vel_energy.dat"), ndmin=2)
        ten = opf.read_state_file(os.path.join(a, "state0001.cdf.dat"), 16, 6, 4)
        _run_dropin("mini4", b, os.path.join(td, "in.cdf"))
        four = os.path.join(b, "state0001.cdf.dat")
        assert abs(opf.state_file_info(four)["t"] - 0.04) <= 1e-15
        _run_dropin("mini6r", c, four)
        rest = opf.read_state_file(os.path.join(c, "state0001.cdf.dat"), 16, 6, 4)
    rs = g["series"]
    assert en.shape == (10, 4)
    assert np.max(np.abs(en[:, 1] - rs[:10, 2]) / rs[:10, 2]) <= 1e-10
    for k, nm in enumerate(("ur", "ut", "uz")):
        assert np.array_equal(rest[nm], ten[nm]), nm
        assert rel_err(ten[nm], g["state_10"][k]) <= 1e-10, nm
    assert abs(rest["t"] - 0.1) <= 1e-12


@pytest.mark.gpu
def test_reference_main_adaptive_matches_reference():
    """main.f90 with d_timestep < 0 (main.f90:24-29): the initial tim_dt
    from the state file, vel_maxtstep on the device, tim_new_tstep and
    vel_matrices on a change; every tim_dt change in tim_step.dat equals
    the reference's, and the final state equals the reference's after 120
    steps (1e-10)."""
    import opf
    g = load_golden("adapt_mini")
    with tempfile.TemporaryDirectory() as td:
        _state_file(opf, os.path.join(td, "in.cdf"), g, dt=0.01)
        _run_dropin("adapt_mini", td, os.path.join(td, "in.cdf"))
        lines = open(os.path.join(td, "tim_step.dat")).read().split("\n")
        end = opf.read_state_file(os.path.join(td, "state0001.cdf.dat"), 16, 6, 4)
    changes = [ln.split() for ln in lines if len(ln.split()) == 6]    # format 11: t step dt corr_dt cfl_dt dir
    ref = g["dtser"]
    ref_changes = ref[ref[:, 5] == 1]
    assert len(changes) == len(ref_changes)
    for got, want in zip(changes, ref_changes):
        assert int(got[1]) == int(want[0])
        assert abs(float(got[2]) - want[1]) <= 1e-8 * want[1]          # printed with 9 digits
    for c, nm in enumerate(("ur", "ut", "uz")):
        assert rel_err(end[nm], g["state_120"][c]) <= 1e-10, nm
