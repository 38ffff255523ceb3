"""Write the golden fixtures from the compiled reference (oracle/_ref).

Run in the build container only (the reference binaries are built there by
oracle/build_ref.sh).  For each workload the seeded initial state is produced
by the package's generator on the oracle's mesh, written in the reference's
raw coll layout, and integrated by the reference executable; the fixture keeps
the initial state, the reference's states after chosen steps, its step-1 stage
dumps (mini only) and its per-step series (t, E, Ek0, Em0, Pr0, Uc, Ufr,
ncorr, dterr); adaptive runs (dt < 0) add the reference's per-step timestep
quantities.  Fixtures are data only: inputs and the reference's outputs.

usage: python tests/golden/make_golden.py [mini c0 s48 c1 c2 mesh ...]
"""
from __future__ import annotations

import os
import resource
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "softx-d-17-00015_amd"))
import opf_oracle as O  # noqa: E402
from opf import ic  # noqa: E402

# name: (N, K, M, Re, alpha, dt, E0, seed, nsteps, state steps saved, keep states fully)
WORKLOADS = {
    "mini": (16, 6, 4, 2400.0, 1.25, 0.01, 1e-3, 1, 100, (1, 2, 10, 100), True),
    "c0": (32, 24, 12, 2400.0, 1.25, 0.01, 1e-3, 1, 100, (1, 100), True),
    "s48": (48, 48, 24, 3000.0, 1.25, 0.01, 1e-3, 7, 50, (1,), False),
    "c1": (64, 64, 48, 5300.0, 0.5, 0.005, 1e-2, 1, 1000, (1,), False),
    "c2": (96, 256, 128, 5300.0, 0.4, 0.0025, 1e-2, 1, 20, (1, 20), False),
    # adaptive time step (dt < 0: |dt| is the initial tim_dt, main.f90:24-29)
    "adapt_mini": (16, 6, 4, 2400.0, 1.25, -0.01, 1e-3, 1, 120, (1, 120), True),
    "adapt_c0": (32, 24, 12, 2400.0, 1.25, -0.01, 1e-3, 1, 200, (200,), True),
    # var_imposesym(0..3) of the initial state (nsteps < 0: no time loop)
    "sym_mini": (16, 6, 4, 2400.0, 1.25, 0.01, 1e-3, 1, -1, (), True),
    # shift-reflect symmetry imposed every step (b_shiftrefl, parameters.f90:27)
    "sr_mini": (16, 6, 4, 2400.0, 1.25, 0.01, 1e-3, 1, 20, (1, 20), True),
    # the variable-viscosity plugins at the var_null(2) points (one corrector
    # per step, tim_pcn = 1), parameters in PLUGIN_PARAMS
    "nn_mini": (16, 6, 4, 2400.0, 1.25, 0.01, 1e-3, 1, 10, (1, 10), True),
    "les_smag_mini": (16, 6, 4, 2400.0, 1.25, 0.01, 1e-3, 1, 10, (1, 10), True),
    "les_smag3_mini": (16, 6, 4, 2400.0, 1.25, 0.01, 1e-3, 1, 10, (1, 10), True),
    "les_vreman_mini": (16, 6, 4, 2400.0, 1.25, 0.01, 1e-3, 1, 10, (1, 10), True),
}
# executable of each workload (the reference is compiled per resolution,
# symmetry flag and plugin: oracle/build_ref.sh mini_sr 16 6 4 1 shiftrefl,
# PLUGIN=les oracle/build_ref.sh les_mini 16 6 4)
EXE = {"adapt_mini": "mini", "adapt_c0": "c0", "sym_mini": "mini", "sr_mini": "mini_sr", "nn_mini": "nn_mini",
       "les_smag_mini": "les_mini", "les_smag3_mini": "les_mini", "les_vreman_mini": "les_mini"}
# params_model.in of the plugin workloads (nonnewtonian.f90:69-80: rho radius /
# model / mu_0 mu_inf / lam_CY n_CY a_CY / fc1; LES.f90:76-78: model /
# C_s filter A_plus) and the same parameters as recorded in the fixture
PLUGIN_PARAMS = {
    "nn_mini": ("1000.0 0.01\n1\n0.056 0.0056\n3e-3 0.6 2.0\n0.0\n",
                dict(rho=1000.0, radius=0.01, mu_0=0.056, mu_inf=0.0056, lam_CY=3e-3, n_CY=0.6, a_CY=2.0)),
    "les_smag_mini": ("1\n0.1 1 25.0\n", dict(model=1, C_s=0.1, filter=1, A_plus=25.0)),
    "les_smag3_mini": ("1\n0.1 3 25.0\n", dict(model=1, C_s=0.1, filter=3, A_plus=25.0)),
    "les_vreman_mini": ("2\n0.1 2 25.0\n", dict(model=2, C_s=0.1, filter=2, A_plus=25.0)),
}
SAMPLE_STRIDE = 37      # harmonics kept from the large states (every 37th)
SAMPLE_COUNT = 400      # at most this many harmonics, evenly spaced, from the largest states

# steporbit (newton.f90:721-775) runs of the reference's procedures from the
# seeded state of a workload: name: (workload, executable, [(ndts, T, sth, sz)]);
# ndts = 1 is multJ's one step with the workload's dt, no shift.
ORBITS = {
    "orbit_mini": ("mini", "mini", [(20, 0.2, 0.3, 0.7), (1, 0.0, 0.0, 0.0), (25, 0.2, 0.0, -0.4)]),
    "orbit_sr_mini": ("sr_mini", "mini_sr", [(20, 0.2, 0.0, 0.5)]),
}


# newtonhook (NewtonHook.f90, GMRESm.f90) of the reference on a small algebraic
# system (oracle/newton_driver.f90; built by oracle/build_newton_ref.sh):
# (n, m, nits, gtol, tol, del, mndl, mxdl, epsJ, amp) per case.
NEWTON_CASES = [
    (10, 8, 30, 1e-3, 1e-12, 0.5, 1e-10, 10.0, 1e-7, 1.5),
    (10, 8, 40, 1e-3, 1e-12, 0.5, 1e-10, 10.0, 1e-7, 4.0),
    (12, 6, 40, 1e-2, 1e-12, -1.0, 1e-10, 10.0, 1e-7, 3.0),
    (10, 20, 40, 1e-3, 1e-12, 0.2, 1e-10, 0.4, 1e-7, 6.0),
]


def make_newton(name):
    exe = os.path.join(REPO, "oracle", "_ref", "newton_ref")
    if not os.path.exists(exe):
        raise SystemExit(f"{exe} missing: run oracle/build_newton_ref.sh")
    out = dict(cases=np.array(NEWTON_CASES, dtype=np.float64))
    with tempfile.TemporaryDirectory() as td:
        for i, c in enumerate(NEWTON_CASES):
            f = os.path.join(td, f"nt{i}.txt")
            env = dict(os.environ, MKL_THREADING_LAYER="SEQUENTIAL", MKL_NUM_THREADS="1")
            subprocess.run([exe, f] + [repr(v) for v in c], check=True, cwd=td, env=env,
                           stdout=subprocess.DEVNULL)
            lines = open(f).read().strip().splitlines()
            out[f"info_{i}"] = np.array(int(lines[-1].split()[1]))
            out[f"iters_{i}"] = np.array([[float(v) for v in ln.split()] for ln in lines[:-1]])
    path = os.path.join(HERE, f"golden_{name}.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


def arnoldi_matrix(n=40, seed=5):
    """A = Q T Q^T with a known dominant spectrum: 0.95 e^{+-0.4i}, 0.9,
    0.8 e^{+-1.1i}, 0.7, the rest in [0, 0.5); T upper block triangular."""
    rng = np.random.default_rng(seed)
    T = np.triu(rng.standard_normal((n, n)) * 0.05, 1)

    def blk(r, th):
        return np.array([[r * np.cos(th), -r * np.sin(th)], [r * np.sin(th), r * np.cos(th)]])

    T[0:2, 0:2] = blk(0.95, 0.4)
    T[2, 2] = 0.9
    T[3:5, 3:5] = blk(0.8, 1.1)
    T[5, 5] = 0.7
    for i in range(6, n):
        T[i, i] = 0.5 * rng.random()
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return Q @ T @ Q.T


def make_arnoldi(name):
    """arnold + logeig of the reference (oracle/arnoldi_driver.f90) on a fixed matrix."""
    exe = os.path.join(REPO, "oracle", "_ref", "arnoldi_ref")
    if not os.path.exists(exe):
        raise SystemExit(f"{exe} missing: run oracle/build_newton_ref.sh")
    n, kmax, ncgd, tol = 40, 30, 4, 1e-10
    A = arnoldi_matrix(n)
    out = dict(A=A, n=n, kmax=kmax, ncgd=ncgd, tol=tol)
    with tempfile.TemporaryDirectory() as td:
        fa = os.path.join(td, "A.bin")
        A.T.astype("<f8").tofile(fa)                       # column-major
        env = dict(os.environ, MKL_THREADING_LAYER="SEQUENTIAL", MKL_NUM_THREADS="1")
        subprocess.run([exe, fa, str(n), str(kmax), str(ncgd), repr(tol), "ar.txt", "ar.bin"], check=True,
                       cwd=td, env=env, stdout=subprocess.DEVNULL)
        lines = open(os.path.join(td, "ar.txt")).read().split("\n")
        out["k"], out["ifail"] = (np.array(int(v)) for v in lines[0].split())
        out["eig"] = np.array([[float(v) for v in ln.split()] for ln in lines[1:1 + ncgd]])
        out["vecs"] = np.fromfile(os.path.join(td, "ar.bin"), dtype="<f8").reshape(2 * ncgd, n)
    path = os.path.join(HERE, f"golden_{name}.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


def make_orbit(name):
    base, exe_name, cases = ORBITS[name]
    N, K, M, Re, alpha, dt, E0, seed = WORKLOADS[base][:8]
    exe = os.path.join(REPO, "oracle", "_ref", f"opf_ref_{exe_name}")
    if not os.path.exists(exe):
        raise SystemExit(f"{exe} missing: run oracle/build_ref.sh")
    mesh = O.build_mesh(N)
    ur, ut, uz = ic.seeded_perturbation(mesh.r, mesh.intrdr, K, M, alpha, seed, E0)
    H = ur.shape[1]
    out = dict(N=N, K=K, M=M, Re=Re, alpha=alpha, dt=dt, E0=E0, seed=seed, r=mesh.r,
               ic=np.stack([ur, ut, uz]), cases=np.array(cases, dtype=np.float64))
    with tempfile.TemporaryDirectory() as td:
        icf = os.path.join(td, "ic.bin")
        O.write_coll3(icf, ur, ut, uz)
        for i, (ndts, T, sth, sz) in enumerate(cases):
            pre = os.path.join(td, f"{name}_{i}")
            env = dict(os.environ, MKL_THREADING_LAYER="SEQUENTIAL", MKL_NUM_THREADS="1",
                       OPF_ORBIT=f"{ndts} {T!r} {sth!r} {sz!r}")
            subprocess.run([exe, icf, pre, repr(Re), repr(alpha), repr(dt), "0"], check=True, cwd=td,
                           env=env, stdout=subprocess.DEVNULL, preexec_fn=_unlimited_stack)
            out[f"orbit_{i}"] = np.stack(O.read_coll3(pre + "_orbit.bin", N, H))
    path = os.path.join(HERE, f"golden_{name}.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


# mes_D (meshs.f90:48-58) of the reference's mes_precompute for the radial
# resolutions it is built at: executable name per N
MESH_EXES = {16: "mini", 32: "c0", 48: "s48", 64: "c1", 96: "c2"}
KL = 4


def make_mesh(name):
    """The reference's radial mesh, finite-difference bands and quadrature
    weights, dumped by oracle/ref_driver.f90 (nsteps = -2)."""
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for N, exe_name in MESH_EXES.items():
            exe = os.path.join(REPO, "oracle", "_ref", f"opf_ref_{exe_name}")
            if not os.path.exists(exe):
                raise SystemExit(f"{exe} missing: run oracle/build_ref.sh")
            K, M = WORKLOADS[{"s48": "s48"}.get(exe_name, exe_name)][1:3]
            H = (2 * K - 1) * M - (K - 1)
            icf = os.path.join(td, f"z{N}.bin")
            np.zeros(6 * N * H).astype("<f8").tofile(icf)
            pre = os.path.join(td, f"m{N}")
            env = dict(os.environ, MKL_THREADING_LAYER="SEQUENTIAL", MKL_NUM_THREADS="1")
            subprocess.run([exe, icf, pre, "1.0", "1.0", "0.01", "-2"], check=True, cwd=td, env=env,
                           stdout=subprocess.DEVNULL, preexec_fn=_unlimited_stack)
            v = np.fromfile(pre + "_mesh.bin", dtype="<f8")
            sizes = [("r", N), ("intrdr", N), ("dr0", (KL + 1) ** 2), ("dr1", (KL + 1) ** 2),
                     ("dr1band", (2 * KL + 1) * N), ("dr2band", (2 * KL + 1) * N),
                     ("radlap", (2 * KL + 1) * N)]
            assert v.size == sum(n for _, n in sizes), (N, v.size)
            o = 0
            for key, n in sizes:
                out[f"{key}_{N}"] = v[o:o + n].copy()
                o += n
    path = os.path.join(HERE, f"golden_{name}.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


def _unlimited_stack():
    # The reference keeps large automatic coll temporaries (e.g. velocity.f90:417)
    # on the stack; its users run it with an unlimited stack.
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def make(name):
    N, K, M, Re, alpha, dt, E0, seed, nsteps, saves, full = WORKLOADS[name]
    exe = os.path.join(REPO, "oracle", "_ref", f"opf_ref_{EXE.get(name, name)}")
    if not os.path.exists(exe):
        raise SystemExit(f"{exe} missing: run oracle/build_ref.sh {name} {N} {K} {M}")
    mesh = O.build_mesh(N)
    ur, ut, uz = ic.seeded_perturbation(mesh.r, mesh.intrdr, K, M, alpha, seed, E0)
    H = ur.shape[1]
    out = dict(N=N, K=K, M=M, Re=Re, alpha=alpha, dt=dt, E0=E0, seed=seed,
               nsteps=nsteps, r=mesh.r)
    with tempfile.TemporaryDirectory() as td:
        icf = os.path.join(td, "ic.bin")
        O.write_coll3(icf, ur, ut, uz)
        pre = os.path.join(td, name)
        if name in PLUGIN_PARAMS:
            text, prm = PLUGIN_PARAMS[name]
            with open(os.path.join(td, "params_model.in"), "w") as f:
                f.write(text)
            out["plugin"] = np.array(list(prm.items()), dtype=object).astype(str)
        env = dict(os.environ, MKL_THREADING_LAYER="SEQUENTIAL", MKL_NUM_THREADS="1")
        subprocess.run([exe, icf, pre, repr(Re), repr(alpha), repr(dt), str(nsteps)]
                       + [str(s) for s in saves], check=True, cwd=td, env=env,
                       stdout=subprocess.DEVNULL, preexec_fn=_unlimited_stack)
        if nsteps < 0:
            out["ic"] = np.stack([ur, ut, uz])
            for i in range(4):
                out[f"sym_{i}"] = np.stack(O.read_coll3(f"{pre}_sym_{i}.bin", N, H))
            path = os.path.join(HERE, f"golden_{name}.npz")
            np.savez_compressed(path, **out)
            print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")
            return
        ser = np.loadtxt(pre + "_series.txt", ndmin=2)
        out["series"] = ser
        if dt < 0:   # step tim_dt tim_cfl_dt tim_cfl_dir tim_corr_dt tim_new_dt
            out["dtser"] = np.loadtxt(pre + "_dt.txt", ndmin=2)
        stride = max(SAMPLE_STRIDE, -(-H // SAMPLE_COUNT))
        sel = np.arange(0, H, stride) if not full else np.arange(H)
        out["sel"] = sel
        if full:
            out["ic"] = np.stack([ur, ut, uz])
        else:
            out["ic_sel"] = np.stack([ur, ut, uz])[:, :, sel]
        for s in saves:
            st = np.stack(O.read_coll3(f"{pre}_state_{s:06d}.bin", N, H))
            out[f"state_{s}"] = st[:, :, sel] if not full else st
            out[f"state_{s}_norm"] = np.array([np.abs(st[c]).max() for c in range(3)])
        if name == "mini":
            stages = sorted(f for f in os.listdir(td) if "_stage_" in f)
            for f in stages:
                tag = f.split("_stage_")[1].split(".")[0]
                out[f"stage_{tag}"] = np.stack(O.read_coll3(os.path.join(td, f), N, H))
        with open(pre + "_timing.txt") as fh:
            out["ref_seconds"] = float(fh.read().split()[0])
    path = os.path.join(HERE, f"golden_{name}.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")


if __name__ == "__main__":
    for nm in (sys.argv[1:] or ["mini", "c0", "s48", "c1", "adapt_mini", "adapt_c0", "sym_mini", "sr_mini",
                                "orbit_mini", "orbit_sr_mini", "newton_alg", "arnoldi_mat", "mesh", "c2",
                                "nn_mini", "les_smag_mini", "les_smag3_mini", "les_vreman_mini"]):
        if nm == "mesh":
            make_mesh(nm)
        elif nm == "newton_alg":
            make_newton(nm)
        elif nm == "arnoldi_mat":
            make_arnoldi(nm)
        else:
            (make_orbit if nm in ORBITS else make)(nm)
