"""Benchmark of the Openpipeflow time step on one MI355X.

Headline workload: one state at (N, K, M) = (64, 64, 48), Mp = 1,
alpha = 0.5, Re = 5300, dt = 0.005, implicitness 0.51, constant flux,
seeded perturbation of energy 1e-2, with the corrector iterated to
convergence as main.f90:35-41 does.  The timed region is `--steps`
consecutive steps of the device-resident loop after `--warmup` steps; the
state is in HBM before timing starts.

Printed: one JSON line with the metric (state-steps/s), the parity check of
the energy / friction-velocity series against the reference's own run
(tests/golden/golden_c1.npz, |dE|/E <= 1e-10), the HBM roofline of the
dominant kernel (HIP events on the library's stream), the reference CPU
code timed on this host's cores on a bounded sample of the same workload,
and two more legs: batched mode (512 mixed states at (32,24,12)) and the
request server (Poisson arrivals of three sizes, p50/p99 latency).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "softx-d-17-00015_amd"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E peak

WORKLOADS = {
    # name: N, K, M, Re, alpha, dt, E0, seed, golden fixture
    "c1": (64, 64, 48, 5300.0, 0.5, 0.005, 1e-2, 1, "golden_c1.npz"),
    "c0": (32, 24, 12, 2400.0, 1.25, 0.01, 1e-3, 1, "golden_c0.npz"),
    # the large size: no reference run to compare with (parity there is
    # checked through size-independent properties by the GPU tests)
    "c2": (96, 256, 128, 5300.0, 0.4, 0.0025, 1e-2, 1, None),
}


def alg_bytes(N, K, M):
    """Algorithmic (compulsory) bytes per launch of each kernel: every array
    the kernel must read or write once.  C = one complex coll field."""
    H = (2 * K - 1) * M - (K - 1)
    P = K * M
    Z = 3 * K
    C = 16 * N * H
    Y = 16 * N * Z * M                       # one field of the z-transformed intermediates
    # LU factors and pivots once per (|k|, m) pair, influence vectors and 4x4 inverses
    F = 4 * P * (8 * 13 * N + 4 * N) + H * (48 * N + 128)
    return dict(C=C, F=F,
                zfwd=6 * C + 6 * Y, theta=9 * Y, zbwd=3 * Y + 3 * C,
                # predictor: read N and u_, write u and curl u; each corrector
                # also reads N_ and the previous iterate
                step_pred=12 * C + F, step_corr=18 * C + F, curl=6 * C)


def alg_bytes_per_step(N, K, M, ncorr):
    """B_alg = (21 + 27 n_corr) C + (1 + n_corr) F, the two-phase compulsory
    traffic model (factors once per +-k pair, influence vectors per harmonic)."""
    H = (2 * K - 1) * M - (K - 1)
    C = 16 * N * H
    F = 4 * K * M * (8 * 13 * N + 4 * N) + H * (48 * N + 128)
    return (21 + 27 * ncorr) * C + (1 + ncorr) * F


def make_ic(ctx, wl):
    from opf import ic
    N, K, M, Re, alpha, dt, E0, seed, _ = wl
    me = ctx.mesh()
    return ic.seeded_perturbation(me["r"], me["intrdr"], K, M, alpha, seed, E0)


def cpu_baseline(name, wl, ic_fields, target_s):
    """Reference Fortran (oracle/_ref, compiled from the reference sources
    with MKL's LAPACK/FFTW3) on one host core, on a bounded number of steps
    of the same workload; falls back to the numpy oracle."""
    N, K, M, Re, alpha, dt, E0, seed, _ = wl
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import opf_oracle as O
    exe = os.path.join(REPO, "oracle", "_ref", f"opf_ref_{name}")
    per_step_guess = {"c1": 0.7, "c0": 0.025, "c2": 10.0}.get(name, 0.5)
    nsteps = max(3, int(target_s / per_step_guess))
    if os.path.exists(exe):
        try:
            with tempfile.TemporaryDirectory() as td:
                icf = os.path.join(td, "ic.bin")
                O.write_coll3(icf, *ic_fields)
                import resource

                def lim():
                    resource.setrlimit(resource.RLIMIT_STACK,
                                       (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
                env = dict(os.environ, MKL_NUM_THREADS="1", MKL_THREADING_LAYER="SEQUENTIAL",
                           OMP_NUM_THREADS="1")
                subprocess.run([exe, icf, os.path.join(td, "b"), repr(Re), repr(alpha), repr(dt),
                                str(nsteps)], check=True, cwd=td, env=env, preexec_fn=lim,
                               stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=300)
                secs = float(open(os.path.join(td, "b_timing.txt")).read().split()[0])
            return dict(value=nsteps / secs, unit="state-steps/s", cores=1, kind="reference",
                        sample=f"{nsteps} steps of the same workload, reference Fortran "
                               f"(amdflang -O2 + MKL sequential), 1 core, {secs:.1f} s")
        except Exception as e:                       # noqa: BLE001
            print(f"cpu_baseline: reference run failed ({e}); timing the oracle port",
                  file=sys.stderr)
    oc = O.Ctx(N, K, M, Re=Re, alpha=alpha, dt=dt)
    oc.set_state(*ic_fields)
    n = max(1, nsteps // 10)
    t0 = time.perf_counter()
    oc.run(n)
    secs = time.perf_counter() - t0
    return dict(value=n / secs, unit="state-steps/s", cores=1, kind="port",
                sample=f"{n} steps of the same workload, numpy oracle, 1 process, {secs:.1f} s")


# batched mode: many (32,24,12) states with their own Re, alpha and initial
# energy, advanced together (every launch covers all slots)
BATCH = dict(N=32, K=24, M=12, dt=0.01, seed=2024)


def batched_leg(nstates, steps):
    """Batched throughput, the same states advanced one at a time (plain
    per-state loop), and the parity of slot 0, which holds the reference's
    (32,24,12) case, against the reference's own energy series."""
    import opf
    from opf import ic
    N, K, M, dt = BATCH["N"], BATCH["K"], BATCH["M"], BATCH["dt"]
    rng = np.random.default_rng(BATCH["seed"])
    params = []
    for i in range(nstates):
        Re = float(rng.uniform(1800.0, 3000.0))
        alpha = float(rng.choice([1.0, 1.25, 1.7]))
        E0 = float(10.0 ** rng.uniform(-5.0, -2.0))
        params.append((Re, alpha, E0, i + 1))
    params[0] = (2400.0, 1.25, 1e-3, 1)           # the reference's (32,24,12) case
    ctx = opf.Context(N, K, M, capacity=nstates)
    gpath = os.path.join(REPO, "tests", "golden", "golden_c0.npz")
    golden = None
    if os.path.exists(gpath):
        with np.load(gpath) as g:
            golden = dict(ic=g["ic"], series=g["series"])
    t0 = time.perf_counter()
    me = None
    for s, (Re, alpha, E0, seed) in enumerate(params):
        ctx.setup(s, Re, alpha, dt)
        if me is None:
            me = ctx.mesh()
        if s == 0 and golden is not None:
            f = golden["ic"]
        else:
            f = ic.seeded_perturbation(me["r"], me["intrdr"], K, M, alpha, seed, E0)
        ctx.put_state(s, f[0], f[1], f[2])
    setup_s = time.perf_counter() - t0
    ctx.run(2, diag_every=10)                       # warm-up (untimed)
    t0 = time.perf_counter()
    ctx.run(steps, diag_every=10)
    secs = time.perf_counter() - t0
    value = nstates * steps / secs
    parity = {"reference": "golden_c0.npz (slot 0)", "checked_records": 0}
    if golden is not None:
        d = ctx.diag(0)
        rows = [r for r in d if int(r[0]) < golden["series"].shape[0]]
        if rows:
            d = np.array(rows)
            ref = golden["series"][d[:, 0].astype(int)]
            dE = np.abs(d[:, 2] - ref[:, 2]) / ref[:, 2]
            parity.update(checked_records=len(rows), max_rel_dE=float(dE.max()),
                          ok=bool(dE.max() <= 1e-10))
    ncorr = [ctx.status(s)["last_ncorr"] for s in range(nstates)]
    ctx.close()
    # plain per-state loop: the same kind of state advanced alone
    one = opf.Context(N, K, M)
    Re, alpha, E0, seed = params[1]
    one.setup(0, Re, alpha, dt)
    f = ic.seeded_perturbation(me["r"], me["intrdr"], K, M, alpha, seed, E0)
    one.put_state(0, *f)
    one.run(5)
    n1 = max(50, steps)
    t0 = time.perf_counter()
    one.run(n1, diag_every=10)
    alone = n1 / (time.perf_counter() - t0)
    one.close()
    return {"metric": "batched state-steps/s", "value": value, "unit": "state-steps/s",
            "config": {"workload": f"{nstates} states at (N,K,M)=({N},{K},{M}), Re~U[1800,3000], "
                                   f"alpha in {{1.0,1.25,1.7}}, E0 log-uniform [1e-5,1e-2], dt={dt}, "
                                   f"{steps} steps, corrector to convergence",
                       "states": nstates, "steps": steps},
            "ms_per_step": 1000.0 * secs / steps, "setup_s": setup_s,
            "plain_per_state_loop": alone, "speedup_vs_plain_loop": value / alone,
            "mean_last_ncorr": float(np.mean(ncorr)), "parity": parity}


# serving: Poisson arrivals of integration requests of three sizes (5:3:2),
# one corrector per step (steporbit), T ~ U[5, 50] at dt = 0.01
SERVE = dict(classes=[(32, 24, 12), (48, 48, 24), (64, 64, 48)], ratio=(5, 3, 2), dt=0.01, seed=77)


def serving_leg(nreq, rate, budget_gb, chunk):
    """p50/p99 request latency and throughput of the request server; parity
    of a served request against the reference, and of a served request
    against the same integration run alone."""
    import opf
    from opf import ic
    classes, ratio, dt = SERVE["classes"], np.array(SERVE["ratio"], float), SERVE["dt"]
    rng = np.random.default_rng(SERVE["seed"])
    per = [opf.bytes_needed(N, K, M, 1) for (N, K, M) in classes]
    w = ratio * np.array(per, float)
    caps = [max(1, int(budget_gb * 1e9 * wi / w.sum() / pi)) for wi, pi in zip(w, per)]
    meshes = [opf.host_mesh(N) for (N, _, _) in classes]
    reqs = []
    t = 0.0
    for i in range(nreq):
        cls = int(rng.choice(len(classes), p=ratio / ratio.sum()))
        Re = float(rng.uniform(2000.0, 5300.0))
        alpha = float(rng.choice([0.5, 1.25]))
        nsteps = int(round(float(rng.uniform(5.0, 50.0)) / dt))
        E0 = float(10.0 ** rng.uniform(-5.0, -3.0))
        t += float(rng.exponential(1.0 / rate))
        N, K, M = classes[cls]
        st = ic.seeded_perturbation(meshes[cls]["r"], meshes[cls]["intrdr"], K, M, alpha, 1000 + i, E0)
        reqs.append(dict(cls=cls, Re=Re, alpha=alpha, nsteps=nsteps, state=st, arrival=t))
    gpath = os.path.join(REPO, "tests", "golden", "golden_c0.npz")
    golden = None
    if os.path.exists(gpath):
        with np.load(gpath) as g:
            golden = dict(ic=g["ic"], state_100=g["state_100"])
        reqs[0].update(cls=0, Re=2400.0, alpha=1.25, nsteps=100, state=tuple(golden["ic"]))
    srv = opf.Server(classes, caps, chunk=chunk)
    t0 = time.perf_counter()
    ids = []
    for r in reqs:
        wait = r["arrival"] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        ids.append(srv.submit(r["cls"], r["Re"], r["alpha"], dt, r["nsteps"], r["state"], corr_fixed=1))
    while not all(srv.poll(i)["done"] for i in ids):
        time.sleep(0.005)
    polls = [srv.poll(i) for i in ids]
    lat = np.array([p["t_done"] - p["t_submit"] for p in polls])
    span = max(p["t_done"] for p in polls) - min(p["t_submit"] for p in polls)
    steps = sum(r["nsteps"] for r in reqs)
    failed = sum(1 for p in polls if p["status"] != 0)
    parity = {}
    if golden is not None:
        u = srv.result(ids[0], 0)
        parity["reference_c0_100_steps_rel_err"] = max(
            float(np.abs(u[k] - golden["state_100"][k]).max() / np.abs(golden["state_100"][k]).max())
            for k in range(3))
    # a request of the mix against the same integration alone
    j = next(k for k in range(1, len(reqs)) if reqs[k]["cls"] == 0 and polls[k]["status"] == 0)
    u = srv.result(ids[j], 0)
    srv.close()
    r = reqs[j]
    N, K, M = classes[0]
    one = opf.Context(N, K, M)
    one.setup(0, r["Re"], r["alpha"], dt)
    one.put_state(0, *r["state"][:3])
    one.run(r["nsteps"], corr_fixed=1)
    v = one.get_state(0)
    parity["alone_vs_in_mix_bitwise"] = bool(all(np.array_equal(u[k], v[k]) for k in range(3)))
    one.close()
    by_cls = {}
    for c, (N, K, M) in enumerate(classes):
        lc = lat[[k for k, rq in enumerate(reqs) if rq["cls"] == c]]
        if lc.size:
            by_cls[f"({N},{K},{M})"] = dict(requests=int(lc.size), p50_s=float(np.percentile(lc, 50)),
                                            p99_s=float(np.percentile(lc, 99)))
    return {"metric": "request latency p50/p99 (s) and served state-steps/s",
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "throughput": steps / span, "unit": "state-steps/s",
            "config": {"workload": f"{nreq} requests, Poisson arrivals at {rate}/s, sizes "
                                   f"{[list(c) for c in classes]} in ratio 5:3:2, Re~U[2000,5300], "
                                   f"alpha in {{0.5,1.25}}, T~U[5,50] at dt={dt}, one corrector per step",
                       "slots": caps, "chunk_steps": chunk, "hbm_budget_gb": budget_gb},
            "failed": failed, "by_class": by_cls, "parity": parity}


def load_traffic(kernel, workload):
    path = os.path.join(REPO, "profiles", f"traffic_{kernel}_{workload}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-states", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-rate", type=float, default=2.0)
    ap.add_argument("--serve-budget-gb", type=float, default=48.0)
    ap.add_argument("--serve-chunk", type=int, default=50)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=40)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes five times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters of the traffic figures) and exit")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        import torch
        a = torch.ones(args.calibrate_copy, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": args.calibrate_copy, "copies": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is a design, not a code path)"}))
        sys.exit(2)
    import opf

    name = args.workload
    wl = WORKLOADS[name]
    N, K, M, Re, alpha, dt, E0, seed, golden = wl
    ctx = opf.Context(N, K, M)
    ctx.setup(0, Re, alpha, dt)
    fields = make_ic(ctx, wl)
    ctx.put_state(0, *fields)

    # warmup, then the timed region (the library synchronises its stream on
    # both sides of opf_run)
    ctx.run(args.warmup, diag_every=10)
    t0 = time.perf_counter()
    ctx.run(args.steps, diag_every=10)
    t1 = time.perf_counter()
    secs = t1 - t0
    value = args.steps / secs
    diag = ctx.diag(0)

    # parity of the energy and friction velocity series with the reference
    parity = {"reference": golden, "checked_records": 0}
    gpath = os.path.join(REPO, "tests", "golden", golden) if golden else None
    ncorr_mean = 1.0
    if gpath and os.path.exists(gpath):
        with np.load(gpath) as g:
            rs = g["series"]
        rows = [r for r in diag if int(r[0]) < rs.shape[0]]
        if rows:
            d = np.array(rows)
            ref = rs[d[:, 0].astype(int)]
            dE = np.abs(d[:, 2] - ref[:, 2]) / ref[:, 2]
            dU = np.abs(d[:, 7] - ref[:, 7]) / ref[:, 7]
            parity.update(checked_records=len(rows), max_rel_dE=float(dE.max()),
                          max_rel_dutau=float(dU.max()),
                          ncorr_equal=bool(np.array_equal(d[:, 8].astype(int), ref[:, 8].astype(int))),
                          ok=bool(dE.max() <= 1e-10 and dU.max() <= 1e-10))
    if len(diag):
        ncorr_mean = float(np.mean(diag[:, 8][diag[:, 8] > 0])) if np.any(diag[:, 8] > 0) else 1.0

    # per-kernel durations with HIP events on the library's stream
    ctx.kernel_timing(True)
    ctx.run(args.timing_steps)
    kt = {}
    for k in ("zfwd", "theta", "zbwd", "step", "curl", "control"):
        ms, n = ctx.kernel_time(k)
        if n:
            kt[k] = dict(ms_total=ms, launches=n, us_avg=1000.0 * ms / n)
    ctx.kernel_timing(False)
    ab = alg_bytes(N, K, M)
    ab["step"] = (ab["step_pred"] + ncorr_mean * ab["step_corr"]) / (1.0 + ncorr_mean)
    dom = max(kt, key=lambda k: kt[k]["ms_total"])
    achieved = ab[dom] / (kt[dom]["us_avg"] * 1e-6) / 1e9
    ms_step = 1000.0 * secs / args.steps
    B_alg = alg_bytes_per_step(N, K, M, ncorr_mean)
    step_gbs = B_alg / (secs / args.steps) / 1e9

    batched = None
    ctx.close()
    if not args.no_batched:
        batched = batched_leg(args.batch_states, args.batch_steps)

    serving = None
    if not args.no_serving:
        serving = serving_leg(args.serve_requests, args.serve_rate, args.serve_budget_gb, args.serve_chunk)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(name, wl, fields, args.cpu_seconds)

    out = {
        "metric": "state-steps/s plain vs batched, p50/p99 request latency, at dE<=1e-10",
        "value": value,
        "unit": "state-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded divergence-free perturbation of Hagen-Poiseuille flow)",
        "config": {"workload": f"plain single-state time step at "
                               f"(N,K,M)=({N},{K},{M}), Re={Re}, alpha={alpha}, dt={dt}, E0={E0}",
                   "N": N, "K": K, "M": M, "Re": Re, "alpha": alpha, "dt": dt,
                   "corrector": "iterated to convergence (main.f90:35-41)",
                   "mean_ncorr": ncorr_mean},
        "parity": parity,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic(dom, name),
                     "alg_bytes_per_launch": ab[dom], "us_per_launch": kt[dom]["us_avg"]},
        "step_roofline": {"B_alg_bytes": B_alg, "achieved_GBs": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS},
        "kernels_us": {k: round(v["us_avg"], 3) for k, v in kt.items()},
        "batched": batched,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
