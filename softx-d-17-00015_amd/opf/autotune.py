"""Autotuner of the launch shapes of the time step.

For each resolution class and capacity it times the device-resident loop
(opf_run, corrector iterated to convergence) under candidate launch shapes
(opf_ctx_create_tiles): the fused step kernel's (|k|, m) pairs per workgroup
(1, 2, 4), the transforms' tile widths (m columns per z-transform
workgroup, z-lines per theta workgroup, each at most the kernels' default),
the number of loop iterations per captured HIP graph (0: eager launches)
and, from four slots up, one or two streams for a run's slots, by
coordinate descent in that order, and writes the
fastest shape per (N, K, M, capacity) to the table that opf_ctx_create reads
(tuning.tsv beside libopf.so).  Results do not depend on the shapes, so the
table only changes speed.

usage: python -m opf.autotune [--out PATH] [--only N,K,M,CAP ...]
"""
from __future__ import annotations

import argparse
import datetime
import os
import sys
import time

# (N, K, M, capacity, Re, alpha, dt, steps timed): the plain sizes, the
# serving classes with a few slots, and the batched ensemble
WORKLOADS = [
    (32, 24, 12, 1, 2400.0, 1.25, 0.01, 200),
    (48, 48, 24, 1, 3000.0, 1.25, 0.01, 100),
    (64, 64, 48, 1, 5300.0, 0.5, 0.005, 100),
    (96, 256, 128, 1, 5300.0, 0.4, 0.0025, 10),
    (32, 24, 12, 64, 2400.0, 1.25, 0.01, 40),
    (48, 48, 24, 32, 3000.0, 1.25, 0.01, 20),
    (64, 64, 48, 16, 5300.0, 0.5, 0.005, 10),
    (32, 24, 12, 512, 2400.0, 1.25, 0.01, 20),
]

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.tsv")


def _states(N, K, M, alpha, cap, mesh):
    from .ic import seeded_perturbation
    return [seeded_perturbation(mesh["r"], mesh["intrdr"], K, M, alpha, 100 + s, 10.0 ** (-5 + 3 * (s % 4) / 3))
            for s in range(cap)]


def time_shape(wl, shape, states):
    """Seconds per state-step of the loop under one launch shape (pairs, tm,
    tz, graph iterations, streams): best of 3 timed runs after a warm-up."""
    from .engine import Context
    N, K, M, cap, Re, alpha, dt, steps = wl
    ctx = Context(N, K, M, capacity=cap, pairs=shape[0], tm=shape[1], tz=shape[2])
    ctx.launch_mode(graph=shape[3], streams=shape[4])
    try:
        for s in range(cap):
            ctx.setup(s, Re, alpha, dt)
            ctx.put_state(s, *states[s])
        ctx.run(3)
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            ctx.run(steps)
            best = min(best, time.perf_counter() - t0)
        return best / (steps * cap), ctx.tile_config()
    finally:
        ctx.close()


def tune(wl, log=print):
    """Coordinate descent over (pairs, tm, tz, graph, streams) from the
    built-in shape; returns (shape, seconds per state-step, default's seconds
    per state-step)."""
    from .engine import Context, host_mesh
    N, K, M, cap, Re, alpha, dt, steps = wl
    os.environ["OPF_TUNING"] = "none"       # the built-in shape, not an older table's
    states = _states(N, K, M, alpha, cap, host_mesh(N))
    probe = Context(N, K, M, capacity=cap)
    base = probe.tile_config() + (0, 2 if cap >= 4 else 1)
    # a grid whose evaluation is the whole-plane kernel has no transform
    # tiles to choose
    probe.setup(0, Re, alpha, dt)
    probe.put_state(0, *states[0])
    probe.kernel_timing(True)
    probe.run(1, slots=[0])
    plane = probe.kernel_time("plane")[1] > 0
    probe.close()
    cache = {}

    def cost(shape):
        if shape not in cache:
            try:
                cache[shape] = time_shape(wl, shape, states)[0]
            except Exception as e:              # noqa: BLE001  (a shape the device cannot take)
                log(f"  {shape}: rejected ({e})")
                cache[shape] = float("inf")
            log(f"  (N,K,M,cap)=({N},{K},{M},{cap}) pairs={shape[0]} tm={shape[1]} tz={shape[2]} "
                f"graph={shape[3]} streams={shape[4]}: {cache[shape] * 1e6:.2f} us per state-step")
        return cache[shape]

    best = tuple(base)
    default_cost = cost(best)
    pair_opts = [p for p in (1, 2, 4) if p != 4 or N <= 96]
    for p in pair_opts:
        cand = (p,) + best[1:]
        if cost(cand) < cost(best):
            best = cand
    # theta tiles stay even (the N_z transforms pair z-lines 2q and 2q + 1)
    cands = {1: lambda d: {d, max(1, d - 1), max(1, (d + 1) // 2)},
             2: lambda d: {d} if d % 2 else {w for w in (d, d - 2, 2 * (d // 4)) if w >= 2},
             3: lambda d: {0, 4, 16},
             4: lambda d: {1, 2} if cap >= 4 else {1}}
    for i, dflt in ((1, base[1]), (2, base[2]), (3, 0), (4, 1)):
        if plane and i in (1, 2):
            continue
        for w in sorted(cands[i](dflt)):
            cand = list(best)
            cand[i] = w
            cand = tuple(cand)
            if cost(cand) < cost(best):
                best = cand
    return best, cost(best), default_cost


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--only", nargs="*", default=None, help="N,K,M,CAP of the workloads to tune")
    args = ap.parse_args(argv)
    wls = WORKLOADS
    if args.only:
        keep = {tuple(int(v) for v in s.split(",")) for s in args.only}
        wls = [w for w in wls if tuple(w[:4]) in keep]
    rows = []
    for wl in wls:
        print(f"tuning (N,K,M)=({wl[0]},{wl[1]},{wl[2]}) capacity {wl[3]}", flush=True)
        shape, c, c0 = tune(wl, log=lambda s: print(s, flush=True))
        rows.append((wl[:4], shape, c, c0))
    with open(args.out, "w") as f:
        f.write("# launch shapes of the time step chosen by opf/autotune.py (python -m opf.autotune)\n")
        f.write(f"# measured {datetime.date.today().isoformat()} on one MI355X (gfx950); opf_ctx_create takes the\n")
        f.write("# row of its (N, K, M) with the largest min_capacity not above its capacity\n")
        f.write("# N K M min_capacity pairs tm tz graph streams us_per_state_step default_us_per_state_step\n")
        for (N, K, M, cap), (p, tm, tz, g, ns), c, c0 in sorted(rows):
            f.write(f"{N} {K} {M} {cap} {p} {tm} {tz} {g} {ns} {c * 1e6:.3f} {c0 * 1e6:.3f}\n")
    print(f"wrote {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
