"""Benchmark: ns/day of the mixed path vs the plain FP64 path on one MI355X.

Workload (default): the ~100,000-atom solvated bead chain -- a 10,000-bead
chain plus 30,000 flexible 3-site waters in a 10.0 nm cubic box, PME cutoff
1.0 nm, grid 96^3, order 5, dt 0.5 fs, Verlet buffer 0.1 nm.  A "step" is one
full force evaluation (bonded, real-space pairs, exclusion correction, PME)
plus the leapfrog update, with the neighbour list rebuilt on the device
whenever an atom has moved more than half the buffer.

Prints ONE JSON line.  `value` is the mixed-precision ns/day over the timed
region; the FP64 path, the force errors of both paths against the FP64 CPU
oracle on identical positions, per-kernel times (HIP events recorded inside
the captured step graphs), the pair-kernel roofline and a CPU baseline (the
oracle, one core, a bounded sample of the same workload) ride along.

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu] [--serving]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ns/day mixed vs FP64 plain (same GPU, stated force err); batched ns/day + p99 step"
PEAK_FP32_VALU = 157.3   # TFLOP/s, MI355X FP32 vector (packed); 78.6 without packed FP32
PEAK_FP64_VALU = 78.6    # TFLOP/s
PEAK_HBM = 8000.0        # GB/s
PAIR_FLOP = 75           # FLOP per within-cutoff pair (FMA = 2, rsqrt/exp/rcp = 1)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def ns_per_day(steps, dt_ps, seconds):
    return 86400.0 * steps * dt_ps * 1e-3 / seconds


def rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b * b)))


def make_workload(name):
    from nexusmd_amd import systems as S
    if name == "solvated_chain_100000":
        return S.solvated_chain()
    if name == "water_24000":
        return S.water_box()
    if name == "bead_chain_1500":
        return S.bead_chain()
    if name == "water_1000002":
        return S.by_name("water_1000002")
    raise SystemExit("unknown workload " + name)


def kernel_bytes(sysm, prec_bytes, stats, n_excl):
    """Algorithmic bytes per launch of the HBM-bound kernels."""
    N = sysm.n
    Kx, Ky, Kz = stats["grid"]
    G = Kx * Ky * Kz
    Gc = Kx * Ky * (Kz // 2 + 1)
    r = prec_bytes
    pos = 4 * r                                   # x, y, z, q
    nterms = len(sysm.bond_idx) * 2 + len(sysm.angle_idx) * 3 + len(sysm.torsion_idx) * 4
    return {
        "pme_spread": N * pos + G * 8,
        "fft": 2 * (G * r + Gc * 2 * r),
        "pme_gather": N * pos + G * r + N * 24,
        "integrate": N * 128,
        "bonded": (len(sysm.bond_idx) * 24 + len(sysm.angle_idx) * 32 + len(sysm.torsion_idx) * 48) + nterms * 48,
        "exclusion": n_excl * (8 + 2 * 32 + 2 * 24),
    }


def time_path(sysm, prec, steps, warmup, skin):
    """Warm up, then the timed region: `steps` steps with HIP events around every
    kernel family (plain launches).  The same count is also run as graph replays
    without events, reported beside it."""
    from nexusmd_amd import engine as E
    ctx = E.Context(sysm, precision=prec, skin=skin)
    ctx.step(warmup, sysm.dt)
    t0 = time.perf_counter()
    ctx.step(steps, sysm.dt)                       # hipGraph replays, no events
    graph_s = time.perf_counter() - t0
    pairs0 = ctx.pair_count()
    ctx.enable_timing(True)
    t0 = time.perf_counter()
    ctx.step(steps, sysm.dt)                       # timed region (events per kernel)
    t1 = time.perf_counter()
    kern = {}
    for name in ("pair", "pme_spread", "fft", "pme_gather", "bonded", "exclusion", "integrate", "nlist", "pack"):
        ms, n = ctx.kernel_time(name)
        kern[name] = {"ms_per_launch": ms, "launches": n}
    stats = ctx.stats()
    ctx.enable_timing(False)
    pairs1 = ctx.pair_count()
    return ctx, (t1 - t0), kern, stats, 0.5 * (pairs0 + pairs1), graph_s


def cpu_baseline(sysm, budget_s=20.0):
    """FP64 oracle, one thread, on a bounded sample of the same workload."""
    from oracle import pyoracle as O
    prep = O.Prepared(sysm)
    t0 = time.perf_counter()
    O.simulate(prep, 1, sysm.dt)
    one = time.perf_counter() - t0
    nsteps = max(1, min(20, int(budget_s / max(one, 1e-3)) - 1))
    t0 = time.perf_counter()
    O.simulate(prep, nsteps, sysm.dt)
    wall = time.perf_counter() - t0
    return {
        "value": ns_per_day(nsteps, sysm.dt, wall),
        "unit": "ns/day",
        "cores": 1,
        "kind": "port",
        "sample": "%d leapfrog steps of the same %d-atom system (FP64, neighbour list rebuilt every step, "
                  "full PME every step), %.1f s wall after one warm-up step" % (nsteps, sysm.n, wall),
        "s_per_step": wall / nsteps,
        "host": platform.processor() or platform.machine(),
    }


def serving_bench(n_requests, steps_per_request, seed=20261205):
    """Continuous batching of ligand-in-water requests vs one at a time."""
    from nexusmd_amd import _lib as L
    from nexusmd_amd import engine as E
    from nexusmd_amd import systems as S
    sizes, seeds = S.serving_batch(n_requests, seed=seed)
    reqs = [S.ligand_in_water(n, sd, steps=steps_per_request) for n, sd in zip(sizes, seeds)]
    dt = reqs[0].dt
    # one at a time (each system alone on the GPU)
    t0 = time.perf_counter()
    sample = reqs[: min(8, len(reqs))]
    for r in sample:
        ctx = E.Context(r, precision=L.MIXED, skin=0.1)
        ctx.step(steps_per_request, dt)
        ctx.close()
    alone_s = time.perf_counter() - t0
    alone_ns_day = ns_per_day(steps_per_request * len(sample), dt, alone_s)
    # batched, Poisson arrivals at 0.8x the measured one-at-a-time service rate
    rate = 0.8 * len(sample) / alone_s * 4.0
    arrivals = S.poisson_arrivals(len(reqs), rate, seed + 1)
    srv = E.Server(precision=L.MIXED, budget_bytes=200 << 30, max_active=256)
    t0 = time.perf_counter()
    nxt = 0
    ids = []
    while True:
        now = time.perf_counter() - t0
        while nxt < len(reqs) and arrivals[nxt] <= now:
            ids.append(srv.submit(reqs[nxt], steps_per_request, dt))
            nxt += 1
        st = srv.stats()
        if nxt >= len(reqs) and st["finished"] + st["failed"] >= len(reqs):
            break
        if st["active"] == 0 and st["submitted"] == st["finished"] + st["failed"]:
            if nxt < len(reqs):
                time.sleep(max(0.0, min(0.01, arrivals[nxt] - now)))
            continue
        srv.run(50)
    wall = time.perf_counter() - t0
    st = srv.stats()
    total_ns = steps_per_request * dt * 1e-3 * st["finished"]
    return {
        "requests": len(reqs),
        "atoms_min_median_max": [int(min(sizes)), int(np.median(sizes)), int(max(sizes))],
        "steps_per_request": steps_per_request,
        "aggregate_ns_per_day": 86400.0 * total_ns / wall,
        "device_aggregate_ns_per_day": st["aggregate_ns_per_day"],
        "one_at_a_time_ns_per_day": alone_ns_day,
        "batched_over_alone": (86400.0 * total_ns / wall) / alone_ns_day,
        "p50_step_ms": st["p50_step_ms"],
        "p99_step_ms": st["p99_step_ms"],
        "failed": st["failed"],
        "wall_s": wall,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="solvated_chain_100000")
    ap.add_argument("--skin", type=float, default=0.1)
    ap.add_argument("--relax", type=int, default=300, help="steepest-descent iterations before timing")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--serving", action="store_true", help="also run the batched-serving measurement")
    ap.add_argument("--serving-requests", type=int, default=256)
    ap.add_argument("--serving-steps", type=int, default=500)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every run here uses a single "
              "MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from nexusmd_amd import _lib as L
    from nexusmd_amd import engine as E

    t = time.perf_counter()
    sysm = make_workload(args.workload)
    log("workload %s: %d atoms, box %.3f nm, grid %s, generated in %.1f s" % (
        args.workload, sysm.n, sysm.L[0], sysm.grid, time.perf_counter() - t))
    from nexusmd_amd.prepare import relax
    t = time.perf_counter()
    sysm = relax(sysm, iterations=args.relax, log=log)
    log("relaxation %.1f s" % (time.perf_counter() - t))

    # force error of both paths against the FP64 oracle on identical positions
    from oracle import pyoracle as O
    t = time.perf_counter()
    prep = O.Prepared(sysm)
    f_ref, e_ref = O.compute_forces(prep)
    log("oracle forces in %.1f s" % (time.perf_counter() - t))
    errs = {}
    for name, prec in (("mixed", L.MIXED), ("fp64", L.FP64)):
        c = E.Context(sysm, precision=prec, skin=args.skin)
        f, e = c.compute_forces()
        errs[name] = {"force_rel_rms": rel_rms(f, f_ref),
                      "energy_rel": abs(e["potential"] - e_ref["potential"]) / abs(e_ref["potential"])}
        c.close()
    log("force errors", errs)

    ctx_m, sec_m, kern_m, stats_m, pairs, graph_m = time_path(sysm, L.MIXED, args.steps, args.warmup, args.skin)
    ctx_m.close()
    nsd_m = ns_per_day(args.steps, sysm.dt, sec_m)
    log("mixed: %.3f ms/step, %.2f ns/day" % (1e3 * sec_m / args.steps, nsd_m))
    fp64 = None
    if not args.no_fp64:
        steps64 = max(10, args.steps // 4)
        ctx_d, sec_d, kern_d, stats_d, _, graph_d = time_path(sysm, L.FP64, steps64, max(5, args.warmup // 4), args.skin)
        ctx_d.close()
        nsd_d = ns_per_day(steps64, sysm.dt, sec_d)
        fp64 = {"ns_per_day": nsd_d, "ms_per_step": 1e3 * sec_d / steps64, "steps": steps64,
                "graph_replay_ns_per_day": ns_per_day(steps64, sysm.dt, graph_d), "kernels": kern_d}
        log("fp64: %.3f ms/step, %.2f ns/day" % (1e3 * sec_d / steps64, nsd_d))

    # pair-kernel roofline (FP32 VALU; mixed path)
    pair_ms = kern_m["pair"]["ms_per_launch"]
    flops = PAIR_FLOP * pairs
    achieved = flops / (pair_ms * 1e-3) / 1e12 if pair_ms > 0 else 0.0
    tiles = stats_m["tiles"]
    roofline = {
        "kernel": "pair (real-space LJ + erfc Coulomb, mixed)",
        "bound": "valu",
        "achieved": achieved,
        "peak": PEAK_FP32_VALU,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_VALU,
        "traffic": None,
        "algorithmic_flop_per_launch": flops,
        "pairs_within_cutoff": pairs,
        "pairs_evaluated": tiles * 64,
        "ms_per_launch": pair_ms,
    }
    ex = O.exclusions(sysm)
    n_excl = int(len(ex[1]) // 2)
    kb = kernel_bytes(sysm, 4, stats_m, n_excl)
    kernels = {}
    for k, v in kern_m.items():
        d = dict(v)
        if k in kb and v["ms_per_launch"] > 0:
            d["bytes"] = kb[k]
            d["GB_per_s"] = kb[k] / (v["ms_per_launch"] * 1e-3) / 1e9
            d["hbm_frac"] = d["GB_per_s"] / PEAK_HBM
        kernels[k] = d

    out = {
        "metric": METRIC,
        "value": nsd_m,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec_m / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (forces in 64-bit fixed point, fp64 state and energy sums)",
        "data": "synthetic (seeded generator; random-walk chain + jittered-lattice water)",
        "config": {"workload": args.workload, "atoms": sysm.n, "box_nm": float(sysm.L[0]),
                   "pme_grid": list(sysm.grid), "pme_order": 5, "cutoff_nm": sysm.cutoff,
                   "ewald_tol": sysm.ewald_tol, "dt_fs": sysm.dt * 1e3, "skin_nm": args.skin,
                   "precision": "mixed"},
        "fp64": fp64,
        "graph_replay_ns_per_day": ns_per_day(args.steps, sysm.dt, graph_m),
        "mixed_over_fp64": (nsd_m / fp64["ns_per_day"]) if fp64 else None,
        "force_error_vs_fp64_oracle": errs,
        "roofline": roofline,
        "kernels_mixed": kernels,
        "pair_list": {"entries": stats_m["list_entries"], "tiles": tiles, "rebuilds": stats_m["rebuilds"]},
    }
    if not args.no_cpu:
        cb = cpu_baseline(sysm)
        out["cpu_baseline"] = cb
        log("cpu baseline %.4f ns/day" % cb["value"])
    if args.serving:
        out["serving"] = serving_bench(args.serving_requests, args.serving_steps)
        log("serving", out["serving"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
