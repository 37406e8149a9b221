"""Benchmark: ns/day of the mixed path vs the plain FP64 path on one MI355X.

Workload (default): the ~100,000-atom solvated bead chain -- a 10,000-bead
chain plus 30,000 flexible 3-site waters in a 10.0 nm cubic box, PME cutoff
1.0 nm, grid 96^3, order 5, dt 0.5 fs, Verlet buffer 0.1 nm.  A "step" is one
full force evaluation (bonded, real-space pairs, exclusion correction, PME)
plus the leapfrog update, with the neighbour list rebuilt on the device
whenever an atom has moved more than half the buffer.

Prints ONE JSON line.  `value` is the mixed-precision ns/day over exactly K
steps (--steps) replayed as step graphs, with nothing else on the stream and
no events inside; the FP64 path is timed over the same K steps.  Beside it
ride, all measured outside those K steps:
  - the force errors of both paths against the FP64 CPU oracle on identical
    positions;
  - the pair kernel timed over 64 launches as it runs in a step (PME beside
    it) and over 32 launches alone, and the roofline figure from the former;
  - a per-kernel breakdown (every kernel family alone, events around each);
  - a steady-state comparison of both paths over 2,000 steps;
  - the total-energy drift of both paths over 5,000 steps, per ns;
  - a CPU baseline: the reference's own functions, compiled from its sources
    (numpy's FFT standing where FFTW would), on one pinned core, 10 steps of
    the same system;
  - the batched-serving measurement (256 ligand-in-water requests,
    continuous batching vs one at a time, per-step latency from device
    events).

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--no-cpu] [--no-fp64] [--no-serving]
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
PAIR_SAMPLE_STEPS = 64   # steps whose pair kernel is bracketed by events (beside PME)
PAIR_ALONE_STEPS = 32    # steps of per-kernel events (every kernel family alone)
STEADY_STEPS = 2000
DRIFT_EQUIL, DRIFT_STEPS, DRIFT_EVERY = 4000, 5000, 250
TUNE_STEPS = 50


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


def workload_config(name, sysm, skin):
    cfg = {"workload": name, "atoms": sysm.n, "box_nm": float(sysm.L[0]), "pme_grid": list(sysm.grid),
           "pme_order": 5, "cutoff_nm": sysm.cutoff, "ewald_tol": sysm.ewald_tol, "dt_fs": sysm.dt * 1e3,
           "skin_nm": skin, "precision": "mixed"}
    if name == "solvated_chain_100000":
        # generator choices the configuration leaves open, stated with the numbers
        cfg.update({"chain": "free self-avoiding walk from the box centre (no globule)",
                    "h_core": "water H sig 0.04 nm, eps 0.192 kJ/mol beside charged beads",
                    "relaxed": "300 steps of GPU steepest descent, then Maxwell-Boltzmann 300 K"})
    return cfg


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
        "bonded": (len(sysm.bond_idx) * 24 + len(sysm.angle_idx) * 32 + len(sysm.torsion_idx) * 48)
                  + nterms * (24 + 3 * prec_bytes),
        "atom_forces": N * (8 + 24) + 2 * n_excl * (4 + 24 + 8) + nterms * (4 + 3 * prec_bytes) + N * 24,
    }


def time_path(sysm, prec, steps, warmup, skin):
    """Warm-up, then the timed region: exactly `steps` steps as step-graph
    replays (nx_step returns after its stream has drained, so the wall clock
    around it covers exactly those steps).  Then, outside the timed region,
    PAIR_SAMPLE_STEPS steps with events around the pair kernel as it runs
    beside PME, and PAIR_ALONE_STEPS steps with events around every kernel
    family launched alone."""
    from nexusmd_amd import _lib as L
    from nexusmd_amd import engine as E
    graphs = os.environ.get("NEXUSMD_GRAPHS", "1") != "0"
    ctx = E.Context(sysm, precision=prec, skin=skin)
    ctx.step(warmup, sysm.dt)
    t0 = time.perf_counter()
    ctx.step(steps, sysm.dt)                       # timed region
    wall = time.perf_counter() - t0
    stats = ctx.stats()
    # the pair kernel as it runs in a step: graphs around it, PME beside it on the
    # side stream, events on the main stream (direct launches with NEXUSMD_GRAPHS=0)
    pairs0 = ctx.pair_count()
    ctx.enable_timing(L.TIMING_GRAPH_PAIR)
    ctx.step(PAIR_SAMPLE_STEPS, sysm.dt)
    beside_ms, beside_n = ctx.kernel_time("pair")
    step_ev_ms, step_ev_n = ctx.kernel_time("step")
    segs = {}
    for name in ("side", "list", "update"):
        try:
            segs[name] = ctx.kernel_time(name)[0]
        except Exception:          # not timed (direct launches, or no second stream)
            segs[name] = None
    pairs1 = ctx.pair_count()
    # every kernel family alone, one stream, events around each
    ctx.enable_timing(L.TIMING_KERNELS)
    ctx.step(PAIR_ALONE_STEPS, sysm.dt)
    kern = {}
    for name in ("pair", "pme_spread", "fft", "pme_gather", "bonded", "atom_forces", "integrate", "nlist", "pack"):
        ms, n = ctx.kernel_time(name)
        kern[name] = {"ms_per_launch": ms, "launches": n}
    ctx.enable_timing(L.TIMING_OFF)
    pair = {"beside_pme_ms_per_launch": beside_ms, "beside_pme_launches": beside_n,
            "alone_ms_per_launch": kern["pair"]["ms_per_launch"], "alone_launches": kern["pair"]["launches"],
            "sampled_step_ms": step_ev_ms if step_ev_n else None,
            # the same sampled steps: the terms and PME on the second stream (the
            # chain beside the pair kernel), the list graph before the fork and
            # the update after the join
            "side_chain_ms_per_step": segs["side"], "list_ms_per_step": segs["list"],
            "update_ms_per_step": segs["update"],
            "launch": "step graphs" if graphs else "direct launches (NEXUSMD_GRAPHS=0)"}
    return ctx, wall, kern, stats, 0.5 * (pairs0 + pairs1), pair


def chains(pair):
    """The two chains of a sampled step, in microseconds: the pair kernel on the
    main stream and the terms and PME on the second stream beside it."""
    us = lambda v: None if v is None else round(1e3 * v, 1)
    return {"pair_beside_pme_us": us(pair["beside_pme_ms_per_launch"]),
            "side_chain_us": us(pair["side_chain_ms_per_step"]), "list_us": us(pair["list_ms_per_step"]),
            "update_us": us(pair["update_ms_per_step"])}


def autotune_leg(sysm, prec, skin, steps_per_trial=TUNE_STEPS):
    """nx_autotune on the system: the chosen launch shapes and the measured
    ms/step of the default and of the choice."""
    from nexusmd_amd import engine as E
    ctx = E.Context(sysm, precision=prec, skin=skin)
    t0 = time.perf_counter()
    chosen, ms_default, ms_best = ctx.autotune(steps_per_trial)
    sec = time.perf_counter() - t0
    ctx.close()
    return {"atoms": sysm.n, "chosen": chosen, "ms_per_step_default": ms_default, "ms_per_step_chosen": ms_best,
            "steps_per_trial": steps_per_trial, "tuning_s": sec}


def ingest_bench(sysm):
    """Host ingest as TaskDispatcher::Simulate runs it, stage by stage, on the
    workload written as system/state XML (one report step, no timing of the
    step itself): the system file in one pass, the state file, the topology
    with its exclusions, box and state, and the device set-up."""
    import tempfile
    from nexusmd_amd import build as Bd
    from nexusmd_amd import xmlio
    sys.path.insert(0, os.path.dirname(Bd.module_path()))
    import NexusMD
    with tempfile.TemporaryDirectory() as d:
        t = time.perf_counter()
        xmlio.write_system_xml(sysm, os.path.join(d, "system.xml"))
        xmlio.write_state_xml(sysm, os.path.join(d, "state.xml"))
        xmlio.write_pdb(sysm, os.path.join(d, "input.pdb"))
        written = time.perf_counter() - t
        sizes = {f: os.path.getsize(os.path.join(d, f)) for f in ("system.xml", "state.xml")}
        disp = NexusMD.TaskDispatcher.CreateDispatcher("GPU_MIXED")
        disp.echo_steps = False
        for en in ("EnableHarmonicBondForce", "EnableHarmonicAngleForce", "EnablePeriodicTorsionForce",
                   "EnableNonbondedForce"):
            getattr(disp, en)()
        cwd = os.getcwd()
        os.chdir(d)
        # the dispatcher prints the reference's console lines; stdout carries
        # only the bench's JSON line, so they go to stderr
        sys.stdout.flush()
        saved = os.dup(1)
        os.dup2(2, 1)
        try:
            disp.Simulate("system.xml", "state.xml", "input.pdb", "out.pdb", sysm.dt, 1, 1)
        finally:
            sys.stdout.flush()
            os.dup2(saved, 1)
            os.close(saved)
            os.chdir(cwd)
        times = dict(disp.IngestTimes())
    times["total_s"] = sum(v for k, v in times.items() if k.endswith("_s"))
    times["xml_bytes"] = sizes
    times["xml_written_in_s"] = written
    return times


def serving_median_system():
    """A ligand-in-water request of the serving mix's median size (2,500 atoms)."""
    from nexusmd_amd import systems as S
    return S.ligand_in_water(2500, 7)


def steady_state(sysm, prec, skin, steps=STEADY_STEPS):
    """ns/day of `steps` graph-replayed steps after 100 warm-up steps."""
    from nexusmd_amd import engine as E
    ctx = E.Context(sysm, precision=prec, skin=skin)
    ctx.step(100, sysm.dt)
    t0 = time.perf_counter()
    ctx.step(steps, sysm.dt)
    wall = time.perf_counter() - t0
    rebuilds = ctx.stats()["rebuilds"]
    ctx.close()
    return {"steps": steps, "ms_per_step": 1e3 * wall / steps, "ns_per_day": ns_per_day(steps, sysm.dt, wall),
            "rebuilds": rebuilds}


def energy_drift(sysm, prec, skin, steps=DRIFT_STEPS, every=DRIFT_EVERY, equil=DRIFT_EQUIL):
    """Total energy (kinetic with the reference's half-step shift + physical
    potential) every `every` steps over `steps` steps, after `equil` steps in
    which the freshly thermalised system shares its energy between kinetic and
    potential (the total is not flat there: the first picoseconds move a large
    share of it between the stiff modes, where the step's integration error is
    largest); drift = least-squares slope in kJ/mol per ns, also per atom."""
    from nexusmd_amd import engine as E
    ctx = E.Context(sysm, precision=prec, skin=skin)
    if equil:
        ctx.step(equil, sysm.dt)
    t, etot = [], []
    done = 0
    while done <= steps:
        _, _, _, e = ctx.step_report(sysm.dt)
        t.append(done * sysm.dt * 1e-3)                      # ns
        etot.append(e["kinetic"] + e["potential"])
        done += 1
        if done > steps:
            break
        ctx.step(every - 1, sysm.dt)
        done += every - 1
    ctx.close()
    t, etot = np.asarray(t), np.asarray(etot)
    slope = float(np.polyfit(t, etot, 1)[0])
    return {"equilibration_steps": equil, "steps": steps, "samples": len(t), "kj_mol_per_ns": slope,
            "kj_mol_per_ns_per_atom": slope / sysm.n, "std_kj_mol": float(np.std(etot)), "e0_kj_mol": float(etot[0])}


def host_info():
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_allowed": affinity}


def cpu_baseline(sysm, nsteps=10):
    """The reference's own functions (compiled from its sources into
    oracle/_ref/libnxref.so; numpy's unnormalised FFT where FFTW would run),
    composed in Engine::RunSimulation's order (neighbour list rebuilt and full
    PME every step), on one core pinned with sched_setaffinity, for `nsteps`
    steps of the same system after one warm-up step.  Without the compiled
    reference, the C oracle's restatement in its place (kind "port")."""
    from oracle import pyoracle as O
    info = host_info()
    cpus = sorted(os.sched_getaffinity(0))
    pin = cpus[0]
    os.sched_setaffinity(0, {pin})
    try:
        if O.ref_available():
            from oracle import pyref as R
            kind = "reference"
            R.simulate(sysm, 1, sysm.dt)
            t0 = time.perf_counter()
            R.simulate(sysm, nsteps, sysm.dt)
            wall = time.perf_counter() - t0
            what = ("the reference's functions compiled from its sources (oracle/_ref), numpy pocketfft in "
                    "FFTW's place")
        else:
            kind = "port"
            prep = O.Prepared(sysm)
            O.simulate(prep, 1, sysm.dt)
            t0 = time.perf_counter()
            O.simulate(prep, nsteps, sysm.dt)
            wall = time.perf_counter() - t0
            what = "the C oracle's restatement of the reference (oracle/libnxo.so)"
    finally:
        os.sched_setaffinity(0, set(cpus))
    return {
        "value": ns_per_day(nsteps, sysm.dt, wall),
        "unit": "ns/day",
        "cores": 1,
        "kind": kind,
        "sample": "%d leapfrog steps of the same %d-atom system after one warm-up step (FP64, neighbour list "
                  "rebuilt and full PME every step, grid %s), %.1f s wall; %s; pinned to CPU %d"
                  % (nsteps, sysm.n, "x".join(map(str, sysm.grid)), wall, what, pin),
        "s_per_step": wall / nsteps,
        "host": info,
    }


def _make_request(args):
    from nexusmd_amd import systems as S
    n, sd, steps = args
    return S.ligand_in_water(n, sd, steps=steps)


def serving_bench(n_requests, steps_per_request, alone_steps, alone_full_sample=4, seed=20261205, workers=16):
    """Continuous batching of ligand-in-water requests (atom counts lognormal,
    median 2,500, clipped to [600, 8000]) against the same requests run one at
    a time, all in the mixed path.

    Batched: requests arrive as a seeded Poisson process at 0.8x the service
    rate measured in a short saturated calibration; the server admits them by
    HBM footprint as they arrive and advances every active system together.
    `aggregate_ns_per_day` = sum of simulated time / wall-clock makespan.
    Per-step latency: for every request, the interval between the device
    events that close its successive steps (the batch's step events), with
    the batch re-layout and admission that precede a step inside the interval
    of that step.
    One at a time: every request runs alone (its own context, graph replays)
    for `alone_steps` steps after warm-up; its per-step time times the
    requested step count is its service time, and the sum of those is the
    sequential makespan.  A sample of requests is also run alone for the
    full step count, to check the extrapolation."""
    import multiprocessing as mp

    from nexusmd_amd import _lib as L
    from nexusmd_amd import engine as E
    from nexusmd_amd import systems as S
    sizes, seeds = S.serving_batch(n_requests, seed=seed)
    t = time.perf_counter()
    jobs = [(n, sd, steps_per_request) for n, sd in zip(sizes, seeds)]
    with mp.get_context("fork").Pool(max(1, min(workers, os.cpu_count() or 1))) as pool:
        reqs = pool.map(_make_request, jobs, chunksize=4)
    log("serving: %d requests generated in %.1f s (%d atoms total)" % (len(reqs), time.perf_counter() - t,
                                                                      sum(r.n for r in reqs)))
    dt = reqs[0].dt
    # one at a time
    alone_step_s = []
    for r in reqs:
        ctx = E.Context(r, precision=L.MIXED, skin=0.1)
        ctx.step(20, dt)
        t0 = time.perf_counter()
        ctx.step(alone_steps, dt)
        alone_step_s.append((time.perf_counter() - t0) / alone_steps)
        ctx.close()
    alone_makespan = steps_per_request * float(np.sum(alone_step_s))
    sim_ns = len(reqs) * steps_per_request * dt * 1e-3
    alone_ns_day = 86400.0 * sim_ns / alone_makespan
    log("serving: one at a time %.1f ns/day aggregate (median step %.3f ms)" % (
        alone_ns_day, 1e3 * float(np.median(alone_step_s))))
    # full-length alone runs of a sample (smallest, median, largest, one more)
    order = np.argsort(sizes)
    pick = sorted(set(int(order[i]) for i in (0, len(order) // 2, len(order) - 1, len(order) // 4)))[:alone_full_sample]
    full = []
    for i in pick:
        ctx = E.Context(reqs[i], precision=L.MIXED, skin=0.1)
        ctx.step(20, dt)
        t0 = time.perf_counter()
        ctx.step(steps_per_request, dt)
        s_full = (time.perf_counter() - t0) / steps_per_request
        ctx.close()
        full.append({"atoms": reqs[i].n, "ms_per_step_full": 1e3 * s_full,
                     "ms_per_step_extrapolated": 1e3 * alone_step_s[i]})
    # service-rate calibration: the first 64 requests together, saturated
    cal = E.Server(precision=L.MIXED, budget_bytes=200 << 30, max_active=n_requests)
    cal_reqs = reqs[: min(64, len(reqs))]
    for r in cal_reqs:
        cal.submit(r, 200, dt)
    cal.run(20)
    t0 = time.perf_counter()
    cal.run(100)
    cal_s = time.perf_counter() - t0
    cal.close()
    atom_steps_per_s = 100 * sum(r.n for r in cal_reqs) / cal_s
    service_rate = atom_steps_per_s / (float(np.mean([r.n for r in reqs])) * steps_per_request)
    rate = 0.8 * service_rate
    arrivals = S.poisson_arrivals(len(reqs), rate, seed + 1)
    # batched
    srv = E.Server(precision=L.MIXED, budget_bytes=200 << 30, max_active=n_requests)
    t0 = time.perf_counter()
    nxt = 0
    max_active = 0
    while True:
        now = time.perf_counter() - t0
        while nxt < len(reqs) and arrivals[nxt] <= now:
            srv.submit(reqs[nxt], steps_per_request, dt)
            nxt += 1
        st = srv.stats()
        max_active = max(max_active, st["active"])
        if nxt >= len(reqs) and st["finished"] + st["failed"] >= len(reqs):
            break
        if st["active"] == 0 and st["submitted"] == st["finished"] + st["failed"]:
            if nxt < len(reqs):
                time.sleep(max(0.0, min(0.01, arrivals[nxt] - now)))
            continue
        srv.run(50)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    done_ns = steps_per_request * dt * 1e-3 * st["finished"]
    agg = 86400.0 * done_ns / wall
    return {
        "requests": len(reqs),
        "atoms_min_median_max": [int(min(sizes)), int(np.median(sizes)), int(max(sizes))],
        "atoms_total": int(sum(r.n for r in reqs)),
        "steps_per_request": steps_per_request,
        "dt_fs": dt * 1e3,
        "arrival_rate_per_s": rate,
        "max_active": int(max_active),
        "aggregate_ns_per_day": agg,
        "one_at_a_time_ns_per_day": alone_ns_day,
        "batched_over_one_at_a_time": agg / alone_ns_day,
        "p50_step_ms": st["p50_step_ms"],
        "p99_step_ms": st["p99_step_ms"],
        "latency_sampling": "per (request, step): interval between the device events closing the request's "
                            "successive batched steps; a step that follows a batch re-layout (admission or "
                            "retirement) includes the re-layout",
        "finished": st["finished"],
        "failed": st["failed"],
        "makespan_s": wall,
        "server_step_s": st["wall_s"],
        "one_at_a_time_makespan_s": alone_makespan,
        "one_at_a_time_sample": "every request alone for %d steps after 20 warm-up steps; its per-step time "
                                "x %d steps is its service time" % (alone_steps, steps_per_request),
        "one_at_a_time_full_runs": full,
    }


def reports_bench(steps_c0=400, steps_c1=60, workdir=None):
    """TaskDispatcher::Simulate with a report on every step (interval 1, the
    reference's own script's setting) through the NexusMD module, on the
    1,500-bead chain and the 24,000-atom water box: ns/day with the report
    steps enqueued without waiting and files written by a writer thread
    (default), and with the synchronous report path."""
    import tempfile

    from nexusmd_amd import build as B
    from nexusmd_amd import systems as S
    from nexusmd_amd import xmlio
    sys.path.insert(0, os.path.dirname(B.module_path()))
    import NexusMD
    out = {}
    for name, mk, steps, dt in (("bead_chain_1500", S.bead_chain, steps_c0, 0.001),
                                ("water_24000", S.water_box, steps_c1, 0.0005)):
        s = mk()
        d = tempfile.mkdtemp(prefix="nx_reports_", dir=workdir)
        xmlio.write_system_xml(s, os.path.join(d, "system.xml"))
        xmlio.write_state_xml(s, os.path.join(d, "state.xml"))
        xmlio.write_pdb(s, os.path.join(d, "input.pdb"))
        row = {"atoms": s.n, "steps": steps, "interval": 1, "dt_fs": dt * 1e3}
        for mode in (True, False):
            disp = NexusMD.TaskDispatcher.CreateDispatcher("GPU_MIXED")
            disp.echo_steps = False
            disp.async_reports = mode
            for en in (disp.EnableHarmonicBondForce, disp.EnableHarmonicAngleForce,
                       disp.EnablePeriodicTorsionForce, disp.EnableNonbondedForce):
                en()
            cwd = os.getcwd()
            os.chdir(d)
            try:
                disp.Simulate("system.xml", "state.xml", "input.pdb", "out.pdb", dt, 5, 1)     # warm-up
                t0 = time.perf_counter()
                disp.Simulate("system.xml", "state.xml", "input.pdb", "out.pdb", dt, steps, 1)
                wall = time.perf_counter() - t0
            finally:
                os.chdir(cwd)
            key = "async" if mode else "sync"
            row[key + "_ns_per_day"] = ns_per_day(steps, dt, wall)
            row[key + "_ms_per_step"] = 1e3 * wall / steps
        row["report_bytes_per_step"] = sum(os.path.getsize(os.path.join(d, f))
                                           for f in ("out.pdb", "PVFReportout.pdb", "TotalEnergyout.pdb")) // steps
        import shutil
        shutil.rmtree(d, ignore_errors=True)
        log("reports %s: %s" % (name, row))
        out[name] = row
    return out


PROFILE = os.path.join(ROOT, "profiles", "kernels_solvated_chain_100000.md")


def profiled_traffic(kernel="pair_kernel_mixed<false>"):
    """HBM bytes per launch of `kernel` as recorded by the two rocprofv3 --pmc
    passes (FETCH_SIZE, WRITE_SIZE, scaled by a 256 MiB copy) of this bench
    command, from the committed summary; None when it is absent."""
    try:
        for line in open(PROFILE):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 4 and cells[0].endswith(kernel):
                return float(cells[3]) * 1e6
    except (OSError, ValueError):
        pass
    return None


VALU_PROFILE = os.path.join(ROOT, "profiles", "valu_pair_mixed_solvated_chain_100000.md")


def profiled_valu_busy(kernel="pair_kernel_mixed<false>"):
    """Fraction of the kernel's cycles with a VALU instruction issuing
    (rocprofiler's VALUBusy from SQ_ACTIVE_INST_VALU and GRBM_GUI_ACTIVE), from
    the committed counter summary of this bench command; None when absent."""
    try:
        for line in open(VALU_PROFILE):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 3 and cells[0] == kernel and cells[1] == "VALU busy":
                return float(cells[2])
    except (OSError, ValueError):
        pass
    return None


def copy_calibration(mib=256, reps=3):
    """A plain copy kernel of known size: 256 MiB read and 256 MiB written per
    launch, to scale the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copy calibration: %d launches of a %d MiB device copy" % (reps, mib))
    del a, b
    torch.cuda.empty_cache()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="solvated_chain_100000")
    ap.add_argument("--skin", type=float, default=0.1)
    ap.add_argument("--relax", type=int, default=300, help="steepest-descent iterations before timing")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-drift", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner leg (default launch shapes)")
    ap.add_argument("--no-ingest", action="store_true", help="skip the host-ingest leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the batched-serving measurement")
    ap.add_argument("--serving-only", action="store_true", help="run only the batched-serving measurement")
    ap.add_argument("--reports-only", action="store_true",
                    help="run only Simulate with a report on every step (c0 and c1 sizes), async and sync reports")
    ap.add_argument("--serving-requests", type=int, default=256)
    ap.add_argument("--serving-steps", type=int, default=5000)
    ap.add_argument("--alone-steps", type=int, default=200)
    ap.add_argument("--no-graphs", action="store_true",
                    help="direct kernel launches instead of step graphs (sets NEXUSMD_GRAPHS=0)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 256 MiB device-to-device copy kernel (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every run here uses a single "
              "MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from nexusmd_amd import _lib as L
    from nexusmd_amd import engine as E

    if args.no_graphs:
        os.environ["NEXUSMD_GRAPHS"] = "0"
    if args.calibrate_copy:
        copy_calibration()
    if args.serving_only:
        sv = serving_bench(args.serving_requests, args.serving_steps, args.alone_steps)
        print(json.dumps({"serving": sv}), flush=True)
        return
    if args.reports_only:
        print(json.dumps({"reports": reports_bench()}), flush=True)
        return

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

    ingest = None
    if not args.no_ingest:
        try:
            ingest = ingest_bench(sysm)
            log("ingest", ingest)
        except Exception as exc:          # reported, not fatal to the headline
            ingest = {"error": str(exc)}

    # launch shapes chosen by the autotuner (timed trials, outside every timed
    # region); the engine keeps the choice per system shape for the process, so
    # the timed paths below run with it
    tuned = {}
    if not args.no_tune:
        for name, prec in (("mixed", L.MIXED), ("fp64", L.FP64)):
            if name == "fp64" and args.no_fp64:
                continue
            tuned[name] = autotune_leg(sysm, prec, args.skin)
        try:
            tuned["serving_median_request"] = autotune_leg(serving_median_system(), L.MIXED, args.skin)
        except Exception as exc:          # reported, not fatal to the headline
            tuned["serving_median_request"] = {"error": str(exc)}
        log("autotune", tuned)

    ctx_m, sec_m, kern_m, stats_m, pairs, pair_m = time_path(sysm, L.MIXED, args.steps, args.warmup, args.skin)
    ctx_m.close()
    nsd_m = ns_per_day(args.steps, sysm.dt, sec_m)
    log("mixed: %.3f ms/step, %.2f ns/day" % (1e3 * sec_m / args.steps, nsd_m), chains(pair_m))
    fp64 = None
    if not args.no_fp64:
        ctx_d, sec_d, kern_d, stats_d, _, pair_d = time_path(sysm, L.FP64, args.steps, args.warmup, args.skin)
        ctx_d.close()
        nsd_d = ns_per_day(args.steps, sysm.dt, sec_d)
        fp64 = {"ns_per_day": nsd_d, "ms_per_step": 1e3 * sec_d / args.steps, "steps": args.steps,
                "pair": pair_d, "kernels": kern_d}
        log("fp64: %.3f ms/step, %.2f ns/day" % (1e3 * sec_d / args.steps, nsd_d), chains(pair_d))
    steady = {"mixed": steady_state(sysm, L.MIXED, args.skin)}
    if fp64:
        steady["fp64"] = steady_state(sysm, L.FP64, args.skin)
        steady["mixed_over_fp64"] = steady["mixed"]["ns_per_day"] / steady["fp64"]["ns_per_day"]
    log("steady state", steady)
    drift = None
    if not args.no_drift:
        drift = {"mixed": energy_drift(sysm, L.MIXED, args.skin)}
        if fp64:
            drift["fp64"] = energy_drift(sysm, L.FP64, args.skin)
        log("drift", drift)

    # pair-kernel roofline (FP32 VALU; mixed path), from the launches beside PME
    pair_ms = pair_m["beside_pme_ms_per_launch"]
    flops = PAIR_FLOP * pairs
    achieved = flops / (pair_ms * 1e-3) / 1e12 if pair_ms > 0 else 0.0
    alone_ms = pair_m["alone_ms_per_launch"]
    tiles = stats_m["tiles"]
    roofline = {
        "kernel": "pair (real-space LJ + erfc Coulomb, mixed)",
        "bound": "valu",
        "achieved": achieved,
        "peak": PEAK_FP32_VALU,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_VALU,
        "traffic": profiled_traffic(),
        "traffic_source": "profiles/kernels_solvated_chain_100000.md (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; "
                          "scaled by a 256 MiB copy)",
        "frac_uses": "the pair kernel as it runs in a step, beside PME: %d launches with events on its stream"
                     % pair_m["beside_pme_launches"],
        "ms_per_launch": pair_ms,
        "ms_per_launch_alone": alone_ms,
        "achieved_alone": flops / (alone_ms * 1e-3) / 1e12 if alone_ms > 0 else 0.0,
        "peak_without_packed_fp32": PEAK_FP32_VALU / 2,
        "valu_busy": profiled_valu_busy(),
        "valu_busy_source": "profiles/valu_pair_mixed_solvated_chain_100000.md (rocprofv3 --pmc "
                            "SQ_ACTIVE_INST_VALU, GRBM_GUI_ACTIVE)",
        "algorithmic_flop_per_launch": flops,
        "pairs_within_cutoff": pairs,
        "pairs_evaluated": tiles * 64,
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
        "config": workload_config(args.workload, sysm, args.skin),
        "fp64": fp64,
        "mixed_over_fp64": (nsd_m / fp64["ns_per_day"]) if fp64 else None,
        "steady_state": steady,
        "autotune": tuned or None,
        "ingest": ingest,
        "energy_drift": drift,
        "force_error_vs_fp64_oracle": errs,
        "roofline": roofline,
        "pair_timing": pair_m,
        "kernels_mixed": kernels,
        "pair_list": {"entries": stats_m["list_entries"], "tiles": tiles, "rebuilds_in_timed_region": stats_m["rebuilds"],
                      # over the context's life (warm-up and timed steps): steps run as the
                      # graph without the list kernels, and steps run again after one froze
                      "steps_without_list_kernels": stats_m.get("steps_without_list"),
                      "steps_repeated": stats_m.get("steps_repeated")},
    }
    if not args.no_cpu:
        cb = cpu_baseline(sysm, args.cpu_steps)
        out["cpu_baseline"] = cb
        log("cpu baseline %.4f ns/day (%s)" % (cb["value"], cb["kind"]))
    if not args.no_serving:
        out["serving"] = serving_bench(args.serving_requests, args.serving_steps, args.alone_steps)
        log("serving", out["serving"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
