"""Calibration of the CPU baselines: the compiled reference (oracle/_ref,
the reference's own functions built from its sources) against the oracle's C
restatement, one core, on the 1,500-bead chain (c0, grid 40^3) and the
24,000-atom water box (c1, grid 64^3), each run in the reference's step order
(neighbour list rebuilt and full PME every step) after one warm-up step.

    taskset -c 0 python -m oracle.calibrate [--c1-steps 3]

TEST INFRASTRUCTURE ONLY: it times the two checkers, nothing of the engine,
and nothing on the product path imports it.
"""
from __future__ import annotations

import argparse
import json
import os
import time

from nexusmd_amd import systems as S
from oracle import pyoracle as O
from oracle import pyref as R


def time_both(sysm, nsteps):
    prep = O.Prepared(sysm)
    O.simulate(prep, 1, sysm.dt)
    t0 = time.perf_counter()
    O.simulate(prep, nsteps, sysm.dt)
    t_oracle = (time.perf_counter() - t0) / nsteps
    R.simulate(sysm, 1, sysm.dt)
    t0 = time.perf_counter()
    R.simulate(sysm, nsteps, sysm.dt)
    t_ref = (time.perf_counter() - t0) / nsteps
    return {"atoms": sysm.n, "grid": list(sysm.grid), "steps": nsteps, "oracle_s_per_step": t_oracle,
            "reference_s_per_step": t_ref, "reference_over_oracle": t_ref / t_oracle}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--c0-steps", type=int, default=10)
    ap.add_argument("--c1-steps", type=int, default=3)
    args = ap.parse_args()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = None
    out = {"cores_allowed": cores,
           "c0": time_both(S.bead_chain(), args.c0_steps),
           "c1": time_both(S.water_box(), args.c1_steps)}
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
