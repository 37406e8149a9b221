"""Preparation of generated systems: steepest-descent relaxation on the GPU.

Lattice-built boxes start with compressed contacts; released into NVE they
heat by hundreds of kelvin and bare hydrogens can fall onto charged beads.
`relax` removes the worst contacts with a force-capped steepest descent
(positions move along F by at most `max_step` nm per iteration, the step
grows after an energy decrease and halves after an increase), using the
engine's own force evaluation (mixed path: deterministic fixed-point sums).
A trial step is also rejected when any bond leaves +-15% of its rest
length: with LJ on water oxygens only, the potential is unbounded below along
a stretched O-H bond whose bare hydrogen falls onto a charged partner, and a
descent method would otherwise slide into that well.
Velocities are then redrawn from Maxwell-Boltzmann.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from . import engine as E
from .systems import maxwell_boltzmann


def _bonds_ok(sys, pos, tol=0.15):
    if len(sys.bond_idx) == 0:
        return True
    d = pos[sys.bond_idx[:, 1]] - pos[sys.bond_idx[:, 0]]
    d -= sys.L * np.round(d / sys.L)
    r = np.sqrt((d * d).sum(1))
    return bool(np.all(np.abs(r - sys.bond_par[:, 0]) <= tol * sys.bond_par[:, 0]))


def relax(sys, iterations=300, max_step=0.01, precision=L.MIXED, temperature=300.0, seed=7, log=None):
    ctx = E.Context(sys, precision=precision, skin=0.1)
    pos = np.array(sys.pos, np.float64)
    f, e = ctx.compute_forces()
    energy = e["potential"]
    step = max_step
    accepted = 0
    for it in range(iterations):
        fmax = float(np.sqrt((f * f).sum(1)).max())
        if fmax == 0.0 or not np.isfinite(fmax):
            break
        trial = pos + (step / fmax) * f
        trial = sys.lb + np.mod(trial - sys.lb, sys.L)
        if not _bonds_ok(sys, trial):
            step *= 0.5
            if step < 1e-7:
                break
            continue
        ctx.set_state(trial)
        ft, et = ctx.compute_forces()
        if np.isfinite(et["potential"]) and et["potential"] < energy:
            pos, f, energy = trial, ft, et["potential"]
            step = min(step * 1.2, max_step)
            accepted += 1
        else:
            step *= 0.5
            ctx.set_state(pos)
            if step < 1e-7:
                break
    ctx.close()
    out = sys.copy()
    out.pos = pos
    out.vel = maxwell_boltzmann(sys.mass, temperature, np.random.default_rng(seed))
    if log:
        log("relaxed %s: %d/%d steps accepted, potential %.1f kJ/mol" % (sys.name, accepted, iterations, energy))
    return out
