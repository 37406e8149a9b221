"""Shared helpers of the test suite."""
from __future__ import annotations

import os

import numpy as np

from nexusmd_amd.systems import System

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_rms(a, b):
    """sqrt(sum |a - b|^2 / sum |b|^2)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    den = float(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2) / den)) if den > 0 else float(np.sqrt(np.sum(a * a)))


def golden(name):
    """Loads tests/golden/<name>.npz (plain arrays; no pickled objects)."""
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def system_from(d, name="fixture"):
    """A System from the input arrays of a golden fixture."""
    return System(name=name, pos=d["pos"].copy(), vel=d["vel"].copy(), mass=d["mass"].copy(),
                  bond_idx=d["bond_idx"], bond_par=d["bond_par"], angle_idx=d["angle_idx"],
                  angle_par=d["angle_par"], torsion_idx=d["torsion_idx"], torsion_par=d["torsion_par"],
                  torsion_per=d["torsion_per"], q=d["q"], sig=d["sig"], eps=d["eps"], lb=d["lb"], ub=d["ub"],
                  cutoff=float(d["cutoff"]), ewald_tol=float(d["ewald_tol"]),
                  grid=tuple(int(g) for g in d["grid"]))


def bare_system(pos, lb, ub, q=None, sig=None, eps=None, mass=None, cutoff=1.0, ewald_tol=5e-4, grid=(0, 0, 0),
                name="bare"):
    """Charged LJ particles without bonded terms."""
    n = len(pos)
    z2, z3, z4 = np.zeros((0, 2), np.int32), np.zeros((0, 3), np.int32), np.zeros((0, 4), np.int32)
    f2 = np.zeros((0, 2))
    return System(name=name, pos=np.ascontiguousarray(pos, np.float64), vel=np.zeros((n, 3)),
                  mass=np.ones(n) if mass is None else mass, bond_idx=z2, bond_par=f2, angle_idx=z3,
                  angle_par=f2, torsion_idx=z4, torsion_par=f2, torsion_per=np.zeros(0, np.int32),
                  q=np.zeros(n) if q is None else np.asarray(q, np.float64),
                  sig=np.full(n, 0.3) if sig is None else np.asarray(sig, np.float64),
                  eps=np.zeros(n) if eps is None else np.asarray(eps, np.float64),
                  lb=np.asarray(lb, np.float64), ub=np.asarray(ub, np.float64), cutoff=cutoff,
                  ewald_tol=ewald_tol, grid=tuple(grid))
