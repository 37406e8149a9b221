"""Seeded synthetic systems for the NexusMD force-and-integrate path.

The workloads: a bead chain (1,500 beads in a 4.0 nm box), a flexible 3-site
water box (24,000 atoms, 6.2 nm), a bead chain in water (~100,000 atoms,
10.0 nm), batches of small ligand-in-water systems (600-8,000 atoms) and a
1,000,002-atom water box (21.5 nm).

Topology follows the OpenMM-style System XML that the reference reads
(core/common/src/SystemXMLParser.cpp:18-196): per-particle masses,
HarmonicBondForce (p1, p2, d, k), HarmonicAngleForce (p1, p2, p3, a, k),
PeriodicTorsionForce (p1..p4, k, phase, periodicity) and NonbondedForce
particles (q, sig, eps) with cutoff and ewaldTolerance.

Generator choices that the configurations leave open are fixed here:
bead mass 12.011 amu, minimum non-excluded bead distance 0.3 nm in the
free chain, velocities Maxwell-Boltzmann with the centre-of-mass motion
removed.  In the pure water box the hydrogens carry sig = 0.1 nm, eps = 0
(LJ on oxygen only); next to charged beads they carry the CHARMM TIP3P core
sig = 0.04 nm, eps = 0.192 kJ/mol (see water_topology).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

KB = 0.0083144626  # kJ/mol/K
BEAD_MASS = 12.011
O_MASS, H_MASS = 15.999, 1.008
OH_R0, OH_K = 0.09572, 462750.0
HOH_A0, HOH_K = 1.824, 836.8
Q_O, Q_H = -0.834, 0.417
SIG_O, EPS_O = 0.315, 0.636
BOND_R0, BOND_K = 0.15, 2.5e5
ANGLE_A0, ANGLE_K = 1.91, 400.0


@dataclasses.dataclass
class System:
    """One periodic system: FP64 arrays, atom-major (N, 3) positions."""
    name: str
    pos: np.ndarray                 # (N, 3) nm
    vel: np.ndarray                 # (N, 3) nm/ps
    mass: np.ndarray                # (N,) amu
    bond_idx: np.ndarray            # (nb, 2) int32
    bond_par: np.ndarray            # (nb, 2) float64: d, k
    angle_idx: np.ndarray           # (na, 3) int32
    angle_par: np.ndarray           # (na, 2) float64: a, k
    torsion_idx: np.ndarray         # (nt, 4) int32
    torsion_par: np.ndarray         # (nt, 2) float64: k, phase
    torsion_per: np.ndarray         # (nt,) int32
    q: np.ndarray
    sig: np.ndarray
    eps: np.ndarray
    lb: np.ndarray                  # (3,) box lower bound
    ub: np.ndarray                  # (3,) box upper bound
    cutoff: float = 1.0
    ewald_tol: float = 5e-4
    grid: tuple = (0, 0, 0)         # PME grid, (0,0,0) -> reference rule
    dt: float = 0.0005              # ps
    steps: int = 1000

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def L(self) -> np.ndarray:
        return np.abs(self.ub - self.lb)

    def copy(self) -> "System":
        return dataclasses.replace(self, **{f.name: (getattr(self, f.name).copy()
                                                     if isinstance(getattr(self, f.name), np.ndarray)
                                                     else getattr(self, f.name))
                                            for f in dataclasses.fields(self)})


def _empty_i(k):
    return np.zeros((0, k), np.int32)


def _empty_f(k):
    return np.zeros((0, k), np.float64)


def maxwell_boltzmann(mass: np.ndarray, temperature: float, rng: np.random.Generator) -> np.ndarray:
    sd = np.sqrt(KB * temperature / mass)[:, None]
    v = rng.normal(size=(mass.size, 3)) * sd
    v -= (mass[:, None] * v).sum(0) / mass.sum()
    return v


def _place(a, b, c, bond, angle, dihedral):
    """Natural-extension reference frame: atom d from a, b, c (vectorised over dihedral)."""
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    nn = np.linalg.norm(n)
    if nn < 1e-12:
        n = np.cross(bc, np.array([1.0, 0.0, 0.0]) if abs(bc[0]) < 0.9 else np.array([0.0, 1.0, 0.0]))
        nn = np.linalg.norm(n)
    n /= nn
    m = np.cross(n, bc)
    dih = np.atleast_1d(dihedral)
    d2 = np.stack([-bond * math.cos(angle) * np.ones_like(dih),
                   bond * math.sin(angle) * np.cos(dih),
                   bond * math.sin(angle) * np.sin(dih)], axis=1)
    return c[None, :] + d2[:, 0:1] * bc[None, :] + d2[:, 1:2] * m[None, :] + d2[:, 2:3] * n[None, :]


class _Grid:
    """Periodic cell hash of placed points for overlap rejection."""

    def __init__(self, L, cell):
        self.L = np.asarray(L, float)
        self.nc = np.maximum((self.L / cell).astype(int), 1)
        self.cells: dict = {}
        self.pts: list = []

    def key(self, p):
        return tuple((np.floor((p % self.L) / self.L * self.nc).astype(int)) % self.nc)

    def add(self, p, idx):
        self.cells.setdefault(self.key(p), []).append(idx)
        self.pts.append(p)

    def near(self, p):
        k = np.array(self.key(p))
        out = []
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    kk = tuple((k + (dx, dy, dz)) % self.nc)
                    out.extend(self.cells.get(kk, ()))
        return out


def _min_dist2(cands, pts, L):
    d = cands[:, None, :] - pts[None, :, :]
    d -= L * np.round(d / L)
    return (d * d).sum(-1)


def grow_chain(n: int, L, rng, min_dist=0.3, center=None, radius=None, avoid=None):
    """Self-avoiding bead chain (bond 0.15 nm, angle 1.91 rad, random dihedrals).

    Beads more than three bonds apart stay >= min_dist apart (periodic);
    with `radius`, beads stay inside a sphere about `center`.
    """
    L = np.asarray(L, float)
    cell = max(min_dist, 0.3)
    grid = _Grid(L, cell)
    if center is None:
        center = L / 2
    pos = np.zeros((n, 3))
    pos[0] = center + (rng.uniform(-0.2, 0.2, 3) if radius is not None else rng.uniform(0, 1, 3) * L - L / 2)
    pos[1] = pos[0] + np.array([BOND_R0, 0, 0])
    ang = ANGLE_A0
    pos[2] = pos[1] + BOND_R0 * np.array([-math.cos(ang), math.sin(ang), 0.0])
    for i in range(3):
        grid.add(pos[i], i)
    md2 = min_dist * min_dist
    i = 3
    fails = 0
    while i < n:
        dih = rng.uniform(-math.pi, math.pi, 24)
        cands = _place(pos[i - 3], pos[i - 2], pos[i - 1], BOND_R0, ANGLE_A0, dih)
        ok = np.ones(len(cands), bool)
        if radius is not None:
            ok &= np.linalg.norm(cands - center, axis=1) < radius
        near = [j for j in grid.near(pos[i - 1]) if j < i - 3]
        if near:
            ok &= (_min_dist2(cands, pos[near], L) >= md2).all(1)
        if avoid is not None:
            ok &= avoid(cands)
        sel = np.flatnonzero(ok)
        if sel.size:
            pos[i] = cands[sel[0]]
            grid.add(pos[i], i)
            i += 1
            fails = 0
        else:
            fails += 1
            back = min(i - 3, 2 + fails)
            # backtrack: drop the last `back` beads
            for j in range(i - back, i):
                k = grid.key(pos[j])
                grid.cells[k].remove(j)
            del grid.pts[i - back:]
            i -= back
            if fails > 200:
                raise RuntimeError("chain growth stuck")
    return pos


def chain_topology(n: int, rng, offset=0):
    """Bonds, angles and a torsion on every dihedral for an n-bead chain."""
    b = np.stack([np.arange(n - 1), np.arange(1, n)], 1).astype(np.int32) + offset
    bp = np.tile([BOND_R0, BOND_K], (n - 1, 1))
    a = np.stack([np.arange(n - 2), np.arange(1, n - 1), np.arange(2, n)], 1).astype(np.int32) + offset
    ap = np.tile([ANGLE_A0, ANGLE_K], (n - 2, 1))
    t = np.stack([np.arange(n - 3), np.arange(1, n - 2), np.arange(2, n - 1), np.arange(3, n)], 1).astype(np.int32) + offset
    tk = rng.uniform(0.0, 10.0, n - 3)
    tph = rng.choice([0.0, math.pi], n - 3)
    tper = rng.choice([1, 2, 3], n - 3).astype(np.int32)
    return b, bp, a, ap, t, np.stack([tk, tph], 1), tper


def chain_nonbonded(n: int, rng):
    sig = rng.uniform(0.25, 0.35, n)
    eps = rng.uniform(0.1, 1.0, n)
    q = rng.normal(0.0, 0.4, n)
    q -= q.mean()
    return q, sig, eps


def bead_chain(n=1500, L=4.0, seed=20261202, cutoff=1.0, ewald_tol=5e-4, grid=(40, 40, 40),
               temperature=300.0, dt=0.001, steps=10) -> System:
    """Bead chain in a cubic periodic box (1,500 beads / 4.0 nm / grid 40^3 by default)."""
    rng = np.random.default_rng(seed)
    Lv = np.array([L, L, L])
    pos = grow_chain(n, Lv, rng) % Lv
    b, bp, a, ap, t, tp, tper = chain_topology(n, rng)
    q, sig, eps = chain_nonbonded(n, rng)
    mass = np.full(n, BEAD_MASS)
    vel = maxwell_boltzmann(mass, temperature, rng)
    return System("bead_chain_%d" % n, pos, vel, mass, b, bp, a, ap, t, tp, tper, q, sig, eps,
                  np.zeros(3), Lv.copy(), cutoff, ewald_tol, tuple(grid), dt, steps)


def _orient(m, rng):
    u = rng.normal(size=(m, 4))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    w, x, y, z = u.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)


def _water_molecules(sites: np.ndarray, rng, L=None, others=None, min_h=0.15, rounds=40) -> np.ndarray:
    """O at each site, two H at r0 with angle a0, each molecule in a random orientation.

    Orientations are re-drawn (up to `rounds` times) for molecules whose
    hydrogens land closer than `min_h` to an atom of another molecule (or to a
    point of `others`), so that no bare hydrogen starts on top of a partner.
    """
    m = sites.shape[0]
    h = HOH_A0 / 2
    loc = np.array([[0.0, 0.0, 0.0],
                    [OH_R0 * math.sin(h), OH_R0 * math.cos(h), 0.0],
                    [-OH_R0 * math.sin(h), OH_R0 * math.cos(h), 0.0]])
    R = _orient(m, rng)
    xyz = sites[:, None, :] + np.einsum("mij,kj->mki", R, loc)
    if L is None or m == 0:
        return xyz.reshape(-1, 3)
    from scipy.spatial import cKDTree
    Lv = np.broadcast_to(np.asarray(L, float), (3,))
    for _ in range(rounds):
        flat = xyz.reshape(-1, 3) % Lv
        mol = np.repeat(np.arange(m), 3)
        tree = cKDTree(flat, boxsize=Lv)
        hs = np.concatenate([np.arange(m) * 3 + 1, np.arange(m) * 3 + 2])
        pairs = tree.query_ball_point(flat[hs], r=min_h)
        bad = set()
        for a, lst in zip(hs, pairs):
            for b in lst:
                if mol[b] != mol[a]:
                    bad.add(mol[a])
                    break
        if others is not None and len(others):
            ot = cKDTree(np.asarray(others) % Lv, boxsize=Lv)
            d, _ = ot.query(flat[hs], k=1)
            bad.update(mol[hs[d < min_h]].tolist())
        if not bad:
            break
        idx = np.fromiter(bad, int)
        R[idx] = _orient(idx.size, rng)
        xyz[idx] = sites[idx, None, :] + np.einsum("mij,kj->mki", R[idx], loc)
    return xyz.reshape(-1, 3)


def _lattice_sites(L, count_min, rng, jitter=0.02, keep=None):
    """Jittered cubic lattice filling [0, L)^3 with at least count_min admissible sites."""
    L = float(L)
    k = max(1, int(math.ceil(count_min ** (1 / 3))))
    while True:
        a = L / k
        g = (np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 3) + 0.5) * a
        g += rng.uniform(-jitter, jitter, g.shape)
        if keep is not None:
            g = g[keep(g)]
        if g.shape[0] >= count_min:
            return g
        k += 1


H_SIG_CORE, H_EPS_CORE = 0.04, 0.192   # CHARMM TIP3P hydrogen


def water_topology(n_mol: int, offset: int, h_core: bool = False):
    """Flexible 3-site water.  h_core gives the hydrogens the small LJ core of the
    CHARMM TIP3P model; used beside charged beads, whose N(0, 0.4) charges reach
    |q| ~ 1.5 and would otherwise pull a bare hydrogen onto the bead (with LJ on
    oxygen only the potential is unbounded below along that path)."""
    o = offset + 3 * np.arange(n_mol, dtype=np.int32)
    b = np.concatenate([np.stack([o, o + 1], 1), np.stack([o, o + 2], 1)]).astype(np.int32)
    order = np.argsort(np.concatenate([o * 2, o * 2 + 1]), kind="stable")
    b = b[order]
    bp = np.tile([OH_R0, OH_K], (2 * n_mol, 1))
    a = np.stack([o + 1, o, o + 2], 1).astype(np.int32)
    ap = np.tile([HOH_A0, HOH_K], (n_mol, 1))
    q = np.tile([Q_O, Q_H, Q_H], n_mol)
    hs, he = (H_SIG_CORE, H_EPS_CORE) if h_core else (0.1, 0.0)
    sig = np.tile([SIG_O, hs, hs], n_mol)
    eps = np.tile([EPS_O, he, he], n_mol)
    mass = np.tile([O_MASS, H_MASS, H_MASS], n_mol)
    return b, bp, a, ap, q, sig, eps, mass


def water_box(n_mol=8000, L=6.2, seed=20261203, cutoff=1.0, ewald_tol=5e-4, grid=(64, 64, 64),
              temperature=300.0, dt=0.0005, steps=2000, name=None) -> System:
    """Flexible 3-site water on a jittered lattice with random orientations."""
    rng = np.random.default_rng(seed)
    sites = _lattice_sites(L, n_mol, rng)
    sites = sites[rng.permutation(sites.shape[0])[:n_mol]]
    sites = sites[np.lexsort((sites[:, 2], sites[:, 1], sites[:, 0]))]
    pos = _water_molecules(sites, rng, L=L) % L
    b, bp, a, ap, q, sig, eps, mass = water_topology(n_mol, 0)
    vel = maxwell_boltzmann(mass, temperature, rng)
    Lv = np.array([L, L, L], float)
    return System(name or "water_%d" % (3 * n_mol), pos, vel, mass, b, bp, a, ap, _empty_i(4), _empty_f(2),
                  np.zeros(0, np.int32), q, sig, eps, np.zeros(3), Lv, cutoff, ewald_tol, tuple(grid), dt, steps)


def solvated_chain(n_chain=10000, n_water=30000, L=10.0, seed=20261204, cutoff=1.0, ewald_tol=5e-4,
                   grid=(96, 96, 96), temperature=300.0, dt=0.0005, steps=5000, chain_min=0.3,
                   water_gap=0.24, name=None) -> System:
    """Bead chain grown as a self-avoiding random walk from the box centre, then waters.

    The chain keeps >= chain_min between non-excluded beads (periodic); water
    oxygens keep >= water_gap from every bead; the water lattice is refined
    until n_water sites fit.  (A compact globule of 10,000 beads at a 0.3 nm
    contact distance cannot leave room for 30,000 waters at liquid density in
    a 10 nm box, so the walk is left free and the waters fill around it.)
    """
    rng = np.random.default_rng(seed)
    Lv = np.array([L, L, L], float)
    chain = grow_chain(n_chain, Lv, rng, min_dist=chain_min, center=Lv / 2)
    from scipy.spatial import cKDTree
    tree = cKDTree(chain % Lv, boxsize=Lv)

    def keep(g):
        d, _ = tree.query(g % Lv, k=1)
        return d >= water_gap + 0.05

    sites = _lattice_sites(L, n_water, rng, keep=keep)
    sites = sites[np.sort(rng.permutation(sites.shape[0])[:n_water])]
    wat = _water_molecules(sites, rng, L=Lv, others=chain)
    pos = np.concatenate([chain, wat]) % Lv
    b, bp, a, ap, t, tp, tper = chain_topology(n_chain, rng)
    q, sig, eps = chain_nonbonded(n_chain, rng)
    wb, wbp, wa, wap, wq, wsig, weps, wmass = water_topology(n_water, n_chain, h_core=True)
    mass = np.concatenate([np.full(n_chain, BEAD_MASS), wmass])
    vel = maxwell_boltzmann(mass, temperature, rng)
    return System(name or "solvated_chain_%d" % pos.shape[0], pos, vel, mass,
                  np.concatenate([b, wb]), np.concatenate([bp, wbp]),
                  np.concatenate([a, wa]), np.concatenate([ap, wap]), t, tp, tper,
                  np.concatenate([q, wq]), np.concatenate([sig, wsig]), np.concatenate([eps, weps]),
                  np.zeros(3), Lv, cutoff, ewald_tol, tuple(grid), dt, steps)


def ligand_in_water(n_total: int, seed: int, cutoff=0.9, ewald_tol=5e-4, spacing=0.12,
                    temperature=300.0, dt=0.0005, steps=5000) -> System:
    """A 20-60 bead chain in a water box sized to ~100 atoms/nm^3 (one serving request)."""
    rng = np.random.default_rng(seed)
    n_lig = int(rng.integers(20, 61))
    n_wat = max(1, (n_total - n_lig) // 3)
    n = n_lig + 3 * n_wat
    L = (n / 100.0) ** (1 / 3)
    L = max(L, 2.0 * cutoff + 0.02)
    Lv = np.array([L, L, L])
    lig = grow_chain(n_lig, Lv, rng, min_dist=0.3, center=Lv / 2, radius=L / 2 - 0.1)
    from scipy.spatial import cKDTree
    tree = cKDTree(lig % Lv, boxsize=Lv)

    def keep(g):
        d, _ = tree.query(g % Lv, k=1)
        return d >= 0.29

    sites = _lattice_sites(L, n_wat, rng, keep=keep)
    sites = sites[np.sort(rng.permutation(sites.shape[0])[:n_wat])]
    pos = np.concatenate([lig, _water_molecules(sites, rng, L=Lv, others=lig)]) % Lv
    b, bp, a, ap, t, tp, tper = chain_topology(n_lig, rng)
    q, sig, eps = chain_nonbonded(n_lig, rng)
    wb, wbp, wa, wap, wq, wsig, weps, wmass = water_topology(n_wat, n_lig, h_core=True)
    mass = np.concatenate([np.full(n_lig, BEAD_MASS), wmass])
    vel = maxwell_boltzmann(mass, temperature, rng)
    k = int(2 * math.ceil(L / spacing / 2))
    k = max(k, 24)
    k = _fft_friendly(k)
    return System("ligand_%d" % n, pos, vel, mass, np.concatenate([b, wb]), np.concatenate([bp, wbp]),
                  np.concatenate([a, wa]), np.concatenate([ap, wap]), t, tp, tper,
                  np.concatenate([q, wq]), np.concatenate([sig, wsig]), np.concatenate([eps, weps]),
                  np.zeros(3), Lv, cutoff, ewald_tol, (k, k, k), dt, steps)


def _fft_friendly(k: int) -> int:
    """Smallest even k' >= k whose prime factors are 2, 3 and 5."""
    while True:
        m = k
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1 and k % 2 == 0:
            return k
        k += 1


def serving_batch(count=256, seed=20261205, lo=600, hi=8000, median=2500, sigma=0.6):
    """Atom counts lognormal(ln median, sigma) clipped to [lo, hi], one seed per request."""
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.exp(rng.normal(math.log(median), sigma, count)), lo, hi).astype(int)
    return [int(s) for s in sizes], [seed * 1000 + i for i in range(count)]


def poisson_arrivals(count: int, rate_per_s: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, count))


def exclusions_csr(n: int, bond_idx: np.ndarray):
    """1-2/1-3/1-4 exclusions (graph distance <= 3 over the bond graph) as CSR.

    Same set as the reference's two-pass expansion with bondCutoff = 3
    (core/baseline/src/Forces/Exclusions.cpp:7-44).
    """
    adj = [[] for _ in range(n)]
    for p1, p2 in bond_idx:
        adj[p1].append(int(p2))
        adj[p2].append(int(p1))
    off = np.zeros(n + 1, np.int32)
    idx = []
    for i in range(n):
        seen = {i: 0}
        frontier = [i]
        for depth in range(1, 4):
            nxt = []
            for u in frontier:
                for v in adj[u]:
                    if v not in seen:
                        seen[v] = depth
                        nxt.append(v)
            frontier = nxt
        ex = sorted(k for k in seen if k != i)
        idx.extend(ex)
        off[i + 1] = len(idx)
    return off, np.asarray(idx, np.int32)


def by_name(name: str, **kw) -> System:
    table = {
        "bead_chain_1500": lambda: bead_chain(**kw),
        "water_24000": lambda: water_box(**kw),
        "solvated_chain_100000": lambda: solvated_chain(**kw),
        "water_1000002": lambda: water_box(n_mol=333334, L=21.5, grid=(192, 192, 192), steps=1000, seed=20261206, **kw),
    }
    return table[name]()
