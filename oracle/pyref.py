"""ctypes access to the reference's own functions, compiled from its sources by
`make -C oracle ref` into oracle/_ref/libnxref.so (see ref_driver.cpp).

TEST INFRASTRUCTURE ONLY: used to write the golden fixtures under tests/golden/
and, where the library is built, to check the oracle against the reference
directly.  Nothing here ships or runs on the product path.

FFTW (the reference's one third-party arithmetic dependency,
NonbondedForce.cpp:264-290, fftw 3.3.5) is absent here, so `recip()` composes
the reference's own spread / convolution / gather with numpy's unnormalised
forward (sign -1) and backward (sign +1) 3D DFT in its place.
"""
from __future__ import annotations

import numpy as np

from . import pyoracle as O

_d, _i = O._d, O._i


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def available() -> bool:
    return O.ref_available()


def box(pos):
    """PeriodicBoundaryCondition::extractBoxBoundaries: bbox of the positions +- 1%."""
    pos = _f64(pos)
    lb, ub, L = np.zeros(3), np.zeros(3), np.zeros(3)
    O.ref().nxr_box_from_positions(pos.shape[0], _d(pos), _d(lb), _d(ub), _d(L))
    return lb, ub, L


def exclusions(n, bond_idx):
    """Exclusions::createExclusions (bond cutoff 3) as CSR."""
    b = _i32(bond_idx).reshape(-1, 2)
    p1, p2 = _i32(b[:, 0]), _i32(b[:, 1])
    off = np.zeros(n + 1, np.int32)
    cap = max(1, 16 * n)
    idx = np.zeros(cap, np.int32)
    tot = O.ref().nxr_exclusions(n, len(b), _i(p1), _i(p2), _i(off), _i(idx), cap)
    if tot > cap:
        idx = np.zeros(tot, np.int32)
        O.ref().nxr_exclusions(n, len(b), _i(p1), _i(p2), _i(off), _i(idx), tot)
    return off, idx[:tot].copy()


def pairs(pos, lb, ub, excl, cutoff):
    """obtainNeighborList: pairs in the order the reference emits them (first > second)."""
    pos = _f64(pos)
    off, idx = _i32(excl[0]), _i32(excl[1]) if len(excl[1]) else np.zeros(1, np.int32)
    lb, ub = _f64(lb), _f64(ub)
    cap = 1 << 16
    while True:
        out = np.zeros(2 * cap, np.int32)
        n = O.ref().nxr_pairs(pos.shape[0], _d(pos), _d(lb), _d(ub), _i(off), _i(idx), float(cutoff), _i(out), cap)
        if n <= cap:
            return out[:2 * n].reshape(-1, 2)
        cap = int(n)


def bonds(sys, pos=None):
    pos = _f64(sys.pos if pos is None else pos)
    f = np.zeros_like(pos)
    ij = _i32(sys.bond_idx).reshape(-1, 2)
    par = _f64(sys.bond_par).reshape(-1, 2)
    p1, p2, d, k = _i32(ij[:, 0]), _i32(ij[:, 1]), _f64(par[:, 0]), _f64(par[:, 1])
    e = O.ref().nxr_bonds(pos.shape[0], _d(pos), _d(_f64(sys.lb)), _d(_f64(sys.ub)), len(ij), _i(p1), _i(p2),
                          _d(d), _d(k), _d(f))
    return f, e


def angles(sys, pos=None):
    pos = _f64(sys.pos if pos is None else pos)
    f = np.zeros_like(pos)
    ij = _i32(sys.angle_idx).reshape(-1, 3)
    par = _f64(sys.angle_par).reshape(-1, 2)
    cols = [_i32(ij[:, c]) for c in range(3)]
    a, k = _f64(par[:, 0]), _f64(par[:, 1])
    e = O.ref().nxr_angles(pos.shape[0], _d(pos), _d(_f64(sys.lb)), _d(_f64(sys.ub)), len(ij), _i(cols[0]),
                           _i(cols[1]), _i(cols[2]), _d(a), _d(k), _d(f))
    return f, e


def torsions(sys, pos=None):
    pos = _f64(sys.pos if pos is None else pos)
    f = np.zeros_like(pos)
    ij = _i32(sys.torsion_idx).reshape(-1, 4)
    par = _f64(sys.torsion_par).reshape(-1, 2)
    cols = [_i32(ij[:, c]) for c in range(4)]
    k, ph, per = _f64(par[:, 0]), _f64(par[:, 1]), _i32(sys.torsion_per)
    e = O.ref().nxr_torsions(pos.shape[0], _d(pos), _d(_f64(sys.lb)), _d(_f64(sys.ub)), len(ij), _i(cols[0]),
                             _i(cols[1]), _i(cols[2]), _i(cols[3]), _d(k), _d(ph), _i(per), _d(f))
    return f, e


def alpha(cutoff, tol):
    return O.ref().nxr_alpha(float(cutoff), float(tol))


def grid_rule(lb, ub):
    g = np.zeros(3, np.int32)
    O.ref().nxr_grid_rule(_d(_f64(lb)), _d(_f64(ub)), _i(g))
    return g


def bspline(pos, lb, ub, grid):
    """computeBSplineParameters: moduli[3][K], theta/dtheta [3][n][5], index [n][3], frac [n][3]."""
    pos = _f64(pos)
    n = pos.shape[0]
    g = _i32(grid)
    mods = [np.zeros(int(k)) for k in g]
    th = np.zeros((3, n, 5))
    dth = np.zeros((3, n, 5))
    idx = np.zeros((n, 3), np.int32)
    frac = np.zeros((n, 3))
    O.ref().nxr_bspline(n, _d(pos), _d(_f64(lb)), _d(_f64(ub)), _i(g), _d(mods[0]), _d(mods[1]), _d(mods[2]),
                        _d(th), _d(dth), _i(idx), _d(frac))
    return mods, th, dth, idx, frac


def spread(pos, q, lb, ub, grid):
    """gridSpreadCharge_Reference -> complex grid [Kx, Ky, Kz]."""
    pos = _f64(pos)
    g = _i32(grid)
    c = np.zeros(tuple(int(k) for k in g), np.complex128)
    O.ref().nxr_pme_spread(pos.shape[0], _d(pos), _d(_f64(q)), _d(_f64(lb)), _d(_f64(ub)), _i(g),
                           c.ctypes.data_as(O.D))
    return c


def convolve(cgrid, lb, ub, cutoff, tol):
    """reciprocalConvolutionAndEnergy on a transformed grid; returns (grid, energy)."""
    c = np.ascontiguousarray(cgrid, np.complex128).copy()
    g = _i32(c.shape)
    e = O.ref().nxr_pme_convolve(_d(_f64(lb)), _d(_f64(ub)), _i(g), float(cutoff), float(tol),
                                 c.ctypes.data_as(O.D))
    return c, e


def gather(pos, q, lb, ub, cgrid, f=None):
    """PMEGridForces: f -= q * grad(phi) (accumulates into a copy of f)."""
    pos = _f64(pos)
    f = np.zeros_like(pos) if f is None else _f64(f).copy()
    c = np.ascontiguousarray(cgrid, np.complex128)
    g = _i32(c.shape)
    O.ref().nxr_pme_gather(pos.shape[0], _d(pos), _d(_f64(q)), _d(_f64(lb)), _d(_f64(ub)), _i(g),
                           c.ctypes.data_as(O.D), _d(f))
    return f


def recip(pos, q, lb, ub, grid, cutoff, tol):
    """Reciprocal-space forces and energy: the reference's spread, convolution and
    gather around numpy's unnormalised 3D DFT (FFTW's transform)."""
    c = spread(pos, q, lb, ub, grid)
    c = np.fft.fftn(c)
    c, e = convolve(c, lb, ub, cutoff, tol)
    c = np.fft.ifftn(c) * c.size
    return gather(pos, q, lb, ub, c), e


def nonbonded_real(pos, q, sig, eps, lb, ub, excl, cutoff, tol, use_switch=0, rsw=0.0):
    """Self energy, direct (Coulomb + LJ) and exclusion correction: (f, self, direct, -exclusion)."""
    pos = _f64(pos)
    f = np.zeros_like(pos)
    e = np.zeros(3)
    off, idx = _i32(excl[0]), _i32(excl[1]) if len(excl[1]) else np.zeros(1, np.int32)
    O.ref().nxr_nonbonded_real(pos.shape[0], _d(pos), _d(_f64(q)), _d(_f64(sig)), _d(_f64(eps)), _d(_f64(lb)),
                               _d(_f64(ub)), _i(off), _i(idx), float(cutoff), float(tol), int(use_switch),
                               float(rsw), _d(f), _d(e))
    return f, e[0], e[1], e[2]


def forces(sys, pos=None, excl=None, grid=None, mask=O.ALL):
    """Engine::CalculateForces composed from the reference's functions.

    Returns (F, energies) with energies = dict(bond, angle, torsion, recip, self,
    direct, exclusion, reference_pe); reference_pe is what Engine reports: with nonbonded on,
    the reciprocal energy overwrites the bonded sum (NonbondedForce.cpp:387)."""
    pos = _f64(sys.pos if pos is None else pos)
    excl = exclusions(sys.n, sys.bond_idx) if excl is None else excl
    grid = sys.grid if grid is None else grid
    F = np.zeros_like(pos)
    e = dict(bond=0.0, angle=0.0, torsion=0.0, recip=0.0, self=0.0, direct=0.0, exclusion=0.0)
    pe = 0.0
    if mask & O.BOND and len(sys.bond_idx):
        f, e["bond"] = bonds(sys, pos)
        F += f
        pe += e["bond"]
    if mask & O.ANGLE and len(sys.angle_idx):
        f, e["angle"] = angles(sys, pos)
        F += f
        pe += e["angle"]
    if mask & O.TORSION and len(sys.torsion_idx):
        f, e["torsion"] = torsions(sys, pos)
        F += f
        pe += e["torsion"]
    if mask & O.NONBONDED:
        fr, e["recip"] = recip(pos, sys.q, sys.lb, sys.ub, grid, sys.cutoff, sys.ewald_tol)
        fd, e["self"], e["direct"], e["exclusion"] = nonbonded_real(pos, sys.q, sys.sig, sys.eps, sys.lb, sys.ub,
                                                                    excl, sys.cutoff, sys.ewald_tol)
        F += fr + fd
        pe = e["recip"] + e["self"] + e["direct"] + e["exclusion"]
    e["reference_pe"] = pe
    return F, e


def leapfrog(pos, vel, f, mass, dt, lb, ub):
    pos, vel = _f64(pos).copy(), _f64(vel).copy()
    O.ref().nxr_leapfrog(pos.shape[0], _d(pos), _d(vel), _d(_f64(f)), _d(_f64(mass)), float(dt), _d(_f64(lb)),
                         _d(_f64(ub)))
    return pos, vel


def kinetic(vel, mass, f, dt):
    vel = _f64(vel)
    return O.ref().nxr_kinetic(vel.shape[0], _d(vel), _d(_f64(mass)), _d(_f64(f)), float(dt))


def simulate(sys, nsteps, dt, interval=1, mask=O.ALL, grid=None):
    """Engine::RunSimulation order (Engine.cpp:162-188): forces, report (KE with
    the half-step shift, PE, total), then leapfrog.  Returns the final state,
    the last forces and the report rows."""
    excl = exclusions(sys.n, sys.bond_idx)
    pos, vel = _f64(sys.pos).copy(), _f64(sys.vel).copy()
    rows = []
    F = None
    for step in range(nsteps):
        F, e = forces(sys, pos, excl, grid, mask)
        if step % interval == 0:
            ke = kinetic(vel, sys.mass, F, dt)
            rows.append((ke, e["reference_pe"], ke + e["reference_pe"]))
        pos, vel = leapfrog(pos, vel, F, sys.mass, dt, sys.lb, sys.ub)
    return pos, vel, F, np.array(rows)


def report(out_name, pos, vel, f, ke, pe, step):
    """Reporter::TotalEnergyReport + TestPVFReport; files land in the working directory."""
    import ctypes as C
    pos, vel, f = _f64(pos), _f64(vel), _f64(f)
    O.ref().nxr_report(out_name.encode(), C.c_int(pos.shape[0]), _d(pos), _d(vel), _d(f), C.c_double(ke),
                       C.c_double(pe), C.c_int(step))


def pdb_report(in_pdb, out_pdb, pos, model, first):
    """Reporter::pdbOutputGeneratorPart1 (when first) and Part2."""
    pos = _f64(pos)
    O.ref().nxr_pdb_report(in_pdb.encode(), out_pdb.encode(), pos.shape[0], _d(pos), int(model), int(bool(first)))


def parse_counts(system_xml, state_xml):
    """(masses, bonds, angles, torsions, nonbonded particles, positions) as the reference's parsers count them."""
    import ctypes as C
    c = (C.c_int * 6)()
    O.ref().nxr_parse_counts(system_xml.encode(), state_xml.encode() if state_xml else None, c)
    return tuple(c[:])
