"""ctypes access to the CPU oracle (libnxo.so) and, where built, the compiled
reference (_ref/libnxref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


class Box(C.Structure):
    _fields_ = [("lb", C.c_double * 3), ("ub", C.c_double * 3), ("L", C.c_double * 3)]


class Bond(C.Structure):
    _fields_ = [("p1", C.c_int), ("p2", C.c_int), ("d", C.c_double), ("k", C.c_double)]


class Angle(C.Structure):
    _fields_ = [("p1", C.c_int), ("p2", C.c_int), ("p3", C.c_int), ("a", C.c_double), ("k", C.c_double)]


class Torsion(C.Structure):
    _fields_ = [("p1", C.c_int), ("p2", C.c_int), ("p3", C.c_int), ("p4", C.c_int),
                ("k", C.c_double), ("phase", C.c_double), ("periodicity", C.c_int)]


class SystemC(C.Structure):
    _fields_ = [("n", C.c_int), ("mass", D),
                ("n_bonds", C.c_int), ("bonds", C.POINTER(Bond)),
                ("n_angles", C.c_int), ("angles", C.POINTER(Angle)),
                ("n_torsions", C.c_int), ("torsions", C.POINTER(Torsion)),
                ("q", D), ("sig", D), ("eps", D),
                ("excl_off", I), ("excl_idx", I),
                ("cutoff", C.c_double), ("ewald_tol", C.c_double),
                ("grid", C.c_int * 3), ("use_switch", C.c_int), ("switch_distance", C.c_double),
                ("switch_quirk", C.c_int), ("box", Box)]


class Energies(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("bond", "angle", "torsion", "coulomb", "lj", "recip", "self",
                                         "exclusion", "potential", "reference_pe")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


BOND, ANGLE, TORSION, NONBONDED = 1, 2, 4, 8
ALL = 15


def _d(a):
    return a.ctypes.data_as(D)


def _i(a):
    return a.ctypes.data_as(I)


def _load(path):
    return C.CDLL(path)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libnxo.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = _load(path)
        L = _lib
        L.nxo_pairs_voxel.restype = C.c_int64
        L.nxo_pairs_brute.restype = C.c_int64
        for f in ("nxo_bonds", "nxo_angles", "nxo_torsions", "nxo_alpha", "nxo_pme_convolve", "nxo_pme_recip",
                  "nxo_self_energy", "nxo_direct", "nxo_exclusion_correction", "nxo_kinetic"):
            getattr(L, f).restype = C.c_double
        L.nxo_alpha.argtypes = [C.c_double, C.c_double]
        L.nxo_pme_recip.argtypes = [C.c_int, D, D, C.POINTER(Box), I, C.c_double, D]
        L.nxo_pme_convolve.argtypes = [D, C.POINTER(Box), I, C.c_double]
        L.nxo_kinetic.argtypes = [C.c_int, D, D, D, C.c_double]
        L.nxo_leapfrog.argtypes = [C.c_int, D, D, D, D, C.c_double, C.POINTER(Box)]
        L.nxo_pairs_voxel.argtypes = [C.c_int, D, C.POINTER(Box), I, I, C.c_double, I, C.c_int64]
        L.nxo_pairs_brute.argtypes = L.nxo_pairs_voxel.argtypes
        L.nxo_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, D, C.c_int]
        L.nxo_bspline_moduli.argtypes = [C.c_int, C.c_int, D]
        L.nxo_bspline_theta.argtypes = [C.c_double, C.c_int, D, D]
        L.nxo_simulate.argtypes = [C.POINTER(SystemC), D, D, C.c_int, C.c_int, C.c_double, C.c_int, D, D]
        L.nxo_compute_forces.argtypes = [C.POINTER(SystemC), D, C.c_int, D, C.POINTER(Energies)]
        L.nxo_exclusions.argtypes = [C.c_int, C.c_int, C.POINTER(Bond), C.c_int, C.POINTER(I), C.POINTER(I)]
        L.nxo_free.argtypes = [C.c_void_p]
        L.nxo_box_from_positions.argtypes = [C.c_int, D, C.POINTER(Box)]
        L.nxo_grid_rule.argtypes = [C.POINTER(Box), I]
        L.nxo_pme_spread.argtypes = [C.c_int, D, D, C.POINTER(Box), I, D]
        L.nxo_pme_gather.argtypes = [C.c_int, D, D, C.POINTER(Box), I, D, D]
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libnxref.so"))


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libnxref.so")
        if not os.path.exists(path):
            raise RuntimeError("reference library not built: run `make -C oracle ref` where the reference exists")
        _ref = _load(path)
        R = _ref
        for f in ("nxr_bonds", "nxr_angles", "nxr_torsions", "nxr_alpha", "nxr_pme_convolve", "nxr_kinetic"):
            getattr(R, f).restype = C.c_double
        R.nxr_exclusions.restype = C.c_long
        R.nxr_pairs.restype = C.c_long
        R.nxr_alpha.argtypes = [C.c_double, C.c_double]
        R.nxr_pme_convolve.argtypes = [D, D, I, C.c_double, C.c_double, D]
        R.nxr_kinetic.argtypes = [C.c_int, D, D, D, C.c_double]
        R.nxr_leapfrog.argtypes = [C.c_int, D, D, D, D, C.c_double, D, D]
        R.nxr_pairs.argtypes = [C.c_int, D, D, D, I, I, C.c_double, I, C.c_long]
        R.nxr_exclusions.argtypes = [C.c_int, C.c_int, I, I, I, I, C.c_long]
        R.nxr_nonbonded_real.argtypes = [C.c_int, D, D, D, D, D, D, I, I, C.c_double, C.c_double, C.c_int,
                                         C.c_double, D, D]
    return _ref


# ---------------------------------------------------------------- helpers

def box_of(sys) -> Box:
    b = Box()
    for k in range(3):
        b.lb[k] = float(sys.lb[k])
        b.ub[k] = float(sys.ub[k])
        b.L[k] = abs(float(sys.ub[k]) - float(sys.lb[k]))
    return b


class Prepared:
    """Keeps the ctypes views of a System alive."""

    def __init__(self, sys, excl=None, grid=None, use_switch=0, switch_distance=0.0, switch_quirk=0):
        self.sys = sys
        n = sys.n
        self.mass = np.ascontiguousarray(sys.mass, np.float64)
        self.q = np.ascontiguousarray(sys.q, np.float64)
        self.sig = np.ascontiguousarray(sys.sig, np.float64)
        self.eps = np.ascontiguousarray(sys.eps, np.float64)
        self.bonds = (Bond * max(1, len(sys.bond_idx)))()
        for t, (ij, p) in enumerate(zip(sys.bond_idx, sys.bond_par)):
            self.bonds[t] = Bond(int(ij[0]), int(ij[1]), float(p[0]), float(p[1]))
        self.angles = (Angle * max(1, len(sys.angle_idx)))()
        for t, (ij, p) in enumerate(zip(sys.angle_idx, sys.angle_par)):
            self.angles[t] = Angle(int(ij[0]), int(ij[1]), int(ij[2]), float(p[0]), float(p[1]))
        self.torsions = (Torsion * max(1, len(sys.torsion_idx)))()
        for t, (ij, p, per) in enumerate(zip(sys.torsion_idx, sys.torsion_par, sys.torsion_per)):
            self.torsions[t] = Torsion(int(ij[0]), int(ij[1]), int(ij[2]), int(ij[3]), float(p[0]), float(p[1]), int(per))
        if excl is None:
            excl = exclusions(sys)
        self.excl_off = np.ascontiguousarray(excl[0], np.int32)
        self.excl_idx = np.ascontiguousarray(excl[1], np.int32) if len(excl[1]) else np.zeros(1, np.int32)
        self.box = box_of(sys)
        g = grid if grid is not None else sys.grid
        self.c = SystemC()
        c = self.c
        c.n = n
        c.mass = _d(self.mass)
        c.n_bonds, c.bonds = len(sys.bond_idx), self.bonds
        c.n_angles, c.angles = len(sys.angle_idx), self.angles
        c.n_torsions, c.torsions = len(sys.torsion_idx), self.torsions
        c.q, c.sig, c.eps = _d(self.q), _d(self.sig), _d(self.eps)
        c.excl_off, c.excl_idx = _i(self.excl_off), _i(self.excl_idx)
        c.cutoff, c.ewald_tol = float(sys.cutoff), float(sys.ewald_tol)
        for k in range(3):
            c.grid[k] = int(g[k])
        c.use_switch, c.switch_distance, c.switch_quirk = int(use_switch), float(switch_distance), int(switch_quirk)
        c.box = self.box


def exclusions(sys):
    """Exclusions::createExclusions restated (nxo_exclusions), as CSR arrays."""
    nb = len(sys.bond_idx)
    bonds = (Bond * max(1, nb))()
    for t, ij in enumerate(sys.bond_idx):
        bonds[t] = Bond(int(ij[0]), int(ij[1]), 0.0, 0.0)
    off, idx = I(), I()
    tot = lib().nxo_exclusions(sys.n, nb, bonds, 3, C.byref(off), C.byref(idx))
    if tot < 0:
        raise ValueError("createExclusions: Illegal particle index in list of bonds")
    o = np.ctypeslib.as_array(off, shape=(sys.n + 1,)).copy()
    x = np.ctypeslib.as_array(idx, shape=(max(tot, 1),))[:tot].copy()
    lib().nxo_free(C.cast(off, C.c_void_p))
    lib().nxo_free(C.cast(idx, C.c_void_p))
    return o, x


def compute_forces(prep: Prepared, pos=None, mask=ALL):
    pos = np.ascontiguousarray(prep.sys.pos if pos is None else pos, np.float64)
    f = np.zeros_like(pos)
    e = Energies()
    lib().nxo_compute_forces(C.byref(prep.c), _d(pos), mask, _d(f), C.byref(e))
    return f, e.as_dict()


def simulate(prep: Prepared, nsteps, dt, interval=None, mask=ALL, pos=None, vel=None):
    pos = np.array(prep.sys.pos if pos is None else pos, np.float64, copy=True)
    vel = np.array(prep.sys.vel if vel is None else vel, np.float64, copy=True)
    interval = interval or nsteps
    nrep = nsteps // interval + 2
    eo = np.zeros(3 * nrep)
    f = np.zeros_like(pos)
    k = lib().nxo_simulate(C.byref(prep.c), _d(pos), _d(vel), mask, nsteps, dt, interval, _d(eo), _d(f))
    return pos, vel, f, eo[:3 * k].reshape(k, 3)


def pairs(sys, excl, cutoff=None, method="voxel", pos=None):
    pos = np.ascontiguousarray(sys.pos if pos is None else pos, np.float64)
    off = np.ascontiguousarray(excl[0], np.int32)
    idx = np.ascontiguousarray(excl[1], np.int32) if len(excl[1]) else np.zeros(1, np.int32)
    rc = float(sys.cutoff if cutoff is None else cutoff)
    box = box_of(sys)
    fn = lib().nxo_pairs_voxel if method == "voxel" else lib().nxo_pairs_brute
    cap = 1 << 16
    while True:
        out = np.zeros(2 * cap, np.int32)
        npairs = fn(pos.shape[0], _d(pos), C.byref(box), _i(off), _i(idx), rc, _i(out), cap)
        if npairs <= cap:
            return out[:2 * npairs].reshape(-1, 2)
        cap = int(npairs)


def fft3d(a: np.ndarray, sign: int) -> np.ndarray:
    a = np.ascontiguousarray(a, np.complex128).copy()
    nx, ny, nz = a.shape
    lib().nxo_fft3d(nx, ny, nz, a.ctypes.data_as(D), sign)
    return a


def alpha(cutoff, tol):
    return lib().nxo_alpha(cutoff, tol)


def leapfrog(sys, pos, vel, f, dt):
    pos = np.array(pos, np.float64, copy=True)
    vel = np.array(vel, np.float64, copy=True)
    f = np.ascontiguousarray(f, np.float64)
    m = np.ascontiguousarray(sys.mass, np.float64)
    box = box_of(sys)
    lib().nxo_leapfrog(pos.shape[0], _d(pos), _d(vel), _d(f), _d(m), dt, C.byref(box))
    return pos, vel


def kinetic(sys, vel, f, dt):
    vel = np.ascontiguousarray(vel, np.float64)
    f = np.ascontiguousarray(f, np.float64)
    m = np.ascontiguousarray(sys.mass, np.float64)
    return lib().nxo_kinetic(vel.shape[0], _d(vel), _d(m), _d(f), dt)


def pme_recip(sys, grid, pos=None):
    pos = np.ascontiguousarray(sys.pos if pos is None else pos, np.float64)
    q = np.ascontiguousarray(sys.q, np.float64)
    g = (C.c_int * 3)(*[int(x) for x in grid])
    box = box_of(sys)
    f = np.zeros_like(pos)
    e = lib().nxo_pme_recip(pos.shape[0], _d(pos), _d(q), C.byref(box), g, alpha(sys.cutoff, sys.ewald_tol), _d(f))
    return f, e


def box_from_positions(pos):
    pos = np.ascontiguousarray(pos, np.float64)
    b = Box()
    lib().nxo_box_from_positions(pos.shape[0], _d(pos), C.byref(b))
    return np.array(b.lb[:]), np.array(b.ub[:]), np.array(b.L[:])


def grid_rule(lb, ub):
    b = Box()
    for k in range(3):
        b.lb[k], b.ub[k] = float(lb[k]), float(ub[k])
        b.L[k] = abs(float(ub[k]) - float(lb[k]))
    g = (C.c_int * 3)()
    lib().nxo_grid_rule(C.byref(b), g)
    return np.array(g[:], np.int32)


def bspline_moduli(K, order=5):
    m = np.zeros(int(K))
    lib().nxo_bspline_moduli(int(K), int(order), _d(m))
    return m


def bspline_theta(frac, order=5):
    th, dth = np.zeros(order), np.zeros(order)
    lib().nxo_bspline_theta(float(frac), int(order), _d(th), _d(dth))
    return th, dth


def particle_index(x, sys, grid):
    x = np.ascontiguousarray(x, np.float64)
    box = box_of(sys)
    g = (C.c_int * 3)(*[int(k) for k in grid])
    idx = (C.c_int * 3)()
    fr = np.zeros(3)
    lib().nxo_particle_index(_d(x), C.byref(box), g, idx, _d(fr))
    return np.array(idx[:], np.int32), fr


def pme_spread(sys, grid, pos=None, q=None):
    pos = np.ascontiguousarray(sys.pos if pos is None else pos, np.float64)
    q = np.ascontiguousarray(sys.q if q is None else q, np.float64)
    g = (C.c_int * 3)(*[int(k) for k in grid])
    c = np.zeros(tuple(int(k) for k in grid), np.complex128)
    box = box_of(sys)
    lib().nxo_pme_spread(pos.shape[0], _d(pos), _d(q), C.byref(box), g, c.ctypes.data_as(D))
    return c


def nonbonded_real(prep: Prepared, pos=None):
    """Self, direct and exclusion terms only (mask NONBONDED minus the reciprocal part)."""
    sys = prep.sys
    pos = np.ascontiguousarray(sys.pos if pos is None else pos, np.float64)
    f_all, e = compute_forces(prep, pos, NONBONDED)
    grid = prep.c.grid[:]
    if min(grid) <= 0:
        grid = grid_rule(sys.lb, sys.ub)
    f_rec, _ = pme_recip(sys, grid, pos)
    return f_all - f_rec, e
