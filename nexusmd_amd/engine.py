"""Python view of the C-ABI: a single-system Context and the batching Server.

Both wrap include/nexusmd.h through ctypes; every call goes to the HIP
library, and a missing library or GPU raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L


def _d(a):
    return a.ctypes.data_as(L.D) if a is not None else None


class TopologyBuffers:
    """Keeps the ctypes arrays behind an nx_topology alive."""

    def __init__(self, sys, skin=0.1, excl=None, use_switch=0, switch_distance=0.0, grid=None):
        n = sys.n
        self.mass = np.ascontiguousarray(sys.mass, np.float64)
        self.q = np.ascontiguousarray(sys.q, np.float64)
        self.sig = np.ascontiguousarray(sys.sig, np.float64)
        self.eps = np.ascontiguousarray(sys.eps, np.float64)
        nb, na, nt = len(sys.bond_idx), len(sys.angle_idx), len(sys.torsion_idx)
        self.bonds = (L.Bond * max(1, nb))()
        for t in range(nb):
            i, j = sys.bond_idx[t]
            self.bonds[t] = L.Bond(int(i), int(j), float(sys.bond_par[t, 0]), float(sys.bond_par[t, 1]))
        self.angles = (L.Angle * max(1, na))()
        for t in range(na):
            i, j, k = sys.angle_idx[t]
            self.angles[t] = L.Angle(int(i), int(j), int(k), float(sys.angle_par[t, 0]), float(sys.angle_par[t, 1]))
        self.torsions = (L.Torsion * max(1, nt))()
        for t in range(nt):
            i, j, k, l = sys.torsion_idx[t]
            self.torsions[t] = L.Torsion(int(i), int(j), int(k), int(l), float(sys.torsion_par[t, 0]),
                                         float(sys.torsion_par[t, 1]), int(sys.torsion_per[t]))
        self.top = L.Topology()
        t = self.top
        t.n_atoms = n
        t.mass = _d(self.mass)
        t.n_bonds, t.bonds = nb, self.bonds
        t.n_angles, t.angles = na, self.angles
        t.n_torsions, t.torsions = nt, self.torsions
        t.charge, t.sigma, t.epsilon = _d(self.q), _d(self.sig), _d(self.eps)
        if excl is not None:
            self.eoff = np.ascontiguousarray(excl[0], np.int32)
            self.eidx = np.ascontiguousarray(excl[1], np.int32) if len(excl[1]) else np.zeros(1, np.int32)
            t.excl_offsets = self.eoff.ctypes.data_as(L.I)
            t.excl_indices = self.eidx.ctypes.data_as(L.I)
        g = grid if grid is not None else sys.grid
        t.nb.cutoff = float(sys.cutoff)
        t.nb.ewald_tolerance = float(sys.ewald_tol)
        for k in range(3):
            t.nb.grid[k] = int(g[k])
        t.nb.pme_order = 5
        t.nb.use_switch = int(use_switch)
        t.nb.switch_distance = float(switch_distance)
        t.nb.skin = float(skin)


class Context:
    """One system on one GPU (nx_ctx)."""

    def __init__(self, sys, precision=L.FP64, device=0, skin=0.1, excl=None, use_switch=0, switch_distance=0.0,
                 grid=None, box=None):
        self.h = None
        self.lib = L.lib()
        self.sys = sys
        self.n = sys.n
        h = C.c_void_p()
        self._check(self.lib.nx_create(device, precision, C.byref(h)), "nx_create", ctx=False)
        self.h = h
        self.tb = TopologyBuffers(sys, skin=skin, excl=excl, use_switch=use_switch, switch_distance=switch_distance,
                                  grid=grid)
        self._check(self.lib.nx_set_topology(self.h, C.byref(self.tb.top)), "nx_set_topology")
        lb, ub = (sys.lb, sys.ub) if box is None else box
        self.lb = np.ascontiguousarray(lb, np.float64)
        self.ub = np.ascontiguousarray(ub, np.float64)
        self._check(self.lib.nx_set_box(self.h, _d(self.lb), _d(self.ub)), "nx_set_box")
        self.set_state(sys.pos, sys.vel)

    def _check(self, st, what, ctx=True):
        if st != L.NX_OK:
            msg = self.lib.nx_last_error(self.h).decode() if ctx and getattr(self, "h", None) else ""
            raise L.NexusError(st, "%s failed: %s" % (what, msg))

    def close(self):
        if getattr(self, "h", None):
            self.lib.nx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, pos, vel=None):
        p = np.ascontiguousarray(pos, np.float64)
        v = None if vel is None else np.ascontiguousarray(vel, np.float64)
        self._check(self.lib.nx_set_state(self.h, _d(p), _d(v)), "nx_set_state")

    def get_state(self, forces=False):
        p = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        f = np.zeros((self.n, 3)) if forces else None
        self._check(self.lib.nx_get_state(self.h, _d(p), _d(v), _d(f)), "nx_get_state")
        return (p, v, f) if forces else (p, v)

    def get_list_state(self):
        """(build positions (n, 3), rebuild due) of the device neighbour list."""
        bp = np.zeros((self.n, 3))
        due = C.c_int(0)
        self._check(self.lib.nx_get_list_state(self.h, _d(bp), C.byref(due)), "nx_get_list_state")
        return bp, int(due.value)

    def set_list_state(self, build_pos, rebuild_due):
        bp = np.ascontiguousarray(build_pos, np.float64)
        self._check(self.lib.nx_set_list_state(self.h, _d(bp), int(rebuild_due)), "nx_set_list_state")

    def compute_forces(self, mask=L.ALL):
        f = np.zeros((self.n, 3))
        e = L.Energies()
        self._check(self.lib.nx_compute_forces(self.h, mask, _d(f), C.byref(e)), "nx_compute_forces")
        return f, e.as_dict()

    def step(self, nsteps, dt, mask=L.ALL):
        self._check(self.lib.nx_step(self.h, int(nsteps), float(dt), mask), "nx_step")

    def step_report(self, dt, mask=L.ALL):
        p = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        f = np.zeros((self.n, 3))
        e = L.Energies()
        self._check(self.lib.nx_step_report(self.h, float(dt), mask, _d(p), _d(v), _d(f), C.byref(e)),
                    "nx_step_report")
        return p, v, f, e.as_dict()

    def neighbor_pairs(self):
        cnt = C.c_int64(0)
        cap = 1 << 16
        while True:
            buf = np.zeros(2 * cap, np.int32)
            self._check(self.lib.nx_neighbor_pairs(self.h, buf.ctypes.data_as(L.I), cap, C.byref(cnt)),
                        "nx_neighbor_pairs")
            if cnt.value <= cap:
                return buf[:2 * cnt.value].reshape(-1, 2)
            cap = int(cnt.value)

    def pair_count(self):
        """Number of within-cutoff non-excluded pairs at the current positions."""
        cnt = C.c_int64(0)
        self._check(self.lib.nx_neighbor_pairs(self.h, None, 0, C.byref(cnt)), "nx_neighbor_pairs")
        return int(cnt.value)

    def enable_timing(self, mode=L.TIMING_KERNELS):
        """mode: L.TIMING_OFF, L.TIMING_KERNELS (every kernel family, direct launches) or
        L.TIMING_GRAPH_PAIR (graph-replayed steps with the pair kernel bracketed)."""
        mode = int(mode) if not isinstance(mode, bool) else (L.TIMING_KERNELS if mode else L.TIMING_OFF)
        self._check(self.lib.nx_enable_timing(self.h, mode), "nx_enable_timing")

    def kernel_time(self, name):
        ms = C.c_double(0)
        n = C.c_int64(0)
        self._check(self.lib.nx_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)), "nx_kernel_time")
        return ms.value, n.value

    def set_tuning(self, pair_per_cu=0, spread_block=0, fft_xblocks=0, fft_lines=0):
        t = L.Tuning(int(pair_per_cu), int(spread_block), int(fft_xblocks), int(fft_lines))
        self._check(self.lib.nx_set_tuning(self.h, C.byref(t)), "nx_set_tuning")

    def get_tuning(self):
        t = L.Tuning()
        self._check(self.lib.nx_get_tuning(self.h, C.byref(t)), "nx_get_tuning")
        return t.as_dict()

    def autotune(self, steps_per_trial=50):
        """Times launch-shape candidates on this system; returns (chosen, ms_default, ms_best)."""
        t = L.Tuning()
        a, b = C.c_double(0), C.c_double(0)
        self._check(self.lib.nx_autotune(self.h, int(steps_per_trial), C.byref(t), C.byref(a), C.byref(b)),
                    "nx_autotune")
        return t.as_dict(), a.value, b.value

    def stats(self):
        s = L.Stats()
        self._check(self.lib.nx_get_stats(self.h, C.byref(s)), "nx_get_stats")
        return {k: (list(getattr(s, k)) if k == "grid" else getattr(s, k)) for k, _ in s._fields_}


def fft3d_r2c(a: np.ndarray, precision=L.FP64, device=0) -> np.ndarray:
    lib = L.lib()
    a = np.ascontiguousarray(a, np.float64)
    nx, ny, nz = a.shape
    out = np.zeros((nx, ny, nz // 2 + 1), np.complex128)
    st = lib.nx_fft3d_r2c(device, precision, nx, ny, nz, _d(a), out.ctypes.data_as(L.D))
    if st != L.NX_OK:
        raise L.NexusError(st, "nx_fft3d_r2c failed")
    return out


def fft3d_c2r(c: np.ndarray, nz: int, precision=L.FP64, device=0) -> np.ndarray:
    lib = L.lib()
    c = np.ascontiguousarray(c, np.complex128)
    nx, ny, _ = c.shape
    out = np.zeros((nx, ny, nz))
    st = lib.nx_fft3d_c2r(device, precision, nx, ny, nz, c.ctypes.data_as(L.D), _d(out))
    if st != L.NX_OK:
        raise L.NexusError(st, "nx_fft3d_c2r failed")
    return out


class Server:
    """Continuous-batching server (nx_server)."""

    def __init__(self, precision=L.MIXED, budget_bytes=64 << 30, max_active=256, device=0):
        self.lib = L.lib()
        h = C.c_void_p()
        st = self.lib.nx_server_create(device, precision, int(budget_bytes), int(max_active), C.byref(h))
        if st != L.NX_OK:
            raise L.NexusError(st, "nx_server_create failed")
        self.h = h
        self._keep = {}
        self._n = {}

    def _check(self, st, what):
        if st != L.NX_OK:
            raise L.NexusError(st, "%s failed: %s" % (what, self.lib.nx_server_last_error(self.h).decode()))

    def submit(self, sys, nsteps, dt, skin=0.1, excl=None):
        tb = TopologyBuffers(sys, skin=skin, excl=excl)
        lb = np.ascontiguousarray(sys.lb, np.float64)
        ub = np.ascontiguousarray(sys.ub, np.float64)
        p = np.ascontiguousarray(sys.pos, np.float64)
        v = np.ascontiguousarray(sys.vel, np.float64)
        rid = C.c_int64(0)
        self._check(self.lib.nx_submit(self.h, C.byref(tb.top), _d(lb), _d(ub), _d(p), _d(v), int(nsteps),
                                       float(dt), C.byref(rid)), "nx_submit")
        self._n[rid.value] = sys.n
        return rid.value

    def run(self, steps=0):
        self._check(self.lib.nx_server_run(self.h, int(steps)), "nx_server_run")

    def poll(self, rid):
        st = C.c_int(0)
        done = C.c_int(0)
        self._check(self.lib.nx_poll(self.h, rid, C.byref(st), C.byref(done)), "nx_poll")
        return st.value, done.value

    def fetch(self, rid):
        n = self._n[rid]
        p = np.zeros((n, 3))
        v = np.zeros((n, 3))
        e = L.Energies()
        self._check(self.lib.nx_fetch(self.h, rid, _d(p), _d(v), C.byref(e)), "nx_fetch")
        return p, v, e.as_dict()

    def stats(self):
        s = L.ServerStats()
        self._check(self.lib.nx_server_get_stats(self.h, C.byref(s)), "nx_server_get_stats")
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.lib.nx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
