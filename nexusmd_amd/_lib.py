"""ctypes binding of include/nexusmd.h (libnexusmd_hip.so, built in-tree).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnexusmd_hip.so")

NX_OK, NX_ERR_ARG, NX_ERR_HIP, NX_ERR_CAPACITY, NX_ERR_STATE, NX_ERR_NOMEM, NX_ERR_NUMERIC = range(7)
FP64, MIXED = 0, 1
TIMING_OFF, TIMING_KERNELS, TIMING_GRAPH_PAIR = 0, 1, 2
BOND, ANGLE, TORSION, NONBONDED, ALL = 1, 2, 4, 8, 15

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
I64 = C.POINTER(C.c_int64)


class Bond(C.Structure):
    _fields_ = [("p1", C.c_int), ("p2", C.c_int), ("d", C.c_double), ("k", C.c_double)]


class Angle(C.Structure):
    _fields_ = [("p1", C.c_int), ("p2", C.c_int), ("p3", C.c_int), ("a", C.c_double), ("k", C.c_double)]


class Torsion(C.Structure):
    _fields_ = [("p1", C.c_int), ("p2", C.c_int), ("p3", C.c_int), ("p4", C.c_int),
                ("k", C.c_double), ("phase", C.c_double), ("periodicity", C.c_int)]


class NonbondedOpts(C.Structure):
    _fields_ = [("cutoff", C.c_double), ("ewald_tolerance", C.c_double), ("grid", C.c_int * 3),
                ("pme_order", C.c_int), ("use_switch", C.c_int), ("switch_distance", C.c_double),
                ("skin", C.c_double)]


class Topology(C.Structure):
    _fields_ = [("n_atoms", C.c_int), ("mass", D),
                ("n_bonds", C.c_int), ("bonds", C.POINTER(Bond)),
                ("n_angles", C.c_int), ("angles", C.POINTER(Angle)),
                ("n_torsions", C.c_int), ("torsions", C.POINTER(Torsion)),
                ("charge", D), ("sigma", D), ("epsilon", D),
                ("excl_offsets", I), ("excl_indices", I),
                ("nb", NonbondedOpts)]


class Energies(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("bond", "angle", "torsion", "coulomb_direct", "lj", "recip", "self",
                                         "exclusion", "potential", "reference_pe", "kinetic")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [("n_slots", C.c_int64), ("n_clusters", C.c_int64), ("n_super", C.c_int64),
                ("list_entries", C.c_int64), ("tiles", C.c_int64), ("rebuilds", C.c_int64), ("steps", C.c_int64),
                ("grid", C.c_int * 3), ("alpha", C.c_double), ("steps_without_list", C.c_int64),
                ("steps_repeated", C.c_int64)]


class Tuning(C.Structure):
    _fields_ = [("pair_per_cu", C.c_int), ("spread_block", C.c_int), ("fft_xblocks", C.c_int), ("fft_lines", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("admitted", C.c_int64), ("finished", C.c_int64), ("failed", C.c_int64),
                ("active", C.c_int64), ("p50_step_ms", C.c_double), ("p99_step_ms", C.c_double),
                ("aggregate_ns_per_day", C.c_double), ("wall_s", C.c_double), ("batched_steps", C.c_int64),
                ("bytes_in_use", C.c_int64), ("bytes_budget", C.c_int64)]


# (name, restype, argtypes) for every entry point of include/nexusmd.h
SIGNATURES = [
    ("nx_create", C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("nx_destroy", C.c_int, [C.c_void_p]),
    ("nx_last_error", C.c_char_p, [C.c_void_p]),
    ("nx_status_string", C.c_char_p, [C.c_int]),
    ("nx_set_topology", C.c_int, [C.c_void_p, C.POINTER(Topology)]),
    ("nx_set_box", C.c_int, [C.c_void_p, D, D]),
    ("nx_box_from_positions", C.c_int, [C.c_int, D, D, D]),
    ("nx_set_state", C.c_int, [C.c_void_p, D, D]),
    ("nx_get_state", C.c_int, [C.c_void_p, D, D, D]),
    ("nx_get_list_state", C.c_int, [C.c_void_p, D, I]),
    ("nx_set_list_state", C.c_int, [C.c_void_p, D, C.c_int]),
    ("nx_compute_forces", C.c_int, [C.c_void_p, C.c_int, D, C.POINTER(Energies)]),
    ("nx_step", C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int]),
    ("nx_step_report", C.c_int, [C.c_void_p, C.c_double, C.c_int, D, D, D, C.POINTER(Energies)]),
    ("nx_step_report_async", C.c_int, [C.c_void_p, C.c_double, C.c_int, D, D, D, I64]),
    ("nx_report_wait", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Energies)]),
    ("nx_host_alloc", C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    ("nx_host_free", C.c_int, [C.c_void_p]),
    ("nx_neighbor_pairs", C.c_int, [C.c_void_p, I, C.c_int64, I64]),
    ("nx_enable_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("nx_kernel_time", C.c_int, [C.c_void_p, C.c_char_p, D, I64]),
    ("nx_get_stats", C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    ("nx_set_tuning", C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    ("nx_get_tuning", C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    ("nx_autotune", C.c_int, [C.c_void_p, C.c_int, C.POINTER(Tuning), D, D]),
    ("nx_fft3d_r2c", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, D, D]),
    ("nx_fft3d_c2r", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, D, D]),
    ("nx_fft3d_r2c_noalloc", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, D, D]),
    ("nx_server_create", C.c_int, [C.c_int, C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_void_p)]),
    ("nx_server_destroy", C.c_int, [C.c_void_p]),
    ("nx_server_last_error", C.c_char_p, [C.c_void_p]),
    ("nx_submit", C.c_int, [C.c_void_p, C.POINTER(Topology), D, D, D, D, C.c_int, C.c_double, I64]),
    ("nx_server_run", C.c_int, [C.c_void_p, C.c_int]),
    ("nx_poll", C.c_int, [C.c_void_p, C.c_int64, I, I]),
    ("nx_fetch", C.c_int, [C.c_void_p, C.c_int64, D, D, C.POINTER(Energies)]),
    ("nx_server_get_stats", C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
]

_lib = None


def lib():
    """The loaded library; raises if it is not built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libnexusmd_hip.so is not built: run `python -m nexusmd_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class NexusError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (status %d)" % (msg, status))
        self.status = status
