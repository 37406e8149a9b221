"""Writes the golden fixtures of tests/golden/ from the reference's own code.

The reference's force, neighbour-list, PME, integrator and kinetic-energy
functions are compiled from its sources into oracle/_ref/libnxref.so
(`make -C oracle ref`, where the reference tree exists) and called through
oracle/pyref.py.  FFTW, the reference's third-party transform, is absent, so
its place between spread and convolution is taken by numpy's unnormalised 3D
DFT (oracle/pyref.py, recip()).

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickled objects):
  chain1500.npz  the 1,500-atom bead chain (4.0 nm box, PME grid 40^3): inputs,
                 box rule, exclusions, sorted pair set, alpha, grid rule,
                 B-spline moduli and theta/dtheta of 16 atoms, spread grid of
                 12 atoms, per-term forces and energies, reciprocal forces at
                 the grid rule's 64^3, 10 leapfrog steps (dt 1 fs),
                 per-term forces and energies at the end of those steps, and
                 the report rows and step-9 state of the same system run as
                 the dispatcher runs it (box and grid rules, v(0) = 0)
  water3000.npz  1,000 flexible waters (3.107 nm box, grid 32x32x30):
                 per-term forces and energies, pair count and pair-set digest
  reports.json   sha256 digests of the reference's PDB / TotalEnergy / PVF
                 files for fixed inputs (the byte formats of its Reporter)
  switch.npz     the LJ switch: single pairs across (0, rc) and a 48-atom
                 system whose switch values carry over between pairs
  edge.npz       tiny cases: a pair at exactly r = rc, near-collinear
                 torsions, a near-linear angle, atoms outside [lb, ub] at the
                 leapfrog wrap, and a box shorter than 2 rc
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from nexusmd_amd import systems as S  # noqa: E402
from oracle import pyref as R  # noqa: E402


def pair_digest(pairs: np.ndarray) -> str:
    """sha256 of the pair set as sorted (first, second) int32 rows."""
    p = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    order = np.lexsort((p[:, 1], p[:, 0]))
    return hashlib.sha256(np.ascontiguousarray(p[order]).tobytes()).hexdigest()


def sorted_pairs(pairs):
    p = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    return np.ascontiguousarray(p[np.lexsort((p[:, 1], p[:, 0]))])


def inputs(sys_):
    return dict(pos=sys_.pos, vel=sys_.vel, mass=sys_.mass, q=sys_.q, sig=sys_.sig, eps=sys_.eps,
                bond_idx=np.asarray(sys_.bond_idx, np.int32).reshape(-1, 2),
                bond_par=np.asarray(sys_.bond_par, np.float64).reshape(-1, 2),
                angle_idx=np.asarray(sys_.angle_idx, np.int32).reshape(-1, 3),
                angle_par=np.asarray(sys_.angle_par, np.float64).reshape(-1, 2),
                torsion_idx=np.asarray(sys_.torsion_idx, np.int32).reshape(-1, 4),
                torsion_par=np.asarray(sys_.torsion_par, np.float64).reshape(-1, 2),
                torsion_per=np.asarray(sys_.torsion_per, np.int32),
                lb=np.asarray(sys_.lb, np.float64), ub=np.asarray(sys_.ub, np.float64),
                cutoff=np.float64(sys_.cutoff), ewald_tol=np.float64(sys_.ewald_tol),
                grid=np.asarray(sys_.grid, np.int32))


def term_outputs(sys_, excl, prefix=""):
    out = {}
    if len(sys_.bond_idx):
        out["f_bond"], out["e_bond"] = R.bonds(sys_)
    if len(sys_.angle_idx):
        out["f_angle"], out["e_angle"] = R.angles(sys_)
    if len(sys_.torsion_idx):
        out["f_torsion"], out["e_torsion"] = R.torsions(sys_)
    out["f_recip"], out["e_recip"] = R.recip(sys_.pos, sys_.q, sys_.lb, sys_.ub, sys_.grid, sys_.cutoff,
                                             sys_.ewald_tol)
    (out["f_real"], out["e_self"], out["e_direct"],
     out["e_exclusion"]) = R.nonbonded_real(sys_.pos, sys_.q, sys_.sig, sys_.eps, sys_.lb, sys_.ub, excl,
                                            sys_.cutoff, sys_.ewald_tol)
    out["f_total"], e = R.forces(sys_, excl=excl)
    out["e_reference_pe"] = e["reference_pe"]
    return {prefix + k: np.asarray(v) for k, v in out.items()}


def chain1500():
    s = S.bead_chain()          # 1,500 beads, 4.0 nm, grid 40^3, seed 20261202
    excl = R.exclusions(s.n, s.bond_idx)
    lb, ub, L = R.box(s.pos)
    out = inputs(s)
    out.update(box_lb=lb, box_ub=ub, box_L=L, excl_off=excl[0], excl_idx=excl[1])
    p = R.pairs(s.pos, s.lb, s.ub, excl, s.cutoff)
    out["pairs"] = sorted_pairs(p)
    out["alpha"] = np.float64(R.alpha(s.cutoff, s.ewald_tol))
    out["grid_rule"] = R.grid_rule(lb, ub)
    mods, th, dth, idx, frac = R.bspline(s.pos, s.lb, s.ub, s.grid)
    out.update(moduli_x=mods[0], theta=th[:, :16], dtheta=dth[:, :16], pindex=idx[:16], pfrac=frac[:16])
    mods64, _, _, _, _ = R.bspline(s.pos[:1], s.lb, s.ub, (64, 64, 64))
    out["moduli_64"] = mods64[0]
    out["spread12_grid16"] = R.spread(s.pos[:12], s.q[:12], s.lb, s.ub, (16, 16, 16)).real.copy()
    out.update(term_outputs(s, excl))
    out["f_recip_64"], out["e_recip_64"] = R.recip(s.pos, s.q, s.lb, s.ub, (64, 64, 64), s.cutoff, s.ewald_tol)
    pos, vel, f, rows = R.simulate(s, 10, 0.001, interval=1)
    out.update(sim_pos=pos, sim_vel=vel, sim_f=f, sim_rows=rows, sim_dt=np.float64(0.001))
    # per-term forces at the end of the 10 steps, where no bond sits exactly at
    # its rest length (at the generated geometry bond forces are rounding noise)
    end = s.copy()
    end.pos = pos
    out.update(term_outputs(end, excl, prefix="end_"))
    # what TaskDispatcher::Simulate reports for this system read from XML: box by
    # the positions rule, grid by the grid rule, zero initial velocities
    # (Engine.cpp:53-69, Initializer.cpp:21), all forces, dt 1 fs, interval 1
    d = s.copy()
    d.lb, d.ub = lb, ub
    d.grid = tuple(int(g) for g in R.grid_rule(lb, ub))
    d.vel = np.zeros_like(s.vel)
    _, _, _, rows = R.simulate(d, 10, 0.001, interval=1)
    p9, v9, _, _ = R.simulate(d, 9, 0.001, interval=1)
    f9, _ = R.forces(d, p9, excl=excl)
    out.update(dispatch_rows=rows, dispatch_pos9=p9, dispatch_vel9=v9, dispatch_f9=f9)
    return out


def water3000():
    s = S.water_box(n_mol=1000, L=3.107, grid=(32, 32, 30), seed=12)
    rng = np.random.default_rng(2)
    s.pos = (s.pos + rng.normal(0.0, 0.005, s.pos.shape)) % s.L
    excl = R.exclusions(s.n, s.bond_idx)
    out = inputs(s)
    out.update(excl_off=excl[0], excl_idx=excl[1])
    p = R.pairs(s.pos, s.lb, s.ub, excl, s.cutoff)
    out["pair_count"] = np.int64(len(p))
    out["pair_digest"] = np.array(pair_digest(p))
    out.update(term_outputs(s, excl))
    return out


def edge():
    out = {}
    lb, ub = np.zeros(3), np.full(3, 3.0)
    # a pair exactly at the cutoff (strict r^2 < rc^2 drops it) and one just inside
    pos = np.array([[0.5, 0.5, 0.5], [1.5, 0.5, 0.5], [0.5, 1.4999999, 0.5]])
    none = (np.zeros(4, np.int32), np.zeros(0, np.int32))
    out["rc_pos"] = pos
    out["rc_pairs"] = sorted_pairs(R.pairs(pos, lb, ub, none, 1.0))
    # near-collinear torsions and a near-linear angle (the asin branch, |cos| > 0.99)
    tors_pos = np.array([[1.0, 1.0, 1.0], [1.1, 1.0, 1.0], [1.2, 1.0005, 1.0], [1.3, 1.0009, 1.0002],
                         [2.0, 2.0, 2.0], [2.1, 2.0, 2.0], [2.2, 2.1, 2.0], [2.3, 2.1, 2.1]])
    out["tors_pos"] = tors_pos
    out["tors_idx"] = np.array([[0, 1, 2, 3], [3, 2, 1, 0], [4, 5, 6, 7]], np.int32)
    out["tors_par"] = np.array([[3.0, 0.0], [2.0, np.pi], [5.0, 0.3]])
    out["tors_per"] = np.array([1, 2, 3], np.int32)
    out["ang_idx"] = np.array([[0, 1, 2], [4, 5, 6], [1, 2, 3]], np.int32)
    out["ang_par"] = np.array([[1.91, 400.0], [1.824, 836.8], [3.14, 100.0]])
    sysl = _tiny(tors_pos, lb, ub, out)
    out["tors_f"], out["tors_e"] = R.torsions(sysl)
    out["ang_f"], out["ang_e"] = R.angles(sysl)
    # leapfrog across both faces of the box and at the faces
    lp = np.array([[0.001, 1.0, 2.999], [2.9995, 0.0004, 1.5], [1.5, 1.5, 1.5], [0.0, 3.0, 0.5]])
    lv = np.array([[-2.0, 0.0, 1.5], [1.5, -1.0, 0.0], [0.0, 0.0, 0.0], [-0.1, 0.1, 0.0]])
    lf = np.array([[10.0, -5.0, 0.0], [0.0, 0.0, 1.0], [3.0, 2.0, 1.0], [0.0, 0.0, 0.0]])
    lm = np.array([1.008, 15.999, 12.011, 1.008])
    out.update(lf_pos=lp, lf_vel=lv, lf_f=lf, lf_mass=lm)
    out["lf_pos_out"], out["lf_vel_out"] = R.leapfrog(lp, lv, lf, lm, 0.002, lb, ub)
    out["lf_ke"] = np.float64(R.kinetic(lv, lm, lf, 0.002))
    # box shorter than 2 rc: 1.5 nm box, cutoff 1.0
    rng = np.random.default_rng(99)
    small = rng.uniform(0.0, 1.5, (60, 3))
    out["small_pos"] = small
    out["small_pairs"] = sorted_pairs(R.pairs(small, np.zeros(3), np.full(3, 1.5), (np.zeros(61, np.int32),
                                                                                      np.zeros(0, np.int32)), 1.0))
    # box rule on a skewed cloud
    cloud = rng.normal(0.0, 1.0, (200, 3)) * np.array([1.0, 2.0, 0.5]) + np.array([-3.0, 4.0, 0.1])
    out["cloud"] = cloud
    out["cloud_lb"], out["cloud_ub"], out["cloud_L"] = R.box(cloud)
    out["cloud_grid_rule"] = R.grid_rule(out["cloud_lb"], out["cloud_ub"])
    return out


def switch():
    """The LJ switch of DirectForcesAndEnergy (NonbondedForce.cpp:530-585) from
    the reference itself, switch distance 0.6 nm, cutoff 1.0 nm:
      single pairs at distances across (0, rc): with one pair per system the
      switch values carried over from the previous pair play no part, so these
      pin the polynomial itself;
      a 48-atom system of charged LJ particles in the reference's pair order:
      its switch values carry over from pair to pair (switchValue and
      switchDeriv are assigned only when r > switchingDistance), which the
      oracle reproduces with switch_quirk = 1."""
    out = {}
    rsw, rc = 0.6, 1.0
    lb, ub = np.zeros(3), np.full(3, 5.0)
    none2 = (np.zeros(3, np.int32), np.zeros(0, np.int32))
    rs = np.array([0.35, 0.5, 0.6, 0.61, 0.7, 0.8, 0.9, 0.97, 0.999])
    q = np.array([0.3, -0.45])
    sig = np.array([0.3, 0.34])
    eps = np.array([0.5, 0.8])
    f_all, e_all = [], []
    for r in rs:
        pos = np.array([[2.0, 2.0, 2.0], [2.0 + r * 0.6, 2.0 + r * 0.8, 2.0]])
        f, es, ed, ex = R.nonbonded_real(pos, q, sig, eps, lb, ub, none2, rc, 5e-4, use_switch=1, rsw=rsw)
        f_all.append(f)
        e_all.append([es, ed, ex])
    out.update(pair_r=rs, pair_q=q, pair_sig=sig, pair_eps=eps, pair_f=np.array(f_all), pair_e=np.array(e_all),
               rsw=np.float64(rsw), rc=np.float64(rc))
    # the same pairs uncharged: the direct forces are the switched LJ alone (no
    # reciprocal part to separate from them)
    f0_all, e0_all = [], []
    for r in rs:
        pos = np.array([[2.0, 2.0, 2.0], [2.0 + r * 0.6, 2.0 + r * 0.8, 2.0]])
        f, es, ed, ex = R.nonbonded_real(pos, np.zeros(2), sig, eps, lb, ub, none2, rc, 5e-4, use_switch=1, rsw=rsw)
        f0_all.append(f)
        e0_all.append([es, ed, ex])
    out.update(pair0_f=np.array(f0_all), pair0_e=np.array(e0_all))
    rng = np.random.default_rng(77)
    n = 48
    g = np.array([[x, y, z] for x in range(4) for y in range(4) for z in range(3)], float)[:n] * 0.62 + 0.2
    pos = g + rng.normal(0.0, 0.05, g.shape)
    qm = rng.normal(0.0, 0.2, n)
    qm -= qm.mean()
    sm = rng.uniform(0.28, 0.34, n)
    em = rng.uniform(0.2, 1.0, n)
    lbm, ubm = np.zeros(3), np.full(3, 2.6)
    nonem = (np.zeros(n + 1, np.int32), np.zeros(0, np.int32))
    f, es, ed, ex = R.nonbonded_real(pos, qm, sm, em, lbm, ubm, nonem, rc, 5e-4, use_switch=1, rsw=rsw)
    out.update(multi_pos=pos, multi_q=qm, multi_sig=sm, multi_eps=em, multi_lb=lbm, multi_ub=ubm, multi_f=f,
               multi_e=np.array([es, ed, ex]))
    return out


class _Tiny:
    pass


def _tiny(pos, lb, ub, d):
    t = _Tiny()
    t.pos, t.lb, t.ub = pos, lb, ub
    t.torsion_idx, t.torsion_par, t.torsion_per = d["tors_idx"], d["tors_par"], d["tors_per"]
    t.angle_idx, t.angle_par = d["ang_idx"], d["ang_par"]
    return t


def reports():
    """sha256 digests of the reference's report files (Reporter.cpp) for fixed inputs:
    the Sample_Protein PDB template with the positions of its state XML (models 1
    and 2, the REMARK date replaced by a token), and the TotalEnergy / PVF files
    of tests/_util.report_inputs() over steps 0, 1, 2."""
    import json
    import tempfile
    import xml.etree.ElementTree as ET

    sys.path.insert(0, os.path.dirname(HERE))
    from _util import gunzip_to, normalize_remark_date, report_inputs
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        pdb = gunzip_to(os.path.join(HERE, "Sample_Protein.pdb.gz"), os.path.join(tmp, "in.pdb"))
        st = gunzip_to(os.path.join(HERE, "state_Sample_Protein.xml.gz"), os.path.join(tmp, "st.xml"))
        pos = np.array([[float(p.get(k)) for k in "xyz"] for p in ET.parse(st).getroot().find("Positions")])
        outp = os.path.join(tmp, "out.pdb")
        R.pdb_report(pdb, outp, pos, 1, True)
        R.pdb_report(pdb, outp, pos * 1.01 - 0.05, 2, False)
        data = normalize_remark_date(open(outp, "rb").read())
        out["sample_protein_pdb_sha256"] = hashlib.sha256(data).hexdigest()
        out["sample_protein_pdb_bytes"] = len(data)
        p, v, f, rows = report_inputs()
        cwd = os.getcwd()
        os.chdir(tmp)
        try:
            for step, ke, pe in rows:
                R.report("rep.pdb", p * (1 + step), v, f, ke, pe, step)
        finally:
            os.chdir(cwd)
        for name in ("TotalEnergyrep.pdb", "PVFReportrep.pdb"):
            data = open(os.path.join(tmp, name), "rb").read()
            out[name + "_sha256"] = hashlib.sha256(data).hexdigest()
            out[name + "_bytes"] = len(data)
    path = os.path.join(HERE, "reports.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print("wrote %s" % path)


def main():
    if not R.available():
        raise SystemExit("oracle/_ref/libnxref.so is not built: run `make -C oracle ref` where the reference exists")
    only = sys.argv[1:]        # fixture names to write (default: all)
    for name, fn in (("chain1500", chain1500), ("water3000", water3000), ("edge", edge), ("switch", switch)):
        if only and name not in only:
            continue
        data = {k: np.asarray(v) for k, v in fn().items()}
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **data)
        print("wrote %s (%d arrays, %.0f KB)" % (path, len(data), os.path.getsize(path) / 1024))
    if not only or "reports" in only:
        reports()


if __name__ == "__main__":
    main()
