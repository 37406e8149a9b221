"""Writers for the OpenMM-style input files the dispatcher reads.

write_system_xml  <System>: <Particles mass>, HarmonicBondForce, HarmonicAngleForce,
                  PeriodicTorsionForce and NonbondedForce (PME) sections, the
                  schema SystemXMLParser reads (core/common/src/SystemXMLParser.cpp:18-196)
write_state_xml   <State><Positions> (StateXMLParser.cpp:17-37), nm
write_pdb         ATOM records used as the report template (Reporter.cpp:99-133), Angstrom

Numbers are written with Python's shortest round-trip repr, so a reader that
parses them with strtod recovers the exact doubles.
"""
from __future__ import annotations

import numpy as np


def _f(x) -> str:
    return repr(float(x))


def write_system_xml(sys_, path: str, use_switch: int = 0, switch_distance: float = 0.0) -> None:
    L = sys_.L
    out = ['<?xml version="1.0" ?>',
           '<System openmmVersion="8.1.1" type="System" version="1">',
           "\t<PeriodicBoxVectors>",
           '\t\t<A x="%s" y="0" z="0"/>' % _f(L[0]),
           '\t\t<B x="0" y="%s" z="0"/>' % _f(L[1]),
           '\t\t<C x="0" y="0" z="%s"/>' % _f(L[2]),
           "\t</PeriodicBoxVectors>",
           "\t<Particles>"]
    out += ['\t\t<Particle mass="%s"/>' % _f(m) for m in sys_.mass]
    out += ["\t</Particles>", "\t<Constraints/>", "\t<Forces>"]
    out.append('\t\t<Force forceGroup="0" name="HarmonicBondForce" type="HarmonicBondForce" usesPeriodic="0" '
               'version="2">')
    out.append("\t\t\t<Bonds>")
    out += ['\t\t\t\t<Bond d="%s" k="%s" p1="%d" p2="%d"/>' % (_f(p[0]), _f(p[1]), ij[0], ij[1])
            for ij, p in zip(sys_.bond_idx, sys_.bond_par)]
    out += ["\t\t\t</Bonds>", "\t\t</Force>"]
    out.append('\t\t<Force forceGroup="0" name="HarmonicAngleForce" type="HarmonicAngleForce" usesPeriodic="0" '
               'version="2">')
    out.append("\t\t\t<Angles>")
    out += ['\t\t\t\t<Angle a="%s" k="%s" p1="%d" p2="%d" p3="%d"/>' % (_f(p[0]), _f(p[1]), ij[0], ij[1], ij[2])
            for ij, p in zip(sys_.angle_idx, sys_.angle_par)]
    out += ["\t\t\t</Angles>", "\t\t</Force>"]
    out.append('\t\t<Force forceGroup="0" name="PeriodicTorsionForce" type="PeriodicTorsionForce" '
               'usesPeriodic="0" version="2">')
    out.append("\t\t\t<Torsions>")
    out += ['\t\t\t\t<Torsion k="%s" p1="%d" p2="%d" p3="%d" p4="%d" periodicity="%d" phase="%s"/>'
            % (_f(p[0]), ij[0], ij[1], ij[2], ij[3], int(n), _f(p[1]))
            for ij, p, n in zip(sys_.torsion_idx, sys_.torsion_par, sys_.torsion_per)]
    out += ["\t\t\t</Torsions>", "\t\t</Force>"]
    g = sys_.grid
    out.append('\t\t<Force alpha="0" cutoff="%s" dispersionCorrection="0" ewaldTolerance="%s" '
               'exceptionsUsePeriodic="0" forceGroup="0" includeDirectSpace="1" ljAlpha="0" ljnx="0" ljny="0" '
               'ljnz="0" method="4" name="NonbondedForce" nx="%d" ny="%d" nz="%d" recipForceGroup="-1" '
               'rfDielectric="78.3" switchingDistance="%s" type="NonbondedForce" useSwitchingFunction="%d" '
               'version="4">' % (_f(sys_.cutoff), _f(sys_.ewald_tol), g[0], g[1], g[2], _f(switch_distance),
                                 int(use_switch)))
    out += ["\t\t\t<GlobalParameters/>", "\t\t\t<ParticleOffsets/>", "\t\t\t<ExceptionOffsets/>", "\t\t\t<Particles>"]
    out += ['\t\t\t\t<Particle eps="%s" q="%s" sig="%s"/>' % (_f(e), _f(q), _f(s))
            for q, s, e in zip(sys_.q, sys_.sig, sys_.eps)]
    out += ["\t\t\t</Particles>", "\t\t\t<Exceptions/>", "\t\t</Force>", "\t</Forces>", "</System>", ""]
    with open(path, "w") as f:
        f.write("\n".join(out))


def write_state_xml(sys_, path: str, pos=None, velocities=False) -> None:
    """<State> with <PeriodicBoxVectors> and <Positions>; with velocities=True
    also <Velocities> (OpenMM's layout), which the dispatcher reads only when
    EnableStateVelocities is on."""
    pos = np.asarray(sys_.pos if pos is None else pos, np.float64)
    L = sys_.L
    out = ['<?xml version="1.0" ?>',
           '<State openmmVersion="8.1.1" stepCount="0" time="0" type="State" version="1">',
           "\t<PeriodicBoxVectors>",
           '\t\t<A x="%s" y="0" z="0"/>' % _f(L[0]),
           '\t\t<B x="0" y="%s" z="0"/>' % _f(L[1]),
           '\t\t<C x="0" y="0" z="%s"/>' % _f(L[2]),
           "\t</PeriodicBoxVectors>",
           "\t<Positions>"]
    out += ['\t\t<Position x="%s" y="%s" z="%s"/>' % (_f(p[0]), _f(p[1]), _f(p[2])) for p in pos]
    out += ["\t</Positions>"]
    if velocities:
        out += ["\t<Velocities>"]
        out += ['\t\t<Velocity x="%s" y="%s" z="%s"/>' % (_f(v[0]), _f(v[1]), _f(v[2]))
                for v in np.asarray(sys_.vel, np.float64)]
        out += ["\t</Velocities>"]
    out += ["</State>", ""]
    with open(path, "w") as f:
        f.write("\n".join(out))


def write_pdb(sys_, path: str, residue="LIG", water_residue="HOH") -> None:
    """ATOM records (Angstrom): chain beads as C in residue `residue`, 3-site waters as O/H1/H2.
    Atom numbers run 1..N (the reports index positions by them); fields widen
    beyond the PDB columns for very large systems, which the whitespace-split
    reader accepts."""
    pos = np.asarray(sys_.pos, np.float64) * 10.0
    n = sys_.n
    is_water = np.zeros(n, bool)
    o = np.nonzero(np.isclose(sys_.q, -0.834))[0]
    for i in o:
        if i + 2 < n:
            is_water[i:i + 3] = True
    lines = ["REMARK    written by nexusmd_amd.xmlio"]
    res = 0
    for i in range(n):
        if is_water[i]:
            k = (i - o[np.searchsorted(o, i, side="right") - 1])
            name = ("O", "H1", "H2")[k]
            if k == 0:
                res += 1
            rn = water_residue
        else:
            name, rn = "C", residue
            res = 1 if res == 0 else res
        lines.append("ATOM  %5d %-4s %3s  %4d    %8.3f%8.3f%8.3f%6.2f%6.2f          %2s"
                     % (i + 1, name, rn, res, pos[i, 0], pos[i, 1], pos[i, 2], 1.0, 0.0,
                        name[0]))
    lines += ["END", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))
