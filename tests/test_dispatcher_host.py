"""Host side of the dispatcher (no GPU): the XML readers and the report writers
of the `NexusMD` module against the reference's own parsers and reporters.

Byte formats are pinned two ways: digests of the reference's files for fixed
inputs (tests/golden/reports.json, written by tests/golden/make_golden.py from
the reference's compiled Reporter) and, where oracle/_ref is built, a direct
file-for-file comparison.  The reference's own TotalEnergy sample file
(test/TotalEnergyoutput_SimRun_Sample_Protein.pdb of the reference) must be
reproduced byte for byte from the numbers it holds.
"""
import hashlib
import json
import os
import sys
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from nexusmd_amd import build as B
from nexusmd_amd import systems as S
from nexusmd_amd import xmlio
from oracle import pyref as R

from _util import GOLDEN, golden, gunzip_to, normalize_remark_date, report_inputs, system_from

MODULE = B.module_path()
pytestmark = pytest.mark.skipif(not os.path.exists(MODULE), reason="NexusMD module not built")
needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (reference sources absent)")


@pytest.fixture(scope="module")
def nx():
    sys.path.insert(0, os.path.dirname(MODULE))
    import NexusMD
    return NexusMD


@pytest.fixture(scope="module")
def sample(tmp_path_factory):
    d = tmp_path_factory.mktemp("sample")
    pdb = gunzip_to(os.path.join(GOLDEN, "Sample_Protein.pdb.gz"), str(d / "Sample_Protein.pdb"))
    st = gunzip_to(os.path.join(GOLDEN, "state_Sample_Protein.xml.gz"), str(d / "state_Sample_Protein.xml"))
    return pdb, st


@pytest.fixture(scope="module")
def c0_files(tmp_path_factory):
    d = tmp_path_factory.mktemp("c0")
    s = system_from(golden("chain1500"), "chain1500")
    xmlio.write_system_xml(s, str(d / "system.xml"))
    xmlio.write_state_xml(s, str(d / "state.xml"))
    xmlio.write_pdb(s, str(d / "input.pdb"))
    return s, str(d / "system.xml"), str(d / "state.xml"), str(d / "input.pdb")


# ------------------------------------------------------------------ parsers

def test_state_parser_sample_protein(nx, sample):
    _, st = sample
    got = np.array(nx.StateXMLParser.parseStateXML(st))
    ref = np.array([[float(p.get(k)) for k in "xyz"] for p in ET.parse(st).getroot().find("Positions")])
    assert got.shape == (23558, 3)
    np.testing.assert_array_equal(got, ref)


def test_system_parser_round_trip(nx, c0_files):
    s, sysxml, _, _ = c0_files
    np.testing.assert_array_equal(nx.SystemXMLParser.MassesParser(sysxml), s.mass)
    b = nx.SystemXMLParser.HBondParser(sysxml)
    np.testing.assert_array_equal([[x.p1, x.p2] for x in b], s.bond_idx)
    np.testing.assert_array_equal([[x.d, x.k] for x in b], s.bond_par)
    a = nx.SystemXMLParser.HAngleParser(sysxml)
    np.testing.assert_array_equal([[x.p1, x.p2, x.p3] for x in a], s.angle_idx)
    np.testing.assert_array_equal([[x.a, x.k] for x in a], s.angle_par)
    t = nx.SystemXMLParser.PTorsionParser(sysxml)
    np.testing.assert_array_equal([[x.p1, x.p2, x.p3, x.p4] for x in t], s.torsion_idx)
    np.testing.assert_array_equal([[x.k, x.phase] for x in t], s.torsion_par)
    np.testing.assert_array_equal([x.periodicity for x in t], s.torsion_per)
    nb = nx.SystemXMLParser.NonBondedParser(sysxml)
    assert nb.cutoff == s.cutoff and nb.ewaldTolerance == s.ewald_tol and nb.useSwitchingFunction == 0
    np.testing.assert_array_equal([[p.q, p.sig, p.eps] for p in nb.particles], np.stack([s.q, s.sig, s.eps], 1))


TRICKY = """<?xml version="1.0" encoding="UTF-8"?>
<!-- a comment with <Particle mass="99"/> inside -->
<!DOCTYPE System>
<System type='System'>
  <Particles>
    <Particle mass="1.5"/>
    <Particle mass=" 2.25 "></Particle>
    <Particle/>
  </Particles>
  <Particles><Particle mass="7"/></Particles>
  <Forces>
    <Force type="HarmonicBondForce"><Bonds><Bond p1="0" p2="1" d="0.1" k="1e3"/></Bonds>
      <Bonds><Bond p1="1" p2="2" d="9" k="9"/></Bonds></Force>
    <Force type="NonbondedForce" cutoff="0.9" ewaldTolerance="5e-4" useSwitchingFunction="1" switchingDistance="0.8">
      <Particles><Particle q="-1" sig="0.3" eps="0.5"/><Particle q="1" sig="0.2" eps="0"/><Particle q="0" sig="0.1" eps="0"/></Particles>
      <Exceptions><Exception p1="0" p2="1" q="0" sig="1" eps="0"/></Exceptions>
    </Force>
    <Force type="HarmonicBondForce"><Bonds><Bond p1="1" p2="2" d="0.2" k="2e3"/></Bonds></Force>
    <Force type="NonbondedForce" cutoff="5"><Particles><Particle q="9"/></Particles></Force>
    <Force type="PeriodicTorsionForce"><Torsions><Torsion p1="0" p2="1" p3="2" p4="0" k="&#49;.5" phase="3.1" periodicity="2"/></Torsions></Force>
  </Forces>
</System>
"""


def test_system_parser_selection_rules(nx, tmp_path):
    """First <Particles>, every matching <Force> but only its first section, first NonbondedForce only."""
    p = tmp_path / "tricky.xml"
    p.write_text(TRICKY)
    f = str(p)
    assert nx.SystemXMLParser.MassesParser(f) == [1.5, 2.25, 0.0]
    b = nx.SystemXMLParser.HBondParser(f)
    assert [(x.p1, x.p2, x.d, x.k) for x in b] == [(0, 1, 0.1, 1000.0), (1, 2, 0.2, 2000.0)]
    nb = nx.SystemXMLParser.NonBondedParser(f)
    assert nb.cutoff == 0.9 and nb.useSwitchingFunction == 1 and nb.switchingDistance == 0.8
    assert [(q.q, q.sig, q.eps) for q in nb.particles] == [(-1.0, 0.3, 0.5), (1.0, 0.2, 0.0), (0.0, 0.1, 0.0)]
    assert len(nb.exceptions) == 1
    t = nx.SystemXMLParser.PTorsionParser(f)
    assert [(x.k, x.periodicity, x.phase) for x in t] == [(1.5, 2, 3.1)]


@needs_ref
def test_system_parser_counts_match_reference(nx, tmp_path, c0_files):
    p = tmp_path / "tricky.xml"
    p.write_text(TRICKY)
    _, sysxml, stxml, _ = c0_files
    for f, st in ((str(p), None), (sysxml, stxml)):
        ref = R.parse_counts(f, st or f)
        got = (len(nx.SystemXMLParser.MassesParser(f)), len(nx.SystemXMLParser.HBondParser(f)),
               len(nx.SystemXMLParser.HAngleParser(f)), len(nx.SystemXMLParser.PTorsionParser(f)),
               len(nx.SystemXMLParser.NonBondedParser(f).particles))
        assert got == ref[:5]
        if st:
            assert len(nx.StateXMLParser.parseStateXML(st)) == ref[5]


def test_missing_file_reports_and_returns_empty(nx, tmp_path, capfd):
    assert nx.SystemXMLParser.MassesParser(str(tmp_path / "absent.xml")) == []
    assert "parsed with errors" in capfd.readouterr().err


def test_malformed_xml_is_an_error(nx, tmp_path, capfd):
    p = tmp_path / "bad.xml"
    p.write_text("<System><Particles><Particle mass='1'></Particles></System>")
    assert nx.SystemXMLParser.MassesParser(str(p)) == []
    assert "mismatched closing tag" in capfd.readouterr().err


# ----------------------------------------------------------------- reporters

def _digests():
    with open(os.path.join(GOLDEN, "reports.json")) as f:
        return json.load(f)


def _write_reports(nx, directory):
    p, v, f, rows = report_inputs()
    cwd = os.getcwd()
    os.chdir(directory)
    try:
        rep = nx.Reporter()
        for step, ke, pe in rows:
            rep.TotalEnergyReport("rep.pdb", ke, pe, ke + pe, step)
            rep.TestPVFReport("rep.pdb", (p * (1 + step)).tolist(), v.tolist(), f.tolist(), step)
    finally:
        os.chdir(cwd)


def test_total_energy_and_pvf_bytes(nx, tmp_path):
    _write_reports(nx, tmp_path)
    d = _digests()
    for name in ("TotalEnergyrep.pdb", "PVFReportrep.pdb"):
        data = (tmp_path / name).read_bytes()
        assert len(data) == d[name + "_bytes"], name
        assert hashlib.sha256(data).hexdigest() == d[name + "_sha256"], name


@needs_ref
def test_total_energy_and_pvf_bytes_live(nx, tmp_path):
    ours, theirs = tmp_path / "ours", tmp_path / "ref"
    ours.mkdir()
    theirs.mkdir()
    _write_reports(nx, ours)
    p, v, f, rows = report_inputs()
    cwd = os.getcwd()
    os.chdir(theirs)
    try:
        for step, ke, pe in rows:
            R.report("rep.pdb", p * (1 + step), v, f, ke, pe, step)
    finally:
        os.chdir(cwd)
    for name in ("TotalEnergyrep.pdb", "PVFReportrep.pdb"):
        assert (ours / name).read_bytes() == (theirs / name).read_bytes(), name


def test_reference_total_energy_sample_reproduced(nx, tmp_path):
    """Re-print the numbers of the reference's own TotalEnergy sample: same bytes.
    The sample was written on Windows (text-mode CRLF line ends); the reference
    built on Linux writes LF, as this build does, so line ends are compared as LF."""
    src = open(os.path.join(GOLDEN, "TotalEnergyoutput_SimRun_Sample_Protein.pdb"), "rb").read()
    src = src.replace(b"\r\n", b"\n")
    rows = [line.split() for line in src.decode().splitlines()[1:]]
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        rep = nx.Reporter()
        for step, ke, pe, te in rows:
            rep.TotalEnergyReport("out.pdb", float(ke), float(pe), float(te), int(step))
    finally:
        os.chdir(cwd)
    assert (tmp_path / "TotalEnergyout.pdb").read_bytes() == src


def test_sample_protein_pdb_bytes(nx, sample, tmp_path):
    pdb, st = sample
    pos = np.array(nx.StateXMLParser.parseStateXML(st))
    out = str(tmp_path / "out.pdb")
    rep = nx.Reporter()
    rep.pdbReport(pdb, out, pos.tolist(), 1, True)
    rep.pdbReport(pdb, out, (pos * 1.01 - 0.05).tolist(), 2, False)
    data = normalize_remark_date(open(out, "rb").read())
    d = _digests()
    assert len(data) == d["sample_protein_pdb_bytes"]
    assert hashlib.sha256(data).hexdigest() == d["sample_protein_pdb_sha256"]


def test_dispatcher_factory(nx, capfd):
    assert nx.TaskDispatcher.CreateDispatcher("GPU") is not None
    assert nx.TaskDispatcher.CreateDispatcher("GPU_MIXED") is not None
    d = nx.TaskDispatcher.CreateDispatcher("CPU")
    assert "Invalid processor type" in capfd.readouterr().err
    d.EnableHarmonicBondForce()
    d.EnableNonbondedForce()


def test_simulate_rejects_bad_arguments_before_touching_a_device(nx, c0_files):
    s, sysxml, stxml, pdb = c0_files
    d = nx.TaskDispatcher.CreateDispatcher("GPU")
    d.echo_steps = False
    with pytest.raises(ValueError):
        d.Simulate(sysxml, stxml, pdb, "out.pdb", 0.0, 10, 1)
    with pytest.raises(ValueError):
        d.Simulate(sysxml, stxml, pdb, "out.pdb", 0.001, 10, 0)


def test_coords3d_has_the_reference_operations(nx):
    """Coords3D as the SWIG wrapper exports it (wrapper/NexusMD.i:18;
    core/common/headers/Coords3D.h:15-151): components by index, + - * /,
    dot, cross, length, normalize (zero stays zero)."""
    a, b = nx.Coords3D(1.0, 2.0, 3.0), nx.Coords3D(-2.0, 0.5, 4.0)
    assert (a[0], a[1], a[2], a[-1]) == (1.0, 2.0, 3.0, 3.0)
    assert list(a + b) == [-1.0, 2.5, 7.0] and list(a - b) == [3.0, 1.5, -1.0]
    assert list(-a) == [-1.0, -2.0, -3.0] and list(a * 2.0) == [2.0, 4.0, 6.0]
    assert list(nx.Coords3D(3.0, 6.0, 9.0) / 3.0) == [1.0, 2.0, 3.0]
    assert a.dot(b) == 1.0 * -2.0 + 2.0 * 0.5 + 3.0 * 4.0
    assert list(a.cross(b)) == list(np.cross([1.0, 2.0, 3.0], [-2.0, 0.5, 4.0]))
    assert a.length() == pytest.approx(np.sqrt(14.0))
    assert list(nx.Coords3D(0.0, 3.0, 4.0).normalize()) == [0.0, 0.6, 0.8]
    assert list(nx.Coords3D().normalize()) == [0.0, 0.0, 0.0]
    c = nx.Coords3D(1.0, 1.0, 1.0)
    c += a
    c *= 2.0
    assert c == nx.Coords3D(4.0, 6.0, 8.0) and c != a
    c[1] = -1.0
    assert c[1] == -1.0
    with pytest.raises(IndexError):
        a[3]


def test_state_file_round_trip(nx, tmp_path):
    """writeState then parseState gives back every double exactly (17
    significant digits), the OpenMM sections and the restart extension."""
    rng = np.random.default_rng(5)
    n = 40
    st = nx.StateData()
    st.positions = rng.normal(size=(n, 3)).tolist()
    st.velocities = (rng.normal(size=(n, 3)) * 1e-7).tolist()
    st.buildPositions = (rng.normal(size=(n, 3)) + 1.0 / 3.0).tolist()
    st.hasVelocities = st.hasBoxVectors = st.hasBoxBounds = st.hasListState = True
    st.boxA, st.boxB, st.boxC = [2.1, 0.0, 0.0], [0.0, 3.0000000000000004, 0.0], [0.0, 0.0, np.pi]
    st.lb, st.ub = [-0.1, 0.0, 1e-17], [2.0, 3.0000000000000004, np.pi + 1e-17]
    st.rebuildDue = 1
    st.stepCount = 123456789012
    st.time = 61728.394506
    path = str(tmp_path / "ck.xml")
    nx.StateXMLParser.writeState(path, st)
    r = nx.StateXMLParser.parseState(path)
    for k in ("positions", "velocities", "buildPositions"):
        assert np.array_equal(np.array(getattr(r, k)), np.array(getattr(st, k))), k
    for k in ("boxA", "boxB", "boxC", "lb", "ub"):
        assert list(getattr(r, k)) == list(getattr(st, k)), k
    assert (r.hasVelocities, r.hasBoxVectors, r.hasBoxBounds, r.hasListState) == (True, True, True, True)
    assert (r.rebuildDue, r.stepCount, r.time) == (1, 123456789012, 61728.394506)
    # the reference's readers see an ordinary OpenMM state
    assert np.array_equal(np.array(nx.StateXMLParser.parseStateXML(path)), np.array(st.positions))
    assert list(nx.StateXMLParser.extractBoxSize(path)) == [2.1, 3.0000000000000004, np.pi]
    root = ET.parse(path).getroot()
    assert root.tag == "State" and root.find("Positions") is not None and root.find("Velocities") is not None


def test_openmm_state_with_velocities(nx, tmp_path):
    """A plain OpenMM state (no extension): positions, velocities and box vectors read,
    no box bounds and no list state."""
    s = S.bead_chain(n=50, L=2.5, grid=(24, 24, 24), seed=4)
    path = str(tmp_path / "s.xml")
    xmlio.write_state_xml(s, path, velocities=True)
    r = nx.StateXMLParser.parseState(path)
    assert np.array_equal(np.array(r.positions), s.pos) and np.array_equal(np.array(r.velocities), s.vel)
    assert r.hasVelocities and r.hasBoxVectors and not r.hasBoxBounds and not r.hasListState
    assert list(r.boxA) == [2.5, 0.0, 0.0] and r.stepCount == 0


def test_pvf_number_format_on_extreme_values(nx, tmp_path):
    """TestPVFReport's numbers are the stream's fixed notation, 10 decimals,
    width 10 (Reporter.cpp:45-54): checked against printf-style "%10.10f" on
    values across magnitudes, signed zeros, ties in the last decimal and
    non-finite values, in a system large enough to be formatted in parallel
    blocks."""
    rng = np.random.default_rng(11)
    n = 20000
    vals = rng.normal(size=(n, 9)) * 10.0 ** rng.integers(-14, 16, size=(n, 9))
    special = [0.0, -0.0, 0.5e-10, 1.5e-10, 2.5e-10, -2.5e-10, 1e300, -1e-300, 123456789.00000000005,
               float("inf"), float("-inf"), float("nan"), 5e-324, 0.12345678905, 9.99999999995]
    vals.ravel()[: len(special)] = special
    p, v, f = vals[:, 0:3].tolist(), vals[:, 3:6].tolist(), vals[:, 6:9].tolist()
    rep = nx.Reporter()
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        rep.TestPVFReport("x.pdb", p, v, f, 0)
    finally:
        os.chdir(cwd)
    lines = open(tmp_path / "PVFReportx.pdb").read().split("\n")
    assert lines[0] == "Step0" and lines[-1] == "" and lines[-2] == ""

    def fx(x):
        return "%10.10f" % x

    for i in range(n):
        r = vals[i]
        want = "%d\t(%s, %s, %s)\t(%s, %s, %s)\t(%s, %s, %s)" % ((i + 1,) + tuple(fx(x) for x in r))
        assert lines[2 + i] == want, (i, lines[2 + i], want)
