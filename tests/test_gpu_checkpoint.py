"""Checkpoint and restart: N steps, a checkpoint, a restart and M more steps
end where N + M uninterrupted steps end -- bit for bit in both paths
(fixed-point force sums over the same neighbour list).  The reference has no
restart (v(0) = 0,
Initializer.cpp:21; box from the positions, PeriodicBoundaryCondition.cpp:
11-88), so these are properties of the extension, not parity checks.

Two levels: the C-ABI (positions, velocities and the neighbour-list state
moved between contexts) and TaskDispatcher::Simulate with EnableCheckpoint /
EnableRestart (the state file written and read back, and the report files of
the two halves appended into the bytes of the uninterrupted run).
"""
import os
import sys

import numpy as np
import pytest

from nexusmd_amd import _lib as L
from nexusmd_amd import build as B
from nexusmd_amd import engine as E
from nexusmd_amd import systems as S
from nexusmd_amd import xmlio

from _util import golden, system_from

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def water():
    s = S.water_box(n_mol=1000, L=3.107, grid=(32, 32, 30), seed=12)
    rng = np.random.default_rng(4)
    s.pos = (s.pos + rng.normal(0.0, 0.005, s.pos.shape)) % s.L
    return s


def _same(prec, a, b):
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_restart_equals_continuous_run(water, prec):
    n1, n2 = 137, 163
    print("[gpu] checkpoint C-ABI precision=%d n=%d" % (prec, water.n), flush=True)
    a = E.Context(water, precision=prec, skin=0.1)
    a.step(n1, water.dt)
    pos, vel = a.get_state()
    bp, due = a.get_list_state()
    a.step(n2, water.dt)
    pa, va = a.get_state()
    a.close()
    b = E.Context(water, precision=prec, skin=0.1)
    b.set_state(pos, vel)
    b.set_list_state(bp, due)
    assert np.array_equal(b.get_list_state()[0], bp)
    b.step(n2, water.dt)
    pb, vb = b.get_state()
    b.close()
    _same(prec, pb, pa)
    _same(prec, vb, va)


@pytest.fixture(scope="module")
def nx():
    sys.path.insert(0, os.path.dirname(B.module_path()))
    import NexusMD
    return NexusMD


def _simulate(nx, d, choice, state, steps, interval, ckpt=None, restart=False, velocities=False):
    disp = nx.TaskDispatcher.CreateDispatcher(choice)
    disp.echo_steps = False
    disp.EnableHarmonicBondForce()
    disp.EnableHarmonicAngleForce()
    disp.EnablePeriodicTorsionForce()
    disp.EnableNonbondedForce()
    if ckpt:
        disp.EnableCheckpoint(ckpt, 0)
    if restart:
        disp.EnableRestart()
    if velocities:
        disp.EnableStateVelocities()
    cwd = os.getcwd()
    os.chdir(d)
    try:
        print("[gpu] Simulate %s %s steps=%d restart=%s" % (choice, state, steps, restart), flush=True)
        disp.Simulate("system.xml", state, "input.pdb", "out.pdb", 0.001, steps, interval)
    finally:
        os.chdir(cwd)


def _state_arrays(path):
    pos, vel, bp = [], [], []
    for line in open(path):
        t = line.strip()
        for tag, dst in (("<Position ", pos), ("<Velocity ", vel), ("<BuildPosition ", bp)):
            if t.startswith(tag):
                parts = dict(kv.split("=") for kv in t[len(tag):].rstrip("/>").split())
                dst.append([float(parts[k].strip('"')) for k in ("x", "y", "z")])
    return np.array(pos), np.array(vel), np.array(bp)


@pytest.mark.parametrize("choice", ["GPU", "GPU_MIXED"])
def test_simulate_checkpoint_restart(nx, tmp_path, choice):
    """Simulate 12 steps with a checkpoint at the end, against 7 steps, a
    checkpoint, and a restart for 5 more (reports every 3 steps): the final
    state files are the same text, and the restart appended its reports so
    that the report files equal the uninterrupted run's byte for byte."""
    c0 = golden("chain1500")
    s = system_from(c0, "chain1500")
    full, half = tmp_path / "full", tmp_path / "half"
    for d in (full, half):
        d.mkdir()
        xmlio.write_system_xml(s, str(d / "system.xml"))
        xmlio.write_state_xml(s, str(d / "state.xml"))
        xmlio.write_pdb(s, str(d / "input.pdb"))
    _simulate(nx, full, choice, "state.xml", 12, 3, ckpt="end.xml")
    _simulate(nx, half, choice, "state.xml", 7, 3, ckpt="mid.xml")
    text = open(half / "mid.xml").read()
    assert 'stepCount="7"' in text and "<NexusMDNeighborList" in text and "<Velocities>" in text
    _simulate(nx, half, choice, "mid.xml", 5, 3, ckpt="end.xml", restart=True)
    pf, vf, bf = _state_arrays(full / "end.xml")
    ph, vh, bh = _state_arrays(half / "end.xml")
    assert pf.shape == (s.n, 3) and bf.shape == (s.n, 3)
    assert open(full / "end.xml").read() == open(half / "end.xml").read()
    for name in ("TotalEnergyout.pdb", "PVFReportout.pdb"):
        assert open(full / name, "rb").read() == open(half / name, "rb").read(), name
    import re
    strip = lambda t: re.sub(rb"Nexaus 1\.0\.0, \d{4}-\d{2}-\d{2}", b"", t)
    assert strip(open(full / "out.pdb", "rb").read()) == strip(open(half / "out.pdb", "rb").read())
    rows = [l.split()[0] for l in open(half / "TotalEnergyout.pdb").read().splitlines()[1:]]
    assert rows == ["0", "2", "5", "8", "11"]


def test_state_velocities_switch(nx, tmp_path):
    """EnableStateVelocities takes v(0) from the state file: one step from a
    state with velocities moves the atoms by about v dt, where the reference's
    default (v(0) = 0) moves them by F/m dt^2 only."""
    c0 = golden("chain1500")
    s = system_from(c0, "chain1500")
    rng = np.random.default_rng(3)
    s.vel = rng.normal(0.0, 0.5, s.pos.shape)
    xmlio.write_system_xml(s, str(tmp_path / "system.xml"))
    xmlio.write_state_xml(s, str(tmp_path / "state.xml"), velocities=True)
    xmlio.write_pdb(s, str(tmp_path / "input.pdb"))
    _simulate(nx, tmp_path, "GPU", "state.xml", 1, 1, ckpt="with_v.xml", velocities=True)
    _simulate(nx, tmp_path, "GPU", "state.xml", 1, 1, ckpt="zero_v.xml")
    p1, v1, _ = _state_arrays(tmp_path / "with_v.xml")
    p0, v0, _ = _state_arrays(tmp_path / "zero_v.xml")
    box = 1.02 * (s.pos.max(0) - s.pos.min(0))      # the reference's box rule (bounding box +- 1%)
    d = p1 - p0
    d -= np.round(d / box) * box
    np.testing.assert_allclose(d, s.vel * 0.001, rtol=0, atol=1e-9)
    np.testing.assert_allclose(v1 - v0, s.vel, rtol=0, atol=1e-9)
