"""GPU parity: the HIP engine (through the C-ABI) against the CPU oracle.

Tolerances (relative RMS force error sqrt(sum|F-F_ref|^2 / sum|F_ref|^2)):
  FP64 path  <= 1e-9 per term (the stated drop-in bound is 1e-5)
  mixed path <= 5e-4 per term and in total
The pair set of the FP64 test is compared bit-exactly with the brute-force
mathematical set and with the reference's voxel algorithm (oracle).
"""
import numpy as np
import pytest

from nexusmd_amd import _lib as L
from nexusmd_amd import engine as E
from nexusmd_amd import systems as S
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
MIXED_TOL = 5e-4


def rel_rms(a, b):
    den = float(np.sum(b * b))
    return float(np.sqrt(np.sum((a - b) ** 2) / den)) if den > 0 else float(np.sqrt(np.sum(a * a)))


def _jiggle(sys, seed, sd=0.005):
    """Displace atoms slightly so that no bonded term sits exactly at its minimum."""
    rng = np.random.default_rng(seed)
    sys.pos = (sys.pos + rng.normal(0.0, sd, sys.pos.shape)) % sys.L
    return sys


@pytest.fixture(scope="module")
def chain():
    return _jiggle(S.bead_chain(n=400, L=2.6, grid=(24, 24, 24), seed=11), 1)


@pytest.fixture(scope="module")
def water():
    return _jiggle(S.water_box(n_mol=1000, L=3.107, grid=(32, 32, 30), seed=12), 2)


@pytest.fixture(scope="module")
def systems(chain, water):
    return {"chain": chain, "water": water}


def _ctx(sys, prec, **kw):
    print("[gpu] context %s n=%d precision=%d %s" % (sys.name, sys.n, prec, kw), flush=True)
    return E.Context(sys, precision=prec, **kw)


@pytest.mark.parametrize("name", ["chain", "water"])
@pytest.mark.parametrize("skin", [0.0, 0.1])
def test_pair_set_bit_exact(systems, name, skin):
    sys = systems[name]
    ctx = _ctx(sys, L.FP64, skin=skin)
    got = ctx.neighbor_pairs()
    ex = O.exclusions(sys)
    brute = O.pairs(sys, ex, method="brute")
    voxel = O.pairs(sys, ex, method="voxel")
    g = set(map(tuple, got.tolist()))
    b = set(map(tuple, brute.tolist()))
    v = set(map(tuple, voxel.tolist()))
    assert len(g) == len(got), "duplicate pairs from the device list"
    assert g == b
    assert g == v
    ctx.close()


TERMS = [("bond", L.BOND), ("angle", L.ANGLE), ("torsion", L.TORSION), ("nonbonded", L.NONBONDED), ("all", L.ALL)]


@pytest.mark.parametrize("name", ["chain", "water"])
@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
@pytest.mark.parametrize("term,mask", TERMS)
def test_forces_vs_oracle(systems, name, prec, term, mask):
    sys = systems[name]
    if term == "torsion" and len(sys.torsion_idx) == 0:
        pytest.skip("no torsions")
    ctx = _ctx(sys, prec, skin=0.1)
    f, e = ctx.compute_forces(mask)
    prep = O.Prepared(sys)
    fr, er = O.compute_forces(prep, mask=mask)
    tol = FP64_TOL if prec == L.FP64 else MIXED_TOL
    err = rel_rms(f, fr)
    print("[gpu] %s %s prec=%d force err %.3e" % (name, term, prec, err), flush=True)
    assert err <= tol
    # energies
    pairs = [("bond", "bond"), ("angle", "angle"), ("torsion", "torsion"), ("coulomb_direct", "coulomb"),
             ("lj", "lj"), ("recip", "recip"), ("self", "self"), ("exclusion", "exclusion"),
             ("potential", "potential"), ("reference_pe", "reference_pe")]
    etol = 1e-9 if prec == L.FP64 else 2e-5
    for a, b in pairs:
        scale = max(1.0, abs(er[b]))
        assert abs(e[a] - er[b]) <= etol * scale + (1e-6 if prec == L.FP64 else 1e-3), (a, e[a], er[b])
    ctx.close()


@pytest.mark.parametrize("passes", ["plane", "lines"])
@pytest.mark.parametrize("shape", [(24, 24, 24), (32, 30, 20), (40, 36, 48), (64, 64, 64), (96, 96, 96)])
@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_fft_vs_numpy(shape, prec, passes, monkeypatch):
    """The plane passes (whole planes in LDS) and the five line passes the engine
    falls back to where a plane does not fit, against numpy's FFT."""
    monkeypatch.setenv("NEXUSMD_FFT_PLANE", "1" if passes == "plane" else "0")
    rng = np.random.default_rng(5)
    a = rng.normal(size=shape)
    got = E.fft3d_r2c(a, precision=prec)
    ref = np.fft.rfftn(a)
    tol = 1e-12 if prec == L.FP64 else 2e-6
    assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref))
    back = E.fft3d_c2r(ref, shape[2], precision=prec)
    refb = np.fft.irfftn(ref, s=shape) * a.size
    assert np.max(np.abs(back - refb)) <= tol * np.max(np.abs(refb))


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_forces_with_line_fft_passes(systems, prec, monkeypatch):
    """The engine on the five line FFT passes (its path where a plane does not
    fit in LDS) gives the oracle's forces and reciprocal energy."""
    monkeypatch.setenv("NEXUSMD_FFT_PLANE", "0")
    sys = systems["water"]
    ctx = _ctx(sys, prec, skin=0.1)
    f, e = ctx.compute_forces(L.ALL)
    fr, er = O.compute_forces(O.Prepared(sys), mask=L.ALL)
    assert rel_rms(f, fr) <= (FP64_TOL if prec == L.FP64 else MIXED_TOL)
    etol = 1e-9 if prec == L.FP64 else 2e-5
    assert abs(e["recip"] - er["recip"]) <= etol * max(1.0, abs(er["recip"])) + (1e-6 if prec == L.FP64 else 1e-3)
    ctx.close()


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_trajectory_vs_oracle(chain, prec):
    """10 leapfrog steps with reports, as Engine::RunSimulation with interval 1."""
    sys = chain
    ctx = _ctx(sys, prec, skin=0.0)
    prep = O.Prepared(sys)
    dt = 0.001
    rows = []
    for _ in range(10):
        p, v, f, e = ctx.step_report(dt)
        rows.append((e["kinetic"], e["reference_pe"]))
    pos, vel = ctx.get_state()
    opos, ovel, of, oe = O.simulate(prep, 10, dt, interval=1)
    tol = 1e-9 if prec == L.FP64 else 2e-4
    dpos = np.abs(pos - opos)
    dpos = np.minimum(dpos, sys.L - dpos)
    assert float(dpos.max()) <= tol * 10
    ke = np.array([r[0] for r in rows])
    pe = np.array([r[1] for r in rows])
    assert np.allclose(ke, oe[:, 0], rtol=tol * 10, atol=1e-6)
    assert np.allclose(pe, oe[:, 1], rtol=tol * 10, atol=1e-4)
    ctx.close()


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_repeated_evaluations_agree(prec):
    """The pair kernel hands its chunks out at run time, and the forward FFT
    clears the spread grid for the next call.  Evaluating the same positions
    again, in the same context and in a second one, gives the same forces and
    energies bit for bit in both paths (every sum that crosses wavefronts is
    fixed point: 2^32 in the mixed path, two 64-bit parts in the FP64 path)."""
    sys = _jiggle(S.water_box(n_mol=8000, L=6.2, grid=(64, 64, 64), seed=21), 3)
    a = _ctx(sys, prec, skin=0.1)
    b = _ctx(sys, prec, skin=0.1)
    f1, e1 = a.compute_forces(L.ALL)
    f2, e2 = a.compute_forces(L.ALL)
    f3, e3 = b.compute_forces(L.ALL)
    for f, e in ((f2, e2), (f3, e3)):
        assert np.array_equal(f, f1)
        assert all(e[k] == e1[k] for k in e1)
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_graph_steps_match_plain(chain, prec):
    """Graph-replayed steps equal step-by-step launches, bit for bit."""
    a = _ctx(chain, prec, skin=0.1)
    b = _ctx(chain, prec, skin=0.1)
    a.step(20, 0.001)               # graph path
    for _ in range(20):
        b.step(1, 0.001)            # plain launches
    pa, va = a.get_state()
    pb, vb = b.get_state()
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_switch_forces_vs_oracle(prec):
    """The LJ switch from 0.6 nm to the 1.0 nm cutoff, on weakly charged LJ
    particles on a jittered 0.55 nm lattice (the switch changes their forces
    by more than 1e-2, far more than the tolerance; the charges keep the
    erfc-Coulomb arithmetic of the mixed path in play beside the switch):
    nonbonded forces and the LJ and Coulomb energies as the oracle's, which
    follows NonbondedForce.cpp:530-585 without the carry-over of the previous
    pair's switch values."""
    from _util import bare_system
    rng = np.random.default_rng(41)
    a, m = 0.55, 7
    g = np.array([[x, y, z] for x in range(m) for y in range(m) for z in range(m)], float) * a + 0.1
    box = a * m
    q = rng.normal(0.0, 0.05, len(g))
    q -= q.mean()
    sysq = bare_system(g + rng.normal(0, 0.04, g.shape), [0, 0, 0], [box, box, box], q=q,
                       sig=rng.uniform(0.28, 0.34, len(g)), eps=rng.uniform(0.2, 1.0, len(g)), grid=(32, 32, 32))
    assert abs(O.compute_forces(O.Prepared(sysq), mask=O.NONBONDED)[1]["coulomb"]) > 1.0
    rsw = 0.6
    ctx = _ctx(sysq, prec, skin=0.1, use_switch=1, switch_distance=rsw)
    f, e = ctx.compute_forces(L.NONBONDED)
    fr, er = O.compute_forces(O.Prepared(sysq, use_switch=1, switch_distance=rsw), mask=O.NONBONDED)
    f0, _ = O.compute_forces(O.Prepared(sysq), mask=O.NONBONDED)
    assert rel_rms(fr, f0) > 1e-2      # the switch changes the forces
    err = rel_rms(f, fr)
    print("[gpu] switch prec=%d force err %.3e" % (prec, err), flush=True)
    assert err <= (FP64_TOL if prec == L.FP64 else MIXED_TOL)
    etol = 1e-9 if prec == L.FP64 else 2e-5
    for a, b in (("lj", "lj"), ("coulomb_direct", "coulomb")):
        assert abs(e[a] - er[b]) <= etol * max(1.0, abs(er[b])) + (1e-6 if prec == L.FP64 else 1e-3), (a, e[a], er[b])
    ctx.close()


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_timed_graph_steps_match_untimed(chain, prec):
    """Steps replayed with the pair kernel bracketed by events (the pair-kernel
    sampling mode of the benchmark: every step split into graphs around the
    pair kernel) end in the state of plain graph replays, bit for bit; the
    step and pair events cover every step once."""
    a = _ctx(chain, prec, skin=0.1)
    b = _ctx(chain, prec, skin=0.1)
    a.step(5, 0.001)
    b.step(5, 0.001)
    b.enable_timing(L.TIMING_GRAPH_PAIR)
    for nsteps in (64, 6):     # the counts cover the last call
        a.step(nsteps, 0.001)
        b.step(nsteps, 0.001)
        ms, n = b.kernel_time("step")
        pms, pn = b.kernel_time("pair")
        assert n == nsteps and ms > 0
        assert pn == nsteps and pms > 0
    b.enable_timing(L.TIMING_OFF)
    pa, va = a.get_state()
    pb, vb = b.get_state()
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_exclusions_outside_the_pair_list(prec):
    """Bonded partners farther apart than the list radius: their exclusion
    correction cannot come from list tiles, so the system falls back to the
    per-pair exclusion kernel.  Forces and energies as the oracle's."""
    rng = np.random.default_rng(21)
    n, L_ = 120, 4.0
    pos = rng.uniform(0, L_, (n, 3))
    q = rng.normal(0, 0.5, n)
    q -= q.mean()
    from _util import bare_system
    s = bare_system(pos, [0, 0, 0], [L_, L_, L_], q=q, sig=rng.uniform(0.25, 0.35, n),
                    eps=rng.uniform(0.1, 1.0, n), grid=(32, 32, 32))
    # spread atoms apart so LJ stays moderate, then bond a few far-apart pairs
    g = np.array([[x, y, z] for x in range(5) for y in range(5) for z in range(5)])[:n] * 0.8 + 0.1
    s.pos = g + rng.normal(0, 0.02, g.shape)
    far = []
    for i in range(0, 40, 2):
        j = (i + 37) % n
        d = s.pos[i] - s.pos[j]
        d -= L_ * np.round(d / L_)
        if np.linalg.norm(d) > 1.3:
            far.append((i, j))
    assert len(far) >= 5
    s.bond_idx = np.array(far, np.int32)
    s.bond_par = np.tile([0.15, 1000.0], (len(far), 1))
    ctx = _ctx(s, prec, skin=0.1)
    f, e = ctx.compute_forces(L.NONBONDED)
    fr, er = O.compute_forces(O.Prepared(s), mask=O.NONBONDED)
    tol = FP64_TOL if prec == L.FP64 else MIXED_TOL
    assert rel_rms(f, fr) <= tol
    assert abs(e["exclusion"] - er["exclusion"]) <= (1e-9 if prec == L.FP64 else 1e-5) * abs(er["exclusion"])
    ctx.close()


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_mid_run_overflow_freezes_the_context(prec):
    """A single context whose list build overflows mid-run returns
    NX_ERR_CAPACITY; its state is that of the last completed step (finite,
    the cloud moving, every atom inside the anchor's column), and a further
    call, which lays the context out again from that state, reports the same
    capacity error."""
    from _util import collapse_system
    s, excl = collapse_system()
    ctx = _ctx(s, prec, skin=0.1, excl=excl)
    with pytest.raises(L.NexusError) as ex:
        ctx.step(80, 0.0005)
    assert ex.value.status == L.NX_ERR_CAPACITY
    pos, vel = ctx.get_state()
    assert np.all(np.isfinite(pos)) and np.all(np.isfinite(vel))
    assert float(np.abs(vel[1:]).max()) > 1.0
    d = pos[1:, :2] - pos[0, :2]
    assert float(np.abs(d - 3.0 * np.round(d / 3.0)).max()) < 0.25    # column width 0.5 nm around the anchor
    with pytest.raises(L.NexusError) as ex2:
        ctx.step(1, 0.0005)
    assert ex2.value.status == L.NX_ERR_CAPACITY
    ctx.close()


def test_switch_distance_is_validated():
    """use_switch with a negative or non-finite switch distance is rejected (the
    device would otherwise leave the switch off without a word)."""
    from _util import bare_system
    s = bare_system(np.random.default_rng(3).uniform(0, 3.0, (50, 3)), [0, 0, 0], [3.0, 3.0, 3.0])
    for sd in (-0.1, float("nan"), float("inf")):
        with pytest.raises(L.NexusError) as ex:
            E.Context(s, precision=L.MIXED, use_switch=1, switch_distance=sd)
        assert ex.value.status == L.NX_ERR_ARG
    E.Context(s, precision=L.MIXED, use_switch=1, switch_distance=0.0).close()


def test_pair_images_match_per_pair(monkeypatch):
    """The mixed path's per-entry periodic images against the nearest image
    taken pair by pair (NEXUSMD_PAIR_IMAGE=0), on a 4.2 nm water box where
    most entries carry an image and the rest do not, after 300 steps so that
    atoms have crossed the box faces since the list builds: the two differ
    only by FP32 rounding of the displacements, and both match the oracle."""
    print("[gpu] pair images vs per pair", flush=True)
    s = _jiggle(S.water_box(n_mol=2500, L=4.217, grid=(40, 40, 40), seed=21), 3)
    ctx = _ctx(s, L.MIXED, skin=0.1)
    ctx.step(300, s.dt)
    pos, _ = ctx.get_state()
    f_img, e_img = ctx.compute_forces(L.ALL)
    rebuilds = ctx.stats()["rebuilds"]
    ctx.close()
    s2 = s.copy()
    s2.pos = pos
    monkeypatch.setenv("NEXUSMD_PAIR_IMAGE", "0")
    ctx = _ctx(s2, L.MIXED, skin=0.1)
    f_pp, e_pp = ctx.compute_forces(L.ALL)
    ctx.close()
    fr, er = O.compute_forces(O.Prepared(s2))
    d = rel_rms(f_img, f_pp)
    print("[gpu] images vs per pair: rel RMS %.3e after %d builds; vs oracle %.3e / %.3e"
          % (d, rebuilds, rel_rms(f_img, fr), rel_rms(f_pp, fr)), flush=True)
    assert rebuilds >= 2
    # FP32 rounding of the displacements, amplified by the steep LJ core (a
    # wrong image would differ by O(1))
    assert d <= 1e-4
    assert rel_rms(f_img, fr) <= MIXED_TOL and rel_rms(f_pp, fr) <= MIXED_TOL
    assert abs(e_img["potential"] - e_pp["potential"]) <= 1e-5 * abs(er["potential"])
