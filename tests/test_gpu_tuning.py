"""Launch shapes and the autotuner (nx_set_tuning, nx_autotune).

Every shape the autotuner chooses among leaves what each kernel computes for
an atom, a pair or a grid line unchanged: the pair kernel does every list
chunk whole on one wavefront however many wavefronts there are, the spread
accumulates fixed-point integers whatever the tile, and each FFT line is
transformed alone whatever the grouping.  So every candidate gives the same
bits as the built-in shapes in the mixed path, and tuning does not change a
trajectory.
"""
import numpy as np
import pytest

from nexusmd_amd import _lib as L
from nexusmd_amd import engine as E
from nexusmd_amd import systems as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def water():
    s = S.water_box(n_mol=2000, L=3.914, grid=(40, 40, 40), seed=14)
    rng = np.random.default_rng(6)
    s.pos = (s.pos + rng.normal(0.0, 0.004, s.pos.shape)) % s.L
    return s


def _run(sys_, tuning, steps=60, fft_plane=None, monkeypatch=None):
    ctx = E.Context(sys_, precision=L.MIXED, skin=0.1)
    if tuning:
        ctx.set_tuning(**tuning)
    ctx.step(steps, sys_.dt)
    f, e = ctx.compute_forces()
    p, v = ctx.get_state()
    got = ctx.get_tuning()
    ctx.close()
    return p, v, f, got


CANDIDATES = [dict(pair_per_cu=c) for c in (1, 2, 4, 5)] + [dict(spread_block=b) for b in (1, 3, 6)] + \
             [dict(fft_xblocks=x) for x in (64, 2048)] + [dict(pair_per_cu=1, spread_block=2, fft_xblocks=128)]


def test_every_candidate_gives_the_default_bits(water):
    p0, v0, f0, t0 = _run(water, None)
    assert t0 == dict(pair_per_cu=0, spread_block=0, fft_xblocks=0, fft_lines=0)
    for cand in CANDIDATES:
        print("[gpu] tuning candidate %s" % cand, flush=True)
        p, v, f, t = _run(water, cand)
        assert all(t[k] == cand[k] for k in cand)
        np.testing.assert_array_equal(p, p0)
        np.testing.assert_array_equal(v, v0)
        np.testing.assert_array_equal(f, f0)


def test_line_fft_candidates_give_the_default_bits(water, monkeypatch):
    monkeypatch.setenv("NEXUSMD_FFT_PLANE", "0")
    p0, v0, f0, _ = _run(water, None)
    for lines in (1, 2, 4):
        print("[gpu] line FFT, %d lines per block" % lines, flush=True)
        p, v, f, _ = _run(water, dict(fft_lines=lines))
        np.testing.assert_array_equal(p, p0)
        np.testing.assert_array_equal(f, f0)


def test_autotune_keeps_the_trajectory_and_caches_its_choice(water, tmp_path, monkeypatch):
    """Autotuning in the middle of a run: the run continues exactly as without
    it; the choice is reported, never slower than the built-in shapes, and a
    later context of the same system starts from it."""
    a = E.Context(water, precision=L.MIXED, skin=0.1)
    b = E.Context(water, precision=L.MIXED, skin=0.1)
    a.step(30, water.dt)
    b.step(30, water.dt)
    chosen, ms_default, ms_best = b.autotune(steps_per_trial=20)
    print("[gpu] autotune: %s, %.4f -> %.4f ms/step" % (chosen, ms_default, ms_best), flush=True)
    assert ms_best <= ms_default and ms_best > 0
    assert b.get_tuning() == chosen
    a.step(40, water.dt)
    b.step(40, water.dt)
    pa, va = a.get_state()
    pb, vb = b.get_state()
    np.testing.assert_array_equal(pb, pa)
    np.testing.assert_array_equal(vb, va)
    c = E.Context(water, precision=L.MIXED, skin=0.1)
    assert c.get_tuning() == chosen
    for x in (a, b, c):
        x.close()
