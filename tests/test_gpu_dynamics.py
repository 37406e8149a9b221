"""GPU dynamics: long leapfrog runs of both precision paths.

The mixed path sums forces in 64-bit fixed point from FP32 pair arithmetic.
Over thousands of steps its total energy must drift and fluctuate like the
FP64 path's: the accumulation adds no drift of its own.  The trajectories
themselves separate (the dynamics are chaotic), so the test compares the
statistics of the total energy, not the states.
"""
import numpy as np
import pytest

from nexusmd_amd import _lib as L
from nexusmd_amd import engine as E
from nexusmd_amd import systems as S

pytestmark = pytest.mark.gpu


def _energy_series(sys, prec, dt=0.0005, settle=200, samples=41, every=100):
    print("[gpu] dynamics n=%d precision=%d" % (sys.n, prec), flush=True)
    ctx = E.Context(sys, precision=prec, skin=0.1)
    ctx.step(settle, dt)
    tot = []
    for _ in range(samples):
        _, _, _, e = ctx.step_report(dt)
        tot.append(e["kinetic"] + e["potential"])
        ctx.step(every - 1, dt)
    ctx.close()
    tot = np.array(tot)
    slope = np.polyfit(np.arange(samples) * every, tot, 1)[0]
    return tot, slope


def test_mixed_energy_drift_matches_fp64():
    """3,000-atom flexible water, dt 0.5 fs, 4,000 steps after 200 to settle:
    total-energy drift per step and fluctuation of the mixed path within 30% /
    50% of the FP64 path's, and the FP64 drift below 1% over the run."""
    rng = np.random.default_rng(2)
    sys = S.water_box(n_mol=1000, L=3.107, grid=(32, 32, 30), seed=12)
    sys.pos = (sys.pos + rng.normal(0.0, 0.005, sys.pos.shape)) % sys.L
    tf, sf = _energy_series(sys, L.FP64)
    tm, sm = _energy_series(sys, L.MIXED)
    print("[gpu] drift per step fp64 %.3e mixed %.3e; std fp64 %.2f mixed %.2f" % (sf, sm, tf.std(), tm.std()),
          flush=True)
    run = 4000
    assert abs(sf) * run <= 0.01 * abs(tf.mean())
    assert abs(sm - sf) <= 0.3 * abs(sf) + 1e-3
    assert tm.std() <= 1.5 * tf.std() + 1.0
