"""GPU: the step graph without the list kernels.

A one-system context runs most steps as a graph without the five list
kernels before the fork, chosen by the host from the update's report of the
largest displacement.  The choice must not change the trajectory: a step
without the list kernels on a system whose build is due freezes it, and the
host runs the step again with the list kernels, so every build happens at the
step the rebuild flag names.  The mixed path is compared bit for bit with a
context that carries the list kernels in every step (NEXUSMD_LITE=0), both
with the forecast (the default) and without it (NEXUSMD_LITE=2, where every
build goes through a frozen step and its repeat).
"""
import numpy as np
import pytest

from nexusmd_amd import _lib as L
from nexusmd_amd import engine as E
from nexusmd_amd import systems as S

pytestmark = pytest.mark.gpu

DT = 0.0005


@pytest.fixture(scope="module")
def water():
    rng = np.random.default_rng(5)
    s = S.water_box(n_mol=1000, L=3.107, grid=(32, 32, 30), seed=14)
    s.pos = (s.pos + rng.normal(0.0, 0.003, s.pos.shape)) % s.L
    return s


def _run(sysm, prec, mode, monkeypatch, calls=(7, 60, 1, 120, 33)):
    """Steps in calls of the given sizes (short calls run as direct launches,
    the others through the step graphs), then the state and counters."""
    monkeypatch.setenv("NEXUSMD_LITE", mode)
    print("[gpu] lite mode=%s precision=%d n=%d" % (mode, prec, sysm.n), flush=True)
    ctx = E.Context(sysm, precision=prec, skin=0.1)
    for n in calls:
        ctx.step(n, DT)
    pos, vel = ctx.get_state()
    st = ctx.stats()
    ctx.close()
    return pos, vel, st


@pytest.mark.parametrize("mode", ["1", "2"])
def test_mixed_lite_steps_give_the_same_bits(water, mode, monkeypatch):
    p0, v0, s0 = _run(water, L.MIXED, "0", monkeypatch)
    p1, v1, s1 = _run(water, L.MIXED, mode, monkeypatch)
    assert s0["steps_without_list"] == 0 and s0["steps_repeated"] == 0
    assert s1["rebuilds"] == s0["rebuilds"] and s0["rebuilds"] >= 5
    assert s1["steps_without_list"] > 0
    if mode == "2":
        # without the forecast the step after each raised flag runs without the
        # list kernels, freezes and is run again
        assert s1["steps_repeated"] > 0
    np.testing.assert_array_equal(p1, p0)
    np.testing.assert_array_equal(v1, v0)


def test_fp64_lite_steps_give_the_same_bits(water, monkeypatch):
    """The FP64 path, without the forecast (every build through a frozen step
    and its repeat): the same bits as with the list kernels in every step."""
    p0, v0, s0 = _run(water, L.FP64, "0", monkeypatch)
    p1, v1, s1 = _run(water, L.FP64, "2", monkeypatch)
    assert s1["steps_without_list"] > 0 and s1["steps_repeated"] > 0
    np.testing.assert_array_equal(p1, p0)
    np.testing.assert_array_equal(v1, v0)


def test_lite_steps_with_reports_between(water, monkeypatch):
    """Report steps (forces, energies, snapshot, update on the engine stream)
    between stepping calls, as Simulate runs them: the same bits as with the
    list kernels in every step."""
    out = []
    for mode in ("0", "1"):
        monkeypatch.setenv("NEXUSMD_LITE", mode)
        ctx = E.Context(water, precision=L.MIXED, skin=0.1)
        es = []
        for _ in range(4):
            ctx.step(29, DT)
            _, _, _, e = ctx.step_report(DT)
            es.append(e["potential"])
        pos, vel = ctx.get_state()
        ctx.close()
        out.append((pos, vel, es))
    np.testing.assert_array_equal(out[1][0], out[0][0])
    np.testing.assert_array_equal(out[1][1], out[0][1])
    assert out[1][2] == out[0][2]
