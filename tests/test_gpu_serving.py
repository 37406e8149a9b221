"""Continuous-batching server (nx_server) on the GPU.

The property under test: a request's result does not depend on what else is
on the GPU.  All force, grid and energy sums that cross wavefronts are fixed
point (64-bit in the mixed path, two 64-bit parts for FP64 forces), and each
system's neighbour list is rebuilt from the same positions whether it runs
alone or in a changing batch, so alone and in the mix agree bit for bit in
both paths.
"""
import numpy as np
import pytest

from nexusmd_amd import _lib as L
from nexusmd_amd import engine as E
from nexusmd_amd import systems as S


pytestmark = pytest.mark.gpu

DT = 0.0005


@pytest.fixture(scope="module")
def requests():
    sizes = [620, 900, 1300, 700, 1100, 800]
    steps = [60, 140, 100, 120, 80, 90]
    return [(S.ligand_in_water(n, seed=900 + i), k) for i, (n, k) in enumerate(zip(sizes, steps))]


def _alone(sys_, nsteps, prec):
    srv = E.Server(precision=prec, budget_bytes=8 << 30, max_active=4)
    rid = srv.submit(sys_, nsteps, DT)
    srv.run()
    st, done = srv.poll(rid)
    assert st == 2 and done == nsteps
    out = srv.fetch(rid)
    srv.close()
    return out


def _mix(reqs, prec, max_active=8, budget=8 << 30):
    srv = E.Server(precision=prec, budget_bytes=budget, max_active=max_active)
    ids = [srv.submit(s, k, DT) for s, k in reqs[:3]]
    srv.run(50)                  # the first three are running when the rest arrive
    for st, _ in (srv.poll(i) for i in ids):
        assert st == 1
    ids += [srv.submit(s, k, DT) for s, k in reqs[3:]]
    srv.run()
    out = []
    for i, (s, k) in zip(ids, reqs):
        st, done = srv.poll(i)
        assert st == 2 and done == k
        out.append(srv.fetch(i))
    stats = srv.stats()
    srv.close()
    return out, stats


def test_mixed_alone_equals_mix_bitwise(requests):
    print("[gpu] serving mixed: %d requests, sizes %s" % (len(requests), [s.n for s, _ in requests]), flush=True)
    mix, stats = _mix(requests, L.MIXED)
    assert stats["finished"] == len(requests) and stats["failed"] == 0
    assert stats["p99_step_ms"] >= stats["p50_step_ms"] > 0
    for (s, k), (p, v, e) in zip(requests, mix):
        pa, va, ea = _alone(s, k, L.MIXED)
        assert np.all(np.isfinite(p))
        np.testing.assert_array_equal(p, pa)
        np.testing.assert_array_equal(v, va)
        assert e["potential"] == ea["potential"]


def test_fp64_alone_equals_mix_bitwise(requests):
    reqs = requests[:4]
    print("[gpu] serving fp64: sizes %s" % [s.n for s, _ in reqs], flush=True)
    mix, _ = _mix(reqs, L.FP64)
    for (s, k), (p, v, e) in zip(reqs, mix):
        pa, va, ea = _alone(s, k, L.FP64)
        assert np.array_equal(p, pa)
        assert np.array_equal(v, va)
        assert e["potential"] == ea["potential"]


def test_fp64_repeated_runs_agree(requests):
    """A request run again and again, with other servers allocating and freeing
    device memory in between, gives the same FP64 result each time, bit for bit
    (new device buffers are cleared in stream order before their uploads, and
    the FP64 sums are fixed point)."""
    s, k = requests[3]
    print("[gpu] serving fp64 repeats: n=%d steps=%d" % (s.n, k), flush=True)
    p0, v0, e0 = _alone(s, k, L.FP64)
    for _ in range(8):
        _mix(requests[:3], L.MIXED)
        p, v, e = _alone(s, k, L.FP64)
        assert np.array_equal(p, p0)
        assert np.array_equal(v, v0)


def test_admission_by_memory_budget(requests):
    """With a budget that holds about two requests, the rest wait and all still finish."""
    per = []
    probe = E.Server(precision=L.MIXED, budget_bytes=8 << 30, max_active=1)
    for s, _ in requests:
        probe.submit(s, 2, DT)
        probe.run(1)                        # admitted and still running: its bytes are in use
        per.append(probe.stats()["bytes_in_use"])
        probe.run()
    probe.close()
    assert min(per) > 0
    budget = int(2.2 * max(per))
    srv = E.Server(precision=L.MIXED, budget_bytes=budget, max_active=64)
    ids = [srv.submit(s, k, DT) for s, k in requests]
    srv.run(10)
    active = srv.stats()["active"]
    assert 1 <= active < len(requests)
    srv.run()
    for i, (s, k) in zip(ids, requests):
        assert srv.poll(i) == (2, k)
    st = srv.stats()
    assert st["finished"] == len(requests) and st["bytes_in_use"] == 0
    srv.close()


def test_bad_request_is_rejected_without_disturbing_others(requests):
    srv = E.Server(precision=L.MIXED, budget_bytes=8 << 30, max_active=8)
    s0, k0 = requests[0]
    good = srv.submit(s0, k0, DT)
    bad = s0.copy()
    bad.bond_idx = bad.bond_idx.copy()
    bad.bond_idx[0, 0] = bad.n + 5            # particle index out of range
    with pytest.raises(L.NexusError) as ex:
        srv.submit(bad, 10, DT)
    assert ex.value.status == L.NX_ERR_ARG
    srv.run()
    assert srv.poll(good) == (2, k0)
    srv.close()


def test_mid_run_overflow_fails_only_that_request(requests):
    """A request whose neighbour-list build overflows mid-run (every atom of a
    collapsing cloud lands in one column after about 34 steps) fails alone:
    the device freezes it at its last completed step, the server names it
    failed, and the requests beside it finish exactly as they do alone
    (mixed path, bit for bit)."""
    from _util import collapse_system
    bad, excl = collapse_system()
    others = [requests[0], requests[2], requests[4]]
    print("[gpu] serving mid-run overflow: collapse n=%d beside %s" % (bad.n, [s.n for s, _ in others]), flush=True)
    srv = E.Server(precision=L.MIXED, budget_bytes=8 << 30, max_active=8)
    ids = [srv.submit(s, k, DT) for s, k in others]
    rid = srv.submit(bad, 80, DT, excl=excl)
    srv.run()
    st, done = srv.poll(rid)
    assert st == 3, "the collapsing request must fail"
    assert 20 <= done < 80
    with pytest.raises(L.NexusError) as ex:
        srv.fetch(rid)
    assert ex.value.status == L.NX_ERR_CAPACITY
    out = []
    for i, (s, k) in zip(ids, others):
        assert srv.poll(i) == (2, k)
        out.append(srv.fetch(i))
    stats = srv.stats()
    assert stats["failed"] == 1 and stats["finished"] == len(others)
    srv.close()
    for (s, k), (p, v, e) in zip(others, out):
        pa, va, ea = _alone(s, k, L.MIXED)
        np.testing.assert_array_equal(p, pa)
        np.testing.assert_array_equal(v, va)
        assert e["potential"] == ea["potential"]


def test_two_overflowing_requests_both_fail(requests):
    """Two requests that overflow in the same chunk are both named failed; the
    third request finishes."""
    from _util import collapse_system
    a, ea = collapse_system(seed=5)
    b, eb = collapse_system(seed=6)
    s0, k0 = requests[1]
    print("[gpu] serving two overflows beside n=%d" % s0.n, flush=True)
    srv = E.Server(precision=L.MIXED, budget_bytes=8 << 30, max_active=8)
    good = srv.submit(s0, k0, DT)
    ra = srv.submit(a, 60, DT, excl=ea)
    rb = srv.submit(b, 60, DT, excl=eb)
    srv.run()
    assert srv.poll(ra)[0] == 3 and srv.poll(rb)[0] == 3
    assert srv.poll(good) == (2, k0)
    assert srv.stats()["failed"] == 2
    srv.close()
