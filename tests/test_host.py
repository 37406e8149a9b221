"""Host-side checks that need no GPU: the C-ABI library loads and exports every
symbol include/nexusmd.h declares, the product package never reaches the
oracle, the build uses the required code-generation flags, and the seeded
workload generators produce the configured systems."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from nexusmd_amd import systems as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nexusmd.h")
LIB = os.path.join(ROOT, "nexusmd_amd", "libnexusmd_hip.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(nx_[a-z0-9_]+)\s*\(", text)))


needs_lib = pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built (python -m nexusmd_amd.build)")


@needs_lib
def test_library_exports_every_declared_symbol():
    import ctypes

    lib = ctypes.CDLL(LIB)
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (nx_[a-z0-9_]+)$", out, flags=re.M))
    assert set(names) <= exported
    # and nothing else leaks out as a C entry point
    assert exported <= set(names), sorted(exported - set(names))


@needs_lib
def test_python_signatures_cover_the_header():
    from nexusmd_amd import _lib

    assert sorted(n for n, _, _ in _lib.SIGNATURES) == declared_symbols()


@needs_lib
def test_status_strings_without_a_device():
    from nexusmd_amd import _lib

    L = _lib.lib()
    for code in range(0, 7):
        assert L.nx_status_string(code)
    assert b"OK" in L.nx_status_string(0).upper()


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "nexusmd_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"\boracle\b|pyoracle|libnxo|nxo_", text), f


def test_build_flags():
    from nexusmd_amd import build as B

    flags = " ".join(B.CFLAGS)
    assert "--offload-arch=gfx950" in flags
    assert "-target-feature -Xclang -packed-fp32-ops" in flags
    srcs = sorted(f for f in os.listdir(B.CSRC) if f.endswith(".hip"))
    assert sorted(B.SOURCES) == srcs


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from nexusmd_amd import _lib

    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"), raising=False)
    monkeypatch.setattr(_lib, "_lib", None, raising=False)
    with pytest.raises(Exception):
        _lib.lib()


# ---------------------------------------------------------------- workloads

def test_bead_chain_matches_its_configuration():
    s = S.bead_chain()
    assert s.n == 1500 and np.allclose(s.L, 4.0) and tuple(s.grid) == (40, 40, 40)
    assert len(s.bond_idx) == 1499 and len(s.angle_idx) == 1498 and len(s.torsion_idx) == 1497
    assert np.allclose(s.bond_par[:, 0], 0.15) and np.allclose(s.bond_par[:, 1], 2.5e5)
    assert np.allclose(s.angle_par[:, 0], 1.91) and np.allclose(s.angle_par[:, 1], 400.0)
    assert set(np.unique(s.torsion_per)) <= {1, 2, 3}
    assert np.all((s.torsion_par[:, 0] >= 0) & (s.torsion_par[:, 0] <= 10))
    assert set(np.round(np.unique(s.torsion_par[:, 1]), 12)) <= {0.0, round(math.pi, 12)}
    assert np.all((s.sig >= 0.25) & (s.sig <= 0.35)) and np.all((s.eps >= 0.1) & (s.eps <= 1.0))
    assert abs(s.q.sum()) < 1e-12
    d = s.pos[s.bond_idx[:, 0]] - s.pos[s.bond_idx[:, 1]]
    d -= s.L * np.round(d / s.L)
    assert np.allclose(np.linalg.norm(d, axis=1), 0.15, atol=1e-9)
    # deterministic
    assert np.array_equal(S.bead_chain().pos, s.pos)


def test_water_box_matches_its_configuration():
    s = S.water_box(n_mol=1000, L=3.107, grid=(32, 32, 32), seed=3)
    assert s.n == 3000 and len(s.bond_idx) == 2000 and len(s.angle_idx) == 1000
    o = np.arange(0, 3000, 3)
    np.testing.assert_allclose(s.q[o], -0.834)
    np.testing.assert_allclose(s.q[o + 1], 0.417)
    np.testing.assert_allclose(s.sig[o], 0.315)
    np.testing.assert_allclose(s.eps[o], 0.636)
    np.testing.assert_allclose(s.bond_par[:, 0], 0.09572)
    np.testing.assert_allclose(s.bond_par[:, 1], 462750.0)
    np.testing.assert_allclose(s.angle_par, np.tile([1.824, 836.8], (1000, 1)))
    d = s.pos[s.bond_idx[:, 1]] - s.pos[s.bond_idx[:, 0]]
    d -= s.L * np.round(d / s.L)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 0.09572, atol=1e-9)
    assert np.all(s.pos >= 0) and np.all(s.pos < s.L)


def test_serving_batch_distribution():
    sizes, seeds = S.serving_batch()
    assert len(sizes) == 256 and len(set(seeds)) == 256
    assert min(sizes) >= 600 and max(sizes) <= 8000
    assert 2000 <= np.median(sizes) <= 3000


def test_ligand_request():
    s = S.ligand_in_water(2500, seed=7)
    rho = s.n / float(np.prod(s.L))
    assert 90 <= rho <= 110
    assert s.cutoff == 0.9 and s.L[0] >= 2 * s.cutoff
    k = s.grid[0]
    assert 24 <= k <= 48 and k % 2 == 0
    m = k
    for p in (2, 3, 5):
        while m % p == 0:
            m //= p
    assert m == 1
    n_lig = s.n - 3 * int(np.sum(s.q == -0.834))
    assert 20 <= n_lig <= 60


def test_poisson_arrivals():
    t = S.poisson_arrivals(1000, 50.0, seed=1)
    assert np.all(np.diff(t) > 0)
    assert 15 <= t[-1] <= 25
