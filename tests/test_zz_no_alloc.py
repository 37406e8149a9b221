"""nx_fft3d_r2c_noalloc: the forward 3D FFT with every device buffer in the
library's static scratch array (no allocation in the call).

Checked against nx_fft3d_r2c, the entry point it is a variant of, which
allocates its device buffers per call: the same kernels on the same input
give the same bits.  Two calls in a row, the second on a smaller grid, so that
the second runs over what the first left in the shared array.
"""
import ctypes as C

import numpy as np
import pytest

from nexusmd_amd import _lib as L

pytestmark = pytest.mark.gpu

SHAPES = [(96, 96, 96), (40, 36, 48)]     # the second smaller than the first


def _r2c(name, a, prec):
    lib = L.lib()
    a = np.ascontiguousarray(a, np.float64)
    nx, ny, nz = a.shape
    out = np.full((nx, ny, nz // 2 + 1), np.nan, np.complex128)
    st = getattr(lib, name)(0, prec, nx, ny, nz, a.ctypes.data_as(L.D), out.ctypes.data_as(L.D))
    return st, out


@pytest.mark.parametrize("passes", ["plane", "lines"])
@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_noalloc_fft_matches_allocating_entry(prec, passes, monkeypatch):
    monkeypatch.setenv("NEXUSMD_FFT_PLANE", "1" if passes == "plane" else "0")
    rng = np.random.default_rng(17)
    for shape in SHAPES:
        print("[gpu] no-alloc fft prec=%d passes=%s shape=%s" % (prec, passes, shape), flush=True)
        a = rng.normal(size=shape)
        st0, want = _r2c("nx_fft3d_r2c", a, prec)
        st1, got = _r2c("nx_fft3d_r2c_noalloc", a, prec)
        assert st0 == L.NX_OK and st1 == L.NX_OK, (st0, st1, L.lib().nx_last_error(None))
        np.testing.assert_array_equal(got, want)
        ref = np.fft.rfftn(a)
        tol = 1e-12 if prec == L.FP64 else 2e-6
        assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref))


def test_noalloc_fft_rejects_grids_beyond_its_scratch():
    a = np.zeros((128, 128, 128))
    st, _ = _r2c("nx_fft3d_r2c_noalloc", a, L.FP64)
    assert st == L.NX_ERR_CAPACITY
    msg = L.lib().nx_last_error(None).decode()
    assert "no-allocation maximum" in msg
    a = np.zeros((300, 4, 4))
    st, _ = _r2c("nx_fft3d_r2c_noalloc", a, L.MIXED)
    assert st == L.NX_ERR_CAPACITY
    st, _ = _r2c("nx_fft3d_r2c_noalloc", np.zeros((14, 8, 8)), L.FP64)    # 7 is not a factor the FFT has
    assert st == L.NX_ERR_ARG
