"""The allocation-free grid evaluation (ads_eval_grid_noalloc_d) through the
C ABI against ads_eval_grid, the entry point it is a variant of: the same
points and values, bit for bit, on two calls in a row on one context (the
second with a smaller grid, so stale table entries of the first would show),
then on a second context of another degree.  Sizes beyond its static tables
are refused before any GPU work."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from ads_mpi_amd import _lib, adss  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()


def grids(L, h, res, lo, hi):
    I3, D3 = ctypes.c_int * 3, ctypes.c_double * 3
    npts = res[0] * res[1] * res[2]
    host = np.empty(npts)
    _lib.check(L.ads_eval_grid(h, I3(*res), D3(*lo), D3(*hi), host.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), h)
    dev = torch.full((npts,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.check(L.ads_eval_grid_noalloc_d(h, I3(*res), D3(*lo), D3(*hi), ctypes.c_void_p(dev.data_ptr())), h)
    _lib.check(L.ads_synchronize(h), h)
    return host, dev.cpu().numpy()


@pytest.mark.parametrize("n,p,calls", [
    ((20, 23, 17), 2, [((31, 40, 29), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
                       ((7, 5, 9), (0.1, -0.2, 0.3), (0.9, 1.1, 0.35))]),
    ((9, 12, 7), 4, [((64, 33, 4096), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)),
                     ((2, 3, 31), (0.25, 0.5, 0.0), (0.75, 0.5, 1.0))]),
])
def test_eval_grid_noalloc_equals_eval_grid(n, p, calls):
    L = _lib.load()
    ads, data = adss.initialize(n, p)
    h = data.handle
    try:
        data.F = np.random.default_rng(5).uniform(-1, 1, data.D)
        for res, lo, hi in calls:
            say("case eval_grid_noalloc", n, p, res, lo, hi)
            host, dev = grids(L, h, res, lo, hi)
            assert np.array_equal(host, dev), float(np.max(np.abs(host - dev)))
        # beyond the static tables: refused before any GPU work
        I3, D3 = ctypes.c_int * 3, ctypes.c_double * 3
        out = torch.empty(8, dtype=torch.float64, device="cuda")
        rc = L.ads_eval_grid_noalloc_d(h, I3(4097, 2, 2), D3(0, 0, 0), D3(1, 1, 1), ctypes.c_void_p(out.data_ptr()))
        assert rc == _lib.ADS_ERR_INVALID
        assert b"4096" in L.ads_last_error(h)
    finally:
        adss.Cleanup(ads, data)
