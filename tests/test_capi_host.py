"""The C-ABI library on the host: it loads, exports what include/ads_mi355x.h
declares, and its host-side tables and generators equal the oracle's."""
import ctypes

import numpy as np
import pytest

from ads_mpi_amd import _lib, adss


@pytest.fixture(scope="module")
def L():
    return _lib.load()


def test_exports_every_declared_symbol(L):
    names = _lib.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("n,p", [(17, 2), (129, 2), (258, 3), (97, 2), (386, 3), (20, 4), (9, 1)])
def test_host_tables_bitwise_equal_oracle(orc, n, p):
    U, O, J, W, X, NN = adss.host_basis(n, p)
    Uo = orc.knots(n, p)
    assert np.array_equal(U, Uo)
    Oo, Jo, Wo, Xo, NNo = orc.basis_data(p, Uo, n - p + 1)
    assert np.array_equal(O, Oo)  # element-to-DOF indexing, bit-exact
    assert np.array_equal(J, Jo)
    assert np.array_equal(W, Wo)
    assert np.array_equal(X, Xo)
    assert np.array_equal(NN, NNo)


def test_host_tables_reject_bad_sizes(L):
    with pytest.raises(_lib.AdsError):
        adss.host_basis(1, 3)  # n < p: no element
    rc = L.ads_host_basis(5, 0, None, None, None, None, None, None)
    assert rc == _lib.ADS_ERR_INVALID


def test_generators_bitwise_equal_oracle(orc):
    for seed in (0, 1, 2, 3, 12345):
        assert np.array_equal(adss.uniform(seed, 1000, -0.01, 0.01), orc.uniform(seed, 1000, -0.01, 0.01))
        assert np.array_equal(adss.lognormal(seed, 1000, 0.0, 0.5), orc.lognormal(seed, 1000, 0.0, 0.5))


def test_init_argument_checks_fail_loudly(L):
    I3 = ctypes.c_int * 3
    h = ctypes.c_void_p()
    assert L.ads_init(I3(9, 9, 9), I3(2, 3, 2), ctypes.byref(h)) == _lib.ADS_ERR_UNSUPPORTED
    assert b"degree" in L.ads_last_error(None)
    assert L.ads_init(I3(9, 9, 9), I3(7, 7, 7), ctypes.byref(h)) == _lib.ADS_ERR_UNSUPPORTED
    assert L.ads_init(I3(1, 9, 9), I3(2, 2, 2), ctypes.byref(h)) == _lib.ADS_ERR_INVALID
    assert not h.value


def test_null_context_is_rejected(L):
    assert L.ads_step(None, 1e-3, 0, 1, None) == _lib.ADS_ERR_INVALID
    assert L.ads_project(None, 0, None) == _lib.ADS_ERR_INVALID


def test_oil_curves_bitwise_equal_oracle(orc, L):
    for seed, cN, cL in ((1, 30, 16), (7, 3, 4)):
        a = orc.Oil.random(seed, cN=cN, cL=cL)
        rhs = adss.OilRHS.with_random_curves(1e-4, seed, cN=cN, cL=cL)
        assert np.array_equal(rhs.cx, a.cx) and np.array_equal(rhs.cy, a.cy) and np.array_equal(rhs.cz, a.cz)
    assert L.ads_oil_curves(1, 0, 4, None, None, None) == _lib.ADS_ERR_INVALID
