"""ComputeMatrix / MKBBT on the caller's knot vector (ads_matrix_create,
src/projection_engine.F90:531-544, 63-132): host-side, no GPU needed.

Pins: the oracle's MKBBT restatement on the same knots (itself pinned to the
LAPACK golden vectors and closed-form Gram rows), the reference's own
BasisData built by flang on graded knots (oracle/_ref), and an independent
assembly with scipy's BSpline and numpy's Gauss-Legendre points."""
import ctypes

import numpy as np
import pytest
from scipy.interpolate import BSpline

from ads_mpi_amd import _lib, adss


@pytest.fixture(scope="module")
def L():
    return _lib.load()


def graded_knots(n, p, power=1.5):
    ne = n - p + 1
    inner = (np.arange(1, ne) / ne) ** power
    return np.concatenate((np.zeros(p + 1), inner, np.ones(p + 1)))


def scipy_matrix(U, p, n, mix):
    """mix1 M + mix2 K + mix3 B + mix4 B^T by Gauss-Legendre quadrature (p+1
    points per span) of scipy B-splines: independent of both builds."""
    N = n + 1
    xg, wg = np.polynomial.legendre.leggauss(p + 1)
    A = np.zeros((N, N))
    spl = BSpline(np.asarray(U, dtype=np.float64), np.eye(N), p)  # all N basis functions
    dspl = spl.derivative()
    for i in range(p, n + 1):
        a, b = U[i], U[i + 1]
        if b <= a:
            continue
        x = 0.5 * (b - a) * xg + 0.5 * (a + b)
        w = 0.5 * (b - a) * wg
        v = spl(x).T  # (N, q)
        d = dspl(x).T
        M = (v * w) @ v.T
        K = (d * w) @ d.T
        B = (d * w) @ v.T
        A += mix[0] * M + mix[1] * K + mix[2] * B + mix[3] * B.T
    return A


MIXES = [(1.0, 0.0, 0.0, 0.0), (0.2, 1.0, 0.0, 0.0), (0.5, 1e-3, 0.0, 0.02), (1.0, 2e-3, 0.02, 0.0)]


@pytest.mark.parametrize("n,p", [(17, 2), (30, 3), (9, 1), (20, 4), (12, 6), (15, 9)])
def test_matrix_open_uniform_equals_oracle(orc, n, p):
    """Degrees 1..9 (Gauss rules up to 10 points, src/gauss.F90:213-257): the
    oracle's MKBBT to rounding; beyond 5 points the two Gauss rules come from
    different computations (closed forms below), so 1e-13 there."""
    U = orc.knots(n, p)
    for mix in MIXES:
        m = adss.ComputeMatrix(U, p, n, n - p + 1, mix)
        A = m.dense()
        m.close()
        O = orc.band_to_dense(orc.mkbbt_band(U, p, n, n - p + 1, mix), p, p)
        tol = 4e-15 if p <= 4 else 1e-13
        assert np.max(np.abs(A - O)) <= tol * np.max(np.abs(O))


@pytest.mark.parametrize("n,p", [(17, 2), (30, 3), (9, 1), (20, 4)])
def test_matrix_graded_knots(orc, n, p):
    """Non-uniform knots: the same matrix as the oracle's MKBBT on these
    knots, and as an independent scipy assembly; the oracle's basis tables on
    them equal the reference's own BasisData (flang) where that is built."""
    U = graded_knots(n, p)
    ne = n - p + 1
    for mix in MIXES:
        m = adss.ComputeMatrix(U, p, n, ne, mix)
        A = m.dense()
        m.close()
        O = orc.band_to_dense(orc.mkbbt_band(U, p, n, ne, mix), p, p)
        S = scipy_matrix(U, p, n, mix)
        scale = np.max(np.abs(S))
        assert np.max(np.abs(A - O)) <= 1e-15 * scale * 4
        assert np.max(np.abs(A - S)) <= 1e-13 * scale
    if orc.ref() is not None:
        mine = orc.basis_data(p, U, ne)
        ref = orc.basis_data(p, U, ne, use_ref=True)
        for a, b in zip(mine, ref):
            assert np.array_equal(a, b)


def test_matrix_errors(orc, L):
    n, p = 10, 2
    U = orc.knots(n, p)
    mix = (ctypes.c_double * 4)(1, 0, 0, 0)
    h = ctypes.c_void_p()
    dp = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    # nelem must be the knot vector's number of non-empty spans
    assert L.ads_matrix_create(dp(U), n, p, n - p, mix, ctypes.byref(h)) == _lib.ADS_ERR_INVALID
    assert b"non-empty spans" in L.ads_matrix_last_error(None)
    bad = U.copy()
    bad[5], bad[6] = bad[6], bad[5] + 0.01
    assert L.ads_matrix_create(dp(bad), n, p, n - p + 1, mix, ctypes.byref(h)) == _lib.ADS_ERR_INVALID
    zero = (ctypes.c_double * 4)(0, 0, 0, 0)
    assert L.ads_matrix_create(dp(U), n, p, n - p + 1, zero, ctypes.byref(h)) == _lib.ADS_ERR_SINGULAR
    assert not h.value
    # repeated interior knots: fewer elements, still a valid space
    Ur = np.concatenate((np.zeros(p + 1), [0.25, 0.5, 0.5, 0.75], np.ones(p + 1)))
    nr = Ur.size - p - 2
    m = adss.ComputeMatrix(Ur, p, nr, 4, (1, 0, 0, 0))
    S = scipy_matrix(Ur, p, nr, (1, 0, 0, 0))
    assert np.max(np.abs(m.dense() - S)) <= 1e-13 * np.max(np.abs(S))
    m.close()
