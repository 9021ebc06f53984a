"""Oracle 1-D matrices (MKBBT, projection_engine.F90:63-132) and banded LU
(dgbtrf/dgbtrs form of the MUMPS solve, ADS.F90:294-379).

Pins: tests/golden/lapack_band.npz (scipy's LAPACK dgbtrf/dgbtrs on mass
matrices built independently with scipy BSpline + numpy Gauss points), the
closed-form interior Gram rows of uniform B-splines, and dense numpy solves.
"""
import os

import numpy as np
import pytest

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "lapack_band.npz"))
CASES = [(17, 2), (97, 2), (258, 3), (3, 3), (20, 4), (6, 1), (2, 1)]


def _band(orc, n, p, mix):
    U = orc.knots(n, p)
    return orc.mkbbt_band(U, p, n, n - p + 1, mix)


@pytest.mark.parametrize("n,p", CASES)
def test_mass_and_stiffness_vs_bruteforce(orc, n, p):
    key = "n%d_p%d" % (n, p)
    M = orc.band_to_dense(_band(orc, n, p, (1, 0, 0, 0)), p, p)
    K = orc.band_to_dense(_band(orc, n, p, (0, 1, 0, 0)), p, p)
    assert np.max(np.abs(M - GOLD[key + "_M"])) <= 1e-15 * np.max(np.abs(M)) * 10
    assert np.max(np.abs(K - GOLD[key + "_K"])) <= 1e-13 * np.max(np.abs(K))


def test_advection_matrices_are_transposes(orc):
    n, p = 12, 2
    B = orc.band_to_dense(_band(orc, n, p, (0, 0, 1, 0)), p, p)
    BT = orc.band_to_dense(_band(orc, n, p, (0, 0, 0, 1)), p, p)
    assert np.max(np.abs(B - BT.T)) < 1e-15
    mix = (0.3, -1.2, 0.7, 2.0)
    A = orc.band_to_dense(_band(orc, n, p, mix), p, p)
    M = orc.band_to_dense(_band(orc, n, p, (1, 0, 0, 0)), p, p)
    K = orc.band_to_dense(_band(orc, n, p, (0, 1, 0, 0)), p, p)
    assert np.max(np.abs(A - (0.3 * M - 1.2 * K + 0.7 * B + 2.0 * BT))) < 1e-13


@pytest.mark.parametrize("p,row", [(2, [1, 26, 66, 26, 1], ), (3, [1, 120, 1191, 2416, 1191, 120, 1])])
def test_interior_gram_rows_closed_form(orc, p, row):
    """Uniform B-spline mass rows are h * (cardinal B-spline of degree 2p+1 at the integers)."""
    ne = 16
    n = ne + p - 1
    M = orc.band_to_dense(_band(orc, n, p, (1, 0, 0, 0)), p, p)
    denom = {2: 120.0, 3: 5040.0}[p]
    expect = np.array(row) / denom / ne
    i = n // 2
    assert np.max(np.abs(M[i, i - p:i + p + 1] - expect)) < 1e-16
    # stiffness rows sum to zero (derivatives of a partition of unity)
    K = orc.band_to_dense(_band(orc, n, p, (0, 1, 0, 0)), p, p)
    assert np.max(np.abs(K.sum(axis=1))) < 1e-12 * np.max(np.abs(K))


@pytest.mark.parametrize("n,p", CASES)
def test_dgbtrf_dgbtrs_vs_lapack_golden(orc, n, p):
    key = "n%d_p%d" % (n, p)
    ab = GOLD[key + "_ab"]
    ldab = ab.shape[0]
    lu, ipiv, info = orc.dgbtrf(ab.T.copy(), p, p)  # oracle stores band columns contiguously
    assert info == 0
    assert np.array_equal(ipiv, GOLD[key + "_ipiv"])
    scale = np.max(np.abs(GOLD[key + "_lu"]))
    assert np.max(np.abs(lu.T - GOLD[key + "_lu"])) <= 4e-16 * scale * ldab
    b = GOLD[key + "_b"]
    xN = orc.dgbtrs("N", lu, p, p, ipiv, b)
    xT = orc.dgbtrs("T", lu, p, p, ipiv, b)
    for x, g in ((xN, GOLD[key + "_xN"]), (xT, GOLD[key + "_xT"])):
        assert np.linalg.norm(x - g) <= 1e-13 * np.linalg.norm(g)
    A = GOLD[key + "_M"]
    assert np.linalg.norm(A @ xN - b) <= 1e-13 * np.linalg.norm(b)
    assert np.linalg.norm(A.T @ xT - b) <= 1e-13 * np.linalg.norm(b)


def test_pivoting_cases_pivot(orc):
    """The cases the sweeps must replay interchanges for: (n, p) = (3, 3), (20, 4)."""
    for n, p in ((3, 3), (20, 4)):
        ipiv = GOLD["n%d_p%d_ipiv" % (n, p)]
        assert np.any(ipiv != np.arange(n + 1))
    for n, p in ((17, 2), (258, 3), (97, 2)):
        ipiv = GOLD["n%d_p%d_ipiv" % (n, p)]
        assert np.array_equal(ipiv, np.arange(n + 1))
