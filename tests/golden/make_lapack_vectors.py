"""Writes tests/golden/lapack_band.npz: golden vectors for the banded LU.

For each (n, p) the 1-D B-spline mass matrix is built independently of the
oracle (scipy.interpolate.BSpline values at numpy Gauss-Legendre points on
knots with REAL(4) interior quotients, knot_vector.F90:89), factored with
scipy.linalg.lapack.dgbtrf and solved with dgbtrs ('N' and 'T') for a fixed
right-hand side.  Rerun with:  python tests/golden/make_lapack_vectors.py
"""
import os

import numpy as np
import scipy.linalg.lapack as la
from scipy.interpolate import BSpline

CASES = [(17, 2), (97, 2), (258, 3), (3, 3), (20, 4), (6, 1), (2, 1)]


def knots(n, p):
    ne = n - p + 1
    inner = (np.arange(1, ne, dtype=np.float32) / np.float32(ne)).astype(np.float64)
    return np.concatenate([np.zeros(p + 1), inner, np.ones(p + 1)])


def mass_matrix(n, p):
    U = knots(n, p)
    xg, wg = np.polynomial.legendre.leggauss(p + 1)
    N = n + 1
    M = np.zeros((N, N))
    K = np.zeros((N, N))
    basis = []
    for j in range(N):
        c = np.zeros(N)
        c[j] = 1.0
        s = BSpline(U, c, p, extrapolate=False)
        basis.append((s, s.derivative()))
    for e in range(n - p + 1):
        a, b = U[p + e], U[p + e + 1]
        x = (xg + 1.0) * (b - a) / 2.0 + a
        w = wg * (b - a) / 2.0
        vals = np.array([np.nan_to_num(s(x)) for s, _ in basis])
        ders = np.array([np.nan_to_num(d(x)) for _, d in basis])
        M += (vals * w) @ vals.T
        K += (ders * w) @ ders.T
    return M, K


def to_band(A, kl, ku):
    n = A.shape[0]
    ab = np.zeros((2 * kl + ku + 1, n))
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            ab[kl + ku + i - j, j] = A[i, j]
    return ab


def main():
    out = {}
    rng = np.random.default_rng(12345)
    for n, p in CASES:
        M, K = mass_matrix(n, p)
        ab = to_band(M, p, p)
        lu, ipiv, info = la.dgbtrf(ab, p, p)
        assert info == 0
        b = rng.standard_normal(n + 1)
        xN, info = la.dgbtrs(lu, p, p, b, ipiv, trans=0)
        xT, info2 = la.dgbtrs(lu, p, p, b, ipiv, trans=1)
        key = "n%d_p%d" % (n, p)
        out[key + "_M"] = M
        out[key + "_K"] = K
        out[key + "_ab"] = ab
        out[key + "_lu"] = lu
        out[key + "_ipiv"] = ipiv.astype(np.int32)
        out[key + "_b"] = b
        out[key + "_xN"] = xN
        out[key + "_xT"] = xT
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lapack_band.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
