"""Oracle quadrature and B-spline tables (gauss.F90, basis.F90).

Pins: closed forms and numpy's Gauss-Legendre rule, scipy's BSpline, a
brute-force Cox-de Boor recursion, partition of unity, and the flang-built
reference leaves (bitwise) for BasisData / DersBasisFuns / FindSpan /
EvalSpline / ReorderRHS.
"""
import numpy as np
import pytest
from scipy.interpolate import BSpline

CONFIG_NP = [(17, 2), (129, 2), (258, 3), (97, 2), (386, 3)]


@pytest.mark.parametrize("q", range(1, 11))
def test_gauss_vs_numpy(orc, q):
    X, W = orc.gauss_rule(q)
    xr, wr = np.polynomial.legendre.leggauss(q)
    assert np.max(np.abs(X - xr)) < 2e-15
    assert np.max(np.abs(W - wr)) < 2e-15
    assert abs(W.sum() - 2.0) < 1e-14


def test_gauss_rejects_unsupported_order(orc):
    with pytest.raises(ValueError):
        orc.gauss_rule(11)


@pytest.mark.parametrize("q", range(1, 6))
def test_gauss_closed_forms_bitwise_vs_reference(orc, ref_leaves, q):
    X, W = orc.gauss_rule(q)
    Xr, Wr = orc.gauss_rule(q, use_ref=True)
    assert np.array_equal(X, Xr) and np.array_equal(W, Wr)


def _cox_de_boor(U, i, p, u):
    """Textbook recursion N_{i,p}(u) on [U_i, U_{i+p+1}); brute force for tiny knots."""
    if p == 0:
        return 1.0 if U[i] <= u < U[i + 1] else 0.0
    a = 0.0
    if U[i + p] != U[i]:
        a = (u - U[i]) / (U[i + p] - U[i]) * _cox_de_boor(U, i, p - 1, u)
    b = 0.0
    if U[i + p + 1] != U[i + 1]:
        b = (U[i + p + 1] - u) / (U[i + p + 1] - U[i + 1]) * _cox_de_boor(U, i + 1, p - 1, u)
    return a + b


@pytest.mark.parametrize("n,p", [(3, 1), (5, 2), (6, 3), (8, 4)])
def test_ders_basis_funs_vs_cox_de_boor(orc, n, p):
    U = orc.knots(n, p)
    rng = np.random.default_rng(0)
    for u in rng.uniform(0.0, 1.0 - 1e-9, 20):
        span = orc.find_span(n, p, u, U)
        ders = orc.ders_basis_funs(span, u, p, 1, U)
        for a in range(p + 1):
            assert abs(ders[0, a] - _cox_de_boor(U, span - p + a, p, u)) < 1e-14


@pytest.mark.parametrize("n,p", [(17, 2), (9, 1), (258, 3), (97, 2), (22, 4)])
def test_basis_vs_scipy_bspline(orc, n, p):
    U = orc.knots(n, p)
    ne = n - p + 1
    O, J, W, X, NN = orc.basis_data(p, U, ne)
    for e in range(0, ne, max(1, ne // 7)):
        for a in range(p + 1):
            j = O[e] + a
            c = np.zeros(n + 1)
            c[j] = 1.0
            spl = BSpline(U, c, p, extrapolate=False)
            dspl = spl.derivative()
            v = spl(X[e])
            dv = dspl(X[e])
            assert np.max(np.abs(v - NN[e, :, a, 0])) < 1e-13
            assert np.max(np.abs(dv - NN[e, :, a, 1])) < 1e-10 * max(1.0, np.max(np.abs(dv)))


@pytest.mark.parametrize("n,p", CONFIG_NP)
def test_partition_of_unity(orc, n, p):
    U = orc.knots(n, p)
    O, J, W, X, NN = orc.basis_data(p, U, n - p + 1)
    assert np.max(np.abs(NN[:, :, :, 0].sum(axis=2) - 1.0)) < 1e-14
    assert np.max(np.abs(NN[:, :, :, 1].sum(axis=2))) < 1e-10
    assert np.all(NN[:, :, :, 0] >= 0)
    assert np.array_equal(O, np.arange(n - p + 1))  # element-to-DOF map: O(e) = e (0-based)
    assert np.allclose(J, np.diff(U[p:n + 2]) / 2.0, rtol=0, atol=0)


@pytest.mark.parametrize("n,p", CONFIG_NP + [(20, 4), (9, 1), (6, 3)])
def test_basis_data_bitwise_vs_reference(orc, ref_leaves, n, p):
    U = orc.knots(n, p)
    ne = n - p + 1
    for a, b in zip(orc.basis_data(p, U, ne), orc.basis_data(p, U, ne, use_ref=True)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n,p", [(17, 2), (97, 2), (386, 3)])
def test_find_span_bitwise_vs_reference(orc, ref_leaves, n, p):
    U = orc.knots(n, p)
    pts = np.concatenate([np.linspace(0, 1, 301), U, [-0.1, 1.1]])
    for u in pts:
        assert orc.find_span(n, p, u, U) == orc.find_span(n, p, u, U, use_ref=True)


def test_eval_spline_vs_reference(orc, ref_leaves):
    n, p = 9, 2
    U = orc.knots(n, p)
    rng = np.random.default_rng(1)
    coeffs = rng.standard_normal((n + 1) ** 3)
    L = orc.lib()
    R = ref_leaves
    for _ in range(30):
        x, y, z = rng.uniform(0, 1, 3)
        a = L.orc_eval_spline(0, orc._d(U), p, n, orc._d(U), p, n, orc._d(U), p, n, orc._d(coeffs), x, y, z)
        b = R.ref_eval_spline(n, p, n, p, n, p, orc._d(U), orc._d(U), orc._d(U), orc._d(coeffs), x, y, z)
        assert a == b


@pytest.mark.parametrize("which", ["y", "z", "x"])
def test_reorder_bitwise_vs_reference(orc, ref_leaves, which):
    s = (7, 5, 3)
    F = np.arange(np.prod(s), dtype=np.float64)
    assert np.array_equal(orc.reorder(which, s, F), orc.reorder(which, s, F, use_ref=True))


def test_reorder_cycle_is_identity(orc):
    s = (6, 4, 5)
    F = np.random.default_rng(2).standard_normal(np.prod(s))
    G = orc.reorder("x", s, orc.reorder("z", s, orc.reorder("y", s, F)))
    assert np.array_equal(F, G)
