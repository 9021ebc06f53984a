"""Oracle knots (knot_vector.F90) against the reference's own known answers.

Golden vectors: tests/test_knot.pf (test_knot1 :1-74, test_knot2 :78-132,
test_knot3 :135-215) and the flang-built reference leaves (oracle/_ref).
"""
import numpy as np
import pytest

# (n, p) -> expected FillOpenKnot output, tests/test_knot.pf:12-73
KNOT1 = [
    ((0, 0), [0.0, 1.0]),
    ((0, 1), [0.0, 1.0, 1.0]),
    ((1, 0), [0.0, 0.5, 1.0]),
    ((3, 0), [0.0, 0.25, 0.5, 0.75, 1.0]),
    ((4, 1), [0.0, 0.0, 0.25, 0.5, 0.75, 1.0, 1.0]),
    ((5, 2), [0.0, 0.0, 0.0, 0.25, 0.5, 0.75, 1.0, 1.0, 1.0]),
    ((6, 3), [0.0, 0.0, 0.0, 0.0, 0.25, 0.5, 0.75, 1.0, 1.0, 1.0, 1.0]),
]

# (n, p, U) -> CountSpans = n - p + 1, tests/test_knot.pf:91-131
KNOT2 = [
    (0, 0, [0.0, 1.0]),
    (0, 1, [0.0, 1.0, 1.0]),
    (1, 0, [0.0, 0.5, 1.0]),
    (3, 0, [0.0, 0.25, 0.5, 0.75, 1.0]),
    (4, 1, [0.0, 0.25, 0.5, 0.75, 1.0]),
    (5, 2, [0.0, 0.0, 0.0, 0.25, 0.5, 0.75, 1.0, 1.0, 1.0]),
    (6, 3, [0.0, 0.0, 0.0, 0.0, 0.25, 0.5, 0.75, 1.0, 1.0, 1.0, 1.0]),
]

CONFIG_NP = [(17, 2), (129, 2), (258, 3), (97, 2), (386, 3)]


@pytest.mark.parametrize("np_, expected", KNOT1)
def test_fill_open_knot_golden(orc, np_, expected):
    n, p = np_
    assert np.array_equal(orc.knots(n, p), np.array(expected))


@pytest.mark.parametrize("n,p,U", KNOT2)
def test_count_spans_golden(orc, n, p, U):
    assert orc.count_spans(n, p, np.array(U)) == n - p + 1


@pytest.mark.parametrize("n,p", [(-1, 0), (0, 0), (0, 1), (1, 0), (3, 0), (4, 1), (5, 2), (6, 3)])
def test_prepare_knot_nelem(orc, n, p):
    """test_knot3: PrepareKnot gives nelem = n - p + 1 (the n = -1 case included)."""
    _, ne = orc.prepare_knot(n, p)
    assert ne == n - p + 1


def test_interior_knots_are_float32_quotients(orc):
    """knot_vector.F90:89 divides in REAL(4): U_k = (double)((float)k / (float)ne)."""
    for n, p in CONFIG_NP:
        ne = n - p + 1
        U = orc.knots(n, p)
        k = np.arange(1, ne, dtype=np.int64)
        expect = (k.astype(np.float32) / np.float32(ne)).astype(np.float64)
        assert np.array_equal(U[p + 1:n + 1], expect)


def test_float32_knot_rounding_magnitude(orc):
    """At 96 and 384 elements the REAL(4) knots move element sizes off 1/ne by
    up to ~3.8e-6 and ~1.5e-5 relative; at 16, 128, 256 they are exact."""
    devs = {}
    for n, p in CONFIG_NP:
        ne = n - p + 1
        h = np.diff(orc.knots(n, p)[p:n + 2])
        devs[ne] = np.max(np.abs(h * ne - 1.0))
    assert devs[16] == 0.0 and devs[128] == 0.0 and devs[256] == 0.0
    assert 1e-6 < devs[96] < 1e-5
    assert 5e-6 < devs[384] < 5e-5


@pytest.mark.parametrize("n,p", CONFIG_NP + [(0, 1), (6, 3), (20, 4), (31, 1)])
def test_knots_bitwise_vs_reference_leaves(orc, ref_leaves, n, p):
    U = orc.knots(n, p)
    R = orc.knots(n, p, use_ref=True)
    assert np.array_equal(U, R)
    assert orc.count_spans(n, p, U) == orc.count_spans(n, p, R, use_ref=True)
