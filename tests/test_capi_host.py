"""The C-ABI library on the host: it loads, exports what include/ads_mi355x.h
declares, and its host-side tables and generators equal the oracle's."""
import ctypes

import numpy as np
from scipy.linalg import solve_triangular
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
    # degrees 1..5 per direction (unequal ones too); beyond that the kernels are not built
    assert L.ads_init(I3(9, 9, 9), I3(2, 6, 2), ctypes.byref(h)) == _lib.ADS_ERR_UNSUPPORTED
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


SEG_MAX, SEG_BLOCK = 32, 16  # outputs per lane at most, block of the recurrences (k_seg_lines)


def seg_emulate(rows, mid, lo, hi, w, p, b):
    """float64 restatement of the segmented solve of the FP32 refinement form
    (k_seg_lines / k_seg_rows): each segment of at most SEG_MAX outputs runs
    L's recurrence from w entries before it and U's from w entries after it,
    both from zero state; blocks of SEG_BLOCK rows inside [lo, hi) use the
    middle row."""
    N = len(b)
    nseg = -(-N // SEG_MAX)
    seg = -(-N // nseg)
    x = np.zeros(N)
    B = SEG_BLOCK

    def coef(j0, j):
        if j0 >= lo and j0 + B <= hi:
            return mid
        return rows[j] if 0 <= j < N else np.zeros(2 * p + 1)

    for s in range(nseg):
        ks, ke = s * seg, min(s * seg + seg, N)
        y = np.zeros(SEG_MAX + w)
        h = [0.0] * p  # h[t] = y(j-1-t)
        for f in range((SEG_MAX + 2 * w) // B):
            j0 = ks - w + f * B
            for i in range(B):
                j = j0 + i
                c = coef(j0, j)
                v = (b[j] if 0 <= j < N else 0.0) - sum(c[t] * h[t] for t in range(p))
                h = [v] + h[:-1]
                if f * B + i >= w:
                    y[f * B + i - w] = v
        g = [0.0] * p  # g[t-1] = x(j+t)
        for bb in reversed(range((SEG_MAX + w) // B)):
            j0 = ks + bb * B
            for i in reversed(range(B)):
                j = j0 + i
                c = coef(j0, j)
                v = y[bb * B + i] - sum(c[p + t] * g[t - 1] for t in range(1, p + 1))
                xj = v * c[p]
                g = [xj] + g[:-1]
                if bb * B < SEG_MAX and j < ke:
                    x[j] = xj
    return x


@pytest.mark.parametrize("n,p", [(258, 3), (128, 2), (20, 4), (9, 1), (385, 3), (150, 4), (200, 1),
                                 (1, 1), (2, 2), (3, 3), (4, 4), (8, 4), (31, 4)])
def test_seg_tables_and_truncated_solve(orc, L, n, p):
    """The segmented FP32 form's tables and algorithm, against the oracle's
    (independently assembled) mass matrix M:
      * the rows are LU factors of M without interchanges: L U = M to FP32
        rounding;
      * the warm-up w fits the decay of the factors' recurrences (about 0.27,
        0.43, 0.54, 0.61 per entry for p = 1..4; 16 for lines of one segment,
        down to 2 entries), and the interior range that uses the middle row
        covers all but the rows near the ends;
      * the segmented solve (restated in float64 with the FP32 tables)
        matches the sequential solve with the same factors to the 1e-8 decay
        tolerance, and the dense solve of M to the FP32 rounding of the
        factors, for smooth and random right-hand sides."""
    N = n + 1
    nt = 2 * p + 1
    tab = np.zeros((N + 1) * nt, dtype=np.float32)
    w, lo, hi = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    assert L.ads_host_seg_tables(n, p, ctypes.byref(w), ctypes.byref(lo), ctypes.byref(hi),
                                 tab.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
    w, lo, hi = w.value, lo.value, hi.value
    t = tab.reshape(N + 1, nt).astype(np.float64)
    rows, mid = t[:N], t[N]
    U = orc.knots(n, p)
    M = orc.band_to_dense(orc.mkbbt_band(U, p, n, n - p + 1, (1, 0, 0, 0)), p, p)
    Lm, Um = np.eye(N), np.zeros((N, N))
    for j in range(N):
        for s_ in range(1, p + 1):
            if j - s_ >= 0:
                Lm[j, j - s_] = rows[j, s_ - 1]
            if j + s_ < N:
                Um[j, j + s_] = rows[j, p + s_]
        Um[j, j] = 1.0 / rows[j, p]
    assert np.abs(Lm @ Um - M).max() <= 1e-6 * np.abs(M).max()
    # a line of at most SEG_MAX entries is one segment, exact for any w
    assert w % SEG_BLOCK == 0 and w <= ({1: 16, 2: 32, 3: 32, 4: 48}[p] if N > SEG_MAX else 16)
    assert 0 <= lo <= N // 2 < hi <= N
    ne = n - p + 1
    if N >= 100 and ne & (ne - 1) == 0:
        # knots of a power-of-two element count are exact (the knot vector is
        # built in single precision, so other counts vary the element lengths
        # by about 1e-5 and few rows match the middle one to 1e-7)
        assert lo <= 40 and hi >= N - 40
    rng = np.random.default_rng(n * 10 + p)
    xs = np.linspace(0.0, 1.0, N)
    for b in (np.sin(3 * xs) + xs ** 2, rng.standard_normal(N)):
        exact = np.linalg.solve(M, b)
        # the same FP32 factors, solved sequentially from the line's start
        seq = solve_triangular(Um, solve_triangular(Lm, b, lower=True, unit_diagonal=True))
        got = seg_emulate(rows, mid, lo, hi, w, p, b)
        assert np.linalg.norm(got - seq) <= 2e-8 * np.linalg.norm(seq)  # warm-up truncation, middle row
        assert np.linalg.norm(got - exact) <= 5e-7 * np.linalg.norm(exact)  # FP32 factors


def part_tables(L, n, p):
    K, S = ctypes.c_int(0), ctypes.c_int(0)
    _lib.check(L.ads_host_part_tables(n, p, ctypes.byref(K), ctypes.byref(S), None, None, None, None))
    K, S = K.value, S.value
    rows = K * S + p + 1
    lrow, urow = np.zeros(rows * p), np.zeros(rows * (p + 1))
    gsp, hsp = np.zeros(rows * p), np.zeros(rows * p)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    _lib.check(L.ads_host_part_tables(n, p, ctypes.byref(ctypes.c_int()), ctypes.byref(ctypes.c_int()), dp(lrow),
                                      dp(urow), dp(gsp), dp(hsp)))
    return K, S, lrow.reshape(rows, p), urow.reshape(rows, p + 1), gsp.reshape(rows, p), hsp.reshape(rows, p)


def test_part_tables_line_limit(L):
    """Lines of 16 partitions: 384 entries take partitions of 24, 385 and 512
    partitions of 32, 513 are refused (the LU sweeps serve them)."""
    for n, S_exp, K_exp in ((383, 24, 16), (384, 32, 13), (511, 32, 16)):
        K, S = ctypes.c_int(0), ctypes.c_int(0)
        _lib.check(L.ads_host_part_tables(n, 3, ctypes.byref(K), ctypes.byref(S), None, None, None, None))
        assert S.value == S_exp and K.value == K_exp
    K, S = ctypes.c_int(0), ctypes.c_int(0)
    assert L.ads_host_part_tables(512, 3, ctypes.byref(K), ctypes.byref(S), None, None, None,
                                  None) == _lib.ADS_ERR_UNSUPPORTED


def part_emulate(K, S, lrow, urow, gsp, hsp, p, b):
    """float64 restatement of the partitioned kernels (k_part_lines): every
    partition from zero state; the two edge-state scans, which take the
    response tables at the partitions' edge rows; the corrections as the
    homogeneous recurrences started from the true edge states."""
    n = b.size
    v = b.astype(np.float64).copy()
    E, X = [], []
    for k in range(K):  # forward, local
        a, ln = k * S, min(S, n - k * S)
        h = np.zeros(p)
        for j in range(a, a + ln):
            y = v[j] - np.dot(lrow[j], h)
            h = np.concatenate(([y], h[:-1]))
            v[j] = y
        E.append(h.copy())
    for k in range(K):  # scan from below, homogeneous forward recurrence
        s = np.zeros(p)
        for kk in range(k):
            b_ = (kk + 1) * S
            s = np.array([E[kk][r] + np.dot(gsp[b_ - 1 - r], s) for r in range(p)])
        a, ln = k * S, min(S, n - k * S)
        d = s.copy()
        for j in range(a, a + ln):
            e = -np.dot(lrow[j], d)
            d = np.concatenate(([e], d[:-1]))
            v[j] += e
    for k in range(K):  # backward, local
        a, ln = k * S, min(S, n - k * S)
        g = np.zeros(p)
        for j in range(a + ln - 1, a - 1, -1):
            x = (v[j] - np.dot(urow[j, 1:], g)) * urow[j, 0]
            g = np.concatenate(([x], g[:-1]))
            v[j] = x
        X.append(g.copy())
    for k in range(K):  # scan from above, homogeneous backward recurrence
        t = np.zeros(p)
        for kk in range(K - 1, k, -1):
            a2 = kk * S
            t = np.array([X[kk][r] + np.dot(hsp[a2 + r], t) for r in range(p)])
        a, ln = k * S, min(S, n - k * S)
        e = t.copy()
        for j in range(a + ln - 1, a - 1, -1):
            w = -np.dot(urow[j, 1:], e) * urow[j, 0]
            e = np.concatenate(([w], e[:-1]))
            v[j] += w
    return v


@pytest.mark.parametrize("n,p", [(258, 3), (129, 2), (20, 4), (9, 1), (500, 3), (40, 2), (5, 4), (3, 3),
                                 (386, 3), (97, 2), (64, 4), (511, 1)])
def test_part_tables_and_partitioned_solve(orc, L, n, p):
    """The partitioned FP64 solve's tables and algorithm, against the oracle's
    (independently assembled) mass matrix M and a dense solve:
      * lrow / urow are LU factors of M without interchanges: L U = M to
        rounding;
      * partitions of 24 entries (32 for lines of 385 to 512 entries), the
        last one padded past the line with zero rows;
      * the partitioned solve (restated in float64 with the tables) equals
        the dense solve of M to rounding (no truncation) for smooth and
        random right-hand sides."""
    N = n + 1
    K, S, lrow, urow, gsp, hsp = part_tables(L, n, p)
    assert S == (24 if N <= 384 else 32) and K <= 16 and (K - 1) * S < N <= K * S
    U = orc.knots(n, p)
    M = orc.band_to_dense(orc.mkbbt_band(U, p, n, n - p + 1, (1, 0, 0, 0)), p, p)
    Lm, Um = np.eye(N), np.zeros((N, N))
    for j in range(N):
        Um[j, j] = 1.0 / urow[j, 0]
        for s_ in range(1, p + 1):
            if j - s_ >= 0:
                Lm[j, j - s_] = lrow[j, s_ - 1]
            if j + s_ < N:
                Um[j, j + s_] = urow[j, s_]
    assert np.abs(Lm @ Um - M).max() <= 1e-14 * np.abs(M).max()
    assert not lrow[N:].any() and not urow[N:].any() and not gsp[N:].any() and not hsp[N:].any()
    rng = np.random.default_rng(n * 10 + p)
    xs = np.linspace(0.0, 1.0, N)
    for b in (np.sin(3 * xs) + xs ** 2, rng.standard_normal(N)):
        exact = np.linalg.solve(M, b)
        got = part_emulate(K, S, lrow, urow, gsp, hsp, p, b)
        assert np.linalg.norm(got - exact) <= 1e-14 * np.linalg.norm(exact)
