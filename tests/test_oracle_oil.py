"""Pins the oracle's restatement of the oil problem (problems/oil/RHS_eq.F90,
input_data.F90, src/math.F90) before it is trusted as the GPU checker:

  * falloff / lerp bitwise, bump3d to 2 ulp, against the reference's own
    math.F90 compiled with flang (oracle/_ref; skipped when not built);
  * dist_from_segment against brute force along the segment and against the
    cross-product distance; dist_from_curves against a vectorised numpy
    minimum over all segments;
  * kq, pumping, draining and initial_state at points where their values are
    known in closed form;
  * the oil right-hand side against an independent numpy quadrature (scipy
    B-spline basis, numpy Gauss-Legendre points, a numpy distance field):
    coefficients, l2norm and drained;
  * the oil step with mi = 0, Kq constant and no pumps or drains against the
    oracle's heat step with that constant diffusivity.
The curves of the reference come from an unseeded random_number, so no run of
the reference has fixed curves to compare with; the generator is checked for
its recurrence (step length 0.05 |dx|) only.
"""
import numpy as np
import pytest
from scipy.interpolate import BSpline

import oracle as O


def _rng_pts(seed, n):
    return np.random.default_rng(seed).uniform(-0.2, 1.2, (n, 3))


def test_math_leaves_against_reference():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    rng = np.random.default_rng(1)
    for _ in range(2000):
        r = rng.uniform(0, 0.2)
        Rr = r + rng.uniform(1e-3, 0.5)
        t = rng.uniform(-0.1, 0.8)
        assert O.oil_falloff(r, Rr, t) == R.ref_falloff(r, Rr, t)
        a, x, y = rng.uniform(0, 1), rng.uniform(-5, 5), rng.uniform(-5, 1000)
        assert O.oil_lerp(a, x, y) == R.ref_lerp(a, x, y)
        X = rng.uniform(0, 1, 3)
        got = O.oil_falloff(0.1, 0.3, float(np.sqrt(((X - 0.5) ** 2).sum())))
        ref = R.ref_bump3d(0.2, 0.6, *X)
        assert abs(got - ref) <= 4.5e-16 * max(1.0, abs(ref))


def test_falloff_closed_forms():
    assert O.oil_falloff(0.1, 0.3, 0.05) == 1.0
    assert O.oil_falloff(0.1, 0.3, 0.31) == 0.0
    assert O.oil_falloff(0.0, 0.06, 0.0) == 1.0          # h = 0: ((-1)(1))^2
    assert O.oil_falloff(0.0, 0.06, 0.06) == 0.0         # h = 1
    assert abs(O.oil_falloff(0.0, 2.0, 1.0) - 0.5625) < 1e-16  # h = 1/2: (3/4)^2
    assert O.oil_lerp(0.0, 3.0, 7.0) == 3.0 and O.oil_lerp(1.0, 3.0, 7.0) == 7.0


def test_dist_from_segment():
    rng = np.random.default_rng(2)
    for _ in range(300):
        a, b, p = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3), rng.uniform(-0.5, 1.5, 3)
        d = O.oil_dist_from_segment(p, a, b)
        s = np.linspace(0.0, 1.0, 200001)
        brute = np.min(((a[None, :] + s[:, None] * (b - a)[None, :] - p[None, :]) ** 2).sum(1))
        assert d <= brute * (1 + 1e-12) + 1e-18
        assert brute - d <= 2 * np.linalg.norm(b - a) * np.sqrt(brute) / 200000 + 1e-12
        proj = np.dot(b - a, p - a) / np.dot(b - a, b - a)
        if 0 < proj < 1:
            cross = np.cross(b - a, p - a)
            assert abs(d - np.dot(cross, cross) / np.dot(b - a, b - a)) <= 1e-13 * max(d, 1e-3)
        elif proj <= 0:
            assert abs(d - np.dot(p - a, p - a)) <= 1e-15
        else:
            assert abs(d - np.dot(p - b, p - b)) <= 1e-15


def _np_dist_curves(cx, cy, cz, cN, cL, pts):
    C = np.stack([cx, cy, cz], 1).reshape(cN, cL, 3)
    A = C[:, :-1].reshape(-1, 3)
    B = C[:, 1:].reshape(-1, 3)
    d = B - A
    best = np.full(len(pts), 1e3)
    for i, p in enumerate(pts):
        t = np.clip(((p - A) * d).sum(1) / (d * d).sum(1), 0.0, 1.0)
        q = A + t[:, None] * d
        best[i] = min(1e3, ((p - q) ** 2).sum(1).min())
    return best


def test_dist_from_curves_against_numpy():
    oil = O.Oil.random(3)
    pts = _rng_pts(4, 200)
    ref = _np_dist_curves(oil.cx, oil.cy, oil.cz, oil.cN, oil.cL, pts)
    got = np.array([oil.dist_from_curves(p) for p in pts])
    assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-12)) <= 1e-12


def test_point_functions_closed_forms():
    cx = np.array([0.2, 0.3, 0.4, 0.9, 0.9, 0.9])
    cy = np.array([0.2, 0.2, 0.2, 0.9, 0.8, 0.7])
    cz = np.array([0.5, 0.5, 0.5, 0.1, 0.1, 0.1])
    oil = O.Oil(cx, cy, cz, 2, 3, pumps=[(0.5, 0.5, 0.5)], drains=[(0.1, 0.1, 0.1), (0.8, 0.1, 0.1)],
                mi=10.0, kq_min=1.0, kq_max=1000.0, radius=0.15, pump_strength=2.0, drain_strength=3.0)
    assert oil.kq(0.3, 0.2, 0.5) == 1000.0             # on a curve: falloff(0, .06, 0) = 1
    assert oil.kq(0.3, 0.8, 0.9) == 1.0                # far from every segment
    # 0.03 from the first curve: falloff(0, 0.06, 0.03) = 9/16
    assert abs(oil.kq(0.3, 0.23, 0.5) - (1.0 * 7 / 16 + 1000.0 * 9 / 16)) < 1e-9
    assert oil.pumping(0.5, 0.5, 0.5) == 2.0
    assert oil.pumping(0.5, 0.5, 0.7) == 0.0           # beyond the radius
    assert oil.draining(0.7, 0.1, 0.1, 0.1) == 3.0 * 0.7
    assert oil.draining(-0.5, 0.45, 0.1, 0.1) == 0.0   # 0.35 from both drains
    # on the first curve, inside bump3d's flat top (|X - .5| < 0.1)
    assert oil.initial_state(0.45, 0.5, 0.5) == 0.0    # 0.3 from the curves: falloff(0, .1, .3) = 0
    on = O.Oil(np.array([0.45, 0.55]), np.array([0.5, 0.5]), np.array([0.5, 0.5]), 1, 2)
    assert on.initial_state(0.5, 0.5, 0.5) == 0.1


def _basis_1d(n, p):
    """Knots (the oracle's, pinned elsewhere), the Gauss points of every
    element, their weights times the element Jacobian, and all basis values
    and first derivatives there from scipy's B-splines."""
    U = O.knots(n, p)
    br = np.unique(U)
    g, w = np.polynomial.legendre.leggauss(p + 1)
    xs, ws = [], []
    for a, b in zip(br[:-1], br[1:]):
        xs.append((a + b) / 2 + (b - a) / 2 * g)
        ws.append((b - a) / 2 * w)
    x = np.concatenate(xs)
    nb = len(U) - p - 1
    N = np.zeros((len(x), nb))
    dN = np.zeros((len(x), nb))
    for i in range(nb):
        c = np.zeros(nb)
        c[i] = 1.0
        s = BSpline(U, c, p, extrapolate=False)
        N[:, i] = s(x)
        dN[:, i] = s.derivative()(x)
    return x, np.concatenate(ws), np.nan_to_num(N), np.nan_to_num(dN)


def _np_oil_rhs(n, p, oil, c, dt):
    (xx, wx, Nx, dNx), (xy, wy, Ny, dNy), (xz, wz, Nz, dNz) = [_basis_1d(n[d], p) for d in range(3)]
    C = c.reshape(Nz.shape[1], Ny.shape[1], Nx.shape[1])  # [z][y][x]
    # values and gradients at every point [qz][qy][qx]
    u = np.einsum("zyx,az,by,cx->abc", C, Nz, Ny, Nx)
    ux = np.einsum("zyx,az,by,cx->abc", C, Nz, Ny, dNx)
    uy = np.einsum("zyx,az,by,cx->abc", C, Nz, dNy, Nx)
    uz = np.einsum("zyx,az,by,cx->abc", C, dNz, Ny, Nx)
    Z, Y, X = np.meshgrid(xz, xy, xx, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    dist = np.sqrt(_np_dist_curves(oil.cx, oil.cy, oil.cz, oil.cN, oil.cL, pts)).reshape(u.shape)
    fo = lambda r, R, t: np.where(t < r, 1.0, np.where(t > R, 0.0, ((((t - r) / (R - r)) - 1) * (((t - r) / (R - r)) + 1)) ** 2))
    kq = (1 - fo(0.0, 0.06, dist)) * oil.c.kq_min + fo(0.0, 0.06, dist) * oil.c.kq_max
    pump = np.zeros(u.shape)
    for q in oil.pumps:
        pump += oil.c.pump_strength * fo(0.0, oil.c.radius, np.sqrt((X - q[0]) ** 2 + (Y - q[1]) ** 2 + (Z - q[2]) ** 2))
    dw = np.zeros(u.shape)
    for q in oil.drains:
        dw += oil.c.drain_strength * fo(0.0, oil.c.radius, np.sqrt((X - q[0]) ** 2 + (Y - q[1]) ** 2 + (Z - q[2]) ** 2))
    vdrain = np.maximum(0.0, dw * u)
    W = np.einsum("a,b,c->abc", wz, wy, wx)
    w0 = W * (u + dt * (pump - vdrain))
    g = -W * dt * kq * np.exp(oil.c.mi * u)
    F = (np.einsum("abc,az,by,cx->zyx", w0, Nz, Ny, Nx) + np.einsum("abc,az,by,cx->zyx", g * ux, Nz, Ny, dNx)
         + np.einsum("abc,az,by,cx->zyx", g * uy, Nz, dNy, Nx) + np.einsum("abc,az,by,cx->zyx", g * uz, dNz, Ny, Nx))
    return F.ravel(), float((W * u * u).sum()), float((W * dt * vdrain).sum())


@pytest.mark.parametrize("n,p", [((5, 4, 6), 2), ((4, 5, 3), 3), ((6, 3, 4), 1)])
def test_oil_rhs_against_numpy_quadrature(n, p):
    oil = O.Oil.random(7, cN=5, cL=6, pumps=[(0.5, 0.45, 0.5), (0.2, 0.8, 0.3)], drains=[(0.6, 0.5, 0.55)],
                       mi=3.0, kq_min=1.0, kq_max=50.0, radius=0.3, pump_strength=1.5, drain_strength=2.0)
    S = O.Setup(n, p)
    c = np.random.default_rng(8).uniform(-0.3, 0.4, S.D)
    kq = S.oil_kq_cache(oil)
    F, l2, dr = S.oil_form_rhs(oil, kq, c, 1e-3)
    Fr, l2r, drr = _np_oil_rhs(n, p, oil, c, 1e-3)
    assert np.linalg.norm(F - Fr) <= 1e-12 * np.linalg.norm(Fr)
    assert abs(l2 - l2r) <= 1e-12 * l2r
    assert drr > 0 and abs(dr - drr) <= 1e-12 * drr


def test_oil_reduces_to_heat():
    n, p, dt, kap = (7, 6, 5), 2, 1e-4, 2.5
    oil = O.Oil.random(9, cN=3, cL=4, mi=0.0, kq_min=kap, kq_max=kap)
    S = O.Setup(n, p)
    c = np.random.default_rng(10).uniform(-1, 1, S.D)
    got, l2, dr = S.oil_step(oil, S.oil_kq_cache(oil), c, dt)
    ref, l2r = S.step(c, dt, kappa=np.full(S.E, kap), mode=O.MODE_FAITHFUL)
    assert np.linalg.norm(got - ref) <= 1e-14 * np.linalg.norm(ref)
    assert abs(l2 - l2r) <= 1e-14 * l2r and dr == 0.0


def test_curve_generator_recurrence():
    a = O.Oil.random(11, cN=4, cL=16)
    b = O.Oil.random(12, cN=4, cL=16)
    assert not np.array_equal(a.cx, b.cx)
    for o in (a, b):
        P = np.stack([o.cx, o.cy, o.cz], 1).reshape(4, 16, 3)
        assert np.all((P[:, 0] >= 0) & (P[:, 0] < 1))
        steps = np.linalg.norm(np.diff(P, axis=1), axis=2)
        assert np.all(np.isfinite(steps)) and np.all(steps > 0)
