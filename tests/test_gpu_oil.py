"""The oil problem (problems/oil) on the GPU against the oracle's faithful
restatement (tests/test_oracle_oil.py pins that restatement):

  * the kq cache at every Gauss point (CacheKqValues): bit-exact, p = 1..4;
  * the t = 0 projection of initial_state, then steps: coefficients, l2norm
    and drained within 1e-12 relative after every step.

The curves are a seeded random set plus one curve through the middle of the
domain (so initial_state, which lives within 0.1 of a curve and 0.3 of the
centre, is not zero); pumps and drains sit where the state is positive.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from ads_mpi_amd import _lib, adss  # noqa: E402

TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def inputs(orc, seed, dt, cN=6, cL=8):
    r = orc.Oil.random(seed, cN=cN - 1, cL=cL)
    mid = np.linspace(0.25, 0.75, cL)
    cx = np.concatenate([r.cx, mid])
    cy = np.concatenate([r.cy, 0.5 + 0.02 * np.sin(7 * mid)])
    cz = np.concatenate([r.cz, np.full(cL, 0.48)])
    kw = dict(pumps=[(0.5, 0.45, 0.55), (0.2, 0.7, 0.3)], drains=[(0.55, 0.5, 0.5)], mi=3.0, kq_min=1.0,
              kq_max=20.0, radius=0.3, pump_strength=1.5, drain_strength=2.0)
    oil = orc.Oil(cx, cy, cz, cN, cL, **kw)
    rhs = adss.OilRHS(dt, cx, cy, cz, cN, cL, **kw)
    return oil, rhs


@pytest.mark.parametrize("n,p", [((9, 8, 7), 2), ((6, 7, 8), 3), ((8, 6, 5), 1), ((7, 6, 8), 4)])
def test_oil_kq_cache_bitwise(orc, n, p):
    say("case oil kq", n, p)
    S = orc.Setup(n, p)
    oil, rhs = inputs(orc, 3, 1e-4)
    ads, data = adss.initialize(n, p)
    try:
        rhs.setup(data)
        L = _lib.load()
        kq = np.zeros(S.E * (p + 1) ** 3)
        _lib.check(L.ads_oil_get_kq(data.handle, kq.ctypes.data_as(adss._dp)), data.handle)
        ref = S.oil_kq_cache(oil)
        assert ref.max() > ref.min()
        assert np.array_equal(kq, ref)
    finally:
        adss.Cleanup(ads, data)


@pytest.mark.parametrize("n,p,steps", [((9, 8, 7), 2, 4), ((6, 7, 8), 3, 3), ((8, 6, 5), 1, 4),
                                       ((24, 24, 24), 2, 2), ((7, 6, 8), 4, 2)])
def test_oil_run_vs_oracle(orc, n, p, steps):
    dt = 1e-4
    say("case oil run", n, p)
    S = orc.Setup(n, p)
    oil, rhs = inputs(orc, 5, dt)
    kq = S.oil_kq_cache(oil)
    ads, data = adss.initialize(n, p)
    try:
        l2 = adss.Step(0, rhs, ads, data)
        c, l2o = S.oil_project(oil)
        say("projection rel err", rel(data.F, c), "l2", l2, l2o)
        assert l2o > 0
        assert rel(data.F, c) <= TOL
        assert abs(l2 - l2o) <= TOL * l2o
        dro_tot = 0.0
        for it in range(1, steps + 1):
            rhs.t += dt
            l2 = adss.Step(it, rhs, ads, data)
            c, l2o, dro = S.oil_step(oil, kq, c, dt)
            dro_tot += dro
            say("step", it, "rel err", rel(data.F, c), "l2", abs(l2 - l2o) / l2o, "drained", rhs.drained, dro_tot)
            assert rel(data.F, c) <= TOL
            assert abs(l2 - l2o) <= TOL * l2o
            assert dro_tot > 0 and abs(rhs.drained - dro_tot) <= TOL * dro_tot
    finally:
        adss.Cleanup(ads, data)


def test_oil_64_cubed_realistic_inputs(orc):
    """The oil problem at 64^3 elements, p = 2, with the benchmark's inputs
    (30 random curves of 16 points, seed 1; two pumps, two drains; mi = 10,
    Kq 1..1000, radius 0.15, dt = 1e-9): the kq cache bit-exact, the
    projection and two steps (coefficients, l2norm, drained) within 1e-12."""
    n, p, dt = (65, 65, 65), 2, 1e-9
    say("case oil 64^3")
    S = orc.Setup(n, p)
    kw = dict(pumps=[(0.3, 0.3, 0.5), (0.7, 0.7, 0.5)], drains=[(0.5, 0.5, 0.2), (0.5, 0.5, 0.8)], mi=10.0,
              kq_min=1.0, kq_max=1000.0, radius=0.15, pump_strength=1.0, drain_strength=1.0)
    oil = orc.Oil.random(1, cN=30, cL=16, **kw)
    rhs = adss.OilRHS(dt, oil.cx, oil.cy, oil.cz, 30, 16, **kw)
    kq_ref = S.oil_kq_cache(oil)
    ads, data = adss.initialize(n, p)
    try:
        l2 = adss.Step(0, rhs, ads, data)
        L = _lib.load()
        kq = np.zeros(S.E * (p + 1) ** 3)
        _lib.check(L.ads_oil_get_kq(data.handle, kq.ctypes.data_as(adss._dp)), data.handle)
        assert kq_ref.max() > kq_ref.min()
        assert np.array_equal(kq, kq_ref)
        c, l2o = S.oil_project(oil)
        say("projection rel err", rel(data.F, c))
        assert l2o > 0 and rel(data.F, c) <= TOL and abs(l2 - l2o) <= TOL * l2o
        dro_tot = 0.0
        for it in (1, 2):
            rhs.t += dt
            l2 = adss.Step(it, rhs, ads, data)
            c, l2o, dro = S.oil_step(oil, kq_ref, c, dt)
            dro_tot += dro
            say("step", it, "rel err", rel(data.F, c), "drained", rhs.drained, dro_tot)
            assert rel(data.F, c) <= TOL and abs(l2 - l2o) <= TOL * l2o
            assert dro_tot > 0 and abs(rhs.drained - dro_tot) <= TOL * dro_tot
    finally:
        adss.Cleanup(ads, data)


def test_oil_entry_points_reject_misuse():
    ads, data = adss.initialize(9, 2)
    L = _lib.load()
    h = data.handle
    try:
        assert L.ads_oil_step(h, 1e-4, None, None) == _lib.ADS_ERR_INVALID  # before ads_oil_setup
        assert L.ads_oil_project(h, None) == _lib.ADS_ERR_INVALID
        x = np.zeros(4)
        pm = (adss.ctypes.c_double * 6)(10.0, 1.0, 1000.0, 0.0, 1.0, 1.0)  # radius 0
        d = x.ctypes.data_as(adss._dp)
        assert L.ads_oil_setup(h, d, d, d, 1, 4, None, 0, None, 0, pm) == _lib.ADS_ERR_INVALID
    finally:
        adss.Cleanup(ads, data)
