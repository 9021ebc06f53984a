"""The reference's finer-grained entry points on the GPU, through the C ABI
(via the Python mirror in ads_mpi_amd/adss.py):
  * FormUn / Form3DRHS / Sub_Step (src/projection_engine.F90:306-441,
    154-284; src/ADS.F90:471-794): a MultiStep written out as the reference
    writes it (FormUn(s); Sub_Step(mix(:, s), s, s) for s = 1..3) against
    ads_multistep and the oracle's MultiStep; Form3DRHS against the oracle's
    right-hand sides; Sub_Step with the mass mix against Step;
  * ComputeMatrix + SolveOneDirection (src/projection_engine.F90:531-544,
    src/ADS.F90:294-379) against dense solves of the matrix, for
    non-symmetric mixes, a pivoting case, graded knots and long lines.
Tolerance 1e-12 relative (FP64)."""
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


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()


MIXES = [[1.0, 2e-3, 0.0, 0.0], [1.0, 0.0, 0.02, 0.0], [0.5, 1e-3, 0.0, 0.02]]


@pytest.mark.parametrize("n,p", [((20, 17, 25), 2), ((14, 12, 16), 3)])
def test_substeps_as_the_reference_writes_them(orc, n, p):
    say("case substeps", n, p)
    S = orc.Setup(n, p)
    dt = 1e-4
    c = np.random.default_rng(43).uniform(-1, 1, S.D)
    kappa = orc.lognormal(6, S.E, 0.0, 0.5)
    ads, data = adss.initialize(n, p)
    try:
        for kap in (None, kappa):
            rhs = adss.HeatRHS(Dt=dt, kappa=kap)
            rhs.t = dt
            data.F = c
            l2s = []
            for s in range(3):
                adss.FormUn(s + 1, ads, data)
                l2s.append(adss.Sub_Step(ads, 1, MIXES[s], s + 1, s + 1, rhs, data))
            got = data.F
            data.F = c
            l2m = adss.MultiStep(1, np.array(MIXES).T, rhs, ads, data)
            ref, l2o = S.multistep(c, MIXES, dt, kappa=kap)
            say("kappa", kap is not None, "rel err vs oracle", rel(got, ref), "vs ads_multistep", rel(got, data.F))
            assert rel(got, ref) <= TOL
            assert rel(got, data.F) <= 1e-14
            # every substep's l2norm is the one of U_n (values from FormUn(1))
            for l2 in l2s:
                assert abs(l2 - l2o) <= TOL * abs(l2o)
            assert abs(l2m - l2o) <= TOL * abs(l2o)
    finally:
        adss.Cleanup(ads, data)


def test_form3drhs_and_mass_substep(orc):
    n, p, dt = (18, 21, 16), 2, 1e-4
    S = orc.Setup(n, p)
    c = np.random.default_rng(44).uniform(-1, 1, S.D)
    kappa = orc.lognormal(7, S.E, 0.0, 0.5)
    ads, data = adss.initialize(n, p)
    try:
        say("case form3drhs")
        rhs = adss.HeatRHS(Dt=dt, kappa=kappa)
        rhs.t = dt
        data.F = c
        adss.FormUn(1, ads, data)
        F, l2 = adss.Form3DRHS(ads, data, 1, 1, rhs)
        Fo, l2o = S.form_rhs_heat(c, dt, kappa=kappa, mode=orc.MODE_FAITHFUL)
        assert rel(F, Fo) <= TOL and abs(l2 - l2o) <= TOL * abs(l2o)
        assert rel(data.F, c) == 0.0  # the state is unchanged
        # the t = 0 branch: the projection of f0
        for f0 in (adss.ADS_F0_CONST, adss.ADS_F0_SINPROD, adss.ADS_F0_BUMP3D):
            r0 = adss.HeatRHS(Dt=dt, f0=f0)
            F, l2 = adss.Form3DRHS(ads, data, 1, 1, r0)
            Fo, l2o = S.form_rhs_project(f0, mode=orc.MODE_FAITHFUL)
            assert rel(F, Fo) <= TOL and abs(l2 - l2o) <= TOL * abs(l2o)
        # Sub_Step with the mass mix is Step (the same kernels; the assembly's
        # global atomic adds make the last bits differ between runs)
        l2a = adss.Sub_Step(ads, 1, [1.0, 0.0, 0.0, 0.0], 1, 1, rhs, data)
        got = data.F
        data.F = c
        l2b = adss.Step(1, rhs, ads, data)
        assert rel(got, data.F) <= 1e-14 and abs(l2a - l2b) <= 1e-14 * abs(l2b)
        exp, _ = S.step(c, dt, kappa=kappa)
        assert rel(got, exp) <= TOL
        # bad substep numbers are refused (the reference prints "wrong substep")
        with pytest.raises(_lib.AdsError):
            adss.FormUn(4, ads, data)
        with pytest.raises(_lib.AdsError):
            adss.Sub_Step(ads, 1, [1.0, 0.0, 0.0, 0.0], 4, 1, rhs, data)
    finally:
        adss.Cleanup(ads, data)


def graded_knots(n, p, power=1.5):
    ne = n - p + 1
    return np.concatenate((np.zeros(p + 1), (np.arange(1, ne) / ne) ** power, np.ones(p + 1)))


@pytest.mark.parametrize("n,p,eqnum,graded", [(17, 2, 324, False), (40, 3, 1000, True), (3, 3, 16, False),
                                              (20, 4, 77, False), (9, 1, 5, True), (1500, 3, 300, False),
                                              (130, 2, 16900, False), (30, 5, 40, False), (300, 6, 64, True),
                                              (45, 7, 120, True), (40, 8, 7, False), (24, 9, 33, False)])
def test_compute_matrix_and_solve_one_direction(orc, n, p, eqnum, graded):
    """SolveOneDirection with the matrix of ComputeMatrix, for the mass
    matrix and non-symmetric mixes, on the GPU: against dense solves; degrees
    1..9 (the reference's Gauss rules reach Q = 10), above the step's 1..4."""
    say("case SolveOneDirection", n, p, eqnum, graded)
    U = graded_knots(n, p) if graded else orc.knots(n, p)
    rng = np.random.default_rng(n + p)
    for mix in ([1.0, 0.0, 0.0, 0.0], [1.0, 2e-3, 0.02, 0.0], [0.5, 1e-3, 0.0, 0.02]):
        m = adss.ComputeMatrix(U, p, n, n - p + 1, mix)
        try:
            A = m.dense()
            RHS = np.asfortranarray(rng.standard_normal((n + 1, eqnum)))
            exp = np.linalg.solve(A, RHS)
            adss.SolveOneDirection(RHS, eqnum, n, p, m)
            err = rel(RHS, exp)
            say("mix", mix, "rel err", err)
            assert err <= TOL
        finally:
            m.close()
