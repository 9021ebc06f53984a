"""Oracle heat step (Form3DRHS + FormUn + the three 1-D solves).

Pins:
  * L2Projection known answer: f = 1 projects to all-ones coefficients with
    l2norm 1 (problems/L2Projection/main.F90:60-81);
  * Kronecker identity: Q = p+1 Gauss points integrate the degree-2p
    integrands exactly, so the assembled heat right-hand side equals
    (Mz(x)My(x)Mx - dt (Mz(x)My(x)Kx + Mz(x)Ky(x)Mx + Kz(x)My(x)Mx)) c with
    1-D matrices built independently (scipy BSpline + numpy Gauss points);
  * the solve equals the dense Kronecker inverse;
  * faithful (per quadrature point, per basis function, as the reference
    loops) and sum-factorised element kernels agree;
  * the adjoint step agrees with central finite differences.
"""
import importlib.util
import os

import numpy as np
import pytest

_spec = importlib.util.spec_from_file_location(
    "mlv", os.path.join(os.path.dirname(__file__), "golden", "make_lapack_vectors.py"))
mlv = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(mlv)


def kron3(Az, Ay, Ax):
    return np.kron(Az, np.kron(Ay, Ax))


def test_l2projection_known_answer(orc):
    for n, p in (((9, 11, 7), (2, 2, 2)), ((17, 17, 17), (2, 2, 2)), ((8, 8, 8), (3, 3, 3)), ((5, 6, 7), (1, 1, 1))):
        S = orc.Setup(n, p)
        for mode in (orc.MODE_FAITHFUL, orc.MODE_SUMFACT):
            c, l2 = S.project(orc.F0_CONST, mode)
            assert np.max(np.abs(c - 1.0)) < 1e-12
            assert abs(l2 - 1.0) < 1e-12


@pytest.mark.parametrize("n,p", [((9, 11, 7), 2), ((6, 7, 8), 3), ((5, 4, 6), 1), ((9, 8, 10), 4)])
def test_heat_rhs_kronecker_identity(orc, n, p):
    S = orc.Setup(n, p)
    mats = [mlv.mass_matrix(nd, p) for nd in n]
    Mx, Kx = mats[0]
    My, Ky = mats[1]
    Mz, Kz = mats[2]
    rng = np.random.default_rng(7)
    c = rng.standard_normal(S.D)
    dt = 3e-3
    F, l2 = S.form_rhs_heat(c, dt)
    A = kron3(Mz, My, Mx) - dt * (kron3(Mz, My, Kx) + kron3(Mz, Ky, Mx) + kron3(Kz, My, Mx))
    Fr = A @ c
    assert np.linalg.norm(F - Fr) <= 1e-13 * np.linalg.norm(Fr)
    # l2norm is the Gauss-quadrature integral of u^2 = c^T M c
    assert abs(l2 - c @ (kron3(Mz, My, Mx) @ c)) <= 1e-12 * abs(l2)
    # the solve is the Kronecker inverse
    x = S.solve(F)
    assert np.linalg.norm(kron3(Mz, My, Mx) @ x - F) <= 1e-12 * np.linalg.norm(F)
    xt = S.solve_T(F)
    assert np.linalg.norm(kron3(Mz, My, Mx).T @ xt - F) <= 1e-12 * np.linalg.norm(F)


def test_faithful_equals_sumfact(orc):
    S = orc.Setup((9, 10, 8), (2, 2, 2))
    rng = np.random.default_rng(3)
    c = rng.standard_normal(S.D)
    kappa = orc.lognormal(11, S.E, 0.0, 0.5)
    Ff, lf = S.form_rhs_heat(c, 1e-3, kappa=kappa, mode=orc.MODE_FAITHFUL)
    Fs, ls = S.form_rhs_heat(c, 1e-3, kappa=kappa, mode=orc.MODE_SUMFACT)
    assert np.linalg.norm(Ff - Fs) <= 1e-14 * np.linalg.norm(Ff)
    assert abs(lf - ls) <= 1e-14 * abs(lf)
    for f0 in (orc.F0_SINPROD, orc.F0_BUMP3D):
        Pf, pf = S.form_rhs_project(f0, orc.MODE_FAITHFUL)
        Ps, ps = S.form_rhs_project(f0, orc.MODE_SUMFACT)
        assert np.linalg.norm(Pf - Ps) <= 1e-14 * np.linalg.norm(Pf)
        assert abs(pf - ps) <= 1e-13 * abs(pf)


def test_faithful_solve_rebuilds_same_factors(orc):
    S = orc.Setup((9, 10, 8), (2, 2, 2))
    F = np.random.default_rng(4).standard_normal(S.D)
    assert np.array_equal(S.solve(F, faithful=True), S.solve(F, faithful=False))


def test_sinprod_projection_l2(orc):
    """||sin(pi x) sin(pi y) sin(pi z)||^2 on [0,1]^3 = 1/8; the quadrature
    value converges to it with the mesh."""
    S = orc.Setup(17, 2)
    _, l2 = S.project(orc.F0_SINPROD)
    assert abs(l2 - 0.125) < 1e-9


def test_adjoint_matches_finite_differences(orc):
    S = orc.Setup((7, 8, 6), (2, 2, 2))
    rng = np.random.default_rng(5)
    c0 = rng.standard_normal(S.D)
    kappa = orc.lognormal(3, S.E, 0.0, 0.5)
    target = rng.standard_normal(S.D)
    dt = 2e-3
    steps = 3

    def loss(c0_, kap_):
        c = c0_
        for _ in range(steps):
            c, _ = S.step(c, dt, kappa=kap_)
        return 0.5 * np.sum((c - target) ** 2), c

    L0, cT = loss(c0, kappa)
    traj = [c0]
    c = c0
    for _ in range(steps):
        c, _ = S.step(c, dt, kappa=kappa)
        traj.append(c)
    lam = traj[-1] - target
    dk = np.zeros(S.E)
    for n in range(steps - 1, -1, -1):
        lam = S.step_adjoint(traj[n], lam, dt, kappa=kappa, dkappa=dk)
    for _ in range(4):
        v = rng.standard_normal(S.D)
        w = rng.standard_normal(S.E)
        eps = 1e-6
        fd_c = (loss(c0 + eps * v, kappa)[0] - loss(c0 - eps * v, kappa)[0]) / (2 * eps)
        fd_k = (loss(c0, kappa + eps * w)[0] - loss(c0, kappa - eps * w)[0]) / (2 * eps)
        assert abs(fd_c - lam @ v) <= 1e-6 * max(1.0, abs(fd_c))
        assert abs(fd_k - dk @ w) <= 1e-6 * max(1.0, abs(fd_k))


def test_generators(orc):
    u = orc.uniform(0, 200000, -0.01, 0.01)
    assert u.min() >= -0.01 and u.max() < 0.01
    assert abs(u.mean()) < 1e-4 and abs(u.std() - 0.02 / np.sqrt(12)) < 1e-4
    k = orc.lognormal(3, 200000, 0.0, 0.5)
    z = np.log(k)
    assert abs(z.mean()) < 5e-3 and abs(z.std() - 0.5) < 5e-3
    assert np.array_equal(orc.uniform(9, 10, 0, 1), orc.uniform(9, 10, 0, 1))
    assert not np.array_equal(orc.uniform(9, 10, 0, 1), orc.uniform(10, 10, 0, 1))


def indep_mkbbt(n, p):
    """M, K, B, B^T of one direction built independently of the oracle (scipy
    BSpline at numpy Gauss points): B(i, j) = int N_i' N_j, the insertion
    order of MKBBT (projection_engine.F90:123-128)."""
    U = mlv.knots(n, p)
    xg, wg = np.polynomial.legendre.leggauss(p + 1)
    N = n + 1
    from scipy.interpolate import BSpline
    vals_d = []
    for j in range(N):
        cc = np.zeros(N)
        cc[j] = 1.0
        s = BSpline(U, cc, p, extrapolate=False)
        vals_d.append((s, s.derivative()))
    M = np.zeros((N, N)); K = np.zeros((N, N)); B = np.zeros((N, N))
    for e in range(n - p + 1):
        a, b = U[p + e], U[p + e + 1]
        x = (xg + 1.0) * (b - a) / 2.0 + a
        w = wg * (b - a) / 2.0
        v = np.array([np.nan_to_num(s(x)) for s, _ in vals_d])
        d = np.array([np.nan_to_num(t(x)) for _, t in vals_d])
        M += (v * w) @ v.T
        K += (d * w) @ d.T
        B += (d * w) @ v.T
    return M, K, B, B.T


MIXES = [[1.0, 2e-3, 0.0, 0.0], [1.0, 0.0, 0.02, 0.0], [0.5, 1e-3, 0.0, 0.02]]


def test_solve_mix_is_dense_kronecker_inverse(orc):
    n, p = (7, 8, 6), 2
    S = orc.Setup(n, p)
    mk = [indep_mkbbt(nd, p) for nd in n]
    F = np.random.default_rng(31).standard_normal(S.D)
    for mix in MIXES:
        A = [sum(m * X for m, X in zip(mix, mk[d])) for d in range(3)]
        x = S.solve_mix(F, mix)
        assert np.linalg.norm(kron3(A[2], A[1], A[0]) @ x - F) <= 1e-12 * np.linalg.norm(F)


def test_multistep_against_dense(orc):
    """MultiStep (ADS.F90:391-429) with the heat integrand: substep s solves
    (A_s (x) A_s (x) A_s) c_s = M c_n - dt K c_{s-1}, where the mass part uses
    the point values of c_n (Un) and the stiffness part the gradients of the
    latest coefficients (dUn); l2norm = c_n^T M c_n."""
    n, p, dt = (7, 8, 6), 2, 2e-3
    S = orc.Setup(n, p)
    mk = [indep_mkbbt(nd, p) for nd in n]
    (Mx, Kx, _, _), (My, Ky, _, _), (Mz, Kz, _, _) = mk
    M3 = kron3(Mz, My, Mx)
    K3 = kron3(Mz, My, Kx) + kron3(Mz, Ky, Mx) + kron3(Kz, My, Mx)
    c = np.random.default_rng(32).standard_normal(S.D)
    out, l2 = S.multistep(c, MIXES, dt)
    prev = c
    for mix in MIXES:
        A = [sum(m * X for m, X in zip(mix, mk[d])) for d in range(3)]
        prev = np.linalg.solve(kron3(A[2], A[1], A[0]), M3 @ c - dt * (K3 @ prev))
    assert np.linalg.norm(out - prev) <= 1e-11 * np.linalg.norm(prev)
    assert abs(l2 - c @ (M3 @ c)) <= 1e-12 * abs(l2)
    # mass matrices in every substep: the first substep is the heat step
    first, _ = S.multistep(c, [[1.0, 0.0, 0.0, 0.0]] * 3, dt)
    one, _ = S.step(c, dt, mode=orc.MODE_FAITHFUL)
    two = S.solve(M3 @ c - dt * (K3 @ one))
    three = S.solve(M3 @ c - dt * (K3 @ two))
    assert np.linalg.norm(first - three) <= 1e-12 * np.linalg.norm(three)
