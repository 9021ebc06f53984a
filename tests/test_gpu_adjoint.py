"""Differentiation form on the GPU: the checkpointed torch.autograd rollout
(ads_mpi_amd/autograd.py over ads_step_d / ads_step_adjoint_d).

  * against the oracle's FP64 adjoint (reference gradient) at a size the CPU
    finishes in seconds, <= 1e-11 relative;
  * against central finite differences (eps = 1e-6) on 32 random directions
    at 96^3 elements, p = 2 (50 steps, checkpoint 10,
    lognormal(0, 0.5) diffusivity from seed 3, loss 0.5 ||c_T - c_obs||^2 with
    c_obs a forward run under a seed-4 diffusivity).  The directions are
    standard normal in every component (not normalised): a unit vector over
    ~10^6 unknowns would move each one by ~1e-9, about where the rounding of
    an FP64 rollout (~1e-12 per component here) already limits the
    difference quotient to ~1e-3.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from ads_mpi_amd import _lib, adss  # noqa: E402
from ads_mpi_amd.autograd import HeatSolver, heat_rollout  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_gradient_vs_oracle_adjoint(orc):
    n, p, dt, steps, ck = 25, 2, 1e-5, 6, 4
    say("case adjoint vs oracle", n, p, steps)
    S = orc.Setup(n, p)
    solver = HeatSolver(n, p)
    try:
        c0n = adss.uniform(11, S.D, -1.0, 1.0)
        kn = adss.lognormal(3, S.E, 0.0, 0.5)
        obs = adss.uniform(12, S.D, -1.0, 1.0)
        c0 = torch.from_numpy(c0n).cuda().requires_grad_(True)
        kap = torch.from_numpy(kn).cuda().requires_grad_(True)
        cT = heat_rollout(c0, kap, solver, dt, steps, ck)
        loss = 0.5 * ((cT - torch.from_numpy(obs).cuda()) ** 2).sum()
        loss.backward()
        # oracle: forward trajectory then the adjoint steps backward
        traj = [c0n]
        for _ in range(steps):
            traj.append(S.step(traj[-1], dt, kappa=kn)[0])
        lam = traj[-1] - obs
        dk = np.zeros(S.E)
        for j in range(steps - 1, -1, -1):
            lam = S.step_adjoint(traj[j], lam, dt, kappa=kn, dkappa=dk)
        assert abs(float(loss) - 0.5 * np.sum((traj[-1] - obs) ** 2)) <= 1e-12 * float(loss)
        gc = c0.grad.cpu().numpy()
        gk = kap.grad.cpu().numpy()
        assert np.linalg.norm(gc - lam) <= 1e-11 * np.linalg.norm(lam)
        assert np.linalg.norm(gk - dk) <= 1e-11 * np.linalg.norm(dk)
    finally:
        solver.close()


def test_gradient_vs_finite_differences_96cubed_p2():
    n, p, dt, steps, ck = 97, 2, 1e-7, 50, 10
    say("case FD 96^3 p=2")
    solver = HeatSolver(n, p)
    try:
        D, E = solver.D, solver.E
        c0n = adss.uniform(3, D, -1.0, 1.0)
        k_true = torch.from_numpy(adss.lognormal(4, E, 0.0, 0.5)).cuda()
        kap0 = torch.from_numpy(adss.lognormal(3, E, 0.0, 0.5)).cuda()
        c0 = torch.from_numpy(c0n).cuda()
        with torch.no_grad():
            c_obs = heat_rollout(c0, k_true, solver, dt, steps, ck)

        def final(c, k):
            with torch.no_grad():
                return heat_rollout(c, k, solver, dt, steps, ck)

        cg = c0.clone().requires_grad_(True)
        kg = kap0.clone().requires_grad_(True)
        cT = heat_rollout(cg, kg, solver, dt, steps, ck)
        loss = 0.5 * ((cT - c_obs) ** 2).sum()
        loss.backward()
        g_c, g_k = cg.grad, kg.grad
        # rounding floor of the forward rollout: the same input twice (the
        # assembly's atomic adds make the summation order vary)
        noise = float(torch.linalg.norm(final(c0, kap0) - final(c0, kap0)))
        gen = torch.Generator(device="cpu").manual_seed(7)
        eps = 1e-6
        worst = 0.0
        for i in range(32):
            vc = torch.randn(D, generator=gen, dtype=torch.float64).cuda()
            vk = torch.randn(E, generator=gen, dtype=torch.float64).cuda()
            # central difference of L = 0.5 ||c_T - c_obs||^2, formed as
            # L+ - L- = 0.5 (r+ - r-) . (r+ + r-) so that the rounding of the
            # two sums of squares does not swamp the O(eps) difference
            rp = final(c0 + eps * vc, kap0 + eps * vk) - c_obs
            rm = final(c0 - eps * vc, kap0 - eps * vk) - c_obs
            fd = float(0.5 * ((rp - rm) * (rp + rm)).sum()) / (2 * eps)
            ad = float((g_c * vc).sum() + (g_k * vk).sum())
            err = abs(fd - ad) / max(abs(ad), 1e-30)
            worst = max(worst, err)
        say("FD worst relative error", worst, "rollout rounding floor (L2 of a repeat difference)", noise)
        assert worst <= 1e-5
    finally:
        solver.close()


def test_fp32_forms_gradient_384x384x256_p3():
    """384x384x256 elements, p = 3, lognormal(0, 0.5) diffusivity from seed 5,
    40 steps, checkpoint 8, dt = 5e-9 (stable for the diffusivity's range):
    the forward in FP32 + refinement and the adjoint in FP32 with FP64
    mass-residual corrections (FP32 LU sweeps, then segmented solves),
    against the all-FP64 rollout and gradient."""
    n, p, dt, steps, ck = (386, 386, 258), 3, 5e-9, 40, 8
    say("case fp32 forms gradient 384x384x256 p=3")
    solver = HeatSolver(n, p)
    try:
        D, E = solver.D, solver.E
        c0 = torch.from_numpy(adss.uniform(5, D, -1.0, 1.0)).cuda()
        kap = torch.from_numpy(adss.lognormal(5, E, 0.0, 0.5)).cuda()
        c_obs = torch.from_numpy(adss.uniform(6, D, -1.0, 1.0)).cuda()
        res = {}
        for name, kw in (("fp64", dict(prec=_lib.ADS_PREC_FP64)),
                         ("fp32", dict(prec=_lib.ADS_PREC_FP32_IR_KRON, refine=2,
                                       adjoint_prec=_lib.ADS_PREC_FP32_IR_KRON, adjoint_refine=2)),
                         ("fp32seg", dict(prec=_lib.ADS_PREC_FP32_IR_SEG, refine=2,
                                          adjoint_prec=_lib.ADS_PREC_FP32_IR_SEG, adjoint_refine=2))):
            cg = c0.clone().requires_grad_(True)
            kg = kap.clone().requires_grad_(True)
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            torch.cuda.synchronize()
            ev[0].record()
            cT = heat_rollout(cg, kg, solver, dt, steps, ck, **kw)
            ev[1].record()
            loss = 0.5 * ((cT - c_obs) ** 2).sum()
            loss.backward()
            ev[2].record()
            res[name] = (cT.detach(), cg.grad, kg.grad)
            torch.cuda.synchronize()
            t_fwd, t_bwd = ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2])
            # forward throughput, and the HBM rate of its minimum traffic per
            # step (read c and kappa, write c': 24 B per DOF) against 8 TB/s
            # gradient/forward: time of loss.backward() (segment recomputation
            # + adjoint steps) over the time of the forward rollout
            say(name, "forward %.2f ms/step, %.3g DOF/s, %.1f GB/s (%.3f of 8 TB/s); gradient/forward %.2f"
                % (t_fwd / steps, D * steps / (t_fwd * 1e-3), 24.0 * D * steps / (t_fwd * 1e-3) / 1e9,
                   24.0 * D * steps / (t_fwd * 1e-3) / 8e12, t_bwd / t_fwd))
        for name in ("fp32", "fp32seg"):
            e_state = float(torch.linalg.norm(res[name][0] - res["fp64"][0]) / torch.linalg.norm(res["fp64"][0]))
            e_gc = float(torch.linalg.norm(res[name][1] - res["fp64"][1]) / torch.linalg.norm(res["fp64"][1]))
            e_gk = float(torch.linalg.norm(res[name][2] - res["fp64"][2]) / torch.linalg.norm(res["fp64"][2]))
            say(name, "rel err: state", e_state, "grad c0", e_gc, "grad kappa", e_gk)
            # each refined step is within 1e-10 of the FP64 step and the step is
            # non-expansive; measured 4.4e-12 after 40 steps
            assert e_state <= 5e-11
            # FP32 arithmetic in the adjoint assembly: ~1e-6 per step; measured
            # 1.2e-5 (c0) and 2.7e-6 (kappa) after 40 steps
            assert e_gc <= 5e-5
            assert e_gk <= 3e-5
    finally:
        solver.close()
