"""Output path on the GPU: ads_eval_grid (EvalSpline on the PlotParams grid,
src/basis.F90:323-369, src/plot.F90:97-145) bit-exact against the oracle's
EvalSpline, which is pinned bitwise to the flang-built reference; and
PrintSolution's step<iter>.vti file."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from ads_mpi_amd import adss  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def oracle_grid(orc, n, p, c, params):
    U = [orc.knots(n[d], p) for d in range(3)]
    axes = []
    for lo, hi, res in ((params.startx, params.endx, params.resx), (params.starty, params.endy, params.resy),
                        (params.startz, params.endz, params.resz)):
        t = np.arange(res) / float(res - 1)
        axes.append((1.0 - t) * lo + t * hi)
    cf = np.ascontiguousarray(c)
    L = orc.lib()
    out = np.empty((params.resz, params.resy, params.resx))
    for k, z in enumerate(axes[2]):
        for j, y in enumerate(axes[1]):
            for i, x in enumerate(axes[0]):
                out[k, j, i] = L.orc_eval_spline(0, orc._d(U[0]), p, n[0], orc._d(U[1]), p, n[1], orc._d(U[2]), p,
                                                 n[2], orc._d(cf), x, y, z)
    return out


@pytest.mark.parametrize("n,p", [((17, 17, 17), 2), ((9, 12, 7), 1), ((20, 14, 11), 3), ((10, 9, 12), 4)])
def test_eval_grid_vs_oracle(orc, n, p):
    say("case eval grid", n, p)
    ads, data = adss.initialize(n, p)
    try:
        data.F = adss.uniform(13, data.D, -1.0, 1.0)
        for params in (adss.PlotParams(), adss.PlotParams(0.05, 0.95, 0.0, 1.0, 0.1, 0.7, 13, 9, 11)):
            got = adss.EvalGrid(data, params)
            exp = oracle_grid(orc, n, p, data.F, params)
            say("max abs diff", np.max(np.abs(got - exp)))
            # same operations in the same order without fused multiply-adds:
            # bit-exact with the oracle (and so with the reference's EvalSpline)
            assert np.array_equal(got, exp)
    finally:
        adss.Cleanup(ads, data)


def test_print_solution_file(orc, tmp_path):
    n, p = (17, 17, 17), 2
    ads, data = adss.initialize(n, p)
    try:
        adss.Step(0, adss.HeatRHS(Dt=1e-5, f0=adss.ADS_F0_SINPROD), ads, data)
        adss.PrintSolution(0, 0.0, ads, data, directory=str(tmp_path))
        lines = (tmp_path / "step0.vti").read_text().splitlines()
        vals = np.array([float(v) for v in lines[6:-5]]).reshape(31, 31, 31)
        exp = oracle_grid(orc, n, p, data.F, adss.PlotParams())
        assert np.max(np.abs(vals - exp)) <= 6e-11
        # the projection of prod sin(pi x_i) is close to it at the grid points
        t = np.arange(31) / 30.0
        s = np.sin(np.pi * t)
        assert np.max(np.abs(vals - s[:, None, None] * s[None, :, None] * s[None, None, :])) < 2e-3
    finally:
        adss.Cleanup(ads, data)
