"""The Fortran drop-in (fortran/: Setup, Interfaces, ADSS, projection_engine
over the C ABI) driving the heat problem, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "fortran", "_build", "heat_driver")


def _has_gpu():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        pytest.skip("fortran/_build/heat_driver not built (needs flang)")
    return DRIVER


def test_driver_fails_loudly_without_gpu(driver):
    if _has_gpu():
        pytest.skip("a GPU is present")
    res = subprocess.run([driver, "17", "2", "2", "1e-5", "1"], capture_output=True, text=True)
    assert res.returncode != 0
    assert "ADS error" in res.stdout


def test_driver_usage_error(driver):
    res = subprocess.run([driver], capture_output=True, text=True)
    assert res.returncode != 0
    assert "usage" in res.stdout


@pytest.mark.gpu
def test_fortran_heat_run_matches_oracle(driver, orc, tmp_path):
    """heat_driver 17 2 5 1e-5 (16^3 elements, p = 2, projection of prod sin,
    5 explicit steps): per-step l2norm and final coefficients vs the oracle."""
    out = tmp_path / "c.bin"
    res = subprocess.run([driver, "17", "2", "5", "1e-5", "1", str(out)], capture_output=True, text=True,
                         timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    l2 = [float(line.split()[-1]) for line in res.stdout.splitlines() if "L2 norm" in line]
    S = orc.Setup(17, 2)
    c, l2o = S.project(orc.F0_SINPROD, orc.MODE_FAITHFUL)
    ref = [l2o]
    for _ in range(5):
        c, l2s = S.step(c, 1e-5, mode=orc.MODE_FAITHFUL)
        ref.append(l2s)
    assert len(l2) == 6
    for a, b in zip(l2, ref):
        assert abs(a - b) <= 1e-12 * abs(b)
    got = np.fromfile(out, dtype=np.float64)
    assert np.linalg.norm(got - c) <= 1e-12 * np.linalg.norm(c)
