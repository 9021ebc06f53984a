"""The Fortran drop-in (fortran/: Setup, Interfaces, ADSS, projection_engine,
plot, vtk over the C ABI) driving the heat problem, against the oracle: the
printed l2norms, the final coefficients and the step<iter>.vti files that
PrintSolution writes (the spline on the 31^3 grid, src/ADS.F90:955-989)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "fortran", "_build", "heat_driver")
REF_PROBLEMS = "/root/reference/problems"
BUILT = os.path.join(ROOT, "oracle", "_ref", "problems")


def _has_gpu():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        pytest.skip("fortran/_build/heat_driver not built (needs flang)")
    return DRIVER


def test_driver_fails_loudly_without_gpu(driver, tmp_path):
    if _has_gpu():
        pytest.skip("a GPU is present")
    res = subprocess.run([driver, "17", "2", "2", "1e-5", "1"], capture_output=True, text=True, cwd=tmp_path)
    assert res.returncode != 0
    assert "ADS error" in res.stdout


def test_driver_usage_error(driver, tmp_path):
    res = subprocess.run([driver], capture_output=True, text=True, cwd=tmp_path)
    assert res.returncode != 0
    assert "usage" in res.stdout


@pytest.mark.gpu
def test_fortran_heat_run_matches_oracle(driver, orc, tmp_path):
    """heat_driver 17 2 5 1e-5 (16^3 elements, p = 2, projection of prod sin,
    5 explicit steps): per-step l2norm and final coefficients vs the oracle."""
    out = tmp_path / "c.bin"
    res = subprocess.run([driver, "17", "2", "5", "1e-5", "1", str(out)], capture_output=True, text=True,
                         timeout=120, cwd=tmp_path)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    l2 = [float(line.split()[-1]) for line in res.stdout.splitlines() if "L2 norm" in line]
    S = orc.Setup(17, 2)
    c, l2o = S.project(orc.F0_SINPROD, orc.MODE_FAITHFUL)
    ref = [l2o]
    states = [c]
    for _ in range(5):
        c, l2s = S.step(c, 1e-5, mode=orc.MODE_FAITHFUL)
        ref.append(l2s)
        states.append(c)
    assert len(l2) == 6
    for a, b in zip(l2, ref):
        assert abs(a - b) <= 1e-12 * abs(b)
    got = np.fromfile(out, dtype=np.float64)
    assert np.linalg.norm(got - c) <= 1e-12 * np.linalg.norm(c)

    # PrintSolution's files: step0.vti .. step5.vti, each the oracle spline of
    # that step's coefficients on the 31^3 grid, printed with 10 decimals
    U = orc.knots(17, 2)
    t = np.arange(31) / 30.0
    for it in (0, 3, 5):
        lines = (tmp_path / ("step%d.vti" % it)).read_text().splitlines()
        assert lines[:6] == ['<?xml version="1.0"?>', '<VTKFile type="ImageData" version="0.1">',
                             '  <ImageData WholeExtent="0    30 0    30 0    30" origin="0 0 0" spacing="1 1 1">',
                             '    <Piece Extent="0    30 0    30 0    30">', '      <PointData Scalars="Result">',
                             '        <DataArray Name="Result" type="Float32">']
        assert lines[-5:] == ['        </DataArray>', '      </PointData>', '    </Piece>', '  </ImageData>',
                              '</VTKFile>']
        vals = np.array([float(v) for v in lines[6:-5]])
        assert vals.size == 31 ** 3 and all(v.startswith("          ") for v in lines[6:8])
        cf = np.ascontiguousarray(states[it])
        exp = np.array([orc.lib().orc_eval_spline(0, orc._d(U), 2, 17, orc._d(U), 2, 17, orc._d(U), 2, 17,
                                                  orc._d(cf), x, y, z)
                        for z in t for y in t for x in t])
        assert np.max(np.abs(vals - exp)) <= 6e-11


@pytest.mark.gpu
def test_fortran_oil_run_matches_oracle(orc, tmp_path):
    """oil_driver 9 2 3 1e-7 seed 1, one pump, one drain (8^3 elements,
    p = 2; the oil problem's constants): per-step l2norm, final coefficients
    and the drained total vs the oracle with the same (seeded) curves."""
    drv = os.path.join(ROOT, "fortran", "_build", "oil_driver")
    if not os.path.exists(drv):
        pytest.skip("fortran/_build/oil_driver not built (needs flang)")
    out = tmp_path / "c.bin"
    pump, drain = (0.5, 0.45, 0.55), (0.55, 0.5, 0.5)
    args = [drv, "9", "2", "3", "1e-7", "1", "1", *map(str, pump), "1", *map(str, drain), str(out)]
    res = subprocess.run(args, capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    l2 = [float(line.split()[-1]) for line in res.stdout.splitlines() if "L2 norm" in line]
    drained = [float(line.split()[-1]) for line in res.stdout.splitlines() if line.startswith("drained:")][0]
    S = orc.Setup(9, 2)
    oil = orc.Oil.random(1, pumps=[pump], drains=[drain])
    kq = S.oil_kq_cache(oil)
    c, l2o = S.oil_project(oil)
    ref, dr_tot = [l2o], 0.0
    for _ in range(3):
        c, l2s, dr = S.oil_step(oil, kq, c, 1e-7)
        ref.append(l2s)
        dr_tot += dr
    assert len(l2) == 4 and ref[0] > 0
    for a, b in zip(l2, ref):
        assert abs(a - b) <= 1e-12 * abs(b)
    got = np.fromfile(out, dtype=np.float64)
    assert np.linalg.norm(got - c) <= 1e-12 * np.linalg.norm(c)
    assert abs(drained - dr_tot) <= 1e-12 * max(abs(dr_tot), 1e-300)


@pytest.mark.parametrize("name,main", [("heat", "heat_main"), ("eriksson", "heat_main"),
                                       ("L2Projection", "l2proj_main")])
def test_reference_problem_sources_compile_unchanged(name, main, tmp_path):
    """The reference's own problems/<name>/input_data.F90 and RHS_eq.F90
    (which import Setup, projection_engine's global2local and math) compile
    unchanged against the drop-in modules and link to the HIP library; on a
    machine without a GPU the program then fails loudly at initialize."""
    src = os.path.join(REF_PROBLEMS, name)
    if not os.path.isdir(src):
        pytest.skip("the reference's sources are not present")
    out = tmp_path / name
    res = subprocess.run(["make", "-C", os.path.join(ROOT, "fortran"), "problem", "PROB=" + src, "MAIN=" + main,
                          "PDIR=" + str(out)], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    exe = out / main
    assert exe.exists()
    if _has_gpu():
        return
    args = ["9", "2", "1", "1e-5", "1", "1", "1"] if main == "heat_main" else ["9", "8", "7", "2", "1", "1", "1"]
    run = subprocess.run([str(exe), *args], capture_output=True, text=True, cwd=tmp_path,
                         env=dict(os.environ, ADS_NO_VTK="1"))
    assert run.returncode != 0
    assert "initialize: ADS error" in run.stdout


def _built(name, main):
    exe = os.path.join(BUILT, name, main)
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/problems/%s not built (needs the reference's sources and flang)" % name)
    return exe


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["heat", "eriksson"])
def test_reference_heat_problem_on_the_dropin(orc, name, tmp_path):
    """The reference's heat (and eriksson) problem, its input_data.F90 and
    RHS_eq.F90 unchanged, on the drop-in: heat_main 17 2 5 1e-5 1 1 1 (16^3
    elements, p = 2, the problem's own initial state 2 bump3d(0.05, 0.4), 5
    explicit steps).  Every printed l2norm against the oracle's faithful run."""
    exe = _built(name, "heat_main")
    res = subprocess.run([exe, "17", "2", "5", "1e-5", "1", "1", "1"], capture_output=True, text=True, timeout=120,
                         cwd=tmp_path, env=dict(os.environ, ADS_NO_VTK="1"))
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    l2 = [float(line.split()[-1]) for line in res.stdout.splitlines() if "L2 norm" in line]
    S = orc.Setup(17, 2)
    c, l2o = S.project(orc.F0_BUMP3D, orc.MODE_FAITHFUL)
    ref = [l2o]
    for _ in range(5):
        c, l2s = S.step(c, 1e-5, mode=orc.MODE_FAITHFUL)
        ref.append(l2s)
    assert len(l2) == 6
    for a, b in zip(l2, ref):
        assert abs(a - b) <= 1e-12 * abs(b)


@pytest.mark.gpu
def test_reference_l2projection_on_the_dropin(tmp_path):
    """The reference's L2Projection problem (f = 1), sources unchanged, on the
    drop-in: its own check (problems/L2Projection/main.F90:60-81) passes."""
    exe = _built("L2Projection", "l2proj_main")
    res = subprocess.run([exe, "9", "11", "7", "2", "1", "1", "1"], capture_output=True, text=True, timeout=120,
                         cwd=tmp_path)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    lines = res.stdout.splitlines()
    l2 = [float(line.split()[-1]) for line in lines if "L2 norm" in line]
    assert len(l2) == 1 and abs(l2[0] - 1.0) <= 1e-12
    assert any(line.strip() == "T" for line in lines)
    assert not any("not OK" in line for line in lines)
