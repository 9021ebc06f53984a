"""Output path on the host (no GPU): the per-direction tables behind
ads_eval_grid (PrintSolution -> SaveSplinePlot -> EvalSpline) and the VTK
text layout of VtkOutput (src/vtk.F90:15-77)."""
import ctypes

import numpy as np
import pytest

from ads_mpi_amd import _lib, adss


@pytest.mark.parametrize("n,p,res,lo,hi", [(17, 2, 31, 0.0, 1.0), (10, 1, 31, 0.0, 1.0), (20, 3, 31, 0.0, 1.0),
                                           (9, 4, 31, 0.0, 1.0), (258, 3, 31, 0.0, 1.0), (40, 3, 7, 0.1, 0.9),
                                           (33, 2, 101, 0.0, 1.0)])
def test_grid_tables_match_oracle_bitwise(orc, n, p, res, lo, hi):
    """first = FindSpan - p (integers, bit-exact) and the p+1 values of
    DersBasisFuns at u = (1-t) lo + t hi, t = i/(res-1) (math.F90:44-51)."""
    L = _lib.load()
    first = np.zeros(res, dtype=np.int32)
    vals = np.zeros((res, p + 1))
    _lib.check(L.ads_host_grid_tables(n, p, res, lo, hi, first.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    U = orc.knots(n, p)
    for i in range(res):
        t = i / float(res - 1)
        u = (1.0 - t) * lo + t * hi
        span = orc.find_span(n, p, u, U)
        assert first[i] == span - p
        ders = orc.ders_basis_funs(span, u, p, 0, U)
        assert np.array_equal(vals[i], ders[0])


def test_grid_tables_reject_bad_arguments():
    L = _lib.load()
    f = np.zeros(4, dtype=np.int32)
    v = np.zeros(12)
    rc = L.ads_host_grid_tables(10, 2, 1, 0.0, 1.0, f.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == _lib.ADS_ERR_INVALID


def test_vtk_text_layout(tmp_path):
    """VtkOutput's file: XML header, extents written as "0 " I5 " 0 " I5 " 0 " I5,
    then one F30.10 value per line (left-adjusted after 10 blanks), x fastest."""
    params = adss.PlotParams(resx=3, resy=2, resz=2)
    vals = np.arange(12, dtype=np.float64).reshape(2, 2, 3) * 0.25 - 1.0
    adss.VtkOutput(tmp_path / "s", vals, params)
    lines = (tmp_path / "s.vti").read_text().splitlines()
    ext = "0     2 0     1 0     1"
    assert lines[:6] == ['<?xml version="1.0"?>', '<VTKFile type="ImageData" version="0.1">',
                         '  <ImageData WholeExtent="%s" origin="0 0 0" spacing="1 1 1">' % ext,
                         '    <Piece Extent="%s">' % ext, '      <PointData Scalars="Result">',
                         '        <DataArray Name="Result" type="Float32">']
    assert lines[6] == "          -1.0000000000"
    assert lines[7] == "          -0.7500000000"
    assert lines[6 + 11] == "          1.7500000000"
    assert lines[-5:] == ['        </DataArray>', '      </PointData>', '    </Piece>', '  </ImageData>',
                          '</VTKFile>']
