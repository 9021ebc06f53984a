"""Python mirror of the reference's ADSS / Setup / Interfaces module API.

The reference (macwozni/ADS_MPI) drives a problem as

    call initialize(n, p, ads, ads_data, mierr)          ! src/ADS.F90:11
    do iter = 0, steps
       call Step(iter, ComputePointForRHS, ads, ads_data, l2norm, mierr)   ! :437
       t = t + Dt
    enddo
    call Cleanup(ads, ads_data, mierr)                   ! :802

(problems/heat/main.F90:46-62).  The same sequence works here:

    ads, ads_data = initialize(n, p)
    rhs = HeatRHS(Dt=1e-5, f0=ADS_F0_SINPROD)
    for it in range(steps + 1):
        l2norm = Step(it, rhs, ads, ads_data)
        rhs.t += rhs.Dt
    Cleanup(ads, ads_data)

The per-point Fortran callback RHS_fun (Interfaces.F90:52-81) cannot run on
the device, so integrands are *registered forms*: HeatRHS restates
problems/heat/RHS_eq.F90:35-88 (t == 0 projection of f0, t > 0 explicit heat
step) and L2ProjectionRHS restates problems/L2Projection/RHS_eq.F90:34-71.
Anything else raises, as the Fortran drop-in returns mierr != 0.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (ADS_F0_BUMP3D, ADS_F0_CONST, ADS_F0_SINPROD, ADS_PREC_FP32_IR, ADS_PREC_FP32_IR_SEG,
                   ADS_PREC_FP32_IR_KRON, ADS_PREC_FP64, ADS_PREC_FP64_KRON, AdsError, check)

__all__ = [
    "ADS_setup", "ADS_compute_data", "initialize", "Step", "MultiStep", "Sub_Step", "Cleanup", "Form3DRHS",
    "FormUn", "ComputeMatrix", "SolveOneDirection", "SparseMatrix", "HeatRHS", "OilRHS", "L2ProjectionRHS", "AdsError", "ADS_F0_CONST", "ADS_F0_SINPROD",
    "ADS_F0_BUMP3D", "ADS_PREC_FP64", "ADS_PREC_FP32_IR", "ADS_PREC_FP32_IR_KRON", "ADS_PREC_FP64_KRON", "ADS_PREC_FP32_IR_SEG", "PlotParams", "EvalGrid", "VtkOutput",
    "PrintSolution",
]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


@dataclass
class ADS_setup:
    """Host mirror of type(ADS_setup) (src/Setup.F90:29-112) on one rank."""

    n: np.ndarray
    p: np.ndarray
    nelem: np.ndarray
    m: np.ndarray
    ng: np.ndarray
    Ux: np.ndarray
    Uy: np.ndarray
    Uz: np.ndarray
    Ox: np.ndarray
    Oy: np.ndarray
    Oz: np.ndarray
    Jx: np.ndarray
    Jy: np.ndarray
    Jz: np.ndarray
    Wx: np.ndarray
    Wy: np.ndarray
    Wz: np.ndarray
    Xx: np.ndarray  # (nelem, q): Fortran X(q, nelem)
    Xy: np.ndarray
    Xz: np.ndarray
    NNx: np.ndarray  # (nelem, q, p+1, 2): Fortran NN(0:1, 0:p, q, nelem)
    NNy: np.ndarray
    NNz: np.ndarray
    # one-rank decomposition values (parallelism.F90:146-171 with rank 0 of 1)
    s: np.ndarray = field(default=None)
    ibeg: np.ndarray = field(default=None)
    iend: np.ndarray = field(default=None)
    mine: np.ndarray = field(default=None)
    maxe: np.ndarray = field(default=None)
    lnelem: np.ndarray = field(default=None)
    nrcpp: np.ndarray = field(default=None)
    tau: float = 0.0

    @property
    def dofs(self):
        return tuple(int(v) + 1 for v in self.n)


class ADS_compute_data:
    """type(ADS_compute_data) (src/Setup.F90:120-140): the solver state.

    The coefficients live on the GPU; ``F`` copies them to the host when read
    (the reference keeps them in ads_data%F, x fastest)."""

    def __init__(self, handle, D, dofs):
        self.handle = handle
        self.D = D
        self.dofs = dofs
        self.t = 0.0

    @property
    def F(self):
        out = np.empty(self.D)
        check(_lib.load().ads_get_coeffs(self.handle, _d(out)), self.handle)
        return out

    @F.setter
    def F(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if v.size != self.D:
            raise ValueError("expected %d coefficients, got %d" % (self.D, v.size))
        check(_lib.load().ads_set_coeffs(self.handle, _d(v)), self.handle)

    def coeffs3d(self):
        """Coefficients as an array [z][y][x] (x fastest in memory)."""
        nx, ny, nz = self.dofs
        return self.F.reshape(nz, ny, nx)

    def device_ptr(self):
        p = ctypes.c_void_p()
        check(_lib.load().ads_coeffs_device(self.handle, ctypes.byref(p)), self.handle)
        return p.value


class HeatRHS:
    """problems/heat/RHS_eq.F90:35-88 as a device form.

    t == 0: ret = J W v f0(X), l2norm = J W v f0^2 (projection of the initial
    state); t > 0: ret = J W (v u + Dt(-kappa grad u . grad v)), l2norm =
    J W v u^2.  ``t`` is advanced by the driver, as main.F90:57 does."""

    def __init__(self, Dt, f0=ADS_F0_BUMP3D, kappa=None, prec=ADS_PREC_FP64, refine=2):
        self.Dt = float(Dt)
        self.f0 = int(f0)
        self.kappa = kappa
        self.prec = int(prec)
        self.refine = int(refine)
        self.t = 0.0


class OilRHS:
    """problems/oil/RHS_eq.F90:76-91 with input_data.F90 as a device form:
    t > 0: ret = J W (v u + Dt (-kq exp(mi u) grad u . grad v + v (pump -
    max(0, drain u)))), l2norm = J W v u^2; t == 0: the projection of
    initial_state.  The curves (cN x cL points), pumps and drains are inputs
    (the reference draws its curves from an unseeded random_number and reads
    pumps and drains from the command line).  ``drained`` is the running total
    of J W v Dt vdrain over the steps, as the reference's module variable."""

    def __init__(self, Dt, cx, cy, cz, cN, cL, pumps=(), drains=(), mi=10.0, kq_min=1.0, kq_max=1000.0,
                 radius=0.15, pump_strength=1.0, drain_strength=1.0):
        self.Dt = float(Dt)
        self.cx = np.ascontiguousarray(cx, dtype=np.float64).ravel()
        self.cy = np.ascontiguousarray(cy, dtype=np.float64).ravel()
        self.cz = np.ascontiguousarray(cz, dtype=np.float64).ravel()
        self.cN, self.cL = int(cN), int(cL)
        self.pumps = np.ascontiguousarray(np.reshape(pumps, (-1, 3)), dtype=np.float64)
        self.drains = np.ascontiguousarray(np.reshape(drains, (-1, 3)), dtype=np.float64)
        self.params = (float(mi), float(kq_min), float(kq_max), float(radius), float(pump_strength),
                       float(drain_strength))
        self.t = 0.0
        self.drained = 0.0
        self._ready_on = None

    @classmethod
    def with_random_curves(cls, Dt, seed, cN=30, cL=16, **kw):
        """Curves by InitInputData's random walk (input_data.F90:139-165) with
        a seeded SplitMix64 stream (ads_oil_curves)."""
        L = _lib.load()
        cx, cy, cz = (np.zeros(cN * cL) for _ in range(3))
        check(L.ads_oil_curves(seed, cN, cL, _d(cx), _d(cy), _d(cz)))
        return cls(Dt, cx, cy, cz, cN, cL, **kw)

    def setup(self, ads_data):
        """InitInputData + PrecomputeKq on the device (ads_oil_setup)."""
        L = _lib.load()
        h = ads_data.handle
        if self._ready_on is h:
            return
        pm = ctypes.c_double * 6
        np_ = len(self.pumps)
        nd = len(self.drains)
        check(L.ads_oil_setup(h, _d(self.cx), _d(self.cy), _d(self.cz), self.cN, self.cL,
                              _d(self.pumps) if np_ else None, np_, _d(self.drains) if nd else None, nd,
                              pm(*self.params)), h)
        self._ready_on = h


class L2ProjectionRHS:
    """problems/L2Projection/RHS_eq.F90:34-71: f = 1."""

    t = 0.0
    f0 = ADS_F0_CONST


def _direction(L, h, d, n, p):
    q = p + 1
    ne = n - p + 1
    U = np.zeros(n + p + 2)
    O = np.zeros(ne, dtype=np.int32)
    J = np.zeros(ne)
    W = np.zeros(q)
    X = np.zeros((ne, q))
    NN = np.zeros((ne, q, p + 1, 2))
    check(L.ads_get_knots(h, d, _d(U)), h)
    check(L.ads_get_basis(h, d, _i(O), _d(J), _d(W), _d(X), _d(NN)), h)
    return U, O, J, W, X, NN


def initialize(n, p):
    """ADSS::initialize (src/ADS.F90:11-107).  Returns (ads, ads_data)."""
    L = _lib.load()
    n3 = (ctypes.c_int * 3)(*([int(n)] * 3 if np.isscalar(n) else [int(v) for v in n]))
    p3 = (ctypes.c_int * 3)(*([int(p)] * 3 if np.isscalar(p) else [int(v) for v in p]))
    h = ctypes.c_void_p()
    rc = L.ads_init(n3, p3, ctypes.byref(h))
    if rc != 0:
        raise AdsError(rc, L.ads_last_error(None).decode())
    nn = np.array(list(n3), dtype=np.int32)
    pp = np.array(list(p3), dtype=np.int32)
    tabs = [_direction(L, h, d, int(nn[d]), int(pp[d])) for d in range(3)]
    nelem = nn - pp + 1
    ads = ADS_setup(
        n=nn, p=pp, nelem=nelem, m=nn + pp + 1, ng=pp + 1,
        Ux=tabs[0][0], Uy=tabs[1][0], Uz=tabs[2][0],
        Ox=tabs[0][1], Oy=tabs[1][1], Oz=tabs[2][1],
        Jx=tabs[0][2], Jy=tabs[1][2], Jz=tabs[2][2],
        Wx=tabs[0][3], Wy=tabs[1][3], Wz=tabs[2][3],
        Xx=tabs[0][4], Xy=tabs[1][4], Xz=tabs[2][4],
        NNx=tabs[0][5], NNy=tabs[1][5], NNz=tabs[2][5],
        s=nn + 1, ibeg=np.ones(3, dtype=np.int32), iend=nn + 1, mine=np.ones(3, dtype=np.int32),
        maxe=nelem, lnelem=nelem, nrcpp=nn + 1,
    )
    D = int(np.prod(nn + 1))
    data = ADS_compute_data(h, D, tuple(int(v) + 1 for v in nn))
    return ads, data


def _set_kappa(ads_data, kappa):
    L = _lib.load()
    if kappa is None:
        check(L.ads_set_kappa(ads_data.handle, None), ads_data.handle)
    else:
        k = np.ascontiguousarray(kappa, dtype=np.float64).ravel()
        check(L.ads_set_kappa(ads_data.handle, _d(k)), ads_data.handle)


def Step(iter, RHS_fun, ads, ads_data):
    """ADSS::Step (src/ADS.F90:437-461) with a registered integrand form.

    Returns l2norm (RHS_eq.F90:81,85).  The step replaces the coefficients in
    ads_data; ``iter`` is accepted for interface parity (the integrand decides
    between projection and time step from its own ``t``)."""
    L = _lib.load()
    h = ads_data.handle
    out = ctypes.c_double(0.0)
    if isinstance(RHS_fun, L2ProjectionRHS):
        check(L.ads_project(h, ADS_F0_CONST, ctypes.byref(out)), h)
        return out.value
    if isinstance(RHS_fun, OilRHS):
        RHS_fun.setup(ads_data)
        if RHS_fun.t == 0.0:
            check(L.ads_oil_project(h, ctypes.byref(out)), h)
        else:
            dr = ctypes.c_double(0.0)
            check(L.ads_oil_step(h, RHS_fun.Dt, ctypes.byref(out), ctypes.byref(dr)), h)
            RHS_fun.drained += dr.value
        ads_data.t = RHS_fun.t
        return out.value
    if not isinstance(RHS_fun, HeatRHS):
        raise AdsError(_lib.ADS_ERR_UNSUPPORTED,
                       "RHS_fun is not a registered device form (HeatRHS, OilRHS, L2ProjectionRHS)")
    if RHS_fun.t == 0.0:
        check(L.ads_project(h, RHS_fun.f0, ctypes.byref(out)), h)
    else:
        _set_kappa(ads_data, RHS_fun.kappa)
        check(L.ads_step(h, RHS_fun.Dt, RHS_fun.prec, RHS_fun.refine, ctypes.byref(out)), h)
    ads_data.t = RHS_fun.t
    return out.value


def MultiStep(iter, mix, RHS_fun, ads, ads_data):
    """ADSS::MultiStep (src/ADS.F90:391-429) with the heat form: mix is the
    Fortran mix(4,3) (mix[:, s] for substep s) or its 3 columns as rows;
    substep s solves with mix(1,s)M + mix(2,s)K + mix(3,s)B + mix(4,s)B^T in
    every direction (ads_multistep).  Returns l2norm."""
    if not isinstance(RHS_fun, HeatRHS):
        raise AdsError(_lib.ADS_ERR_UNSUPPORTED, "MultiStep needs the heat form")
    L = _lib.load()
    m = np.asarray(mix, dtype=np.float64)
    cols = m.T if m.shape == (4, 3) else m.reshape(3, 4)
    arr = (ctypes.c_double * 12)(*cols.ravel())
    _set_kappa(ads_data, RHS_fun.kappa)
    l2 = ctypes.c_double(0.0)
    check(L.ads_multistep(ads_data.handle, arr, float(RHS_fun.Dt), ctypes.byref(l2)), ads_data.handle)
    return l2.value


def _form_of(RHS_fun, ads_data, iter=None):
    """(form, f0 kind, dt) of a registered integrand: the heat and oil forms
    take their t == 0 branch (the projection of their initial state) when t
    is 0, as the reference's integrands test it."""
    if isinstance(RHS_fun, L2ProjectionRHS):
        return _lib.ADS_FORM_PROJECT, ADS_F0_CONST, 0.0
    if isinstance(RHS_fun, OilRHS):
        RHS_fun.setup(ads_data)
        if RHS_fun.t == 0.0:
            return _lib.ADS_FORM_OIL_PROJECT, ADS_F0_CONST, 0.0
        return _lib.ADS_FORM_OIL, ADS_F0_CONST, RHS_fun.Dt
    if isinstance(RHS_fun, HeatRHS):
        _set_kappa(ads_data, RHS_fun.kappa)
        if RHS_fun.t == 0.0:
            return _lib.ADS_FORM_PROJECT, RHS_fun.f0, 0.0
        return _lib.ADS_FORM_HEAT, RHS_fun.f0, RHS_fun.Dt
    raise AdsError(_lib.ADS_ERR_UNSUPPORTED,
                   "RHS_fun is not a registered device form (HeatRHS, OilRHS, L2ProjectionRHS)")


def Form3DRHS(ads, ads_data, direction, substep, RHS_fun):
    """projection_engine::Form3DRHS (projection_engine.F90:154-284) with a
    registered integrand form, on the current coefficients (point values of
    U_n kept by FormUn(1), gradients of the current coefficients).  Returns
    (F, l2norm): F on the host, x fastest (the reference's ads_data%F); the
    state is unchanged."""
    L = _lib.load()
    h = ads_data.handle
    form, f0, dt = _form_of(RHS_fun, ads_data)
    F = np.empty(ads_data.D)
    l2 = ctypes.c_double(0.0)
    check(L.ads_form_rhs_sub(h, form, f0, int(direction), int(substep), dt, _d(F), ctypes.byref(l2)), h)
    return F, l2.value


def FormUn(subun, ads, ads_data):
    """projection_engine::FormUn (:306-441).  The point values and gradients
    are formed inside the assembly kernels: subun = 1 keeps the current
    coefficients as U_n (the point values of the step's later right-hand
    sides, the reference's Un); subun = 2, 3 keep U_n (the gradients always
    come from the current coefficients)."""
    L = _lib.load()
    check(L.ads_form_un(ads_data.handle, int(subun)), ads_data.handle)


def Sub_Step(ads, iter, mix, direction, substep, RHS_fun, ads_data):
    """ADSS::Sub_Step (src/ADS.F90:471-794): Form3DRHS with the registered
    form, then the solves with mix(1)M + mix(2)K + mix(3)B + mix(4)B^T in x,
    y and z; the result becomes the coefficients.  Returns l2norm."""
    L = _lib.load()
    h = ads_data.handle
    form, f0, dt = _form_of(RHS_fun, ads_data)
    m = (ctypes.c_double * 4)(*[float(v) for v in np.asarray(mix, dtype=np.float64).ravel()[:4]])
    l2 = ctypes.c_double(0.0)
    check(L.ads_sub_step(h, form, f0, m, int(direction), int(substep), dt, ctypes.byref(l2)), h)
    if form == _lib.ADS_FORM_OIL:
        dr = ctypes.c_double(0.0)
        check(L.ads_last_drained(h, ctypes.byref(dr)), h)
        RHS_fun.drained += dr.value
    ads_data.t = getattr(RHS_fun, "t", 0.0)
    return l2.value


class SparseMatrix:
    """The 1-D matrix of ComputeMatrix (type(sparse_matrix), src/sparse.F90)
    with its banded LU, held by the library (ads_matrix)."""

    def __init__(self, U, p, n, nelem, mix):
        L = _lib.load()
        self.n, self.p = int(n), int(p)
        u = np.ascontiguousarray(U, dtype=np.float64).ravel()
        if u.size != self.n + self.p + 2:
            raise ValueError("U must hold n + p + 2 knots")
        m = (ctypes.c_double * 4)(*[float(v) for v in np.asarray(mix, dtype=np.float64).ravel()[:4]])
        self.handle = ctypes.c_void_p()
        rc = L.ads_matrix_create(_d(u), self.n, self.p, int(nelem), m, ctypes.byref(self.handle))
        if rc != 0:
            raise AdsError(rc, L.ads_matrix_last_error(None).decode())

    def dense(self):
        N = self.n + 1
        out = np.zeros((N, N))
        check(_lib.load().ads_matrix_dense(self.handle, _d(out)))
        return out

    def close(self):
        if self.handle:
            _lib.load().ads_matrix_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ComputeMatrix(U, p, n, nelem, mix):
    """projection_engine::ComputeMatrix / MKBBT (:531-544, :63-132): the 1-D
    matrix mix(1)M + mix(2)K + mix(3)B + mix(4)B^T of the degree-p B-splines
    on the knot vector U(0:n+p+1) with nelem elements, factored for
    SolveOneDirection."""
    return SparseMatrix(U, p, n, nelem, mix)


def SolveOneDirection(RHS, eqnum, n, p, sprsmtrx):
    """ADSS::SolveOneDirection (src/ADS.F90:294-379): the columns
    RHS[:, 0:eqnum] of the (n+1) x eqnum array RHS (the Fortran RHS(1:n+1,
    1:eqnum)) are overwritten by the solutions, computed on the GPU.  Returns
    RHS."""
    L = _lib.load()
    if sprsmtrx.n != int(n) or sprsmtrx.p != int(p):
        raise ValueError("SolveOneDirection: the matrix was built for another n or p")
    if RHS.ndim != 2 or RHS.shape[0] != int(n) + 1 or RHS.shape[1] < int(eqnum):
        raise ValueError("SolveOneDirection: RHS must be (n+1) x eqnum")
    buf = np.ascontiguousarray(RHS[:, :int(eqnum)].T, dtype=np.float64)  # one system per row
    rc = L.ads_matrix_solve(sprsmtrx.handle, _d(buf), int(eqnum))
    if rc != 0:
        raise AdsError(rc, L.ads_matrix_last_error(sprsmtrx.handle).decode())
    RHS[:, :int(eqnum)] = buf.T
    return RHS


@dataclass
class PlotParams:
    """type(PlotParams) (src/plot.F90:14-22): sample ranges and counts."""

    startx: float = 0.0
    endx: float = 1.0
    starty: float = 0.0
    endy: float = 1.0
    startz: float = 0.0
    endz: float = 1.0
    resx: int = 31
    resy: int = 31
    resz: int = 31


def EvalGrid(ads_data, params):
    """The spline of the current coefficients on the PlotParams grid
    (SaveSplinePlot + EvalSpline, src/plot.F90:97-145, src/basis.F90:323-369),
    evaluated on the GPU (ads_eval_grid).  Returns vals[iz, iy, ix]."""
    L = _lib.load()
    res = (ctypes.c_int * 3)(params.resx, params.resy, params.resz)
    lo = (ctypes.c_double * 3)(params.startx, params.starty, params.startz)
    hi = (ctypes.c_double * 3)(params.endx, params.endy, params.endz)
    vals = np.zeros((params.resz, params.resy, params.resx))
    check(L.ads_eval_grid(ads_data.handle, res, lo, hi, _d(vals)), ads_data.handle)
    return vals


def VtkOutput(filename, vals, params):
    """VtkOutput (src/vtk.F90:15-77): <filename>.vti, text VTK ImageData with
    one F30.10 value per line, x fastest.  vals[iz, iy, ix]."""
    extent = "0 %5d 0 %5d 0 %5d" % (params.resx - 1, params.resy - 1, params.resz - 1)
    lines = ['<?xml version="1.0"?>', '<VTKFile type="ImageData" version="0.1">',
             '  <ImageData WholeExtent="%s" origin="0 0 0" spacing="1 1 1">' % extent,
             '    <Piece Extent="%s">' % extent, '      <PointData Scalars="Result">',
             '        <DataArray Name="Result" type="Float32">']
    lines += ["          %.10f" % v for v in np.asarray(vals, dtype=np.float64).ravel()]
    lines += ['        </DataArray>', '      </PointData>', '    </Piece>', '  </ImageData>', '</VTKFile>']
    with open(str(filename) + ".vti", "w") as f:
        f.write("\n".join(lines) + "\n")


def PrintSolution(iter, t, ads, ads_data, directory="."):
    """ADSS::PrintSolution (src/ADS.F90:955-989): step<iter>.vti with the
    spline on the 31^3 grid of [0,1]^3."""
    import os

    params = PlotParams()
    VtkOutput(os.path.join(directory, "step%d" % int(iter)), EvalGrid(ads_data, params), params)


def Cleanup(ads, ads_data):
    """ADSS::Cleanup (src/ADS.F90:802-863)."""
    if ads_data.handle:
        _lib.load().ads_destroy(ads_data.handle)
        ads_data.handle = None
    return 0


def uniform(seed, n, lo, hi):
    """The library's SplitMix64 uniform generator (ads_fill_uniform)."""
    out = np.zeros(int(n))
    check(_lib.load().ads_fill_uniform(int(seed), int(n), float(lo), float(hi), _d(out)))
    return out


def lognormal(seed, n, mu, sigma):
    out = np.zeros(int(n))
    check(_lib.load().ads_fill_lognormal(int(seed), int(n), float(mu), float(sigma), _d(out)))
    return out


def host_basis(n, p):
    """Tables of one direction computed on the host only (ads_host_basis)."""
    L = _lib.load()
    q = p + 1
    ne = n - p + 1
    if ne < 1 or p < 1:
        check(L.ads_host_basis(int(n), int(p), None, None, None, None, None, None))
    U = np.zeros(n + p + 2)
    O = np.zeros(ne, dtype=np.int32)
    J = np.zeros(ne)
    W = np.zeros(q)
    X = np.zeros((ne, q))
    NN = np.zeros((ne, q, p + 1, 2))
    check(L.ads_host_basis(int(n), int(p), _d(U), _i(O), _d(J), _d(W), _d(X), _d(NN)))
    return U, O, J, W, X, NN
