"""ctypes bindings for the CPU oracle (liboracle.so) and the flang-built
reference leaves (oracle/_ref/libadsref.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product package never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libadsref.so")

F0_CONST, F0_SINPROD, F0_BUMP3D = 0, 1, 2
MODE_FAITHFUL, MODE_SUMFACT = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def build(quiet=True):
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    out = subprocess.run(["make", "-C", HERE], capture_output=quiet, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + (out.stdout or "") + (out.stderr or ""))
    if os.path.isdir("/root/reference/src"):
        out = subprocess.run(["make", "-C", HERE, "ref"], capture_output=quiet, text=True)
        if out.returncode != 0:
            raise RuntimeError("oracle/_ref build failed:\n" + (out.stdout or "") + (out.stderr or ""))


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_fill_open_knot.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_count_spans.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_count_spans.restype = ctypes.c_int
        L.orc_prepare_knot.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _ip]
        L.orc_prepare_knot.restype = ctypes.c_int
        L.orc_gauss_rule.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_gauss_rule.restype = ctypes.c_int
        L.orc_ders_basis_funs.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_find_span.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_find_span.restype = ctypes.c_int
        L.orc_basis_data.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, _ip, _dp, _dp, _dp, _dp]
        L.orc_basis_data.restype = ctypes.c_int
        L.orc_eval_spline.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                      ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_eval_spline.restype = ctypes.c_double
        L.orc_mkbbt_band.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.orc_mkbbt_band.restype = ctypes.c_int
        L.orc_dgbtrf.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip]
        L.orc_dgbtrf.restype = ctypes.c_int
        L.orc_dgbtrs.argtypes = [ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 _dp, ctypes.c_int, _ip, _dp, ctypes.c_int]
        for nm in ("orc_reorder_for_y", "orc_reorder_for_z", "orc_reorder_for_x"):
            getattr(L, nm).argtypes = [_ip, _dp, _dp]
        L.orc_setup_create.argtypes = [_ip, _ip]
        L.orc_setup_create.restype = ctypes.c_void_p
        L.orc_setup_destroy.argtypes = [ctypes.c_void_p]
        L.orc_setup_dofs.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_setup_dofs.restype = ctypes.c_int
        L.orc_setup_nelem.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_setup_nelem.restype = ctypes.c_int
        L.orc_form_rhs_heat.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_double, _dp, _dp,
                                        ctypes.c_int, ctypes.c_int]
        L.orc_form_rhs_project.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_solve_kron.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int]
        L.orc_solve_kron_T.argtypes = [ctypes.c_void_p, _dp]
        L.orc_solve_kron_mix.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_multistep.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, _dp, _dp]
        L.orc_heat_step.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_double, _dp, _dp]
        L.orc_project.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_heat_step_adjoint.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, _dp, _dp, _dp]
        L.orc_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_lognormal.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_splitmix64.argtypes = [ctypes.c_uint64]
        L.orc_splitmix64.restype = ctypes.c_uint64
        for nm, nargs in (("orc_oil_lerp", 3), ("orc_oil_falloff", 3)):
            getattr(L, nm).argtypes = [ctypes.c_double] * nargs
            getattr(L, nm).restype = ctypes.c_double
        L.orc_oil_dist_from_segment.argtypes = [_dp, _dp, _dp]
        L.orc_oil_dist_from_segment.restype = ctypes.c_double
        L.orc_oil_dist_from_curves.argtypes = [ctypes.POINTER(_OilC), _dp]
        L.orc_oil_dist_from_curves.restype = ctypes.c_double
        for nm in ("orc_oil_kq", "orc_oil_pumping", "orc_oil_initial_state"):
            getattr(L, nm).argtypes = [ctypes.POINTER(_OilC)] + [ctypes.c_double] * 3
            getattr(L, nm).restype = ctypes.c_double
        L.orc_oil_draining.argtypes = [ctypes.POINTER(_OilC)] + [ctypes.c_double] * 4
        L.orc_oil_draining.restype = ctypes.c_double
        L.orc_oil_kq_cache.argtypes = [ctypes.c_void_p, ctypes.POINTER(_OilC), _dp]
        L.orc_oil_form_rhs.argtypes = [ctypes.c_void_p, ctypes.POINTER(_OilC), _dp, _dp, ctypes.c_double,
                                       ctypes.c_int, _dp, _dp, _dp]
        L.orc_oil_step.argtypes = [ctypes.c_void_p, ctypes.POINTER(_OilC), _dp, _dp, ctypes.c_double, _dp, _dp, _dp]
        L.orc_oil_project.argtypes = [ctypes.c_void_p, ctypes.POINTER(_OilC), _dp, _dp]
        L.orc_oil_curves.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        _lib = L
    return _lib


class _OilC(ctypes.Structure):
    _fields_ = [("cx", _dp), ("cy", _dp), ("cz", _dp), ("cN", ctypes.c_int), ("cL", ctypes.c_int),
                ("pumps", _dp), ("npumps", ctypes.c_int), ("drains", _dp), ("ndrains", ctypes.c_int),
                ("mi", ctypes.c_double), ("kq_min", ctypes.c_double), ("kq_max", ctypes.c_double),
                ("radius", ctypes.c_double), ("pump_strength", ctypes.c_double),
                ("drain_strength", ctypes.c_double)]


class Oil:
    """Inputs of the oil problem (problems/oil/input_data.F90): curves
    (cN x cL points each of cx, cy, cz), pumps and drains (rows x, y, z) and
    the constants mi, Kqmin, Kqmax, radius, pumping/draining strength."""

    def __init__(self, cx, cy, cz, cN, cL, pumps=(), drains=(), mi=10.0, kq_min=1.0, kq_max=1000.0,
                 radius=0.15, pump_strength=1.0, drain_strength=1.0):
        self.cx = np.ascontiguousarray(cx, dtype=np.float64)
        self.cy = np.ascontiguousarray(cy, dtype=np.float64)
        self.cz = np.ascontiguousarray(cz, dtype=np.float64)
        self.pumps = np.ascontiguousarray(np.reshape(pumps, (-1, 3)), dtype=np.float64)
        self.drains = np.ascontiguousarray(np.reshape(drains, (-1, 3)), dtype=np.float64)
        self.c = _OilC(_d(self.cx), _d(self.cy), _d(self.cz), int(cN), int(cL), _d(self.pumps),
                       len(self.pumps), _d(self.drains), len(self.drains), mi, kq_min, kq_max, radius,
                       pump_strength, drain_strength)
        self.cN, self.cL = int(cN), int(cL)

    @classmethod
    def random(cls, seed, cN=30, cL=16, **kw):
        cx, cy, cz = (np.zeros(cN * cL) for _ in range(3))
        lib().orc_oil_curves(seed, cN, cL, _d(cx), _d(cy), _d(cz))
        return cls(cx, cy, cz, cN, cL, **kw)

    def kq(self, x, y, z):
        return lib().orc_oil_kq(ctypes.byref(self.c), x, y, z)

    def pumping(self, x, y, z):
        return lib().orc_oil_pumping(ctypes.byref(self.c), x, y, z)

    def draining(self, u, x, y, z):
        return lib().orc_oil_draining(ctypes.byref(self.c), u, x, y, z)

    def initial_state(self, x, y, z):
        return lib().orc_oil_initial_state(ctypes.byref(self.c), x, y, z)

    def dist_from_curves(self, p):
        return lib().orc_oil_dist_from_curves(ctypes.byref(self.c), _d(np.ascontiguousarray(p, dtype=np.float64)))


def oil_falloff(r, R, t):
    return lib().orc_oil_falloff(r, R, t)


def oil_lerp(t, x, y):
    return lib().orc_oil_lerp(t, x, y)


def oil_dist_from_segment(p, a, b):
    f = lambda v: _d(np.ascontiguousarray(v, dtype=np.float64))
    return lib().orc_oil_dist_from_segment(f(p), f(a), f(b))


def ref():
    """The flang-built reference leaves, or None when they were not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        R = ctypes.CDLL(REF_PATH)
        R.ref_fill_open_knot.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        R.ref_count_spans.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        R.ref_count_spans.restype = ctypes.c_int
        R.ref_gauss_rule.argtypes = [ctypes.c_int, _dp, _dp]
        R.ref_basis_data.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, _ip, _dp, _dp, _dp, _dp]
        R.ref_ders_basis_funs.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp]
        R.ref_find_span.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
        R.ref_find_span.restype = ctypes.c_int
        R.ref_eval_spline.argtypes = [ctypes.c_int] * 6 + [_dp, _dp, _dp, _dp] + [ctypes.c_double] * 3
        R.ref_eval_spline.restype = ctypes.c_double
        R.ref_reorder.argtypes = [ctypes.c_int, _ip, _dp, _dp]
        for nm, nargs in (("ref_falloff", 3), ("ref_lerp", 3), ("ref_bump3d", 5)):
            getattr(R, nm).argtypes = [ctypes.c_double] * nargs
            getattr(R, nm).restype = ctypes.c_double
        _ref = R
    return _ref


# ---------------------------------------------------------------------------
# numpy conveniences
# ---------------------------------------------------------------------------

def knots(n, p, use_ref=False):
    U = np.zeros(n + p + 2)
    (ref() if use_ref else lib())
    if use_ref:
        ref().ref_fill_open_knot(n, p, _d(U))
    else:
        lib().orc_fill_open_knot(n, p, _d(U))
    return U


def prepare_knot(n, p):
    """PrepareKnot: (U, nelem)."""
    U = np.zeros(max(n + p + 2, 1))
    ne = ctypes.c_int(0)
    rc = lib().orc_prepare_knot(n, p, _d(U), ctypes.byref(ne))
    if rc != 0:
        raise ValueError("prepare_knot failed")
    return U, ne.value


def count_spans(n, p, U, use_ref=False):
    U = np.ascontiguousarray(U, dtype=np.float64)
    if use_ref:
        return ref().ref_count_spans(n, p, _d(U))
    return lib().orc_count_spans(n, p, _d(U))


def gauss_rule(q, use_ref=False):
    X = np.zeros(q)
    W = np.zeros(q)
    if use_ref:
        ref().ref_gauss_rule(q, _d(X), _d(W))
        return X, W
    rc = lib().orc_gauss_rule(q, _d(X), _d(W))
    if rc != 0:
        raise ValueError("unsupported quadrature order %d" % q)
    return X, W


def basis_data(p, U, nelem, use_ref=False):
    """BasisData tables: O (int32), J, W, X(q, r) as [r][q], NN(0:1,0:p,q,r) as [r][q][p+1][2]."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    q = p + 1
    m = len(U) - 1
    O = np.zeros(nelem, dtype=np.int32)
    J = np.zeros(nelem)
    W = np.zeros(q)
    X = np.zeros((nelem, q))
    NN = np.zeros((nelem, q, p + 1, 2))
    if use_ref:
        ref().ref_basis_data(p, m, _d(U), 1, q, nelem, _i(O), _d(J), _d(W), _d(X), _d(NN))
    else:
        rc = lib().orc_basis_data(p, m, _d(U), 1, q, nelem, _i(O), _d(J), _d(W), _d(X), _d(NN))
        if rc != 0:
            raise ValueError("basis_data failed: %d" % rc)
    return O, J, W, X, NN


def ders_basis_funs(i, uu, p, d, U, use_ref=False):
    U = np.ascontiguousarray(U, dtype=np.float64)
    ders = np.zeros((d + 1, p + 1))
    if use_ref:
        ref().ref_ders_basis_funs(i, uu, p, d, _d(U), _d(ders))
    else:
        lib().orc_ders_basis_funs(i, uu, p, d, _d(U), _d(ders))
    return ders


def find_span(n, p, uu, U, use_ref=False):
    U = np.ascontiguousarray(U, dtype=np.float64)
    if use_ref:
        return ref().ref_find_span(n, p, uu, _d(U))
    return lib().orc_find_span(n, p, uu, _d(U))


def mkbbt_band(U, p, n, nelem, mix=(1.0, 0.0, 0.0, 0.0)):
    U = np.ascontiguousarray(U, dtype=np.float64)
    ldab = 3 * p + 1
    ab = np.zeros((n + 1, ldab))  # column-major band: ab[j][row]
    mixa = np.asarray(mix, dtype=np.float64)
    rc = lib().orc_mkbbt_band(_d(U), p, n, nelem, _d(mixa), _d(ab), ldab)
    if rc != 0:
        raise ValueError("mkbbt failed")
    return ab


def band_to_dense(ab, kl, ku):
    ncol, ldab = ab.shape
    A = np.zeros((ncol, ncol))
    for j in range(ncol):
        for i in range(max(0, j - ku), min(ncol, j + kl + 1)):
            A[i, j] = ab[j, kl + ku + i - j]
    return A


def dgbtrf(ab, kl, ku):
    ab = np.array(ab, dtype=np.float64, copy=True)
    n, ldab = ab.shape
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().orc_dgbtrf(n, kl, ku, _d(ab), ldab, _i(ipiv))
    return ab, ipiv, info


def dgbtrs(trans, ab, kl, ku, ipiv, b):
    ab = np.ascontiguousarray(ab, dtype=np.float64)
    n, ldab = ab.shape
    b = np.array(b, dtype=np.float64, copy=True, order="F")
    if b.ndim == 1:
        bb = b.reshape(n, 1, order="F")
    else:
        bb = b
    bb = np.asfortranarray(bb)
    lib().orc_dgbtrs(trans.encode(), n, kl, ku, bb.shape[1], _d(ab), ldab, _i(np.ascontiguousarray(ipiv, dtype=np.int32)),
                     bb.ctypes.data_as(_dp), n)
    return bb.reshape(b.shape, order="F")


def reorder(which, s, F, use_ref=False):
    s3 = np.asarray(s, dtype=np.int32)
    F = np.ascontiguousarray(F, dtype=np.float64)
    F2 = np.zeros_like(F)
    if use_ref:
        ref().ref_reorder({"y": 1, "z": 2, "x": 3}[which], _i(s3), _d(F), _d(F2))
    else:
        getattr(lib(), "orc_reorder_for_" + which)(_i(s3), _d(F), _d(F2))
    return F2


def uniform(seed, n, lo, hi):
    out = np.zeros(int(n))
    lib().orc_uniform(seed, int(n), lo, hi, _d(out))
    return out


def lognormal(seed, n, mu, sigma):
    out = np.zeros(int(n))
    lib().orc_lognormal(seed, int(n), mu, sigma, _d(out))
    return out


class Setup:
    """ADSS::initialize on one rank (ADS.F90:11-107), CPU oracle side."""

    def __init__(self, n, p):
        n3 = np.asarray(n if np.ndim(n) else (n, n, n), dtype=np.int32)
        p3 = np.asarray(p if np.ndim(p) else (p, p, p), dtype=np.int32)
        self.n = tuple(int(v) for v in n3)
        self.p = tuple(int(v) for v in p3)
        self.h = lib().orc_setup_create(_i(n3), _i(p3))
        if not self.h:
            raise ValueError("orc_setup_create failed for n=%s p=%s" % (self.n, self.p))
        self.dofs = tuple(lib().orc_setup_dofs(self.h, d) for d in range(3))
        self.nelem = tuple(lib().orc_setup_nelem(self.h, d) for d in range(3))
        self.D = self.dofs[0] * self.dofs[1] * self.dofs[2]
        self.E = self.nelem[0] * self.nelem[1] * self.nelem[2]

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_setup_destroy(h)
            self.h = None

    def _vec(self):
        return np.zeros(self.D)

    def form_rhs_heat(self, c, dt, kappa=None, mode=MODE_SUMFACT, ez_range=(0, -1)):
        c = np.ascontiguousarray(c, dtype=np.float64).ravel()
        F = self._vec()
        l2 = ctypes.c_double(0.0)
        kp = _d(np.ascontiguousarray(kappa, dtype=np.float64)) if kappa is not None else None
        rc = lib().orc_form_rhs_heat(self.h, mode, _d(c), kp, dt, _d(F), ctypes.byref(l2), ez_range[0], ez_range[1])
        if rc:
            raise RuntimeError("form_rhs_heat rc=%d" % rc)
        return F, l2.value

    def form_rhs_project(self, f0=F0_SINPROD, mode=MODE_SUMFACT):
        F = self._vec()
        l2 = ctypes.c_double(0.0)
        rc = lib().orc_form_rhs_project(self.h, mode, f0, _d(F), ctypes.byref(l2))
        if rc:
            raise RuntimeError("form_rhs_project rc=%d" % rc)
        return F, l2.value

    def solve(self, F, faithful=False, line_limit=0):
        x = np.array(F, dtype=np.float64, copy=True).ravel()
        rc = lib().orc_solve_kron(self.h, 1 if faithful else 0, _d(x), line_limit)
        if rc:
            raise RuntimeError("solve rc=%d" % rc)
        return x

    def solve_T(self, F):
        x = np.array(F, dtype=np.float64, copy=True).ravel()
        rc = lib().orc_solve_kron_T(self.h, _d(x))
        if rc:
            raise RuntimeError("solve_T rc=%d" % rc)
        return x

    def solve_mix(self, F, mix):
        """The three solves with mix1 M + mix2 K + mix3 B + mix4 B^T per direction."""
        x = np.array(F, dtype=np.float64, copy=True).ravel()
        m = np.ascontiguousarray(mix, dtype=np.float64)
        rc = lib().orc_solve_kron_mix(self.h, _d(m), _d(x))
        if rc:
            raise RuntimeError("solve_mix rc=%d" % rc)
        return x

    def multistep(self, c, mix, dt, kappa=None):
        """MultiStep with the heat integrand; mix[s] is substep s's (mix1..mix4)."""
        c = np.ascontiguousarray(c, dtype=np.float64).ravel()
        m = np.ascontiguousarray(np.asarray(mix, dtype=np.float64).reshape(3, 4))
        out = self._vec()
        l2 = ctypes.c_double(0.0)
        kp = _d(np.ascontiguousarray(kappa, dtype=np.float64)) if kappa is not None else None
        rc = lib().orc_multistep(self.h, _d(m), _d(c), kp, dt, _d(out), ctypes.byref(l2))
        if rc:
            raise RuntimeError("multistep rc=%d" % rc)
        return out, l2.value

    def step(self, c, dt, kappa=None, mode=MODE_SUMFACT):
        c = np.ascontiguousarray(c, dtype=np.float64).ravel()
        out = self._vec()
        l2 = ctypes.c_double(0.0)
        kp = _d(np.ascontiguousarray(kappa, dtype=np.float64)) if kappa is not None else None
        rc = lib().orc_heat_step(self.h, mode, _d(c), kp, dt, _d(out), ctypes.byref(l2))
        if rc:
            raise RuntimeError("step rc=%d" % rc)
        return out, l2.value

    def project(self, f0=F0_SINPROD, mode=MODE_SUMFACT):
        out = self._vec()
        l2 = ctypes.c_double(0.0)
        rc = lib().orc_project(self.h, mode, f0, _d(out), ctypes.byref(l2))
        if rc:
            raise RuntimeError("project rc=%d" % rc)
        return out, l2.value

    def oil_kq_cache(self, oil):
        kq = np.zeros(self.E * int(np.prod([p + 1 for p in self.p])))
        lib().orc_oil_kq_cache(self.h, ctypes.byref(oil.c), _d(kq))
        return kq

    def oil_form_rhs(self, oil, kq, c, dt, tzero=False):
        F = self._vec()
        l2 = ctypes.c_double(0.0)
        dr = ctypes.c_double(0.0)
        cc = np.ascontiguousarray(c, dtype=np.float64).ravel() if c is not None else None
        lib().orc_oil_form_rhs(self.h, ctypes.byref(oil.c), _d(kq) if kq is not None else None,
                               _d(cc) if cc is not None else None, dt, int(tzero), _d(F), ctypes.byref(l2),
                               ctypes.byref(dr))
        return F, l2.value, dr.value

    def oil_step(self, oil, kq, c, dt):
        c = np.ascontiguousarray(c, dtype=np.float64).ravel()
        out = self._vec()
        l2 = ctypes.c_double(0.0)
        dr = ctypes.c_double(0.0)
        rc = lib().orc_oil_step(self.h, ctypes.byref(oil.c), _d(kq), _d(c), dt, _d(out), ctypes.byref(l2),
                                ctypes.byref(dr))
        if rc:
            raise RuntimeError("oil step rc=%d" % rc)
        return out, l2.value, dr.value

    def oil_project(self, oil):
        out = self._vec()
        l2 = ctypes.c_double(0.0)
        rc = lib().orc_oil_project(self.h, ctypes.byref(oil.c), _d(out), ctypes.byref(l2))
        if rc:
            raise RuntimeError("oil project rc=%d" % rc)
        return out, l2.value

    def step_adjoint(self, c_n, lam_next, dt, kappa=None, dkappa=None):
        c_n = np.ascontiguousarray(c_n, dtype=np.float64).ravel()
        lam_next = np.ascontiguousarray(lam_next, dtype=np.float64).ravel()
        lam = self._vec()
        kp = _d(np.ascontiguousarray(kappa, dtype=np.float64)) if kappa is not None else None
        dk = _d(dkappa) if dkappa is not None else None
        rc = lib().orc_heat_step_adjoint(self.h, _d(c_n), kp, dt, _d(lam_next), _d(lam), dk)
        if rc:
            raise RuntimeError("step_adjoint rc=%d" % rc)
        return lam
