"""ctypes binding of the C ABI in include/ads_mi355x.h.

The shared library is the in-tree build ads_mpi_amd/_lib/libads_mi355x.so
(python -m ads_mpi_amd.build).  There is no fallback: if the library is
missing or the GPU cannot be initialised, calls raise.
"""
from __future__ import annotations

import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "_lib", "libads_mi355x.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "ads_mi355x.h")

ADS_OK = 0
ADS_ERR_INVALID = 1
ADS_ERR_HIP = 2
ADS_ERR_NOMEM = 3
ADS_ERR_SINGULAR = 4
ADS_ERR_UNSUPPORTED = 5
ADS_ERR_NUMERIC = 6

ADS_F0_CONST = 0
ADS_F0_SINPROD = 1
ADS_F0_BUMP3D = 2

ADS_PREC_FP64 = 0
ADS_PREC_FP32_IR = 1
ADS_PREC_FP32_IR_KRON = 2
ADS_PREC_FP64_KRON = 3
ADS_PREC_FP32_IR_SEG = 4

ADS_KERNEL_HEAT_FP64 = 0
ADS_KERNEL_KRON_RESIDUAL = 1
ADS_KERNEL_OIL = 2
ADS_KERNEL_SOLVE_FP64 = 3
ADS_KERNEL_SOLVE_FP64_LU = 4
ADS_KERNEL_SOLVE_SEG_FP32 = 5

ADS_FORM_HEAT = 0
ADS_FORM_PROJECT = 1
ADS_FORM_OIL = 2
ADS_FORM_OIL_PROJECT = 3

ADS_SOLVER_AUTO = 0
ADS_SOLVER_LU = 1
ADS_SOLVER_PARTITIONED = 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class AdsError(RuntimeError):
    """A non-zero return code of the C ABI, with ads_last_error's text."""

    def __init__(self, code, msg):
        super().__init__("ADS error %d: %s" % (code, msg))
        self.code = code


_lib = None


def header_symbols():
    """Function names declared in include/ads_mi355x.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(ads_\w+)\s*\(", text, re.M)))


def load(path: str | None = None):
    """Load the library (once) and declare the signatures."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            "libads_mi355x.so not built (%s); run python -m ads_mpi_amd.build" % p)
    L = ctypes.CDLL(p)
    I3 = ctypes.c_int * 3
    L.ads_init.argtypes = [I3, I3, ctypes.POINTER(_vp)]
    L.ads_destroy.argtypes = [_vp]
    L.ads_last_error.argtypes = [_vp]
    L.ads_last_error.restype = ctypes.c_char_p
    L.ads_set_stream.argtypes = [_vp, _vp]
    L.ads_synchronize.argtypes = [_vp]
    L.ads_get_sizes.argtypes = [_vp, _ip, _ip, _ip]
    L.ads_get_knots.argtypes = [_vp, ctypes.c_int, _dp]
    L.ads_get_basis.argtypes = [_vp, ctypes.c_int, _ip, _dp, _dp, _dp, _dp]
    L.ads_get_matrix.argtypes = [_vp, ctypes.c_int, _dp, _dp]
    L.ads_set_coeffs.argtypes = [_vp, _dp]
    L.ads_get_coeffs.argtypes = [_vp, _dp]
    L.ads_coeffs_device.argtypes = [_vp, ctypes.POINTER(_vp)]
    L.ads_set_kappa.argtypes = [_vp, _dp]
    L.ads_project.argtypes = [_vp, ctypes.c_int, _dp]
    L.ads_step.argtypes = [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp]
    L.ads_last_l2norm.argtypes = [_vp, _dp]
    L.ads_multistep.argtypes = [_vp, ctypes.c_double * 12, ctypes.c_double, _dp]
    L.ads_form_rhs.argtypes = [_vp, ctypes.c_double, _dp, _dp]
    L.ads_form_rhs_d.argtypes = [_vp, _vp, _vp, ctypes.c_double, _vp, _vp]
    L.ads_solve_d.argtypes = [_vp, _vp, ctypes.c_int]
    L.ads_step_d.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _vp]
    L.ads_step_adjoint_d.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _vp, _vp]
    L.ads_eval_grid.argtypes = [_vp, I3, ctypes.c_double * 3, ctypes.c_double * 3, _dp]
    L.ads_eval_grid_noalloc_d.argtypes = [_vp, I3, ctypes.c_double * 3, ctypes.c_double * 3, _vp]
    L.ads_host_grid_tables.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       _ip, _dp]
    L.ads_time_kernel.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
    L.ads_oil_setup.argtypes = [_vp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int,
                                ctypes.c_double * 6]
    L.ads_oil_project.argtypes = [_vp, _dp]
    L.ads_oil_step.argtypes = [_vp, ctypes.c_double, _dp, _dp]
    L.ads_oil_get_kq.argtypes = [_vp, _dp]
    L.ads_host_seg_tables.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _ip, ctypes.POINTER(ctypes.c_float)]
    L.ads_oil_curves.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
    L.ads_fill_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _dp]
    L.ads_fill_lognormal.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _dp]
    L.ads_host_basis.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _ip, _dp, _dp, _dp, _dp]
    L.ads_form_un.argtypes = [_vp, ctypes.c_int]
    L.ads_form_rhs_sub.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp,
                                   _dp]
    L.ads_sub_step.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double * 4, ctypes.c_int, ctypes.c_int,
                               ctypes.c_double, _dp]
    L.ads_last_drained.argtypes = [_vp, _dp]
    L.ads_set_solver.argtypes = [_vp, ctypes.c_int]
    L.ads_host_part_tables.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, _dp, _dp, _dp]
    L.ads_matrix_create.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double * 4,
                                    ctypes.POINTER(_vp)]
    L.ads_matrix_destroy.argtypes = [_vp]
    L.ads_matrix_last_error.argtypes = [_vp]
    L.ads_matrix_last_error.restype = ctypes.c_char_p
    L.ads_matrix_dense.argtypes = [_vp, _dp]
    L.ads_matrix_solve.argtypes = [_vp, _dp, ctypes.c_int]
    for name in header_symbols():
        if name not in ("ads_last_error", "ads_matrix_last_error"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc, ctx=None):
    if rc != ADS_OK:
        L = load()
        msg = L.ads_last_error(ctx)
        raise AdsError(rc, msg.decode() if msg else "")
    return rc
