"""ctypes binding of libgtps.so (the C ABI declared in include/gtps.h).

The library is built in-tree by :func:`gputensorproductsmoother_amd.build.build`
and loaded from the package directory.  There is no fallback: if the shared
library is missing or does not load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgtps.so")

GTPS_OK = 0
GTPS_EINVAL = 1
GTPS_EHIP = 2
GTPS_ENOMEM = 3
GTPS_ENOTCONV = 4
GTPS_ENAN = 5

GTPS_FP32 = 0
GTPS_FP64 = 1

GTPS_MULT_FWD = 0
GTPS_MULT_REV = 1
GTPS_ADDITIVE = 2

GTPS_MG_SYMMETRIC = 0
GTPS_MG_REFERENCE_ORDER = 1
GTPS_MG_ADDITIVE = 2

GTPS_LAYOUT_CG = 0
GTPS_LAYOUT_DG = 1

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_ll = ctypes.c_longlong
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes); every symbol of include/gtps.h
SIGNATURES = {
    "gtps_ctx_create": (_i, [_i, _vp, ctypes.POINTER(_vp)]),
    "gtps_ctx_destroy": (_i, [_vp]),
    "gtps_ctx_synchronize": (_i, [_vp]),
    "gtps_ctx_stream": (_vp, [_vp]),
    "gtps_last_error": (ctypes.c_char_p, []),
    "gtps_version": (ctypes.c_char_p, []),
    "gtps_set_tuning": (_ll, [ctypes.c_char_p, _ll]),
    "gtps_get_tuning": (_ll, [ctypes.c_char_p]),
    "gtps_level_create": (_i, [_vp, _i, _i, _i, _i, ctypes.POINTER(_vp)]),
    "gtps_level_destroy": (_i, [_vp]),
    "gtps_level_n_dofs": (_ll, [_vp]),
    "gtps_laplace_vmult": (_i, [_vp, _vp, _vp]),
    "gtps_laplace_residual": (_i, [_vp, _vp, _vp, _vp]),
    "gtps_level_tensor_product_data": (_i, [_vp, _pd, _pd, _pd, _pd, _pd, _pd]),
    "gtps_level_set_coefficient": (_i, [_vp, _vp]),
    "gtps_laplace_vmult_transpose": (_i, [_vp, _vp, _vp]),
    "gtps_level_coefficient_gradient": (_i, [_vp, _vp, _vp, _vp]),
    "gtps_smoother_create": (_i, [_vp, _i, _d, ctypes.POINTER(_vp)]),
    "gtps_smoother_destroy": (_i, [_vp]),
    "gtps_smoother_step": (_i, [_vp, _vp, _vp]),
    "gtps_smoother_vmult": (_i, [_vp, _vp, _vp]),
    "gtps_transfer_create": (_i, [_vp, _i, _i, _i, _i, _i, ctypes.POINTER(_vp)]),
    "gtps_transfer_destroy": (_i, [_vp]),
    "gtps_transfer_prolongate_add": (_i, [_vp, _i, _vp, _vp]),
    "gtps_transfer_prolongate": (_i, [_vp, _i, _vp, _vp]),
    "gtps_transfer_restrict_add": (_i, [_vp, _i, _vp, _vp]),
    "gtps_vec_fill": (_i, [_vp, _ll, _i, _vp, _d]),
    "gtps_vec_axpby": (_i, [_vp, _ll, _d, _vp, _d, _vp]),
    "gtps_vec_dot": (_i, [_vp, _ll, _vp, _vp, _pd]),
    "gtps_vec_nrm2": (_i, [_vp, _ll, _vp, _pd]),
    "gtps_vec_dot_noalloc": (_i, [_vp, _ll, _vp, _vp, _vp]),
    "gtps_vec_convert_f64_f32": (_i, [_vp, _ll, _vp, _vp]),
    "gtps_vec_convert_f32_f64": (_i, [_vp, _ll, _vp, _vp]),
    "gtps_vec_set_boundary": (_i, [_vp, _vp, _d]),
    "gtps_mg_create": (_i, [_vp, _i, _i, _i, _i, _i, _d, ctypes.POINTER(_vp)]),
    "gtps_mg_destroy": (_i, [_vp]),
    "gtps_mg_n_dofs": (_ll, [_vp]),
    "gtps_mg_vcycle": (_i, [_vp, _vp, _vp]),
    "gtps_mg_pcg_solve": (_i, [_vp, _vp, _vp, _d, _i, _pi, _pd, _pd]),
    "gtps_mg_matvec": (_i, [_vp, _vp, _vp]),
    "gtps_mg_set_coefficient": (_i, [_vp, _vp]),
    "gtps_mg_fmg_solve": (_i, [_vp, _vp, ctypes.POINTER(_vp), _i]),
    "gtps_mg_matvec_transpose": (_i, [_vp, _vp, _vp]),
    "gtps_mg_coefficient_gradient": (_i, [_vp, _vp, _vp, _vp]),
    "gtps_mg_set_graph": (_i, [_vp, _i]),
    "gtps_mg_set_profiling": (_i, [_vp, _i]),
    "gtps_mg_profile": (_i, [_vp, _i, ctypes.POINTER(_ll), _pd]),
    "gtps_mg_profile_reset": (_i, [_vp]),
    "gtps_dg_level_create": (_i, [_vp, _i, _i, _i, _i, _d, ctypes.POINTER(_vp)]),
    "gtps_dg_level_destroy": (_i, [_vp]),
    "gtps_dg_level_n_dofs": (_ll, [_vp]),
    "gtps_dg_vmult": (_i, [_vp, _vp, _vp]),
    "gtps_dg_residual": (_i, [_vp, _vp, _vp, _vp]),
    "gtps_dg_smoother_step": (_i, [_vp, _vp, _vp, _i, _d]),
    "gtps_dg_prolongate_add": (_i, [_vp, _vp, _vp]),
    "gtps_dg_restrict_add": (_i, [_vp, _vp, _vp]),
    "gtps_dg_mg_create": (_i, [_vp, _i, _i, _i, _i, _i, _d, _i, ctypes.POINTER(_vp)]),
    "gtps_dg_mg_destroy": (_i, [_vp]),
    "gtps_dg_mg_set_graph": (_i, [_vp, _i]),
    "gtps_dg_mg_pcg_solve_batch": (_i, [_vp, _vp, _vp, _i, _d, _i, _pi, _pd]),
    "gtps_dg_mg_n_dofs": (_ll, [_vp]),
    "gtps_dg_mg_matvec": (_i, [_vp, _vp, _vp]),
    "gtps_dg_mg_vcycle": (_i, [_vp, _vp, _vp]),
    "gtps_dg_mg_pcg_solve": (_i, [_vp, _vp, _vp, _d, _i, _pi, _pd, _pd]),
    "gtps_assemble_rhs": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "gtps_rhs_sine": (_i, [_vp, _i, _i, _i, _d, _vp]),
    "gtps_rhs_constant": (_i, [_vp, _i, _i, _i, _d, _vp]),
    "gtps_dg_assemble_rhs": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "gtps_dg_rhs_sine": (_i, [_vp, _i, _i, _i, _d, _vp]),
    "gtps_error_norms": (_i, [_vp, _i, _i, _i, _d, _i, _vp, _pd, _pd]),
    "gtps_dg_error_norms": (_i, [_vp, _i, _i, _i, _d, _i, _vp, _pd, _pd]),
    "gtps_vector_save": (_i, [_vp, ctypes.c_char_p, _vp, _i, _i, _i, _i, _i]),
    "gtps_vector_info": (_i, [ctypes.c_char_p, _pi, _pi, _pi, _pi, _pi,
                              ctypes.POINTER(_ll)]),
    "gtps_vector_load": (_i, [_vp, ctypes.c_char_p, _vp, _ll, _i]),
}

PROF_GROUPS = {
    "smooth_fine": 0,
    "smooth_coarse": 1,
    "residual_fine": 2,
    "transfer": 3,
    "cg_matvec": 4,
    "vcycle": 5,
}

_lib = None


class GtpsError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"gtps error {code}: {what}")
        self.code = code


def load(path: str | None = None):
    """Load libgtps.so (raises if it is missing: there is no CPU fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} not found: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int, what: str = ""):
    if rc != GTPS_OK:
        msg = load().gtps_last_error().decode(errors="replace")
        raise GtpsError(rc, f"{what}: {msg}" if what else msg)
