"""ctypes binding of libfasticehip.so (declarations in include/fastice_hip.h).

The library is the only compute path of this package: if it is missing or
fails to load, every model constructor raises.  There is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfasticehip.so")

FI_ABI_VERSION = 1

FI_OK = 0
FI_ERR_ARG = -1
FI_ERR_OOM = -2
FI_ERR_HIP = -3
FI_ERR_NONFINITE = -4
FI_ERR_UNSUPPORTED = -5
FI_ERR_STATE = -6

FI_BC_TRACTION, FI_BC_VELOCITY, FI_BC_SLIP = 0, 1, 2
FI_RHEO_LINEAR_CONST, FI_RHEO_LINEAR_FIELD, FI_RHEO_POWERLAW = 0, 1, 2
FI_RELAX_NONE, FI_RELAX_LIN, FI_RELAX_LOG, FI_RELAX_LOG10 = 0, 1, 2, 3
FI_PREC_FP64, FI_PREC_FP32_REFINED = 0, 1
FI_LAYOUT_PARENT, FI_LAYOUT_INTERIOR = 0, 1
FI_BCGROUP_STRESS, FI_BCGROUP_VELOCITY, FI_BCGROUP_ETA, FI_BCGROUP_RESIDUAL = 0, 1, 2, 3
FI_PHASE_NA, FI_PHASE_NS = 0, 1
FI_OPT_FUSED, FI_OPT_ZCHUNK, FI_OPT_REFINE_FULL_PHASE, FI_OPT_REFINE_ETA_RATE, FI_OPT_REFINE_PHASE1_DIGITS = 1, 2, 3, 4, 5
FI_OPT_FUSED_VARIANT, FI_OPT_FUSED_VARIANT_FP32, FI_OPT_TILE_ORDER, FI_OPT_REFINE_ANDERSON = 6, 7, 8, 9
FI_OPT_REFINE_MODE, FI_OPT_FUSED_SHELL, FI_OPT_REFINE_COLS = 10, 11, 12
FI_BATCH_RUNNING, FI_BATCH_CONVERGED, FI_BATCH_MAX_ITER, FI_BATCH_FAILED = 0, 1, 2, 3
FI_BOPT_VARIANT, FI_BOPT_ZCHUNK, FI_BOPT_AUTOTUNE, FI_BOPT_GRAPH_ITERS, FI_BOPT_TILE_ORDER = 1, 2, 3, 4, 5

FIELD_NAMES = [
    "P", "TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ", "VX", "VY", "VZ", "ETA", "ETA_NEXT",
    "RP", "RVX", "RVY", "RVZ", "GX", "GY", "GZ", "RHEO", "PSI_NA", "PSI_NS",
]
F = {name: i for i, name in enumerate(FIELD_NAMES)}
FI_F_OM_NA0 = len(FIELD_NAMES)
FI_F_OM_NS0 = FI_F_OM_NA0 + 8
FI_F_COUNT = FI_F_OM_NS0 + 8

# location bits (bit d set = Vertex in dimension d) of every field id
L_CCC, L_VCC, L_CVC, L_VVC, L_CCV, L_VCV, L_CVV, L_VVV = range(8)
MASK3 = ["ccc", "vvv", "vcc", "cvc", "ccv", "vvc", "vcv", "cvv"]
MASK2 = ["cc", "vv", "vc", "cv"]
_MASK_LOC3 = [L_CCC, L_VVV, L_VCC, L_CVC, L_CCV, L_VVC, L_VCV, L_CVV]
_MASK_LOC2 = [L_CCC, L_VVC, L_VCC, L_CVC]


def field_location(ndims, fid):
    """Location bits of a field id."""
    name = FIELD_NAMES[fid] if fid < len(FIELD_NAMES) else None
    if name in ("P", "TXX", "TYY", "TZZ", "ETA", "ETA_NEXT", "RP", "RHEO"):
        return L_CCC
    if name == "TXY":
        return L_VVC
    if name == "TXZ":
        return L_VCV
    if name == "TYZ":
        return L_CVV
    if name in ("VX", "RVX", "GX"):
        return L_VCC
    if name in ("VY", "RVY", "GY"):
        return L_CVC
    if name in ("VZ", "RVZ", "GZ"):
        return L_CCV
    if name in ("PSI_NA", "PSI_NS"):
        return L_VVV if ndims == 3 else L_VVC
    q = (fid - FI_F_OM_NA0) % 8
    return _MASK_LOC3[q] if ndims == 3 else _MASK_LOC2[q]


class fi_rheology(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("has_reg", C.c_int32),
        ("relax", C.c_int32),
        ("A_is_field", C.c_int32),
        ("eta", C.c_double),
        ("A", C.c_double),
        ("n", C.c_double),
        ("eta_reg", C.c_double),
        ("rate", C.c_double),
    ]


class fi_model_desc(C.Structure):
    _fields_ = [
        ("ndims", C.c_int32),
        ("n", C.c_int32 * 3),
        ("origin", C.c_double * 3),
        ("extent", C.c_double * 3),
        ("bc_kind", (C.c_int32 * 2) * 3),
        ("bc_value", ((C.c_double * 3) * 2) * 3),
        ("rheology", fi_rheology),
        ("gravity_is_field", C.c_int32),
        ("gravity", C.c_double * 3),
        ("dtau_P", C.c_double),
        ("dtau_tau", C.c_double * 6),
        ("dtau_V", C.c_double * 3),
        ("immersed_boundary", C.c_int32),
        ("precision", C.c_int32),
    ]


class FastIceError(RuntimeError):
    """Error reported by libfasticehip.so (negative fi_status)."""

    def __init__(self, code, msg):
        super().__init__("fastice HIP error %d: %s" % (code, msg))
        self.code = code


class NonFiniteError(FastIceError):
    """simulation failed: non-finite residual (tm_stokes_chmy.jl:104-106)."""


class fi_heat_desc(C.Structure):
    _fields_ = [("ndims", C.c_int), ("n", C.c_int * 3), ("origin", C.c_double * 3), ("extent", C.c_double * 3),
                ("bc_kind", (C.c_int * 2) * 3), ("bc_value", (C.c_double * 2) * 3), ("lambda_rho_cp", C.c_double),
                ("dtau_q", C.c_double), ("dtau_T", C.c_double)]


FI_HEAT_BC_NONE, FI_HEAT_BC_VALUE, FI_HEAT_BC_FLUX = 0, 1, 2
FI_HF_T, FI_HF_T_OLD, FI_HF_QX, FI_HF_QY, FI_HF_QZ = 0, 1, 2, 3, 4

_lib = None

_DP = C.POINTER(C.c_double)


def _declare(lib):
    H = C.c_void_p
    sig = {
        "fi_abi_version": (C.c_int, []),
        "fi_last_error": (C.c_char_p, []),
        "fi_model_create": (C.c_int, [C.POINTER(fi_model_desc), C.POINTER(H)]),
        "fi_model_destroy": (C.c_int, [H]),
        "fi_field_upload": (C.c_int, [H, C.c_int, _DP, C.c_size_t, C.c_int]),
        "fi_field_download": (C.c_int, [H, C.c_int, _DP, C.c_size_t, C.c_int]),
        "fi_field_fill": (C.c_int, [H, C.c_int, C.c_double]),
        "fi_field_download_noalloc": (C.c_int, [H, C.c_int, _DP, C.c_size_t]),
        "fi_field_size": (C.c_size_t, [H, C.c_int, C.c_int]),
        "fi_apply_bc": (C.c_int, [H, C.c_int]),
        "fi_advance_iterations": (C.c_int, [H, C.c_int]),
        "fi_compute_residuals": (C.c_int, [H, _DP]),
        "fi_compute_tau": (C.c_int, [H]),
        "fi_solve": (C.c_int, [H, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int), _DP]),
        "fi_synchronize": (C.c_int, [H]),
        "fi_time_iterations": (C.c_int, [H, C.c_int, _DP]),
        "fi_set_option": (C.c_int, [H, C.c_int, C.c_int]),
        "fi_fused_region": (C.c_int, [H, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "fi_fused_active_cells": (C.c_int, [H, C.POINTER(C.c_longlong)]),
        "fi_ib_validity": (C.c_int, [H, C.POINTER(C.c_uint8), C.c_size_t]),
        "fi_levelset_from_dem": (C.c_int, [H, C.c_int, _DP, C.c_int, C.c_int, _DP, _DP, _DP]),
        "fi_invert_levelset": (C.c_int, [H, C.c_int]),
        "fi_levelset_from_polygon": (C.c_int, [H, C.c_int, _DP, C.c_int]),
        "fi_omega_from_psi": (C.c_int, [H, C.c_int]),
        "fi_mem_info": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
        "fi_refined_solve": (C.c_int, [H, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int), _DP]),
        "fi_refined_advance_inner": (C.c_int, [H, C.c_int]),
        "fi_refined_time_inner": (C.c_int, [H, C.c_int, _DP]),
        "fi_model_bytes": (C.c_size_t, [C.POINTER(fi_model_desc)]),
        "fi_fused_variants": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.c_int]),
        "fi_batch_create": (C.c_int, [C.POINTER(C.c_void_p)]),
        "fi_batch_destroy": (C.c_int, [C.c_void_p]),
        "fi_batch_add": (C.c_int, [C.c_void_p, H, C.c_int, C.c_int, C.c_double]),
        "fi_batch_remove": (C.c_int, [C.c_void_p, H]),
        "fi_batch_add_refined": (C.c_int, [C.c_void_p, H, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]),
        "fi_batch_refined_counts": (C.c_int, [C.c_void_p, H, C.POINTER(C.c_int)]),
        "fi_batch_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
        "fi_batch_solve": (C.c_int, [C.c_void_p]),
        "fi_batch_status": (C.c_int, [C.c_void_p, H, C.POINTER(C.c_int), C.POINTER(C.c_int), _DP]),
        "fi_batch_set_option": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "fi_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
        "fi_heat_create": (C.c_int, [C.POINTER(fi_heat_desc), C.POINTER(H)]),
        "fi_heat_destroy": (C.c_int, [H]),
        "fi_heat_field_upload": (C.c_int, [H, C.c_int, _DP, C.c_size_t, C.c_int]),
        "fi_heat_field_download": (C.c_int, [H, C.c_int, _DP, C.c_size_t, C.c_int]),
        "fi_heat_apply_bc": (C.c_int, [H]),
        "fi_heat_advance_iterations": (C.c_int, [H, C.c_double, C.c_int]),
        "fi_heat_residual_norms": (C.c_int, [H, C.c_double, _DP]),
        "fi_heat_step": (C.c_int, [H, C.c_double, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int), _DP]),
        "fi_heat_time_iterations": (C.c_int, [H, C.c_double, C.c_int, _DP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib():
    """Load libfasticehip.so once; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libfasticehip.so is not built (expected at %s); run fastice.jl_amd/build.py" % LIB_PATH)
        _lib = _declare(C.CDLL(LIB_PATH))
        v = _lib.fi_abi_version()
        if v != FI_ABI_VERSION:
            raise ImportError("libfasticehip.so ABI %d, expected %d" % (v, FI_ABI_VERSION))
    return _lib


def check(rc):
    if rc == FI_OK:
        return
    msg = lib().fi_last_error().decode(errors="replace")
    if rc == FI_ERR_NONFINITE:
        raise NonFiniteError(rc, msg)
    raise FastIceError(rc, msg)


def fused_variants(fp32=False):
    """[(tile lanes in x, tile rows, min waves per SIMD)] of the compiled
    fused-sweep variants."""
    buf = (C.c_int * 48)()
    n = lib().fi_fused_variants(1 if fp32 else 0, buf, 16)
    return [(buf[3 * q], buf[3 * q + 1], buf[3 * q + 2]) for q in range(n)]


def dptr(a):
    return a.ctypes.data_as(_DP)
