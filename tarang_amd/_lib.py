"""ctypes loader for libtarang_amd.so (the C ABI in include/tarang_gpu.h).

The library is built in-tree by tarang_amd.build.  Loading fails loudly when
it is missing: there is no fallback implementation of any entry point.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtarang_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "tarang_gpu.h")

TG_OK = 0
ERR_ARG, ERR_SHAPE, ERR_OOM, ERR_HIP, ERR_NAN, ERR_STATE = -1, -2, -3, -4, -5, -6
PREC_FP64, PREC_FP32 = 0, 1
DEALIAS_NONE, DEALIAS_REF_CORNER, DEALIAS_TWOTHIRDS = 0, 1, 2
STEP_CONTINUE = 1


class TgConfig(ctypes.Structure):
    _fields_ = [("basis", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("kfactor", ctypes.c_double * 3), ("precision", ctypes.c_int), ("dealias", ctypes.c_int),
                ("strong_reality", ctypes.c_int), ("device", ctypes.c_int)]


class TgFluid(ctypes.Structure):
    _fields_ = [("nu", ctypes.c_double), ("dt", ctypes.c_double), ("force_kind", ctypes.c_int),
                ("force_eps", ctypes.c_double), ("force_r_in", ctypes.c_double),
                ("force_r_out", ctypes.c_double), ("nu_hyper", ctypes.c_double), ("hyper_exponent", ctypes.c_int),
                ("force_helicity", ctypes.c_double), ("force_tau", ctypes.c_double),
                ("force_seed", ctypes.c_ulonglong)]


BASIS_FFF, BASIS_SFF, BASIS_CHFF = 0, 1, 2
SC_FOURIER, SC_SIN, SC_COS, SC_CHEB = 0, 1, 2, 3


class TgRbc(ctypes.Structure):
    _fields_ = [("Ra", ctypes.c_double), ("Pr", ctypes.c_double), ("dt", ctypes.c_double)]


class TarangGpuError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with code {code}: {msg}")
        self.code = code


_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PP = ctypes.POINTER(ctypes.c_void_p)
_DP = ctypes.POINTER(ctypes.c_double)

SIGNATURES = {
    "tg_last_error_string": (ctypes.c_char_p, []),
    "tg_version": (ctypes.c_char_p, []),
    "tg_plan_create": (_I, [ctypes.POINTER(TgConfig), ctypes.POINTER(_P)]),
    "tg_plan_destroy": (_I, [_P]),
    "tg_set_stream": (_I, [_P, _P]),
    "tg_synchronize": (_I, [_P]),
    "tg_get_shapes": (_I, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "tg_plan_device_bytes": (ctypes.c_int64, [_P]),
    "tg_forward": (_I, [_P, _I, _P, _P]),
    "tg_inverse": (_I, [_P, _I, _P, _P]),
    "tg_get_dt": (_I, [_P, _PP, _D, _D, _DP, _DP]),
    "tg_dealias": (_I, [_P, _PP, _I]),
    "tg_strong_reality": (_I, [_P, _PP, _I]),
    "tg_compute_nlin": (_I, [_P, _PP, _PP]),
    "tg_compute_rhs": (_I, [_P, _PP, ctypes.POINTER(TgFluid), _PP]),
    "tg_compute_nlin_giver": (_I, [_P, _PP, _PP, _PP]),
    "tg_rk4_step": (_I, [_P, _PP, ctypes.POINTER(TgFluid), _I, _I]),
    "tg_energy": (_I, [_P, _P, _DP]),
    "tg_Sn": (_I, [_P, _P, _I, _DP]),
    "tg_shell_spectrum": (_I, [_P, _P, _I, _DP, _I]),
    "tg_flux": (_I, [_P, _PP, _DP, _I, _DP]),
    "tg_shell_to_shell": (_I, [_P, _PP, _DP, _I, _DP]),
    "tg_energy_transfer": (_I, [_P, _PP, _DP, _I, _DP, _DP]),
    "tg_ring_spectrum": (_I, [_P, _PP, _I, _I, _I, _DP, _I, _DP]),
    "tg_cheb_xderiv": (_I, [_P, _P, _P]),
    "tg_helmholtz": (_I, [_P, _P, _P, _D, _DP, _DP]),
    "tg_compute_nlin_chff": (_I, [_P, _PP, _PP]),
    "tg_bdf1_step": (_I, [_P, _PP, _P, ctypes.POINTER(TgFluid), _I]),
    "tg_shell_spectrum_noalloc": (_I, [ctypes.POINTER(TgConfig), _P, _I, _P, _I, _P]),
    "tg_cylindrical_ring_spectrum": (_I, [_P, _PP, _I, _I, _DP, _I, _DP, _DP]),
    "tg_ring_to_ring": (_I, [_P, _PP, _I, _DP, _I, _DP, _I, _DP]),
    "tg_cylindrical_ring_to_ring": (_I, [_P, _PP, _I, _DP, _I, _DP, _I, _DP]),
    "tg_init_field": (_I, [_P, _I, _D, _D, _D, _D, ctypes.c_uint64, _PP]),
    "tg_compute_nlin_mhd": (_I, [_P, _PP, _PP, _PP, _PP]),
    "tg_rk4_step_mhd": (_I, [_P, _PP, _PP, ctypes.POINTER(TgFluid), _D, _I]),
    "tg_weak_reality": (_I, [_P, _PP, _I]),
    "tg_compute_nlin_rbc": (_I, [_P, _PP, _P, _PP, _P]),
    "tg_compute_rhs_rbc": (_I, [_P, _PP, _P, ctypes.POINTER(TgRbc), _PP, _P]),
    "tg_rk4_step_rbc": (_I, [_P, _PP, _P, ctypes.POINTER(TgRbc), _I]),
    "tg_nusselt": (_I, [_P, _P, _P, ctypes.POINTER(TgRbc), ctypes.POINTER(_D)]),
    "tg_rhs_vjp": (_I, [_P, _PP, _PP, _PP]),
    "tg_adjoint_forward": (_I, [_P, _PP, ctypes.POINTER(TgFluid), _I, _I]),
    "tg_adjoint_backward": (_I, [_P, _PP]),
    "tg_adjoint_free": (_I, [_P]),
    "tg_cross_helicity": (_I, [_P, _PP, _PP, _DP]),
    "tg_flux_mhd": (_I, [_P, _PP, _PP, _DP, _I, _DP, _DP]),
    "tg_profile_enable": (_I, [_P, _I]),
    "tg_profile_reset": (_I, [_P]),
    "tg_profile_read": (_I, [_P, _I, _DP, ctypes.POINTER(ctypes.c_int64)]),
    "tg_kernel_name": (ctypes.c_char_p, [_I]),
}

KERNELS = ["x_inv", "y_fft", "z_products", "x_nlin_rk", "x_nlin_giver", "reality", "z_r2c", "z_c2r", "x_fwd",
           "reduce", "x_inv_sym", "z_adj", "x_fwd_adj", "x_sc", "z_rbc", "x_rbc"]


def header_symbols():
    """Function names declared in include/tarang_gpu.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(tg_\w+)\s*\(", text, flags=re.M)))


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run tarang_amd.build.build() (hipcc for gfx950)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(fn_name, rc):
    if rc != TG_OK:
        msg = load().tg_last_error_string().decode(errors="replace")
        raise TarangGpuError(fn_name, rc, msg)
    return rc


def ptr_array(ptrs):
    return (ctypes.c_void_p * len(ptrs))(*ptrs)
