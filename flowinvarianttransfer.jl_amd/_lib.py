"""ctypes binding of libfit_hip.so (declared in include/fit_hip.h).

The library is the product path: nothing here falls back to a CPU implementation.  If the
shared library has not been built, or a call fails on the device, an exception is raised.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfit_hip.so")

F32, F64 = 0, 1
DEALIAS_NONE, DEALIAS_TWOTHIRDS, DEALIAS_PADDED = 0, 1, 2
KINETIC_ENERGY, HELICITY, ENSTROPHY, PASSIVE_SCALAR = 0, 1, 2, 3
PTR_HOST, PTR_DEVICE = 0, 1
OPT_PLAIN_PIPELINE, OPT_GIVER_BATCH, OPT_XPASS_GIVERS, OPT_XPASS_BLOCK, OPT_SUBGRID, OPT_COL_TILE = 1, 2, 3, 4, 5, 6
OPT_XPASS_SELF = 7
OPT_XPASS_STAGE = 8
OPT_GRAD_SPLIT = 9
OPT_FUSED_FLUX = 10
OPT_RADIX_ORDER = 11
OPT_COL_F64 = 12
OPT_COL_PIPE = 13
OPT_LAST_BINS = 14
TUNE_FLUX, TUNE_SHELL_TO_SHELL = 1, 2
OPT_NAMES = {1: "plain_pipeline", 2: "giver_batch", 3: "xpass_givers", 4: "xpass_block", 5: "subgrid",
             6: "col_tile", 7: "xpass_self", 8: "xpass_stage", 9: "grad_split", 10: "fused_flux",
             11: "radix_order", 12: "col_f64", 13: "col_pipe", 14: "last_bins"}
ENS_OPT_GRAPH, ENS_OPT_BATCH, ENS_OPT_PLAN = 1, 2, 1000
DECOMP_HELICAL, DECOMP_TOROIDAL_POLOIDAL = 1, 2
K_CLASSES = ("fft_column", "fft_row_inverse", "fft_row_forward", "density_shell", "reduce", "other",
             "fused_x")

# every entry point of include/fit_hip.h with its ctypes signature
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "fit_last_error": (ctypes.c_char_p, []),
    "fit_version": (ctypes.c_char_p, []),
    "fit_plan_create": (_I, [ctypes.POINTER(_P), _I, ctypes.POINTER(ctypes.c_int64), _I, _I, _I, _I,
                             ctypes.POINTER(ctypes.POINTER(ctypes.c_double)),
                             ctypes.POINTER(ctypes.c_int64), _I, _I]),
    "fit_plan_destroy": (_I, [_P]),
    "fit_plan_device_bytes": (ctypes.c_int64, [_P]),
    "fit_plan_live_count": (ctypes.c_int64, []),
    "fit_plan_set_option": (_I, [_P, _I, _I]),
    "fit_plan_get_option": (_I, [_P, _I, ctypes.POINTER(_I)]),
    "fit_plan_autotune": (_I, [_P, _I, _I, _P, _P, _I, _I, _I]),
    "fit_plan_tuning": (_I, [_P, ctypes.c_char_p, _I]),
    "fit_plan_subgrid": (_I, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "fit_plan_subgrid_classes": (_I, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(_I), _I]),
    "fit_nonlinear_term": (_I, [_P, _P, _P, _P, _I, _I]),
    "fit_transfer_density": (_I, [_P, _I, _P, _P, _P, _I, _I]),
    "fit_spectral_flux": (_I, [_P, _I, _P, _P, _D, _D, _I, _I]),
    "fit_spectral_flux_with_nhat": (_I, [_P, _I, _P, _P, _D, _D, _I, _I]),
    "fit_spectral_fluxes": (_I, [_P, _I, ctypes.POINTER(_I), _P, _P, _D, _D, _I, _I]),
    "fit_spectral_fluxes_with_nhat": (_I, [_P, _I, ctypes.POINTER(_I), _P, _P, _D, _D, _I, _I]),
    "fit_spectral_flux_with_nhat_noalloc": (_I, [_P, _I, _P, _P, _P, _P, _I, _I]),
    "fit_shell_to_shell": (_I, [_P, _I, _P, _P, _D, _D, _D, _I, _I, _I]),
    "fit_mode_to_mode": (_I, [_P, _I, _P, _P, _P, _P, _I, _I]),
    "fit_helical_decompose": (_I, [_P, _P, _P, _P, _I, _I]),
    "fit_decompose": (_I, [_P, _I, _P, _P, _P, _I, _I]),
    "fit_physical_to_spectral": (_I, [_P, _P, _I, _P, _I, _I]),
    "fit_band_to_band": (_I, [_P, _I, _D, _I, ctypes.c_double, _D, _P, _P, _D, _D, _D, _I, _I]),
    "fit_profile_enable": (_I, [_P, _I]),
    "fit_profile_query": (_I, [_P, _I, ctypes.POINTER(ctypes.c_int64), _D, _D]),
    "fit_plan_stream": (_P, [_P]),
    "fit_profile_kernel_count": (_I, [_P]),
    "fit_profile_kernel": (_I, [_P, _I, ctypes.c_char_p, _I, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_int64),
                                _D, _D]),
    "fit_ensemble_profile_kernel_count": (_I, [_P]),
    "fit_ensemble_profile_kernel": (_I, [_P, _I, ctypes.c_char_p, _I, ctypes.POINTER(_I),
                                         ctypes.POINTER(ctypes.c_int64), _D, _D]),
    "fit_plan_wait_stream": (_I, [_P, _P]),
    "fit_ensemble_wait_stream": (_I, [_P, _P]),
    "fit_ensemble_create": (_I, [ctypes.POINTER(_P), _I, _I, ctypes.POINTER(ctypes.c_int64), _I, _I,
                                 ctypes.POINTER(ctypes.POINTER(ctypes.c_double)),
                                 ctypes.POINTER(ctypes.POINTER(ctypes.c_int64)), ctypes.POINTER(_I), _I,
                                 _I]),
    "fit_ensemble_destroy": (_I, [_P]),
    "fit_ensemble_groups": (_I, [_P]),
    "fit_ensemble_set_option": (_I, [_P, _I, _I]),
    "fit_ensemble_flux": (_I, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P), _D, _D, _D, _D, _I, _I]),
    "fit_ensemble_profile_enable": (_I, [_P, _I]),
    "fit_ensemble_profile_query": (_I, [_P, _I, ctypes.POINTER(ctypes.c_int64), _D, _D]),
    "fit_ensemble_stream": (_P, [_P]),
}


class DimensionMismatch(ValueError):
    """Shape errors (the reference throws Julia's DimensionMismatch)."""


_ERRORS = {1: ValueError, 2: DimensionMismatch, 3: NotImplementedError, 4: MemoryError,
           5: RuntimeError}

_lib = None


def load_library(path=None):
    """Load libfit_hip.so once; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build the HIP backend first (python -c 'import __graft_entry__ as g; "
            "g.build()' or python flowinvarianttransfer.jl_amd/build_lib.py)")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def kernel_stats(count_fn, kernel_fn, handle):
    """{label: {"class", "launches", "ms", "alg_bytes"}} from the per-kernel profiler."""
    out = {}
    for i in range(count_fn(handle)):
        name = ctypes.create_string_buffer(256)
        cls = ctypes.c_int()
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        by = ctypes.c_double()
        check(kernel_fn(handle, i, name, 256, ctypes.byref(cls), ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by)))
        out[name.value.decode()] = {"class": K_CLASSES[cls.value], "launches": n.value, "ms": ms.value,
                                    "alg_bytes": by.value}
    return out


def check(rc):
    if rc != 0:
        msg = load_library().fit_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, RuntimeError)(msg or f"libfit_hip error code {rc}")
