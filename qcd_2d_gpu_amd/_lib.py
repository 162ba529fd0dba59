"""ctypes binding of libsu2mc.so (include/su2mc.h).

The library is loaded from this package directory only; importing this
module never builds anything.  A missing library raises ImportError-like
RuntimeError with the build command, so no caller can silently fall back to
another implementation.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsu2mc.so")

SU2_OK = 0
SU2_EINVAL = -1
SU2_EODD_EXTENT = -2
SU2_ENOMEM = -3
SU2_EHIP = -4
SU2_EIO = -5
SU2_ENODEV = -6
SU2_EBUDGET = -7
SU2_ECANCELLED = -8

ALGO_METROPOLIS = 0
ALGO_HEATBATH = 1
UPDATE_LINK = 0
UPDATE_OVERRELAX = 1
BC_PERIODIC = 0
BC_PCC = 1
BC_PMC = 2
START_HOT = 0
START_COLD = 1
START_HAAR = 2
PREC_F64 = 0
PREC_F32 = 1
BACKEND_CPU = 0
BACKEND_HIP = 1


class Su2Lattice(C.Structure):
    """struct su2_lattice: same members and order as the reference's struct
    lattice (functions/lattice_setup.h:2-14), full-lattice extents."""

    _fields_ = [
        ("y", C.c_int), ("x", C.c_int), ("t", C.c_int), ("V", C.c_int),
        ("d", C.c_int), ("dg", C.c_int), ("dt", C.c_int),
        ("beta", C.c_float), ("epsilon", C.c_float), ("updates", C.c_int),
    ]


class Su2RunConfig(C.Structure):
    _fields_ = [
        ("Lt", C.c_int), ("Lx", C.c_int), ("Ly", C.c_int),
        ("beta", C.c_float),
        ("boundary", C.c_int), ("algo", C.c_int), ("readconf", C.c_int), ("init", C.c_int),
        ("sconf", C.c_int), ("save", C.c_int),
        ("N", C.c_int), ("N_term", C.c_int), ("N_cor", C.c_int), ("Nsave", C.c_int),
        ("epsilon", C.c_float), ("upd", C.c_int), ("W", C.c_int), ("or_sweeps", C.c_int),
        ("precision", C.c_int), ("normalize_every", C.c_int),
        ("seed", C.c_uint64),
        ("outdir", C.c_char_p), ("nsave", C.c_char_p), ("nconfr", C.c_char_p), ("nconfs", C.c_char_p),
        ("verbose", C.c_int), ("device", C.c_int), ("N_block", C.c_int), ("Nblock_term", C.c_int),
        ("fp64_twin", C.c_int),
    ]


class Su2RunResult(C.Structure):
    _fields_ = [
        ("energy", C.POINTER(C.c_double)), ("plaquette", C.POINTER(C.c_double)),
        ("E_mean", C.c_double), ("E_err", C.c_double), ("P_mean", C.c_double), ("P_err", C.c_double),
        ("accept", C.c_double), ("det_mean", C.c_double),
        ("seconds_therm", C.c_double), ("seconds_meas", C.c_double),
        ("link_updates", C.c_longlong), ("Nterms", C.c_int),
        ("P_fp64_twin", C.c_double), ("P_fp64_twin_err", C.c_double), ("P_err_binned", C.c_double),
        ("P_diff_fp32_fp64", C.c_double), ("P_diff_sigma", C.c_double),
    ]


_P = C.c_void_p
_L = C.POINTER(Su2Lattice)
_U32P = C.c_void_p
_I = C.c_int

# name -> (restype, argtypes); every symbol include/su2mc.h declares
PROTOTYPES = {
    "su2_lattice_init": (_I, [_L, _I, _I, _I, C.c_float, C.c_float, _I]),
    "su2_field_bytes": (C.c_size_t, [_L, _I]),
    "su2_rng_bytes": (C.c_size_t, [_L]),
    "su2_accept_bytes": (C.c_size_t, [_L]),
    "su2_en_av_work_bytes": (C.c_size_t, [_L]),
    "su2_setup_kernel": (_I, [_L, _U32P, C.c_uint64, _P]),
    "su2_id_init2": (_I, [_L, _P, _I, _P]),
    "su2_rand_init2": (_I, [_L, _P, _I, _U32P, _P]),
    "su2_rand_init2_pcc": (_I, [_L, _P, _I, _U32P, _P]),
    "su2_haar_init": (_I, [_L, _P, _I, _U32P, _I, _P]),
    "su2_mc_up_per": (_I, [_L, _P, _I, _P, _U32P, _I, _I, _P]),
    "su2_mc_up_pcc": (_I, [_L, _P, _I, _P, _U32P, _I, _I, _P]),
    "su2_mc_up_pmc": (_I, [_L, _P, _I, _P, _U32P, _I, _I, _P]),
    "su2_mc_update": (_I, [_L, _P, _I, _P, _U32P, _I, _I, _I, _I, _I, _P]),
    "su2_mc_up_measured": (_I, [_L, _P, _I, _P, _U32P, _I, _I, _I, _P, _P, _P, _P]),
    "su2_normalize": (_I, [_L, _P, _I, _P]),
    "su2_en_av": (_I, [_L, _P, _I, _I, _P, _P, _P, _P]),
    "su2_en_av_noalloc": (_I, [_L, _P, _I, _I, _P, _P, _P]),
    "su2_accept_clear": (_I, [_L, _P, _P]),
    "su2_mean_lattice": (_I, [_L, _P, C.c_longlong, C.POINTER(C.c_double), C.POINTER(C.c_double), _P]),
    "su2_field_to_ref": (_I, [_L, _P, _I, _P, _P]),
    "su2_field_from_ref": (_I, [_L, _P, _P, _I, _P]),
    "su2_device_count": (_I, [C.POINTER(C.c_int)]),
    "su2_set_device": (_I, [_I]),
    "su2_malloc": (_I, [C.POINTER(C.c_void_p), C.c_size_t]),
    "su2_free": (_I, [_P]),
    "su2_memcpy_h2d": (_I, [_P, _P, C.c_size_t, _P]),
    "su2_memcpy_d2h": (_I, [_P, _P, C.c_size_t, _P]),
    "su2_memset_zero": (_I, [_P, C.c_size_t, _P]),
    "su2_device_copy": (_I, [_P, _P, C.c_size_t, _P]),
    "su2_stream_probe": (_I, [_P, _P, C.c_size_t, _I, _I, _I, _I, _P]),
    "su2_phase_probe": (_I, [_P, _I, _I, _I, _P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                         C.POINTER(C.c_double)]),
    "su2_stream_create": (_I, [C.POINTER(C.c_void_p)]),
    "su2_stream_destroy": (_I, [_P]),
    "su2_stream_synchronize": (_I, [_P]),
    "su2_event_create": (_I, [C.POINTER(C.c_void_p)]),
    "su2_event_destroy": (_I, [_P]),
    "su2_event_record": (_I, [_P, _P]),
    "su2_event_elapsed_ms": (_I, [_P, _P, C.POINTER(C.c_float)]),
    "su2_strerror": (C.c_char_p, [_I]),
    "su2_last_error": (C.c_char_p, []),
    "su2_abi_version": (_I, []),
    "su2_set_sweep_block": (_I, [_I]),
    "su2_ctx_create": (_I, [_I, _I, C.POINTER(C.c_void_p)]),
    "su2_ctx_destroy": (_I, [_P]),
    "su2_ctx_stream": (_I, [_P, C.POINTER(C.c_void_p)]),
    "su2_ctx_synchronize": (_I, [_P]),
    "su2_lattice_create": (_I, [_P, _I, _I, _I, _I, _I, C.POINTER(C.c_void_p)]),
    "su2_lattice_destroy": (_I, [_P]),
    "su2_lattice_get": (_I, [_P, _L, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "su2_lattice_buffers": (_I, [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "su2_rng_init": (_I, [_P, C.c_uint64]),
    "su2_init": (_I, [_P, _I]),
    "su2_sweep": (_I, [_P, _I, _I, C.c_float, C.c_float, _I]),
    "su2_renormalize": (_I, [_P]),
    "su2_measure": (_I, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "su2_accept_read": (_I, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "su2_accept_reset": (_I, [_P]),
    "su2_field_download": (_I, [_P, _P]),
    "su2_field_upload": (_I, [_P, _P]),
    "su2_rng_download": (_I, [_P, _P]),
    "su2_rng_upload": (_I, [_P, _P]),
    "su2_run_config_default": (None, [C.POINTER(Su2RunConfig)]),
    "su2_run": (_I, [C.POINTER(Su2RunConfig), C.POINTER(Su2RunResult)]),
    "su2_binned_mean_error": (_I, [C.POINTER(C.c_double), _I, _I, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "su2_series_compare": (_I, [C.POINTER(C.c_double), C.POINTER(C.c_double), _I, _I, C.POINTER(C.c_double),
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]),
}

_lib = None


class Su2Error(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        super().__init__(f"{where}: status {code} ({detail})")


def load() -> C.CDLL:
    """Load libsu2mc.so from the package directory (built by
    `python -c 'import __graft_entry__ as g; g.build()'` or
    `python qcd_2d_gpu_amd/build.py`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libsu2mc.so is not built ({LIB_PATH}); run python qcd_2d_gpu_amd/build.py")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.su2_abi_version() != 5:
        raise RuntimeError("libsu2mc ABI version mismatch")
    _lib = lib
    return lib


def check(code: int, where: str) -> None:
    if code != SU2_OK:
        lib = load()
        detail = lib.su2_last_error().decode(errors="replace")
        raise Su2Error(code, where, detail)
