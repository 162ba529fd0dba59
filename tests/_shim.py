"""ctypes loader of tests/shim/libsu2shim.so (host build of the product's
numerics headers)."""
import ctypes as C
import os
import subprocess

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shim")
PATH = os.path.join(HERE, "libsu2shim.so")
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = C.CDLL(PATH)
    P, D, F, I, U32, U64 = C.c_void_p, C.c_double, C.c_float, C.c_int, C.c_uint32, C.c_uint64
    sig = {
        "shim_logf": (F, [F]), "shim_log_d": (D, [D]), "shim_cospif": (F, [F]),
        "shim_sincospif": (None, [F, C.POINTER(F), C.POINTER(F)]),
        "shim_sincospi_d": (None, [D, C.POINTER(D), C.POINTER(D)]), "shim_cospi_d": (D, [D]),
        "shim_expf": (F, [F]), "shim_exp": (D, [D]), "shim_sqrtf": (F, [F]), "shim_rsqrt": (D, [D]),
        "shim_mrg_seed": (None, [U64, P]), "shim_mrg_jump": (None, [P, U32]), "shim_mrg_next": (U32, [P]),
        "shim_uniform": (F, [P]), "shim_uniform_double": (D, [P]), "shim_jump_table": (C.POINTER(U32), [I]),
        "shim_heat_bath": (None, [P, D, P, P]), "shim_metropolis": (I, [P, P, D, D, I, P]),
        "shim_metropolis_accept": (I, [D, D]), "shim_over_relax": (None, [P, P, P]),
        "shim_rand_link": (None, [P, P]), "shim_haar_link": (None, [P, P]), "shim_q_mul": (None, [P, P, P, I]),
        "shim_qf_mul": (None, [P, P, P, I]),
        "shim_q_scalar": (D, [P, P, I]),
        "shim_fmath_max_ulp": (D, [I, U32, U32, I, C.POINTER(U32), C.POINTER(C.c_longlong)]),
    }
    for n, (r, a) in sig.items():
        f = getattr(lib, n)
        f.restype = r
        f.argtypes = a
    _lib = lib
    return lib
