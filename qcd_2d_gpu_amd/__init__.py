"""qcd_2d_gpu_amd - MI355X-native SU(2) Yang-Mills Monte Carlo (2+1 D).

The compute path is libsu2mc.so (hand-written HIP kernels for gfx950 behind
the C ABI of include/su2mc.h).  This package is its Python host mirror:
`Lattice` exposes the reference's per-lattice operations and `run` its
driver (kernel.cu:main).
"""
from . import _lib
from ._lib import (ALGO_HEATBATH, ALGO_METROPOLIS, BC_PCC, BC_PERIODIC, BC_PMC, PREC_F32, PREC_F64, START_COLD,
                   START_HAAR, START_HOT, UPDATE_LINK, UPDATE_OVERRELAX, Su2Error)
from .lattice import DeviceBuffer, Lattice, make_lattice
from .run import run
from .engine import Engine, Job, Server
from . import stats

__all__ = [
    "ALGO_HEATBATH", "ALGO_METROPOLIS", "BC_PCC", "BC_PERIODIC", "BC_PMC", "PREC_F32", "PREC_F64", "START_COLD",
    "START_HAAR", "START_HOT", "UPDATE_LINK", "UPDATE_OVERRELAX", "Su2Error", "DeviceBuffer", "Lattice",
    "make_lattice", "run", "stats", "load_library", "device_count", "Engine", "Job", "Server",
]


def load_library():
    """Load libsu2mc.so; raises if it is not built."""
    return _lib.load()


def device_count() -> int:
    import ctypes

    n = ctypes.c_int(0)
    rc = _lib.load().su2_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0
