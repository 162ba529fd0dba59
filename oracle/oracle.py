"""ctypes front end of the CPU checker liboracle_su2.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg; never by the product package.

`Oracle` holds a lattice in the reference's own layout (double[12 V],
functions/lattice_setup.h:27-32), one MRG32k3a state per half-site, and
int accept[V], and runs the reference's operations on them on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_su2.so")


class OrcLattice(C.Structure):
    """struct lattice (functions/lattice_setup.h:2-14), used as the half
    lattice l2 of kernel.cu:59."""

    _fields_ = [("y", C.c_int), ("x", C.c_int), ("t", C.c_int), ("V", C.c_int), ("d", C.c_int),
                ("dg", C.c_int), ("dt", C.c_int), ("beta", C.c_float), ("epsilon", C.c_float),
                ("updates", C.c_int)]


class OrcRng(C.Structure):
    _fields_ = [("g1", C.c_uint32 * 3), ("g2", C.c_uint32 * 3)]


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    LP = C.POINTER(OrcLattice)
    sig = {
        "orc_mrg_next": (C.c_uint32, [P]),
        "orc_uniform": (C.c_float, [P]),
        "orc_uniform_double": (D, [P]),
        "orc_mrg_init": (None, [C.c_uint64, C.c_uint64, C.c_uint64, P]),
        "orc_setup_kernel": (None, [P, I, C.c_uint64]),
        "orc_init": (None, [P, P, LP, I, I]),
        "orc_mc_update": (None, [P, LP, P, P, I, I, I, I, I, I]),
        "orc_mc_up": (None, [P, LP, P, P, I, I, I, I]),
        "orc_normalize": (None, [P, C.c_longlong, I]),
        "orc_en_av": (None, [P, LP, I, C.POINTER(D), C.POINTER(D)]),
        "orc_time_sweeps": (D, [P, LP, P, P, I, I, I, I, I]),
        "orc_max_threads": (I, []),
        "orc_set_threads": (None, [I]),
        "orc_heat_bath": (I, [P, P, D, P]),
        "orc_metropolis": (I, [P, P, D, D, P]),
        "orc_metropolis_accept": (I, [D, D]),
        "orc_over_relaxation": (I, [P, P]),
        "su2_product": (None, [P, P, P]),
        "su2_product_tc": (None, [P, P, P]),
        "su2_product_tc1": (None, [P, P, P]),
        "su2_product_tc2": (None, [P, P, P]),
        "su2_productf": (None, [P, P, P]),
        "su2_product_tcf": (None, [P, P, P]),
        "su2_product_tc2f": (None, [P, P, P]),
        "su2_prod_trace": (D, [P, P]),
        "su2_det": (D, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Rng:
    """One MRG32k3a stream of the checker."""

    def __init__(self, seed: int, subsequence: int = 0, offset: int = 0):
        self.lib = load()
        self.st = OrcRng()
        self.lib.orc_mrg_init(seed, subsequence, offset, C.byref(self.st))

    @property
    def state(self) -> list[int]:
        return list(self.st.g1) + list(self.st.g2)

    def next(self) -> int:
        return self.lib.orc_mrg_next(C.byref(self.st))

    def uniform(self) -> float:
        return self.lib.orc_uniform(C.byref(self.st))

    def uniform_double(self) -> float:
        return self.lib.orc_uniform_double(C.byref(self.st))


class Oracle:
    """A lattice on the CPU checker, reference layout."""

    def __init__(self, Lt: int, Lx: int, Ly: int, beta: float, epsilon: float = 0.1, updates: int = 1,
                 f32: bool = False):
        if Lt % 2 or Lx % 2 or Ly % 2:
            raise ValueError("extents must be even")
        self.lib = load()
        self.Lt, self.Lx, self.Ly = Lt, Lx, Ly
        self.V = Lt * Lx * Ly
        self.Vh = self.V // 2
        self.f32 = int(bool(f32))
        self.l2 = OrcLattice(Ly // 2, Lx, Lt, self.Vh, 3, 4, 12, beta, epsilon, updates)
        self.phi = np.zeros(12 * self.V, dtype=np.float64)
        self.state = np.zeros((self.Vh, 6), dtype=np.uint32)
        self.accept = np.zeros(self.V, dtype=np.int32)

    def set_params(self, beta=None, epsilon=None, updates=None):
        if beta is not None:
            self.l2.beta = beta
        if epsilon is not None:
            self.l2.epsilon = epsilon
        if updates is not None:
            self.l2.updates = updates

    def setup_kernel(self, seed: int) -> None:
        self.lib.orc_setup_kernel(ptr(self.state), self.Vh, seed)

    def init(self, start: int) -> None:
        """start: 0 rand_init2, 1 id_init2, 2 rand_init2_pcc, 3 Haar, 4 Haar+PCC"""
        self.lib.orc_init(ptr(self.phi), ptr(self.state), C.byref(self.l2), start, self.f32)

    def start(self, start: int, bc: int) -> None:
        """kernel.cu:85-136; start 0 hot, 1 cold, 2 Haar."""
        if start == 1:
            self.init(1)
        elif start == 2:
            self.init(4 if bc == 1 else 3)
        else:
            self.init(2 if bc == 1 else 0)

    def mc_up(self, bc: int, algo: int, update: int = 0) -> None:
        self.lib.orc_mc_up(ptr(self.phi), C.byref(self.l2), ptr(self.accept), ptr(self.state), bc, algo, update,
                           self.f32)

    def mc_update(self, bc: int, algo: int, update: int, mu: int, parity: int) -> None:
        self.lib.orc_mc_update(ptr(self.phi), C.byref(self.l2), ptr(self.accept), ptr(self.state), bc, algo,
                               update, mu, parity, self.f32)

    def normalize(self) -> None:
        self.lib.orc_normalize(ptr(self.phi), 3 * self.V, self.f32)

    def en_av(self, pmc: bool = False) -> tuple[float, float]:
        e, p = C.c_double(0.0), C.c_double(0.0)
        self.lib.orc_en_av(ptr(self.phi), C.byref(self.l2), int(pmc), C.byref(e), C.byref(p))
        return e.value, p.value

    def rng_planes(self) -> np.ndarray:
        """States as uint32[6, V/2] planes, the device layout."""
        return np.ascontiguousarray(self.state.T)

    def time_sweeps(self, bc: int, algo: int, or_per: int, nsweeps: int) -> float:
        return self.lib.orc_time_sweeps(ptr(self.phi), C.byref(self.l2), ptr(self.accept), ptr(self.state), bc,
                                        algo, or_per, nsweeps, self.f32)


def max_threads() -> int:
    return load().orc_max_threads()


def set_threads(n: int) -> None:
    load().orc_set_threads(n)
