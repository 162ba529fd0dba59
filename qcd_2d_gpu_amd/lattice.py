"""Host-side mirror of the reference's interface for the sweep path.

`Lattice` owns the device buffers kernel.cu allocates (field, RNG states,
acceptance counters; kernel.cu:77-82,144-151) and exposes the reference's
functions under their own names -- setup_kernel, rand_init2, id_init2,
rand_init2_pcc, mc_up_per, mc_up_pcc, mc_up_pmc, normalize, en_av,
mean_lattice -- each a thin call through the C ABI of libsu2mc.so.  Every
computation runs in the HIP library; there is no host fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class DeviceBuffer:
    """A device allocation made through su2_malloc."""

    def __init__(self, nbytes: int):
        self.lib = L.load()
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        L.check(self.lib.su2_malloc(C.byref(self.ptr), self.nbytes), "su2_malloc")

    @property
    def value(self) -> int:
        return self.ptr.value or 0

    def upload(self, arr: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise ValueError("upload larger than the buffer")
        L.check(self.lib.su2_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, stream), "su2_memcpy_h2d")

    def download(self, dtype, count: int | None = None, stream=None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else int(count)
        out = np.empty(n, dtype=dt)
        L.check(self.lib.su2_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, stream), "su2_memcpy_d2h")
        return out

    def zero(self, stream=None) -> None:
        L.check(self.lib.su2_memset_zero(self.ptr, self.nbytes, stream), "su2_memset_zero")

    def free(self) -> None:
        if self.ptr and self.ptr.value:
            self.lib.su2_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def make_lattice(Lt: int, Lx: int, Ly: int, beta: float, epsilon: float = 0.1, updates: int = 1) -> L.Su2Lattice:
    """struct lattice for an Lt x Lx x Ly lattice (kernel.cu:58); raises on
    odd extents (kernel.cu:61-64)."""
    lib = L.load()
    lat = L.Su2Lattice()
    L.check(lib.su2_lattice_init(C.byref(lat), Lt, Lx, Ly, beta, epsilon, updates), "su2_lattice_init")
    return lat


class Lattice:
    """Device state of one lattice and the reference's operations on it."""

    def __init__(self, Lt: int, Lx: int, Ly: int, beta: float, epsilon: float = 0.1, updates: int = 1,
                 precision: int = L.PREC_F64, stream=None):
        self.lib = L.load()
        self.l = make_lattice(Lt, Lx, Ly, beta, epsilon, updates)
        self.Lt, self.Lx, self.Ly = Lt, Lx, Ly
        self.V = Lt * Lx * Ly
        self.Vh = self.V // 2
        self.precision = precision
        self.stream = stream
        lp = C.byref(self.l)
        self.phi = DeviceBuffer(self.lib.su2_field_bytes(lp, precision))
        self.states = DeviceBuffer(self.lib.su2_rng_bytes(lp))
        self.accept_buf = DeviceBuffer(self.lib.su2_accept_bytes(lp))
        self.work = DeviceBuffer(self.lib.su2_en_av_work_bytes(lp))
        self.work.zero(stream)
        self.res = DeviceBuffer(2 * 8)
        self.accept_clear()

    # -- parameters ---------------------------------------------------
    def set_params(self, beta: float | None = None, epsilon: float | None = None, updates: int | None = None):
        if beta is not None:
            self.l.beta = beta
        if epsilon is not None:
            self.l.epsilon = epsilon
        if updates is not None:
            self.l.updates = updates

    @property
    def lp(self):
        return C.byref(self.l)

    # -- set-up and starts (functions/lattice_setup.h) -----------------
    def setup_kernel(self, seed: int) -> None:
        L.check(self.lib.su2_setup_kernel(self.lp, self.states.ptr, seed, self.stream), "setup_kernel")

    def id_init2(self) -> None:
        L.check(self.lib.su2_id_init2(self.lp, self.phi.ptr, self.precision, self.stream), "id_init2")

    def rand_init2(self) -> None:
        L.check(self.lib.su2_rand_init2(self.lp, self.phi.ptr, self.precision, self.states.ptr, self.stream),
                "rand_init2")

    def rand_init2_pcc(self) -> None:
        L.check(self.lib.su2_rand_init2_pcc(self.lp, self.phi.ptr, self.precision, self.states.ptr, self.stream),
                "rand_init2_pcc")

    def haar_init(self, pcc: bool = False) -> None:
        L.check(self.lib.su2_haar_init(self.lp, self.phi.ptr, self.precision, self.states.ptr, int(pcc),
                                       self.stream), "haar_init")

    def start(self, start: int, bc: int) -> None:
        """The start of kernel.cu:85-136 for a boundary condition."""
        if start == L.START_COLD:
            self.id_init2()
        elif start == L.START_HAAR:
            self.haar_init(pcc=bc == L.BC_PCC)
        elif bc == L.BC_PCC:
            self.rand_init2_pcc()
        else:
            self.rand_init2()

    # -- sweeps (functions/mc_updates*.h) ------------------------------
    def mc_up_per(self, algo: int = L.ALGO_HEATBATH, update: int = L.UPDATE_LINK) -> None:
        L.check(self.lib.su2_mc_up_per(self.lp, self.phi.ptr, self.precision, self.accept_buf.ptr, self.states.ptr,
                                       algo, update, self.stream), "mc_up_per")

    def mc_up_pcc(self, algo: int = L.ALGO_HEATBATH, update: int = L.UPDATE_LINK) -> None:
        L.check(self.lib.su2_mc_up_pcc(self.lp, self.phi.ptr, self.precision, self.accept_buf.ptr, self.states.ptr,
                                       algo, update, self.stream), "mc_up_pcc")

    def mc_up_pmc(self, algo: int = L.ALGO_HEATBATH, update: int = L.UPDATE_LINK) -> None:
        L.check(self.lib.su2_mc_up_pmc(self.lp, self.phi.ptr, self.precision, self.accept_buf.ptr, self.states.ptr,
                                       algo, update, self.stream), "mc_up_pmc")

    def mc_up(self, bc: int, algo: int = L.ALGO_HEATBATH, update: int = L.UPDATE_LINK) -> None:
        (self.mc_up_per, self.mc_up_pcc, self.mc_up_pmc)[bc](algo, update)

    def mc_up_measured(self, bc: int, res_ptr: int, res2_ptr: int, algo: int = L.ALGO_HEATBATH,
                       update: int = L.UPDATE_LINK) -> None:
        """en_av of the current field into caller device slots, then one sweep
        (su2_mc_up_measured): the results of en_av_into followed by mc_up."""
        L.check(self.lib.su2_mc_up_measured(self.lp, self.phi.ptr, self.precision, self.accept_buf.ptr,
                                            self.states.ptr, bc, algo, update, C.c_void_p(res_ptr),
                                            C.c_void_p(res2_ptr), self.work.ptr, self.stream), "mc_up_measured")

    def mc_update(self, bc: int, algo: int, update: int, mu: int, parity: int) -> None:
        L.check(self.lib.su2_mc_update(self.lp, self.phi.ptr, self.precision, self.accept_buf.ptr, self.states.ptr,
                                       bc, algo, update, mu, parity, self.stream), "mc_update")

    def normalize(self) -> None:
        L.check(self.lib.su2_normalize(self.lp, self.phi.ptr, self.precision, self.stream), "normalize")

    # -- observables (functions/observables.h:99-141) ------------------
    def en_av(self, pmc: bool = False) -> tuple[float, float]:
        """Returns (energy, average plaquette) of the current field."""
        self.res.zero(self.stream)
        e = C.c_void_p(self.res.value)
        p = C.c_void_p(self.res.value + 8)
        L.check(self.lib.su2_en_av(self.lp, self.phi.ptr, self.precision, int(pmc), e, p, self.work.ptr,
                                   self.stream), "en_av")
        r = self.res.download(np.float64, 2, self.stream)
        return float(r[0]), float(r[1])

    def en_av_into(self, pmc: bool, res_ptr: int, res2_ptr: int) -> None:
        """en_av accumulating into caller device slots (kernel.cu:208)."""
        L.check(self.lib.su2_en_av(self.lp, self.phi.ptr, self.precision, int(pmc), C.c_void_p(res_ptr),
                                   C.c_void_p(res2_ptr), self.work.ptr, self.stream), "en_av")

    # -- acceptance (statistics.h:17-28) --------------------------------
    def accept_clear(self) -> None:
        L.check(self.lib.su2_accept_clear(self.lp, self.accept_buf.ptr, self.stream), "accept_clear")

    def accept(self) -> np.ndarray:
        return self.accept_buf.download(np.int32, self.V, self.stream)

    def mean_lattice(self, n: int) -> tuple[float, float]:
        m, e = C.c_double(), C.c_double()
        L.check(self.lib.su2_mean_lattice(self.lp, self.accept_buf.ptr, n, C.byref(m), C.byref(e), self.stream),
                "mean_lattice")
        return m.value, e.value

    # -- field and RNG state access ------------------------------------
    def field_ref(self) -> np.ndarray:
        """The field in the reference layout, double[12 V] (lattice_setup.h:
        27-32)."""
        tmp = DeviceBuffer(self.V * 12 * 8)
        L.check(self.lib.su2_field_to_ref(self.lp, self.phi.ptr, self.precision, tmp.ptr, self.stream), "to_ref")
        out = tmp.download(np.float64, self.V * 12, self.stream)
        tmp.free()
        return out

    def set_field_ref(self, phi: np.ndarray) -> None:
        a = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
        if a.size != self.V * 12:
            raise ValueError("field must hold 12 V doubles")
        tmp = DeviceBuffer(a.nbytes)
        tmp.upload(a, self.stream)
        L.check(self.lib.su2_field_from_ref(self.lp, tmp.ptr, self.phi.ptr, self.precision, self.stream), "from_ref")
        self.synchronize()
        tmp.free()

    def rng_state(self) -> np.ndarray:
        """RNG states as uint32[6, V/2] (g1[0..2], g2[0..2] planes)."""
        return self.states.download(np.uint32, 6 * self.Vh, self.stream).reshape(6, self.Vh)

    def set_rng_state(self, st: np.ndarray) -> None:
        a = np.ascontiguousarray(st, dtype=np.uint32).reshape(6, self.Vh)
        self.states.upload(a, self.stream)

    def synchronize(self) -> None:
        L.check(self.lib.su2_stream_synchronize(self.stream), "stream_synchronize")

    def close(self) -> None:
        for b in (self.phi, self.states, self.accept_buf, self.work, self.res):
            b.free()
