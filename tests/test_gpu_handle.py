"""The handle API of include/su2mc.h (su2_ctx_* / su2_lattice_* / su2_sweep
...) against the CPU checker on the same seeds: links, RNG states and
acceptance sums bit-exact, energy/plaquette to 1e-12 relative, and
field/RNG download-upload round trips that resume a run bit for bit."""
import ctypes as C
import sys

import numpy as np
import pytest

from qcd_2d_gpu_amd import _lib as L

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ok(lib, rc, where):
    assert rc == 0, f"{where}: {rc} {lib.su2_last_error().decode()}"


class Handle:
    def __init__(self, lib, ctx, Lt, Lx, Ly, bc, prec):
        self.lib, self.V, self.bc = lib, Lt * Lx * Ly, bc
        h = C.c_void_p()
        _ok(lib, lib.su2_lattice_create(ctx, Lt, Lx, Ly, bc, prec, C.byref(h)), "lattice_create")
        self.h = h

    def field(self):
        out = np.empty(12 * self.V, dtype=np.float64)
        _ok(self.lib, self.lib.su2_field_download(self.h, out.ctypes.data), "field_download")
        return out

    def rng(self):
        out = np.empty(6 * (self.V // 2), dtype=np.uint32)
        _ok(self.lib, self.lib.su2_rng_download(self.h, out.ctypes.data), "rng_download")
        return out

    def measure(self):
        e, p = C.c_double(), C.c_double()
        _ok(self.lib, self.lib.su2_measure(self.h, C.byref(e), C.byref(p)), "measure")
        return e.value, p.value

    def accept(self):
        a, n = C.c_int64(), C.c_int64()
        _ok(self.lib, self.lib.su2_accept_read(self.h, C.byref(a), C.byref(n)), "accept_read")
        return a.value, n.value

    def close(self):
        if self.h:
            self.lib.su2_lattice_destroy(self.h)
            self.h = None


@pytest.fixture
def ctx(gpu):
    lib = L.load()
    c = C.c_void_p()
    _ok(lib, lib.su2_ctx_create(L.BACKEND_HIP, 0, C.byref(c)), "ctx_create")
    yield lib, c
    _ok(lib, lib.su2_ctx_destroy(c), "ctx_destroy")


def _check(h, o, where):
    f = h.field()
    assert np.array_equal(f.view(np.uint64), np.ascontiguousarray(o.phi, dtype=np.float64).view(np.uint64)), \
        f"{where}: links differ"
    assert np.array_equal(h.rng(), o.rng_planes().ravel()), f"{where}: RNG states differ"
    e, p = h.measure()
    eo, po = o.en_av(h.bc == 2)
    assert abs(e - eo) <= REL_TOL * max(abs(eo), 1e-300) + 1e-300, f"{where}: energy {e!r} vs {eo!r}"
    assert abs(p - po) <= REL_TOL * abs(po), f"{where}: plaquette {p!r} vs {po!r}"
    acc, _ = h.accept()
    assert acc == int(np.asarray(o.accept, dtype=np.int64).sum()), f"{where}: acceptance sums differ"


@pytest.mark.parametrize("bc,algo,prec,hits", [(0, 1, 0, 1), (1, 0, 1, 4), (2, 0, 0, 2), (2, 1, 1, 1)])
def test_handle_sweeps_match_checker(ctx, oracle_mod, bc, algo, prec, hits):
    lib, c = ctx
    Lt, Lx, Ly, beta, eps, seed = 8, 6, 10, 3.5, 0.5, 4321 + bc
    _say(f"[handle] {Lt}x{Lx}x{Ly} bc={bc} algo={algo} prec={prec} hits={hits}")
    h = Handle(lib, c, Lt, Lx, Ly, bc, prec)
    o = oracle_mod.Oracle(Lt, Lx, Ly, beta, eps, hits, f32=(prec == 1))
    try:
        _ok(lib, lib.su2_rng_init(h.h, seed), "rng_init")
        _ok(lib, lib.su2_init(h.h, L.START_HOT), "init")
        o.setup_kernel(seed)
        o.start(L.START_HOT, bc)
        _check(h, o, "start")
        for i in range(3):
            _ok(lib, lib.su2_sweep(h.h, algo, L.UPDATE_LINK, beta, eps, hits), "sweep")
            o.mc_up(bc, algo, 0)
            _ok(lib, lib.su2_sweep(h.h, algo, L.UPDATE_OVERRELAX, beta, eps, hits), "sweep OR")
            o.mc_up(bc, algo, 1)
            _check(h, o, f"sweep {i + 1}")
        _ok(lib, lib.su2_renormalize(h.h), "renormalize")
        o.normalize()
        _check(h, o, "renormalize")
        _, attempts = h.accept()
        per = hits if algo == 0 else 1
        assert attempts == h.V * 3 * (per + 1) * 3
        _ok(lib, lib.su2_accept_reset(h.h), "accept_reset")
        assert h.accept() == (0, 0)
    finally:
        h.close()


def test_handle_checkpoint_resume_is_bit_exact(ctx):
    """field + RNG downloaded mid-run and uploaded into a fresh lattice
    continue exactly as the original run."""
    lib, c = ctx
    _say("[handle] checkpoint/resume 8x8x8 heat-bath")
    a = Handle(lib, c, 8, 8, 8, 0, 0)
    b = Handle(lib, c, 8, 8, 8, 0, 0)
    try:
        _ok(lib, lib.su2_rng_init(a.h, 77), "rng_init")
        _ok(lib, lib.su2_init(a.h, L.START_HOT), "init")
        for _ in range(2):
            _ok(lib, lib.su2_sweep(a.h, 1, 0, 2.5, 0.1, 1), "sweep")
        phi, st = a.field(), a.rng()
        for _ in range(3):
            _ok(lib, lib.su2_sweep(a.h, 1, 0, 2.5, 0.1, 1), "sweep")
        _ok(lib, lib.su2_field_upload(b.h, phi.ctypes.data), "field_upload")
        _ok(lib, lib.su2_rng_upload(b.h, st.ctypes.data), "rng_upload")
        assert np.array_equal(b.field().view(np.uint64), phi.view(np.uint64))
        for _ in range(3):
            _ok(lib, lib.su2_sweep(b.h, 1, 0, 2.5, 0.1, 1), "sweep")
        assert np.array_equal(a.field().view(np.uint64), b.field().view(np.uint64))
        assert np.array_equal(a.rng(), b.rng())
        assert a.measure() == b.measure()
    finally:
        a.close()
        b.close()


def test_handle_cold_start_and_errors(ctx):
    lib, c = ctx
    h = Handle(lib, c, 4, 4, 4, 0, 1)
    try:
        _ok(lib, lib.su2_init(h.h, L.START_COLD), "init")
        assert h.measure() == (0.0, 1.0)
        assert lib.su2_init(h.h, 7) == L.SU2_EINVAL
        assert lib.su2_sweep(h.h, 0, 0, 1.0, 0.1, 0) == L.SU2_EINVAL
        bad = C.c_void_p()
        assert lib.su2_lattice_create(c, 4, 5, 4, 0, 0, C.byref(bad)) == L.SU2_EODD_EXTENT
        assert lib.su2_lattice_create(c, 4, 4, 4, 3, 0, C.byref(bad)) == L.SU2_EINVAL
        assert lib.su2_ctx_destroy(c) == L.SU2_EINVAL  # a lattice is still alive
    finally:
        h.close()


def test_zero_field_sweep_ends(ctx):
    """A heat-bath sweep over an all-zero field (zero staples: the
    Kennedy-Pendleton scale is infinite or NaN) ends after one trial per
    site with NaN links instead of looping for ever; the RNG of every site
    advances by exactly 6 draws per launch it takes part in."""
    lib, c = ctx
    _say("[handle] zero field heat-bath")
    h = Handle(lib, c, 4, 4, 4, 0, 0)
    try:
        _ok(lib, lib.su2_rng_init(h.h, 3), "rng_init")
        zero = np.zeros(12 * h.V)
        _ok(lib, lib.su2_field_upload(h.h, zero.ctypes.data), "field_upload")
        _ok(lib, lib.su2_sweep(h.h, L.ALGO_HEATBATH, L.UPDATE_LINK, 2.0, 0.1, 1), "sweep")
        assert np.isnan(h.field()).all()
        assert h.accept()[0] == 3 * h.V
    finally:
        h.close()
