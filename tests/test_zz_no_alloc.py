"""su2_en_av_noalloc (the en_av entry point that allocates nothing: its
partial sums live in the library's static device array) against su2_en_av
with a caller-allocated work buffer, through the C ABI: identical bits on
two calls in a row, the second on a smaller lattice, whose partials overlap
the first call's leftovers in the shared array; oversized lattices are
refused before anything is launched."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _measure_both(q, lib, Lt, Lx, Ly, bc, prec, seed):
    """One lattice after a few sweeps: (E, P) from su2_en_av and from
    su2_en_av_noalloc, each accumulated into zeroed device slots."""
    from qcd_2d_gpu_amd import _lib as L

    g = q.Lattice(Lt, Lx, Ly, 3.5, precision=prec)
    slots = q.DeviceBuffer(4 * 8)
    work = q.DeviceBuffer(lib.su2_en_av_work_bytes(C.byref(g.l)))
    try:
        g.setup_kernel(seed)
        g.start(q.START_HOT, bc)
        for _ in range(3):
            g.mc_up(bc, q.ALGO_HEATBATH, 0)
        slots.zero()
        base = slots.value
        L.check(lib.su2_en_av(C.byref(g.l), g.phi.ptr, prec, int(bc == q.BC_PMC), C.c_void_p(base),
                              C.c_void_p(base + 8), work.ptr, None), "su2_en_av")
        L.check(lib.su2_en_av_noalloc(C.byref(g.l), g.phi.ptr, prec, int(bc == q.BC_PMC), C.c_void_p(base + 16),
                                      C.c_void_p(base + 24), None), "su2_en_av_noalloc")
        out = slots.download(np.float64, 4)
        return out[:2], out[2:]
    finally:
        work.free()
        slots.free()
        g.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_en_av_equals_en_av(gpu, prec):
    q = gpu
    lib = q.load_library()
    print(f"[noalloc] en_av_noalloc vs en_av prec={prec}", flush=True)
    # first call: 48 x 40 x 64 (61 440 half-sites); second, smaller: 10 x 12 x 14
    for (Lt, Lx, Ly, bc, seed) in ((48, 40, 64, q.BC_PERIODIC, 11), (10, 12, 14, q.BC_PMC, 12)):
        a, b = _measure_both(q, lib, Lt, Lx, Ly, bc, prec, seed)
        print(f"[noalloc]   {Lt}x{Lx}x{Ly} bc={bc}: E = {a[0]:.12f} P = {a[1]:.12f}", flush=True)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (Lt, Lx, Ly, a, b)
        assert -1.0 < a[0] < 0.0 and 0.0 < a[1] < 1.0


def test_noalloc_en_av_refuses_oversized(gpu):
    q = gpu
    from qcd_2d_gpu_amd import _lib as L

    lib = q.load_library()
    l = L.Su2Lattice()
    # 512 x 512 x 258 = 67 633 152 sites > 2^26: refused before any launch
    assert lib.su2_lattice_init(C.byref(l), 512, 512, 258, 3.0, 0.1, 1) == 0
    rc = lib.su2_en_av_noalloc(C.byref(l), C.c_void_p(256), 0, 0, C.c_void_p(256), C.c_void_p(512), None)
    assert rc == L.SU2_EINVAL
    assert b"scratch limit" in lib.su2_last_error()
