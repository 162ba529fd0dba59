"""The largest lattice the library accepts, on the GPU: 1126^3 = 1.43 G sites
(3V = 4.28 G links, just under the 2^32 limit of the kernels' 32-bit link
indices; 160 GB of FP64 links, RNG state and acceptance counters).  Checked
at sampled half-sites, including the last one:
- setup_kernel: the six RNG words equal the checker's MRG32k3a stream h;
- the hot start (functions/su2.h:29-34): each link is the normalised vector
  of four uniform_double - 0.5 draws of stream h, taken for parity 0 then 1
  and mu = y, x, t; recomputed here in IEEE double (the squared norm summed
  with fused multiply-adds, as the kernels do) and compared bit for bit;
- one heat-bath sweep: every sampled site counts 3 accepted updates per
  parity, the links stay unitary after normalize, and en_av is deterministic.
"""
import ctypes as C
import math
import sys
from fractions import Fraction

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1126
SEED = 2024


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _read(lib, base, offset, nbytes, dtype):
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
    rc = lib.su2_memcpy_d2h(out.ctypes.data, C.c_void_p(base + offset), nbytes, None)
    assert rc == 0, lib.su2_last_error()
    return out


def _fma(a, b, c):
    """a * b + c rounded once (exact rationals)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _rsqrt(x):
    """The kernels' reciprocal square root: four Newton steps
    y <- y + y (1/2 - (x/2) y^2) from the bit-level seed, each fused
    multiply-add rounded once (its one-ulp accuracy is checked against mpmath
    in test_fmath.py)."""
    import struct

    bits = struct.unpack("<Q", struct.pack("<d", x))[0]
    y = struct.unpack("<d", struct.pack("<Q", 0x5FE6EB50C7B537A9 - (bits >> 1)))[0]
    h = 0.5 * x
    for _ in range(4):
        y = _fma(y, _fma(-h * y, y, 0.5), y)
    return y


def _hot_link(r):
    a = [r.uniform_double() - 0.5 for _ in range(4)]
    det = _fma(a[3], a[3], _fma(a[2], a[2], _fma(a[1], a[1], a[0] * a[0])))
    inv = _rsqrt(det)
    assert abs(inv * math.sqrt(det) - 1.0) < 3e-16
    return np.array([x * inv for x in a])


def test_largest_lattice(gpu, oracle_mod):
    q = gpu
    from qcd_2d_gpu_amd import _lib as L

    lib = L.load()
    V = N ** 3
    Vh = V // 2
    assert 3 * V < 2 ** 32
    _say(f"[limits] {N}^3: V = {V}, Vh = {Vh}, 3V = {3 * V}")
    try:
        g = q.Lattice(N, N, N, 3.0)
    except L.Su2Error as e:
        if e.code == L.SU2_ENOMEM:
            pytest.skip(f"not enough device memory for {N}^3: {e}")
        raise
    try:
        samples = sorted({0, 1, 12345, Vh // 2, Vh - N * N // 2 - 1, Vh - 2, Vh - 1})
        g.setup_kernel(SEED)
        g.synchronize()
        _say("[limits] setup_kernel done")
        st = g.states.value
        for h in samples:
            dev = [int(_read(lib, st, (i * Vh + h) * 4, 4, np.uint32)[0]) for i in range(6)]
            assert dev == oracle_mod.Rng(SEED, h).state, f"setup state of stream {h}"

        g.rand_init2()
        g.synchronize()
        _say("[limits] hot start done")
        phi = g.phi.value
        for h in samples:
            r = oracle_mod.Rng(SEED, h)
            for p in range(2):
                for mu in range(3):
                    want = _hot_link(r)
                    got = _read(lib, phi, ((p * 3 + mu) * Vh + h) * 32, 32, np.float64)
                    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (h, p, mu, got, want)
            dev = [int(_read(lib, st, (i * Vh + h) * 4, 4, np.uint32)[0]) for i in range(6)]
            assert dev == r.state, f"stream {h} after the hot start"

        g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_LINK)
        g.synchronize()
        _say("[limits] heat-bath sweep done")
        acc = g.accept_buf.value
        for h in samples:
            for p in range(2):
                assert int(_read(lib, acc, (p * Vh + h) * 4, 4, np.int32)[0]) == 3, (h, p)
        e1, p1 = g.en_av()
        e2, p2 = g.en_av()
        assert (e1, p1) == (e2, p2)
        assert 0.0 < p1 < 1.0 and -2.0 < e1 < 0.0, (e1, p1)
        _say(f"[limits] en_av E = {e1:.6f}, P = {p1:.6f}")

        g.normalize()
        g.synchronize()
        for h in samples:
            for k in range(6):
                a = _read(lib, phi, (k * Vh + h) * 32, 32, np.float64)
                assert abs(float(a @ a) - 1.0) < 1e-14, (h, k, a)
        _say("[limits] normalize done")
    finally:
        g.close()
