"""SU(2) algebra in the quaternion basis against 2x2 complex matrices
(functions/su2.h:50-76), for the checker and for the product header."""
import numpy as np

from tests import _naive as N
from tests import _shim


def _rand(rng, n):
    return rng.normal(size=(n, 4))


def test_products_checker_and_product_vs_matrices(oracle_mod):
    lib = oracle_mod.load()
    shim = _shim.load()
    rng = np.random.default_rng(1)
    A, B = _rand(rng, 200), _rand(rng, 200)
    MA, MB = N.quat_to_mat(A), N.quat_to_mat(B)
    want = {
        "su2_product": MA @ MB,
        "su2_product_tc": MA @ N.dag(MB),
        "su2_product_tc1": N.dag(MA) @ MB,
        "su2_product_tc2": N.dag(MA) @ N.dag(MB),
    }
    for kind, (name, W) in enumerate(want.items()):
        for i in range(len(A)):
            a, b = np.ascontiguousarray(A[i]), np.ascontiguousarray(B[i])
            c = np.zeros(4)
            getattr(lib, name)(a.ctypes.data, b.ctypes.data, c.ctypes.data)
            np.testing.assert_allclose(N.quat_to_mat(c), W[i], rtol=0, atol=1e-13)
            c2 = np.zeros(4)
            shim.shim_q_mul(a.ctypes.data, b.ctypes.data, c2.ctypes.data, kind)
            assert np.array_equal(c.view(np.uint64), c2.view(np.uint64)), name  # same rounding
    for i in range(20):
        a, b = np.ascontiguousarray(A[i]), np.ascontiguousarray(B[i])
        assert abs(lib.su2_prod_trace(a.ctypes.data, b.ctypes.data) - N.half_retr(MA[i] @ MB[i])) < 1e-13
        assert abs(lib.su2_det(a.ctypes.data) - np.linalg.det(MA[i]).real) < 1e-12


def test_quaternion_matrices_are_su2_times_norm():
    rng = np.random.default_rng(2)
    A = rng.normal(size=(50, 4))
    A /= np.linalg.norm(A, axis=1, keepdims=True)
    M = N.quat_to_mat(A)
    np.testing.assert_allclose(M @ N.dag(M), np.broadcast_to(np.eye(2), M.shape), atol=1e-14)
    np.testing.assert_allclose(np.linalg.det(M), 1.0, atol=1e-14)


# Term lists of the four products in the reference's order (functions/su2.h:
# 50-73): per component, (sign, index into A, index into B).
_TERMS = {
    "su2_product": [[(1, 0, 0), (-1, 1, 1), (-1, 2, 2), (-1, 3, 3)],
                    [(1, 1, 0), (1, 0, 1), (1, 2, 3), (-1, 3, 2)],
                    [(1, 2, 0), (1, 3, 1), (1, 0, 2), (-1, 1, 3)],
                    [(1, 3, 0), (-1, 2, 1), (1, 0, 3), (1, 1, 2)]],
    "su2_product_tc": [[(1, 0, 0), (1, 1, 1), (1, 2, 2), (1, 3, 3)],
                       [(1, 1, 0), (-1, 0, 1), (-1, 2, 3), (1, 3, 2)],
                       [(1, 2, 0), (-1, 3, 1), (-1, 0, 2), (1, 1, 3)],
                       [(1, 3, 0), (1, 2, 1), (-1, 0, 3), (-1, 1, 2)]],
    "su2_product_tc1": [[(1, 0, 0), (1, 1, 1), (1, 2, 2), (1, 3, 3)],
                        [(-1, 1, 0), (1, 0, 1), (-1, 2, 3), (1, 3, 2)],
                        [(-1, 2, 0), (-1, 3, 1), (1, 0, 2), (1, 1, 3)],
                        [(-1, 3, 0), (1, 2, 1), (1, 0, 3), (-1, 1, 2)]],
    "su2_product_tc2": [[(1, 0, 0), (-1, 1, 1), (-1, 2, 2), (-1, 3, 3)],
                        [(-1, 1, 0), (-1, 0, 1), (1, 2, 3), (-1, 3, 2)],
                        [(-1, 2, 0), (1, 3, 1), (-1, 0, 2), (-1, 1, 3)],
                        [(-1, 3, 0), (-1, 2, 1), (-1, 0, 3), (1, 1, 2)]],
}


def _fused_sum(terms, a, b):
    """First product rounded alone, every later term added with one rounding
    (an exact fused multiply-add, computed with rationals)."""
    from fractions import Fraction as F

    s0, i0, j0 = terms[0]
    t = (s0 * a[i0]) * b[j0]
    for s, i, j in terms[1:]:
        t = float(F(s * a[i]) * F(b[j]) + F(t))
    return t


def _plain_sum(terms, a, b):
    t = 0.0
    for k, (s, i, j) in enumerate(terms):
        p = (s * a[i]) * b[j]
        t = p if k == 0 else t + p
    return t


def test_products_round_as_the_fused_pattern(oracle_mod):
    """The checker's and the product header's products, det, 1/2 Re tr(A B)
    and the first component of A B^dag equal, bit for bit, an exact-rational
    model of the documented rounding (first product alone, then one fused
    multiply-add per term, in the reference's term order).  The unfused
    evaluation differs on these inputs, so a lost fusion would be caught."""
    lib = oracle_mod.load()
    shim = _shim.load()
    rng = np.random.default_rng(3)
    A, B = _rand(rng, 300), _rand(rng, 300)
    differs = 0
    for kind, (name, terms) in enumerate(_TERMS.items()):
        for i in range(len(A)):
            a, b = np.ascontiguousarray(A[i]), np.ascontiguousarray(B[i])
            want = np.array([_fused_sum(t, a, b) for t in terms])
            c = np.zeros(4)
            getattr(lib, name)(a.ctypes.data, b.ctypes.data, c.ctypes.data)
            assert np.array_equal(c.view(np.uint64), want.view(np.uint64)), (name, i)
            c2 = np.zeros(4)
            shim.shim_q_mul(a.ctypes.data, b.ctypes.data, c2.ctypes.data, kind)
            assert np.array_equal(c2.view(np.uint64), want.view(np.uint64)), (name, i)
            differs += any(_plain_sum(t, a, b) != w for t, w in zip(terms, want))
    assert differs > 100
    for i in range(100):
        a, b = np.ascontiguousarray(A[i]), np.ascontiguousarray(B[i])
        det = _fused_sum([(1, k, k) for k in range(4)], a, a)
        retr = _fused_sum(_TERMS["su2_product"][0], a, b)
        nd_re = _fused_sum(_TERMS["su2_product_tc"][0], a, b)
        assert lib.su2_det(a.ctypes.data) == det == shim.shim_q_scalar(a.ctypes.data, b.ctypes.data, 0)
        assert lib.su2_prod_trace(a.ctypes.data, b.ctypes.data) == retr == shim.shim_q_scalar(
            a.ctypes.data, b.ctypes.data, 1)
        assert nd_re == shim.shim_q_scalar(a.ctypes.data, b.ctypes.data, 2)


def _round_f32(x):
    """A rational rounded to the nearest binary32 value (ties to even),
    exactly; the values here are normal."""
    from fractions import Fraction as F

    if x == 0:
        return np.float32(0.0)
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if F(2) ** e > a:
        e -= 1
    scale = F(2) ** (23 - e)  # 24 significant bits
    m = a * scale
    q, r = divmod(m.numerator, m.denominator)
    if 2 * r > m.denominator or (2 * r == m.denominator and q % 2 == 1):
        q += 1
    v = F(q) / scale
    return np.float32(float(v if x > 0 else -v))


def _fused_sum_f32(terms, a, b):
    """The fused pattern in binary32: first product rounded alone, each later
    term added with one rounding, modelled with rationals."""
    from fractions import Fraction as F

    s0, i0, j0 = terms[0]
    t = _round_f32(F(float(s0 * a[i0])) * F(float(b[j0])))
    for s, i, j in terms[1:]:
        t = _round_f32(F(float(s * a[i])) * F(float(b[j])) + F(float(t)))
    return t


def test_float_paths_round_as_the_fused_pattern(oracle_mod):
    """The FP32-link staple paths: the checker's float products and the
    product header's (shim) equal, bit for bit, a rational model of the fused
    pattern rounded to binary32, and the 2x2 matrix products to float
    accuracy.  A dropped fusion or a term in double is caught: the unfused
    binary32 evaluation and the double evaluation both differ on these
    inputs."""
    lib = oracle_mod.load()
    shim = _shim.load()
    rng = np.random.default_rng(5)
    A = rng.normal(size=(200, 4)).astype(np.float32)
    B = rng.normal(size=(200, 4)).astype(np.float32)
    names = {"su2_productf": ("su2_product", 0), "su2_product_tcf": ("su2_product_tc", 1),
             "su2_product_tc2f": ("su2_product_tc2", 3)}
    unfused = in_double = 0
    for name, (dname, kind) in names.items():
        terms = _TERMS[dname]
        for i in range(len(A)):
            a, b = np.ascontiguousarray(A[i]), np.ascontiguousarray(B[i])
            want = np.array([_fused_sum_f32(t, a, b) for t in terms], dtype=np.float32)
            c = np.zeros(4, dtype=np.float32)
            getattr(lib, name)(a.ctypes.data, b.ctypes.data, c.ctypes.data)
            assert np.array_equal(c.view(np.uint32), want.view(np.uint32)), (name, i)
            c2 = np.zeros(4, dtype=np.float32)
            shim.shim_qf_mul(a.ctypes.data, b.ctypes.data, c2.ctypes.data, kind)
            assert np.array_equal(c2.view(np.uint32), want.view(np.uint32)), (name, i)
            ad, bd = a.astype(np.float64), b.astype(np.float64)
            wd = np.array([_fused_sum(t, ad, bd) for t in terms])
            M = {0: N.quat_to_mat(ad) @ N.quat_to_mat(bd), 1: N.quat_to_mat(ad) @ N.dag(N.quat_to_mat(bd)),
                 3: N.dag(N.quat_to_mat(ad)) @ N.dag(N.quat_to_mat(bd))}[kind]
            np.testing.assert_allclose(N.quat_to_mat(c.astype(np.float64)), M, rtol=0, atol=2e-6)
            in_double += any(np.float64(w) != x for w, x in zip(want, wd))
            for t, w in zip(terms, want):
                s0, i0, j0 = t[0]
                p = np.float32(s0 * a[i0]) * b[j0]
                for s, i2, j in t[1:]:
                    p = np.float32(p + np.float32(np.float32(s * a[i2]) * b[j]))
                unfused += p != w
    assert unfused > 50 and in_double > 300
