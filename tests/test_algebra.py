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
