"""Single-link update bodies.

- The product's bodies (host build of the header the kernels use) are bit
  for bit the checker's restatement of functions/mc_updates.h on random
  staples, links and RNG states.
- The checker's heat-bath samples the Kennedy-Pendleton target
  sqrt(1 - x^2) exp(beta a x) (mean I2(beta a)/I1(beta a), KS test) with an
  isotropic direction.
- The Metropolis decision equals exp(-s) > u evaluated exactly (mpmath),
  including inside the +-1e-4 band of the float estimate, for every u the
  generator can produce (u >= 1/m1).
- Over-relaxation keeps 1/2 Re tr(U A), is an involution and stays in SU(2).
"""
import ctypes as C

import mpmath
import numpy as np
import pytest
from scipy import special, stats

from tests import _naive as N
from tests import _shim


def _state(oracle_mod, seed, sub):
    r = oracle_mod.Rng(seed, sub)
    st = np.array(r.state, dtype=np.uint32)
    return r, st


def _rand_staple(rng, scale=3.0):
    return rng.normal(size=4) * scale


def test_heat_bath_product_equals_checker(oracle_mod):
    lib, shim = oracle_mod.load(), _shim.load()
    rng = np.random.default_rng(3)
    for k in range(400):
        A = _rand_staple(rng, scale=rng.uniform(0.05, 4.0))
        beta = float(np.float32(rng.uniform(0.3, 12.0)))
        r, st = _state(oracle_mod, 1000 + k, k)
        a1, out1 = A.copy(), np.zeros(4)
        lib.orc_heat_bath(a1.ctypes.data, out1.ctypes.data, beta, C.byref(r.st))
        out2 = np.zeros(4)
        shim.shim_heat_bath(A.ctypes.data, beta, st.ctypes.data, out2.ctypes.data)
        assert np.array_equal(out1.view(np.uint64), out2.view(np.uint64)), k
        assert list(map(int, st)) == r.state, k


def test_heat_bath_with_zero_or_nan_staple_ends(oracle_mod):
    """A zero staple (a field that was never initialised) or a NaN one makes
    the Kennedy-Pendleton scale infinite or NaN: the loop must end after one
    trial (4 + 2 draws) in the product and the checker alike, not hang."""
    lib, shim = oracle_mod.load(), _shim.load()
    for A in (np.zeros(4), np.full(4, np.nan), np.array([1e-300, 0.0, 0.0, 0.0])):
        for beta in (1.0, 3.0):
            r, st = _state(oracle_mod, 5, 9)
            ref = oracle_mod.Rng(5, 9)
            for _ in range(6):
                ref.next()
            a1, out1 = A.copy(), np.zeros(4)
            lib.orc_heat_bath(a1.ctypes.data, out1.ctypes.data, beta, C.byref(r.st))
            out2 = np.zeros(4)
            shim.shim_heat_bath(A.ctypes.data, beta, st.ctypes.data, out2.ctypes.data)
            assert r.state == ref.state and list(map(int, st)) == ref.state
            assert np.isnan(out1).all() and np.isnan(out2).all()


def test_metropolis_product_equals_checker(oracle_mod):
    lib, shim = oracle_mod.load(), _shim.load()
    rng = np.random.default_rng(4)
    for k in range(400):
        A = _rand_staple(rng)
        U = rng.normal(size=4)
        U /= np.linalg.norm(U)
        beta = float(np.float32(rng.uniform(0.5, 8.0)))
        eps = float(np.float32(rng.uniform(0.05, 0.9)))
        hits = int(rng.integers(1, 6))
        r, st = _state(oracle_mod, 77, k)
        u1 = U.copy()
        acc1 = 0
        for _ in range(hits):
            acc1 += lib.orc_metropolis(A.ctypes.data, u1.ctypes.data, beta, eps, C.byref(r.st))
        u2 = U.copy()
        acc2 = shim.shim_metropolis(A.ctypes.data, u2.ctypes.data, beta, eps, hits, st.ctypes.data)
        assert acc1 == acc2
        assert np.array_equal(u1.view(np.uint64), u2.view(np.uint64))
        assert list(map(int, st)) == r.state


def test_over_relaxation_product_equals_checker(oracle_mod):
    lib, shim = oracle_mod.load(), _shim.load()
    rng = np.random.default_rng(5)
    for k in range(300):
        A = _rand_staple(rng)
        U = rng.normal(size=4)
        U /= np.linalg.norm(U)
        a1, u1 = A.copy(), U.copy()
        lib.orc_over_relaxation(a1.ctypes.data, u1.ctypes.data)
        out = np.zeros(4)
        shim.shim_over_relax(A.ctypes.data, U.ctypes.data, out.ctypes.data)
        assert np.array_equal(u1.view(np.uint64), out.view(np.uint64))


@pytest.mark.parametrize("ba", [(0.01, 2.0), (1.0, 1.5), (3.0, 2.0), (6.0, 3.5)])
def test_kennedy_pendleton_distribution(oracle_mod, ba):
    """beta = 0.01 is the smallest heat-bath coupling the library accepts;
    with beta |A| = 0.02 each update takes some 300 trials and still samples
    the target."""
    beta, a = ba
    beta = float(np.float32(beta))
    lib = oracle_mod.load()
    rng = np.random.default_rng(6)
    Ahat = rng.normal(size=4)
    Ahat /= np.linalg.norm(Ahat)
    r = oracle_mod.Rng(2024, 1)
    n = 40000
    x0 = np.empty(n)
    dirs = np.empty((n, 3))
    MA = N.quat_to_mat(Ahat)
    for i in range(n):
        A = Ahat * a
        out = np.zeros(4)
        lib.orc_heat_bath(A.ctypes.data, out.ctypes.data, beta, C.byref(r.st))
        X = N.mat_to_quat(N.quat_to_mat(out) @ MA)  # X = U A^
        x0[i] = X[0]
        dirs[i] = X[1:] / np.linalg.norm(X[1:])
    ba_ = beta * a
    mean_exact = special.iv(2, ba_) / special.iv(1, ba_)
    assert abs(x0.mean() - mean_exact) < 5 * x0.std() / np.sqrt(n)
    # CDF of sqrt(1 - x^2) exp(ba x) on [-1, 1] by quadrature
    grid = np.linspace(-1, 1, 20001)
    dens = np.sqrt(1 - grid ** 2) * np.exp(ba_ * (grid - 1))
    cdf = np.concatenate([[0], np.cumsum((dens[1:] + dens[:-1]) / 2 * np.diff(grid))])
    cdf /= cdf[-1]
    ks = stats.kstest(x0, lambda x: np.interp(x, grid, cdf))
    assert ks.pvalue > 1e-3
    assert np.all(np.abs(dirs.mean(axis=0)) < 5 / np.sqrt(3 * n))
    assert np.all(np.abs((dirs ** 2).mean(axis=0) - 1 / 3) < 0.02)


def test_metropolis_decision_exact(oracle_mod):
    lib, shim = oracle_mod.load(), _shim.load()
    mpmath.mp.dps = 50
    rng = np.random.default_rng(8)
    cases = []
    for s in np.concatenate([rng.uniform(0, 30, 400), rng.uniform(0, 1e-3, 50), [0.0, 87.0, 88.5, 100.0]]):
        e = float(mpmath.exp(-mpmath.mpf(float(s))))
        for f in (1 - 2e-4, 1 - 1e-4, 1 - 5e-5, 1 - 1e-12, 1.0, 1 + 1e-12, 1 + 5e-5, 1 + 1e-4, 1 + 2e-4):
            u = e * f
            if 0 < u <= 1:
                cases.append((float(s), u))
            for du in (-1, 1):
                v = np.nextafter(e, e + du)
                if 0 < v <= 1:
                    cases.append((float(s), float(v)))
        cases.append((float(s), float(rng.uniform(2.3e-10, 1.0))))
    umin = 1.0 / 4294967087.0  # smallest draw of the generator (mc_updates.h:5-6 relies on it)
    for s, u in cases:
        if u < umin:
            continue
        want = mpmath.exp(-mpmath.mpf(s)) > mpmath.mpf(u)
        # exact up to the 2-ulp accuracy of the double exp
        e = float(mpmath.exp(-mpmath.mpf(s)))
        if abs(u - e) <= 2 * np.spacing(e):
            continue
        assert bool(lib.orc_metropolis_accept(s, u)) == bool(want), (s, u)
        assert bool(shim.shim_metropolis_accept(s, u)) == bool(want), (s, u)


def test_over_relaxation_invariants(oracle_mod):
    lib = oracle_mod.load()
    rng = np.random.default_rng(9)
    for _ in range(200):
        A = rng.normal(size=4) * 2
        U = rng.normal(size=4)
        U /= np.linalg.norm(U)
        before = N.half_retr(N.quat_to_mat(U) @ N.quat_to_mat(A))
        u1, a1 = U.copy(), A.copy()
        lib.orc_over_relaxation(a1.ctypes.data, u1.ctypes.data)
        after = N.half_retr(N.quat_to_mat(u1) @ N.quat_to_mat(A))
        assert abs(after - before) < 1e-13 * max(1, abs(before))
        assert abs(np.dot(u1, u1) - 1) < 1e-14
        u2, a2 = u1.copy(), A.copy()
        lib.orc_over_relaxation(a2.ctypes.data, u2.ctypes.data)
        np.testing.assert_allclose(u2, U, atol=1e-14)


def test_hot_start_links(oracle_mod):
    """su2_rand: four (u - 0.5) draws normalised (functions/su2.h:29-34)."""
    shim = _shim.load()
    r, st = _state(oracle_mod, 1234, 5)
    raw = np.array([r.uniform_double() - 0.5 for _ in range(4)])
    out = np.zeros(4)
    shim.shim_rand_link(st.ctypes.data, out.ctypes.data)
    np.testing.assert_allclose(out, raw / np.linalg.norm(raw), rtol=0, atol=1e-15)


def test_haar_start_is_uniform_on_s3(oracle_mod):
    shim = _shim.load()
    st = np.array(oracle_mod.Rng(9, 0).state, dtype=np.uint32)
    xs = np.empty((20000, 4))
    for i in range(len(xs)):
        shim.shim_haar_link(st.ctypes.data, xs[i].ctypes.data)
    assert np.allclose(np.linalg.norm(xs, axis=1), 1, atol=1e-14)
    # a0 of a Haar SU(2) element has density (2/pi) sqrt(1 - a0^2)
    ks = stats.kstest(xs[:, 0], lambda x: (x * np.sqrt(1 - x * x) + np.arcsin(x)) / np.pi + 0.5)
    assert ks.pvalue > 1e-3
    assert np.all(np.abs(xs.mean(axis=0)) < 0.02)
