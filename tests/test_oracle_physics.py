"""The checker's lattice operations against physics and an independent
full-matrix implementation (tests/_naive.py).

- cold start: E = 0, P = 1 exactly, every boundary condition;
- en_av equals the full-matrix observables and is gauge invariant;
- staples: an over-relaxation launch leaves the full Wilson action
  unchanged for every (direction, parity) -- a wrong neighbour, sign or
  orientation in any staple term breaks this; with PMC the x/t launches
  conserve the PMC action (boundary y-plaquettes absent);
- PCC: the x and t links at y = 0 stay the identity under every update;
- equilibrium: strong coupling <P> = u + 4u^5 (u = I2/I1, d = 3), weak
  coupling <P> = 1 - 1/beta + O(1/beta^2), and Metropolis agrees with
  heat-bath within errors;
- FP32-link mode: the float staple paths keep the action to float accuracy;
- results do not depend on the number of threads.
"""
import os

import numpy as np
import pytest
from scipy import special

from tests import _naive as N


def lattice(oracle_mod, Lt, Lx, Ly, beta, start=0, bc=0, seed=1, eps=0.3, hits=2):
    o = oracle_mod.Oracle(Lt, Lx, Ly, beta, eps, hits)
    o.setup_kernel(seed)
    o.start(start, bc)
    return o


@pytest.mark.parametrize("bc", [0, 1, 2])
def test_cold_start_exact(oracle_mod, bc):
    o = lattice(oracle_mod, 4, 6, 8, 3.0, start=1, bc=bc)
    assert o.en_av(bc == 2) == (0.0, 1.0)


@pytest.mark.parametrize("pmc", [False, True])
def test_en_av_matches_full_matrix_and_is_gauge_invariant(oracle_mod, pmc):
    o = lattice(oracle_mod, 6, 4, 10, 2.0, start=0, bc=2 if pmc else 0)
    for _ in range(3):
        o.mc_up(2 if pmc else 0, 1)
    e, p = o.en_av(pmc)
    U = N.field_to_links(o.phi, 6, 4, 10)
    en, pn = N.observables(U, pmc)
    assert abs(e - en) < 1e-13 and abs(p - pn) < 1e-13
    g = N.random_su2((6, 4, 10), np.random.default_rng(0))
    o.phi[:] = N.links_to_field(N.gauge_transform(U, g), 6, 4, 10)
    e2, p2 = o.en_av(pmc)
    assert abs(e2 - e) < 1e-13 and abs(p2 - p) < 1e-13


def test_layout_round_trip(oracle_mod):
    o = lattice(oracle_mod, 4, 6, 2, 2.0)
    U = N.field_to_links(o.phi, 4, 6, 2)
    assert np.array_equal(N.links_to_field(U, 4, 6, 2), o.phi)


@pytest.mark.parametrize("shape", [(4, 4, 4), (2, 2, 2), (6, 4, 10), (4, 8, 2)])
@pytest.mark.parametrize("bc", [0, 1, 2])
def test_overrelaxation_conserves_action(oracle_mod, shape, bc):
    Lt, Lx, Ly = shape
    o = lattice(oracle_mod, Lt, Lx, Ly, 2.0, start=0, bc=0, seed=3)
    for mu in range(3):
        for par in range(2):
            # the PMC action is the one the x/t updates see; y links use the
            # periodic staple in every boundary condition
            pmc_act = bc == 2 and mu != 0
            before = N.action(N.field_to_links(o.phi, Lt, Lx, Ly), 1.0, pmc_act)
            old = o.phi.copy()
            o.mc_update(bc, 1, 1, mu, par)
            after = N.action(N.field_to_links(o.phi, Lt, Lx, Ly), 1.0, pmc_act)
            assert not np.array_equal(old, o.phi)
            assert abs(after - before) < 1e-11 * max(1.0, abs(before)), (mu, par)


def test_heatbath_changes_only_the_launched_links(oracle_mod):
    o = lattice(oracle_mod, 4, 4, 6, 3.0, start=0, seed=4)
    for mu in range(3):
        for par in range(2):
            old = o.phi.reshape(2, -1, 3, 4).copy()
            o.mc_update(0, 1, 0, mu, par)
            new = o.phi.reshape(2, -1, 3, 4)
            changed = np.any(old != new, axis=-1)
            assert changed[par, :, mu].all()
            mask = np.ones_like(changed)
            mask[par, :, mu] = False
            assert not changed[mask].any()


@pytest.mark.parametrize("algo", [0, 1])
def test_pcc_boundary_links_stay_identity(oracle_mod, algo):
    Lt, Lx, Ly = 4, 6, 8
    o = lattice(oracle_mod, Lt, Lx, Ly, 3.0, start=0, bc=1, seed=5)
    for _ in range(3):
        o.mc_up(1, algo, 0)
        o.mc_up(1, algo, 1)
        U = N.field_to_links(o.phi, Lt, Lx, Ly)
        ident = np.broadcast_to(np.eye(2), U[1][:, :, 0].shape)
        assert np.array_equal(U[1][:, :, 0], ident) and np.array_equal(U[2][:, :, 0], ident)
        assert not np.allclose(U[0][:, :, 0], np.eye(2))


def _equilibrium(oracle_mod, beta, algo, L=8, therm=60, meas=200, seed=11, eps=0.5, hits=4):
    o = oracle_mod.Oracle(L, L, L, beta, eps, hits)
    o.setup_kernel(seed)
    o.init(0)
    for _ in range(therm):
        o.mc_up(0, algo, 0)
    ps = []
    for _ in range(meas):
        o.mc_up(0, algo, 0)
        ps.append(o.en_av()[1])
    ps = np.array(ps)
    nb = 20
    blocks = ps[: len(ps) // nb * nb].reshape(nb, -1).mean(axis=1)
    return ps.mean(), blocks.std(ddof=1) / np.sqrt(nb)


def test_strong_coupling(oracle_mod):
    beta = 0.5
    p, err = _equilibrium(oracle_mod, beta, 1)
    u = special.iv(2, beta) / special.iv(1, beta)
    assert abs(p - (u + 4 * u ** 5)) < 4 * err + 2e-4


def test_weak_coupling(oracle_mod):
    beta = 20.0
    p, err = _equilibrium(oracle_mod, beta, 1, therm=100, meas=100)
    assert abs((1 - p) * beta - 1.0) < 0.1


def test_metropolis_agrees_with_heatbath(oracle_mod):
    beta = 2.5
    p_hb, e_hb = _equilibrium(oracle_mod, beta, 1, meas=200)
    p_m, e_m = _equilibrium(oracle_mod, beta, 0, therm=150, meas=300, eps=0.6, hits=4)
    assert abs(p_hb - p_m) < 4 * np.hypot(e_hb, e_m) + 1e-3


def test_thread_count_independence(oracle_mod):
    res = []
    for threads in (1, 4):
        oracle_mod.set_threads(threads)
        o = lattice(oracle_mod, 8, 6, 10, 2.2, start=0, bc=2, seed=8)
        for _ in range(3):
            o.mc_up(2, 0, 0)
            o.mc_up(2, 1, 0)
        res.append((o.phi.copy(), o.state.copy(), o.accept.copy()))
    oracle_mod.set_threads(os.cpu_count() or 1)
    assert np.array_equal(res[0][0].view(np.uint64), res[1][0].view(np.uint64))
    assert np.array_equal(res[0][1], res[1][1]) and np.array_equal(res[0][2], res[1][2])


@pytest.mark.parametrize("bc", [0, 2])
def test_fp32_link_staples_multiply_in_float(oracle_mod, bc):
    """FP32-link mode multiplies the staple paths in float: an over-relaxation
    launch on a float-valued field keeps the action to float accuracy (a
    wrong neighbour, sign or orientation in a float path moves it by O(1)),
    and its links differ from the FP64-mode launch on the same field, rounded
    to float, by float rounding alone, but do differ (the float paths are
    taken)."""
    Lt, Lx, Ly = 4, 6, 8
    o32 = oracle_mod.Oracle(Lt, Lx, Ly, 2.0, 0.3, 2, f32=True)
    o32.setup_kernel(3)
    o32.start(0, bc)
    o64 = oracle_mod.Oracle(Lt, Lx, Ly, 2.0, 0.3, 2)
    assert np.array_equal(o32.phi, o32.phi.astype(np.float32).astype(np.float64))
    differ = 0
    for mu in range(3):
        for par in range(2):
            pmc_act = bc == 2 and mu != 0
            before = N.action(N.field_to_links(o32.phi, Lt, Lx, Ly), 1.0, pmc_act)
            o64.phi[:] = o32.phi
            o32.mc_update(bc, 1, 1, mu, par)
            o64.mc_update(bc, 1, 1, mu, par)
            after = N.action(N.field_to_links(o32.phi, Lt, Lx, Ly), 1.0, pmc_act)
            assert abs(after - before) < 2e-5 * abs(before), (mu, par)
            # with double paths o32 would equal o64 rounded to float
            d = np.abs(o32.phi - o64.phi.astype(np.float32).astype(np.float64)).max()
            assert d < 1e-5, (mu, par)
            differ += d > 0
    assert differ == 6
