"""GPU parity: the HIP path of libsu2mc against the CPU checker, through the
C ABI, on the same seeds.

Bar (the RNG-stream and accept-decision contract): after every compared
sweep the links, the RNG states and the acceptance counters are BIT-EXACT
equal to the checker's; the fused energy/plaquette reduction agrees to
1e-12 relative (its summation order differs from the checker's block
order).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)
    b = np.ascontiguousarray(b, dtype=np.float64).view(np.uint64)
    return np.array_equal(a, b)


def _first_diff(a, b):
    d = np.nonzero(np.ascontiguousarray(a).view(np.uint64) != np.ascontiguousarray(b).view(np.uint64))[0]
    return (int(d.size), int(d[0]) if d.size else -1)


def _compare(g, o, where):
    phi = g.field_ref()
    assert _bits_equal(phi, o.phi), f"{where}: links differ {_first_diff(phi, o.phi)}"
    st = g.rng_state()
    assert np.array_equal(st, o.rng_planes()), f"{where}: RNG states differ"
    acc = g.accept()
    assert np.array_equal(acc, o.accept), f"{where}: acceptance counters differ"


def _en_av_close(g, o, pmc, where):
    e, p = g.en_av(pmc)
    eo, po = o.en_av(pmc)
    assert abs(e - eo) <= REL_TOL * max(abs(eo), 1e-300) + 1e-300, f"{where}: energy {e!r} vs {eo!r}"
    assert abs(p - po) <= REL_TOL * abs(po), f"{where}: plaquette {p!r} vs {po!r}"
    return e, p


def run_pair(q, O, name, Lt, Lx, Ly, beta, bc, algo, start, seed, nsweeps, prec=0, eps=0.1, hits=1,
             updates=(0,), compare_every=1, normalize_at=()):
    _say(f"[parity] {name}: {Lt}x{Lx}x{Ly} bc={bc} algo={algo} start={start} prec={prec} seed={seed}")
    g = q.Lattice(Lt, Lx, Ly, beta, eps, hits, precision=prec)
    o = O.Oracle(Lt, Lx, Ly, beta, eps, hits, f32=(prec == 1))
    try:
        g.setup_kernel(seed)
        o.setup_kernel(seed)
        assert np.array_equal(g.rng_state(), o.rng_planes()), f"{name}: setup_kernel states differ"
        g.start(start, bc)
        o.start(start, bc)
        _compare(g, o, f"{name} start")
        pmc = bc == 2
        _en_av_close(g, o, pmc, f"{name} start")
        for i in range(1, nsweeps + 1):
            for upd in updates:
                g.mc_up(bc, algo, upd)
                o.mc_up(bc, algo, upd)
            if i in normalize_at:
                g.normalize()
                o.normalize()
            if i % compare_every == 0 or i == nsweeps:
                _compare(g, o, f"{name} sweep {i}")
                _en_av_close(g, o, pmc, f"{name} sweep {i}")
        return g, o
    finally:
        g.close()


def test_c1_periodic_heatbath_cold(gpu, oracle_mod):
    """16^3 periodic, heat-bath, beta 3.0, cold start, seed 1234, 10 sweeps."""
    run_pair(gpu, oracle_mod, "C1", 16, 16, 16, 3.0, bc=0, algo=1, start=1, seed=1234, nsweeps=10)


@pytest.mark.parametrize("bc", [1, 2])
@pytest.mark.parametrize("prec", [0, 1])
def test_c3_metropolis_pcc_pmc(gpu, oracle_mod, bc, prec):
    """32x32x48 PCC / PMC, Metropolis eps 0.5, 4 hits, beta 4.0, hot start."""
    run_pair(gpu, oracle_mod, "C3", 32, 32, 48, 4.0, bc=bc, algo=0, start=0, seed=4321, nsweeps=10, prec=prec,
             eps=0.5, hits=4, compare_every=5)


def test_c2_shape_heatbath_overrelax_hot(gpu, oracle_mod):
    """Heat-bath + one over-relaxation sweep, beta 5.0, hot start, seed 7
    (the 128^3 workload's update sequence at 16x24x32)."""
    run_pair(gpu, oracle_mod, "C2-small", 16, 24, 32, 5.0, bc=0, algo=1, start=0, seed=7, nsweeps=10,
             updates=(0, 1), compare_every=5)


def test_headline_workload_at_full_size(gpu, oracle_mod):
    """The measured workload itself: 128^3, periodic, beta 5, Haar-random hot
    start, seed 7, heat-bath + over-relaxation steps with a reunitarisation,
    every link, stream and counter of the 2 097 152-site lattice compared
    after each step."""
    run_pair(gpu, oracle_mod, "128^3", 128, 128, 128, 5.0, bc=0, algo=1, start=2, seed=7, nsweeps=3,
             updates=(0, 1), normalize_at=(2,))


def test_256cube_at_full_size(gpu, oracle_mod):
    """256^3 (1.6 GB of links, the DRAM-bound lattice): one heat-bath +
    over-relaxation step, all 16.8 M sites compared."""
    run_pair(gpu, oracle_mod, "256^3", 256, 256, 256, 5.0, bc=0, algo=1, start=0, seed=7, nsweeps=1,
             updates=(0, 1))


@pytest.mark.parametrize("bc", [0, 1, 2])
@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("prec", [0, 1])
def test_all_update_kinds(gpu, oracle_mod, bc, algo, prec):
    """Every boundary x algorithm x storage precision, with an over-relaxation
    sweep after each update sweep (over-relaxation for PCC/PMC is an
    extension) and a reunitarisation at sweep 3."""
    run_pair(gpu, oracle_mod, f"kinds bc{bc} a{algo} p{prec}", 8, 10, 12, 2.5, bc=bc, algo=algo, start=0,
             seed=99 + bc, nsweeps=6, prec=prec, eps=0.3, hits=3, updates=(0, 1), compare_every=3,
             normalize_at=(3,))


@pytest.mark.parametrize("shape", [(2, 2, 2), (4, 2, 6), (10, 6, 2), (6, 14, 10), (2, 8, 34)])
def test_ragged_and_minimal_shapes(gpu, oracle_mod, shape):
    """Extents down to 2 and half-rows of odd length (Ly/2 = 1, 3, 5, 17)."""
    Lt, Lx, Ly = shape
    run_pair(gpu, oracle_mod, f"shape {shape}", Lt, Lx, Ly, 3.5, bc=0, algo=1, start=0, seed=5, nsweeps=4)
    run_pair(gpu, oracle_mod, f"shape {shape} pmc", Lt, Lx, Ly, 3.5, bc=2, algo=0, start=0, seed=6, nsweeps=3,
             eps=0.4, hits=2)


@pytest.mark.parametrize("shape,bc", [((8, 64, 64), 1), ((32, 64, 64), 0), ((16, 64, 128), 2)])
def test_thin_lattices_column_block_order(gpu, oracle_mod, shape, bc):
    """Lattices of at most 32 t-planes whose planes are whole multiples of 8
    blocks take the x-column block order (each XCD a column of x-rows through
    all t); FP64 links, link updates and over-relaxation."""
    run_pair(gpu, oracle_mod, "columns", *shape, 4.0, bc, 1, 0, 31, 2, updates=(0, 1))


@pytest.mark.parametrize("shape,bc", [((64, 64, 64), 0), ((48, 128, 64), 2), ((40, 64, 96), 1),
                                      ((40, 256, 128), 0)])
def test_strip_and_slab_block_orders(gpu, oracle_mod, shape, bc):
    """FP64 lattices of more than 32 t-planes: planes of 8 blocks walk
    8-block x-strips (64^3), planes of 16 blocks 16-block strips
    (48x128x64, like 128^3's 32), planes of 12 blocks the slab order
    (40x64x96), planes of 64 blocks 32-block strips (40x256x128, like
    256^3's 128); link updates and over-relaxation."""
    run_pair(gpu, oracle_mod, "strips", *shape, 4.5, bc, 1, 0, 37, 2, updates=(0, 1))


@pytest.mark.parametrize("bc", [1, 2])
def test_beyond_the_infinity_cache_pcc_pmc(gpu, oracle_mod, bc):
    """128x128x160 FP64 (283 MB of links and RNG state, more than the 256 MB
    Infinity Cache): the heat-bath launches stream their RNG state with
    nontemporal hints; PCC and PMC, one heat-bath + over-relaxation step."""
    run_pair(gpu, oracle_mod, f"stream-rng bc {bc}", 128, 128, 160, 4.5, bc=bc, algo=1, start=0, seed=43,
             nsweeps=1, updates=(0, 1))


@pytest.mark.parametrize("start", [1, 2])
def test_cold_and_haar_starts(gpu, oracle_mod, start):
    for bc in (0, 1, 2):
        run_pair(gpu, oracle_mod, f"start {start} bc {bc}", 8, 8, 8, 6.0, bc=bc, algo=1, start=start, seed=11,
                 nsweeps=2)


def test_large_seed_and_many_streams(gpu, oracle_mod):
    """2^17 streams with a seed whose scramble exercises the 64-bit wrap."""
    q, O = gpu, oracle_mod
    _say("[parity] setup_kernel 64^3 seed 2^63+17")
    g = q.Lattice(64, 64, 64, 4.0)
    o = O.Oracle(64, 64, 64, 4.0)
    try:
        seed = (1 << 63) + 17
        g.setup_kernel(seed)
        o.setup_kernel(seed)
        assert np.array_equal(g.rng_state(), o.rng_planes())
        g.rand_init2()
        o.init(0)
        assert _bits_equal(g.field_ref(), o.phi)
        g.mc_up_per(1, 0)
        o.mc_up(0, 1, 0)
        _compare(g, o, "64^3 sweep 1")
        _en_av_close(g, o, False, "64^3 sweep 1")
    finally:
        g.close()


def test_single_launch_entry(gpu, oracle_mod):
    """su2_mc_update runs exactly one (mu, parity) launch."""
    q, O = gpu, oracle_mod
    _say("[parity] single launches")
    g = q.Lattice(8, 8, 12, 3.0, 0.2, 2)
    o = O.Oracle(8, 8, 12, 3.0, 0.2, 2)
    try:
        g.setup_kernel(3)
        o.setup_kernel(3)
        g.rand_init2()
        o.init(0)
        for mu in (2, 0, 1):
            for par in (1, 0):
                g.mc_update(2, 0, 0, mu, par)
                o.mc_update(2, 0, 0, mu, par)
                _compare(g, o, f"launch mu={mu} parity={par}")
    finally:
        g.close()


def _random_cases(n, seed=20261203):
    rng = np.random.default_rng(seed)
    cases = []
    for i in range(n):
        Lt, Lx, Ly = (int(2 * rng.integers(1, 13)) for _ in range(3))
        algo = int(rng.integers(0, 2))
        beta = float(np.float32(rng.uniform(0.3, 8.0) if algo == 1 else rng.uniform(0.1, 8.0)))
        cases.append(dict(i=i, shape=(Lt, Lx, Ly), bc=int(rng.integers(0, 3)), algo=algo, beta=beta,
                          eps=float(np.float32(rng.uniform(0.0, 1.0))), hits=int(rng.integers(1, 5)),
                          prec=int(rng.integers(0, 2)), start=int(rng.integers(0, 3)),
                          seed=int(rng.integers(0, 2 ** 63)),
                          updates=((0,), (0, 1))[int(rng.integers(0, 2))],
                          block=int((64, 128, 192, 256)[int(rng.integers(0, 4))])))
    return cases


@pytest.mark.parametrize("case", _random_cases(24), ids=lambda c: f"case{c['i']}")
def test_randomised_cases(gpu, oracle_mod, case):
    """Seeded random draws over extents (even, 2..24), boundary, algorithm,
    couplings (beta, epsilon in [0, 1], 1-4 hits), start, link precision,
    over-relaxation and the sweep block size (64..256 threads): three sweeps
    each, bit-exact against the checker."""
    lib = gpu.load_library()
    c = case
    _say(f"[parity] random case {c}")
    assert lib.su2_set_sweep_block(c["block"]) == 0
    try:
        run_pair(gpu, oracle_mod, f"random {c['i']}", *c["shape"], c["beta"], bc=c["bc"], algo=c["algo"],
                 start=c["start"], seed=c["seed"], nsweeps=3, prec=c["prec"], eps=c["eps"], hits=c["hits"],
                 updates=c["updates"], normalize_at=(2,))
    finally:
        lib.su2_set_sweep_block(0)


def _res_slots(q, n):
    buf = q.DeviceBuffer(16 * n)
    buf.zero()
    return buf


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("shape,bc,algo,update", [
    ((16, 12, 20), 0, 1, 0),   # heat-bath, padded last block (V/2 = 1920)
    ((16, 12, 20), 1, 0, 0),   # PCC Metropolis
    ((16, 12, 20), 2, 1, 0),   # PMC heat-bath
    ((16, 12, 20), 2, 0, 1),   # PMC over-relaxation
    ((64, 64, 64), 0, 1, 0),   # FP64: strip block order
    ((8, 64, 64), 2, 1, 0),    # FP64: x-column block order
])
def test_measured_sweep_equals_en_av_then_sweep(gpu, oracle_mod, prec, shape, bc, algo, update):
    """su2_mc_up_measured (the measurement carried by the sweep's first
    launch) against su2_en_av followed by the same sweep: links, RNG states,
    acceptance counters and both measured values bit for bit, three passes in
    a row; the measurement also against the checker's en_av to 1e-12, and
    with a sweep block of 128 threads (the measurement then runs on its own)
    the same bits again."""
    q = gpu
    Lt, Lx, Ly = shape
    pmc = bc == 2
    _say(f"[measured] {shape} bc={bc} algo={algo} update={update} prec={prec}")
    gs = [q.Lattice(Lt, Lx, Ly, 3.0, 0.4, 2, precision=prec) for _ in range(3)]
    o = oracle_mod.Oracle(Lt, Lx, Ly, 3.0, 0.4, 2, f32=(prec == 1))
    slots = [_res_slots(q, 3) for _ in gs]
    lib = q.load_library()
    try:
        for g in gs:
            g.setup_kernel(77)
            g.start(0, bc)
        for r in range(3):
            ga, gb, gc = gs
            ga.en_av_into(pmc, slots[0].value + 16 * r, slots[0].value + 16 * r + 8)
            ga.mc_up(bc, algo, update)
            gb.mc_up_measured(bc, slots[1].value + 16 * r, slots[1].value + 16 * r + 8, algo, update)
            lib.su2_set_sweep_block(128)
            try:
                gc.mc_up_measured(bc, slots[2].value + 16 * r, slots[2].value + 16 * r + 8, algo, update)
            finally:
                lib.su2_set_sweep_block(0)
            pa = ga.field_ref()
            for g, name in ((gb, "measured"), (gc, "measured, block 128")):
                assert _bits_equal(g.field_ref(), pa), f"pass {r} {name}: links differ"
                assert np.array_equal(g.rng_state(), ga.rng_state()), f"pass {r} {name}: RNG differs"
                assert np.array_equal(g.accept(), ga.accept()), f"pass {r} {name}: acceptance differs"
        ra, rb, rc = (s.download(np.float64, 6).reshape(3, 2) for s in slots)
        assert _bits_equal(rb, ra) and _bits_equal(rc, ra), (ra, rb, rc)
        # the measurement of pass 0 is of the start field
        g0 = q.Lattice(Lt, Lx, Ly, 3.0, 0.4, 2, precision=prec)
        try:
            g0.setup_kernel(77)
            g0.start(0, bc)
            o.phi[:] = g0.field_ref()
        finally:
            g0.close()
        eo, po = o.en_av(pmc)
        assert abs(ra[0, 0] - eo) <= REL_TOL * abs(eo) and abs(ra[0, 1] - po) <= REL_TOL * abs(po), (ra[0], eo, po)
    finally:
        for g in gs:
            g.close()
