"""GPU tests of the batched engine and the server: every job's results are
bit-identical to running that job alone through the single-lattice C ABI
(the reference's run loop, kernel.cu:161-208), whatever else is resident,
and whether jobs join, leave or the shapes are autotuned in between."""
import sys
import time

import numpy as np
import pytest

from _jobs import check_against_oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def run_alone(q, job, prec):
    """The reference schedule on one lattice with the single-lattice entry
    points; returns E, P series, final field (reference layout), RNG planes
    and mean acceptance."""
    g = q.Lattice(job.Lt, job.Lx, job.Ly, job.beta, job.epsilon, job.hits, precision=prec)
    try:
        g.setup_kernel(job.seed)
        g.start(job.start, job.bc)
        nev = job.normalize_every or 100

        def sweep():
            g.mc_up(job.bc, job.algo, 0)
            for _ in range(job.or_sweeps):
                g.mc_up(job.bc, job.algo, 1)

        for i in range(1, job.n_therm + 1):
            if i % nev == 0:
                g.normalize()
            sweep()
        g.normalize()
        g.accept_clear()
        E, P = [], []
        for i in range(1, job.n_meas + 1):
            for _ in range(job.n_cor):
                sweep()
            if i % nev == 0:
                g.normalize()
            e, p = g.en_av(job.bc == 2)
            E.append(e)
            P.append(p)
        acc = g.accept().astype(np.float64).sum()
        hits = job.hits if job.algo == 0 else 1
        n = job.n_meas * job.n_cor * 3 * (hits + job.or_sweeps)
        acc = acc / g.V / n
        if job.bc == 1:
            acc = acc / ((job.Ly - 1) / job.Ly + 1 / job.Ly / 3)
        return np.array(E), np.array(P), g.field_ref(), g.rng_state(), acc
    finally:
        g.close()


def mixed_jobs(q, n, seed0=1000, small=True):
    rng = np.random.default_rng(seed0)
    jobs = []
    for i in range(n):
        L = int(rng.choice([4, 6, 8, 10])) if small else int(rng.choice(range(8, 49, 2)))
        Ly = int(rng.choice([2, 4, 6, 8, 12])) if small else int(rng.choice(range(8, 65, 2)))
        jobs.append(q.Job(L, L, Ly, float(np.float32(rng.uniform(2, 8))), bc=int(rng.integers(0, 3)),
                          algo=int(rng.integers(0, 2)), start=int(rng.integers(0, 3)), seed=seed0 + i,
                          n_therm=int(rng.integers(1, 5)), n_meas=int(rng.integers(2, 6)),
                          n_cor=int(rng.integers(1, 3)), or_sweeps=int(rng.integers(0, 2)),
                          epsilon=float(np.float32(rng.uniform(0.1, 0.6))), hits=int(rng.integers(1, 4)),
                          normalize_every=int(rng.integers(2, 4))))
    return jobs


def check_same(q, eng, slot, job, prec, tag, oracle=None):
    """The job's engine results bit for bit against the same job run alone
    on the single-lattice entry points; with `oracle` (the checker module)
    also against the checker's run loop."""
    got = eng.read(slot, field=True, rng=True)
    E, P, phi, st, acc = run_alone(q, job, prec)
    assert np.array_equal(got["E"].view(np.uint64), E.view(np.uint64)), f"{tag}: energies differ"
    assert np.array_equal(got["P"].view(np.uint64), P.view(np.uint64)), f"{tag}: plaquettes differ"
    assert np.array_equal(got["field"].view(np.uint64), phi.view(np.uint64)), f"{tag}: links differ"
    assert np.array_equal(got["rng"], st), f"{tag}: RNG states differ"
    assert abs(got["accept"] - acc) <= 1e-12 * max(1.0, abs(acc)), f"{tag}: acceptance {got['accept']} vs {acc}"
    if oracle is not None:
        check_against_oracle(oracle, job, prec, got["E"], got["P"], got["field"], got["rng"], got["accept"],
                             tag=f"{tag} vs checker")


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_alone(gpu, oracle_mod, prec, use_graph):
    """Ten mixed jobs (every boundary, algorithm, start, over-relaxation
    count and reunitarisation cadence) in one engine: each job's series,
    final links, RNG states and acceptance equal, bit for bit, the same job
    run alone, and agree with the checker's run loop (kernel.cu:161-208):
    links and RNG states bit for bit, series to 1e-12."""
    q = gpu
    _say(f"[engine] batch vs alone prec={prec} graph={use_graph}")
    jobs = mixed_jobs(q, 10, seed0=2000 + 10 * prec)
    eng = q.Engine(precision=prec, max_jobs=32, budget_bytes=256 << 20, use_graph=use_graph)
    try:
        slots = [eng.add(j) for j in jobs]
        eng.run()
        for s, j in zip(slots, jobs):
            d, t = eng.status(s)
            assert d == t == j.n_therm + j.n_meas * j.n_cor
            check_same(q, eng, s, j, prec, f"slot {s} {j}", oracle=oracle_mod)
    finally:
        eng.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_without_carried_measurement(gpu, prec):
    """Sweep launches of 128-thread workgroups: no measurement rides on the
    next tick's first update launch (that needs 256), each runs as en_av at
    the end of its tick, and every job still gives the bits it gives alone."""
    q = gpu
    _say(f"[engine] block 128, measurements at the end of their ticks, prec={prec}")
    jobs = mixed_jobs(q, 6, seed0=2500 + 10 * prec)
    eng = q.Engine(precision=prec, max_jobs=8, budget_bytes=64 << 20, use_graph=True, block=128)
    try:
        slots = [eng.add(j) for j in jobs]
        eng.run()
        for s, j in zip(slots, jobs):
            check_same(q, eng, s, j, prec, f"block 128 slot {s} {j}")
    finally:
        eng.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_shape_changes_with_measurements_pending(gpu, prec):
    """The sweep launches' workgroup size changed between ticks, 256 -> 128
    -> 256 -> 64: measurements deferred to the next tick's first launch
    (possible only at 256) are taken by an en_av pass before that tick's
    sweeps, and none is taken twice; every job still gives the bits it gives
    alone."""
    q = gpu
    _say(f"[engine] shape changes mid-run, prec={prec}")
    jobs = mixed_jobs(q, 8, seed0=2700 + 10 * prec)
    for j in jobs:
        j.n_meas, j.n_cor = 6, 1  # a measurement after every sweep past thermalisation
    eng = q.Engine(precision=prec, max_jobs=8, budget_bytes=64 << 20, use_graph=True)
    try:
        slots = [eng.add(j) for j in jobs]
        for block in (128, 256, 64, 256):
            eng.tick(2)
            eng.set_shape(block)
        eng.run()
        for s, j in zip(slots, jobs):
            check_same(q, eng, s, j, prec, f"shape changes slot {s} {j}")
    finally:
        eng.close()


def test_jobs_join_and_leave(gpu):
    """Continuous batching on the engine: jobs added while others run, one
    removed unfinished, results of the rest unchanged."""
    q = gpu
    _say("[engine] dynamic membership")
    jobs = mixed_jobs(q, 8, seed0=3000)
    eng = q.Engine(precision=0, max_jobs=8, budget_bytes=64 << 20, use_graph=True)
    try:
        first = [eng.add(j) for j in jobs[:4]]
        eng.tick(2)
        victim = first[1]
        eng.remove(victim)
        later = [eng.add(j) for j in jobs[4:]]
        eng.tick(3)
        eng.run()
        for s, j in zip([first[0], first[2], first[3]] + later, [jobs[0], jobs[2], jobs[3]] + jobs[4:]):
            check_same(q, eng, s, j, 0, f"slot {s}")
    finally:
        eng.close()


def test_budget_admission(gpu):
    q = gpu
    _say("[engine] budget admission")
    job = q.Job(16, 16, 16, 3.0, n_therm=2, n_meas=2)
    need = q.load_library().su2_job_bytes(__import__("ctypes").byref(job.c()), 0)
    eng = q.Engine(precision=0, max_jobs=4, budget_bytes=int(need * 2.5), use_graph=False)
    try:
        a = eng.add(job)
        b = eng.add(q.Job(16, 16, 16, 3.0, seed=5, n_therm=2, n_meas=2))
        with pytest.raises(q.Su2Error) as e:
            eng.add(q.Job(16, 16, 16, 3.0, seed=6, n_therm=2, n_meas=2))
        assert e.value.code == -7
        eng.run()
        eng.remove(a)
        c = eng.add(q.Job(16, 16, 16, 3.0, seed=6, n_therm=2, n_meas=2))
        eng.run()
        check_same(q, eng, c, q.Job(16, 16, 16, 3.0, seed=6, n_therm=2, n_meas=2), 0, "readmitted")
        used, cap = eng.memory()
        assert used <= cap
    finally:
        eng.close()


def test_autotune_keeps_results(gpu):
    q = gpu
    _say("[engine] autotune")
    jobs = mixed_jobs(q, 6, seed0=4000)
    eng = q.Engine(precision=0, max_jobs=8, budget_bytes=64 << 20, use_graph=True)
    try:
        slots = [eng.add(j) for j in jobs]
        eng.tick(1)
        b, g = eng.autotune()
        assert b in (64, 128, 256) and g > 0
        eng.run()
        for s, j in zip(slots, jobs):
            check_same(q, eng, s, j, 0, f"slot {s} after autotune")
    finally:
        eng.close()


def test_server_continuous_batching(gpu):
    """Requests submitted while others run are admitted as memory frees up;
    each result equals the job run alone."""
    q = gpu
    _say("[server] continuous batching")
    jobs = mixed_jobs(q, 12, seed0=5000)
    need = max(q.load_library().su2_job_bytes(__import__("ctypes").byref(j.c()), 0) for j in jobs)
    srv = q.Server(precision=0, max_jobs=16, budget_bytes=need * 3, use_graph=True)
    try:
        ids = []
        for i, j in enumerate(jobs):
            ids.append(srv.submit(j))
            if i % 3 == 2:
                time.sleep(0.002)
        srv.drain()
        series = {i: srv.series(i) for i in ids}
        res = {r.id: r for r in srv.poll()}
        assert sorted(res) == sorted(ids)
        for i, j in zip(ids, jobs):
            r = res[i]
            assert r.status == 0
            assert r.t_submit <= r.t_admit <= r.t_done
            E, P, _, _, acc = run_alone(q, j, 0)
            assert np.array_equal(series[i][0].view(np.uint64), E.view(np.uint64)), f"job {i} energies"
            assert np.array_equal(series[i][1].view(np.uint64), P.view(np.uint64)), f"job {i} plaquettes"
            assert abs(r.E_mean - E.mean()) < 1e-12 and abs(r.P_mean - P.mean()) < 1e-12
            assert abs(r.accept - acc) < 1e-12
        # the mix really happened: some jobs were resident together, and some
        # waited for memory (three fit in the budget, twelve were submitted)
        iv = sorted((res[i].t_admit, res[i].t_done) for i in ids)
        assert any(nxt[0] < cur[1] for cur, nxt in zip(iv, iv[1:])), iv
        assert max(res[i].t_admit - res[i].t_submit for i in ids) > 0.0
    finally:
        srv.close()


def test_server_rejects_invalid_and_oversized(gpu):
    q = gpu
    _say("[server] invalid requests")
    srv = q.Server(precision=0, max_jobs=4, budget_bytes=8 << 20, use_graph=False)
    try:
        with pytest.raises(q.Su2Error):
            srv.submit(q.Job(15, 16, 16, 3.0))
        with pytest.raises(q.Su2Error):
            srv.submit(q.Job(64, 64, 64, 3.0))  # larger than the whole budget
        jid = srv.submit(q.Job(4, 4, 4, 3.0, n_therm=1, n_meas=2))
        srv.drain()
        (r,) = srv.poll()
        assert r.id == jid and r.status == 0
    finally:
        srv.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_ensemble_of_64_at_full_size(gpu, oracle_mod, prec):
    """The measured ensemble itself (64 lattices, 1.89 M sites, Lt = Lx over
    even 8..48, Ly over even 8..64, beta in [2, 8], every boundary, seeds
    1000 + i, 20 + 80 sweeps) in one engine with graphs and the autotuned
    shape: every job's 80 energies and plaquettes equal, bit for bit, those
    of the same job run alone through the C driver su2_run; four jobs (the
    smallest of each boundary, and the largest) agree with the checker's
    run loop: links and RNG states bit for bit, series to 1e-12."""
    import ctypes as C
    import os

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench

    q = gpu
    _say(f"[engine] 64-lattice ensemble prec={prec}")
    lib = q.load_library()
    jobs = bench.ensemble_jobs(q, 64, 20, 80)
    need = sum(lib.su2_job_bytes(C.byref(j.c()), prec) for j in jobs)
    eng = q.Engine(precision=prec, max_jobs=64, budget_bytes=int(need * 1.05) + (16 << 20), use_graph=True)
    try:
        slots = [eng.add(j) for j in jobs]
        eng.sync()
        eng.autotune()
        eng.run()
        for s, j in zip(slots, jobs):
            got = eng.read(s)
            r = q.run(Lt=j.Lt, Lx=j.Lx, Ly=j.Ly, beta=j.beta, boundary=j.bc, algo=j.algo, init=q.START_HOT,
                      N=j.n_meas, N_term=j.n_therm, seed=j.seed, precision=prec, save=1, sconf=1)
            assert np.array_equal(got["E"].view(np.uint64), r.energy.view(np.uint64)), f"job {s}: energies"
            assert np.array_equal(got["P"].view(np.uint64), r.plaquette.view(np.uint64)), f"job {s}: plaquettes"
        picks = {min((j.V, k) for k, j in enumerate(jobs) if j.bc == bc)[1] for bc in (0, 1, 2)}
        picks.add(max((j.V, k) for k, j in enumerate(jobs))[1])
        for k in sorted(picks):
            j = jobs[k]
            _say(f"[engine]   job {k} {j.Lt}x{j.Lx}x{j.Ly} bc={j.bc} against the checker")
            got = eng.read(slots[k], field=True, rng=True)
            check_against_oracle(oracle_mod, j, prec, got["E"], got["P"], got["field"], got["rng"], got["accept"],
                                 tag=f"ensemble job {k}")
    finally:
        eng.close()


def c5_requests(q, n, rng):
    """Run requests of the serving configuration: Lt = Lx in {16, 24, 32,
    48, 64}, Ly in {16, 32, 64}, beta uniform in [2, 8] (rounded to float as
    the reference stores it), boundary uniform over periodic / PCC / PMC,
    heat-bath, hot start, 50 thermalisation + 200 measured sweeps."""
    L = rng.choice([16, 24, 32, 48, 64], n)
    Ly = rng.choice([16, 32, 64], n)
    beta = rng.uniform(2.0, 8.0, n)
    bc = rng.integers(0, 3, n)
    return [q.Job(int(L[i]), int(L[i]), int(Ly[i]), float(np.float32(beta[i])), bc=int(bc[i]),
                  algo=q.ALGO_HEATBATH, start=q.START_HOT, seed=9000 + i, n_therm=50, n_meas=200)
            for i in range(n)]


def test_server_c5_poisson_stream(gpu, oracle_mod):
    """The serving configuration as a stream: 40 requests of the C5 shape,
    the first 12 at once and the rest as a Poisson process (mean 200/s, so
    that several are in flight at once), into a server whose memory budget
    holds two of the largest jobs: the burst alone needs more than the
    budget, so requests wait for memory and join running ticks as others
    retire.  Every served series equals, bit for bit, the same job run alone
    through the C driver su2_run (the reference's run loop,
    kernel.cu:161-236), and the four smallest jobs agree with the checker's
    run loop to 1e-12 per measurement."""
    import ctypes as C

    q = gpu
    _say("[server] C5 burst + Poisson stream, 40 requests, tight budget")
    rng = np.random.default_rng(777)
    jobs = c5_requests(q, 40, rng)
    burst = 12
    arrivals = np.concatenate([np.zeros(burst), np.cumsum(rng.exponential(1.0 / 200.0, size=len(jobs) - burst))])
    lib = q.load_library()
    nbytes = [lib.su2_job_bytes(C.byref(j.c()), 0) for j in jobs]
    budget = 2 * max(nbytes) + (1 << 20)
    assert sum(nbytes[:burst]) > budget  # the burst cannot be resident at once
    srv = q.Server(precision=0, max_jobs=64, budget_bytes=budget, use_graph=True)
    try:
        base = srv.now()
        ids = []
        for t, j in zip(arrivals, jobs):
            d = base + t - srv.now()
            if d > 0:
                time.sleep(d)
            ids.append(srv.submit(j))
        srv.drain()
        series = {i: srv.series(i) for i in ids}
        res = {r.id: r for r in srv.poll()}
        assert sorted(res) == sorted(ids)
        for i, j in zip(ids, jobs):
            assert res[i].status == 0, f"request {i}: status {res[i].status}"
            r = q.run(Lt=j.Lt, Lx=j.Lx, Ly=j.Ly, beta=j.beta, boundary=j.bc, algo=j.algo, init=q.START_HOT,
                      N=j.n_meas, N_term=j.n_therm, seed=j.seed, save=1, sconf=1)
            assert np.array_equal(series[i][0].view(np.uint64), r.energy.view(np.uint64)), f"request {i}: energies"
            assert np.array_equal(series[i][1].view(np.uint64), r.plaquette.view(np.uint64)), \
                f"request {i}: plaquettes"
            assert abs(res[i].P_mean - r.P_mean) <= 1e-12 and abs(res[i].accept - r.accept) <= 1e-12
        for k in sorted(range(len(jobs)), key=lambda k: jobs[k].V)[:4]:
            j = jobs[k]
            _say(f"[server]   request {ids[k]} {j.Lt}x{j.Lx}x{j.Ly} bc={j.bc} against the checker")
            E, P = series[ids[k]]
            check_against_oracle(oracle_mod, j, 0, E, P, accept=res[ids[k]].accept, tag=f"served request {ids[k]}")
        waited = [res[i].t_admit - res[i].t_submit for i in ids]
        iv = sorted((res[i].t_admit, res[i].t_done) for i in ids)
        _say(f"[server]   longest wait for admission {max(waited) * 1e3:.1f} ms")
        assert any(nxt[0] < cur[1] for cur, nxt in zip(iv, iv[1:])), "no two requests were resident together"
        # some request of the burst was admitted only after another had retired
        first_done = min(res[i].t_done for i in ids[:burst])
        assert any(res[i].t_admit >= first_done for i in ids[:burst]), "no request waited for memory"
    finally:
        srv.close()


def test_server_fp32_jobs_with_fp64_twins(gpu):
    """An FP32-link server with requests that ask for an FP64 twin and
    requests that do not: every FP32 series equals the job run alone; each
    twin's plaquette and binned error equal the same job run with FP64 links
    through su2_run, and the stated difference and combined error follow
    from the two series."""
    q = gpu
    _say("[server] FP32 jobs with FP64 twins")
    rng = np.random.default_rng(31)
    jobs = c5_requests(q, 6, rng)
    for k, j in enumerate(jobs):
        j.n_therm, j.n_meas = 20, 60
        j.fp64_twin = int(k % 3 != 2)
    srv = q.Server(precision=q.PREC_F32, max_jobs=16, budget_bytes=256 << 20, use_graph=True)
    try:
        ids = [srv.submit(j) for j in jobs]
        srv.drain()
        series = {i: srv.series(i) for i in ids}
        res = {r.id: r for r in srv.poll()}
        for i, j in zip(ids, jobs):
            r = res[i]
            assert r.status == 0
            kw = dict(Lt=j.Lt, Lx=j.Lx, Ly=j.Ly, beta=j.beta, boundary=j.bc, algo=j.algo, init=q.START_HOT,
                      N=j.n_meas, N_term=j.n_therm, seed=j.seed)
            r32 = q.run(precision=q.PREC_F32, **kw)
            assert np.array_equal(series[i][1].view(np.uint64), r32.plaquette.view(np.uint64)), f"request {i}"
            if not j.fp64_twin:
                assert r.P_fp64_twin == 0.0 and r.P_diff_sigma == 0.0
                continue
            r64 = q.run(precision=q.PREC_F64, **kw)
            m64, e64 = q.stats.binned_error(r64.plaquette, 20)
            m32, e32 = q.stats.binned_error(r32.plaquette, 20)
            assert abs(r.P_fp64_twin - m64) <= 1e-12 and abs(r.P_fp64_twin_err - e64) <= 1e-9 * e64
            assert abs(r.P_err_binned - e32) <= 1e-9 * e32
            assert abs(r.P_diff - (m32 - m64)) <= 1e-12
            assert abs(r.P_diff_sigma - np.hypot(e32, e64)) <= 1e-9 * r.P_diff_sigma
            _say(f"[server]   request {i} {j.Lt}x{j.Lx}x{j.Ly}: FP32 - FP64 = {r.P_diff / r.P_diff_sigma:+.2f} sigma")
    finally:
        srv.close()
