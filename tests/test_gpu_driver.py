"""GPU tests of the host driver su2_run (kernel.cu:49-268): the measurement
series against the CPU checker run on the reference's loop, the result
files in the reference's text formats (functions/output.h), checkpoint and
bit-exact resume, acceptance normalisation, and the su2mc executable."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL = 1e-12


def oracle_run(O, Lt, Lx, Ly, beta, bc, algo, init, N_term, N, seed, eps=0.1, upd=1):
    o = O.Oracle(Lt, Lx, Ly, beta, eps, upd)
    o.setup_kernel(seed)
    o.start(init, bc)
    for i in range(1, N_term + 1):
        if i % 100 == 0:
            o.normalize()
        o.mc_up(bc, algo, 0)
    o.normalize()
    after_therm = o.phi.copy()
    o.accept[:] = 0
    E, P = [], []
    for i in range(1, N + 1):
        o.mc_up(bc, algo, 0)
        if i % 100 == 0:
            o.normalize()
        e, p = o.en_av(bc == 2)
        E.append(e)
        P.append(p)
    return np.array(E), np.array(P), after_therm, o


def close(a, b):
    return np.all(np.abs(a - b) <= REL * np.maximum(np.abs(b), 1e-300))


def test_c1_run_matches_checker_and_writes_reference_files(gpu, oracle_mod, tmp_path):
    q = gpu
    print("[driver] C1 16^3 cold beta 3 seed 1234", flush=True)
    r = q.run(Lt=16, Lx=16, Ly=16, beta=3.0, boundary=0, algo=1, init=q.START_COLD, N=100, N_term=100, seed=1234,
              save=0, sconf=0, outdir=str(tmp_path))
    E, P, after_therm, o = oracle_run(oracle_mod, 16, 16, 16, 3.0, 0, 1, 1, 100, 100, 1234)
    assert close(r.energy, E) and close(r.plaquette, P)
    assert r.accept == 1.0 and r.Nterms == 100
    assert abs(r.det_mean - 1.0) < 1e-14
    tail = "16_16_16_B_3.00_per_1"
    conf = np.fromfile(tmp_path / f"Config_{tail}.txt", dtype=np.float64)
    assert np.array_equal(conf.view(np.uint64), after_therm.view(np.uint64))
    assert (tmp_path / f"Config_{tail}_data.txt").read_text() == "N_termalization 100 Type_termalization 1"
    m, v = q.stats.mean_variance(r.energy)
    lines = (tmp_path / f"Energy_{tail}.txt").read_text().splitlines()
    assert lines[0] == ("N_termalizacion 100 N 100 N_correlation 1  D_t 16 D_x 16 D_ y 16 deltaphi 0.10 "
                        "accept 1.00")
    assert lines[1] == "M Energy Error"
    assert lines[2] == f"3.000000 {m:.10f} {np.sqrt(v / 99):.10f}"
    chk = (tmp_path / f"Check_{tail}.txt").read_text().splitlines()
    assert chk[1] == "I Energy Avpl" and len(chk) == 102
    for i in (0, 57, 99):
        assert chk[i + 2] == f"{i} {r.energy[i]:.10f} {r.plaquette[i]:.10f}"
    av = (tmp_path / f"Energy_av_{tail}.txt").read_text().splitlines()
    assert len(av) == 4 and av[1] == "M Energy/Avplaquette Error"
    ac = (tmp_path / f"Autocorrelation_{tail}.txt").read_text().splitlines()
    assert len(ac) == 82 and ac[1] == "N Autocorrelation"
    rho = q.stats.autocorrelation(r.energy, 80, m, v)
    assert ac[2] == f"1 {rho[0]:.10f}"


def test_long_run_matches_checker(gpu, oracle_mod, tmp_path):
    """A long run through su2_run: 16x16x24 PMC heat-bath, beta 2.5, hot
    start, 200 thermalisation + 2000 measured sweeps (20 reunitarisations).
    Every one of the 2000 energies and plaquettes agrees with the checker's
    run loop to 1e-12 relative (no drift over the run), and so do the
    binned means, far inside their binned statistical errors."""
    q = gpu
    print("[driver] long run 16x16x24 PMC beta 2.5, 200 + 2000 sweeps", flush=True)
    r = q.run(Lt=16, Lx=16, Ly=24, beta=2.5, boundary=2, algo=1, init=q.START_HOT, N=2000, N_term=200, seed=99,
              save=1, sconf=1, outdir=str(tmp_path))
    E, P, _, _ = oracle_run(oracle_mod, 16, 16, 24, 2.5, 2, 1, 0, 200, 2000, 99)
    assert close(r.energy, E) and close(r.plaquette, P)
    for gpu_series, cpu_series in ((r.energy, E), (r.plaquette, P)):
        mg, eg = q.stats.binned_error(gpu_series, 20)
        mc, ec = q.stats.binned_error(cpu_series, 20)
        assert eg > 0 and abs(mg - mc) <= 1e-12 * abs(mc) and abs(mg - mc) < 0.01 * eg


@pytest.mark.parametrize("bc", [1, 2])
@pytest.mark.parametrize("prec", [0, 1])
def test_c3_run_matches_checker(gpu, oracle_mod, bc, prec):
    """The full 32x32x48 PCC / PMC Metropolis run (eps 0.5, 4 hits, beta
    4.0, hot start, seed 4321, 100 + 100 sweeps) through su2_run: every
    measurement against the checker's run loop, FP64 and FP32 links, and the
    acceptance against the checker's counters."""
    q = gpu
    print(f"[driver] C3 32x32x48 bc={bc} prec={prec}", flush=True)
    r = q.run(Lt=32, Lx=32, Ly=48, beta=4.0, boundary=bc, algo=0, init=q.START_HOT, N=100, N_term=100, seed=4321,
              epsilon=0.5, upd=4, precision=prec)
    o = oracle_mod.Oracle(32, 32, 48, 4.0, 0.5, 4, f32=(prec == 1))
    o.setup_kernel(4321)
    o.start(0, bc)
    for i in range(1, 101):
        if i % 100 == 0:
            o.normalize()
        o.mc_up(bc, 0, 0)
    o.normalize()
    o.accept[:] = 0
    E, P = [], []
    for i in range(1, 101):
        o.mc_up(bc, 0, 0)
        if i % 100 == 0:
            o.normalize()
        e, p = o.en_av(bc == 2)
        E.append(e)
        P.append(p)
    assert close(r.energy, np.array(E)) and close(r.plaquette, np.array(P))
    acc = o.accept.astype(np.float64).sum() / o.V / (100 * 3 * 4)
    if bc == 1:
        acc /= (47 / 48 + 1 / 48 / 3)
    assert abs(r.accept - acc) <= 1e-12


def test_resume_from_checkpoint_is_bit_exact(gpu, tmp_path):
    q = gpu
    print("[driver] checkpoint/resume", flush=True)
    a = tmp_path / "a"
    a.mkdir()
    kw = dict(Lt=8, Lx=10, Ly=12, beta=2.5, boundary=2, algo=0, init=q.START_HOT, epsilon=0.4, upd=3, seed=77)
    full = q.run(N_term=60, N=40, sconf=0, save=1, outdir=str(a), **kw)
    resumed = q.run(N_term=0, N=40, readconf=2, save=1, outdir=str(a), **kw)
    assert np.array_equal(full.energy.view(np.uint64), resumed.energy.view(np.uint64))
    assert np.array_equal(full.plaquette.view(np.uint64), resumed.plaquette.view(np.uint64))
    assert resumed.Nterms == 60
    # without the RNG file the run reseeds (the reference's behaviour) and diverges
    os.remove(a / "Config_8_10_12_B_2.50_pmc_1_rng.bin")
    reseeded = q.run(N_term=0, N=40, readconf=2, save=1, outdir=str(a), **kw)
    assert not np.array_equal(reseeded.energy, full.energy)


@pytest.mark.parametrize("bc", [0, 1, 2])
def test_acceptance_normalisation(gpu, bc):
    q = gpu
    print(f"[driver] acceptance bc={bc}", flush=True)
    hb = q.run(Lt=8, Lx=8, Ly=10, beta=3.0, boundary=bc, algo=1, init=q.START_HOT, N=20, N_term=10, seed=3)
    assert abs(hb.accept - 1.0) < 1e-12  # heat-bath always accepts; PCC rescaled (kernel.cu:226-228)
    me = q.run(Lt=8, Lx=8, Ly=10, beta=3.0, boundary=bc, algo=0, init=q.START_HOT, N=20, N_term=10, seed=3,
               epsilon=0.5, upd=4)
    assert 0.05 < me.accept < 0.95


def test_fp32_links_run(gpu):
    q = gpu
    print("[driver] FP32 links", flush=True)
    r = q.run(Lt=16, Lx=16, Ly=16, beta=4.0, boundary=0, algo=1, init=q.START_HOT, N=50, N_term=50, seed=9,
              precision=q.PREC_F32)
    assert abs(r.det_mean - 1.0) < 1e-6
    assert 0.5 < r.P_mean < 0.9


def test_cli(gpu, tmp_path):
    exe = os.path.join(os.path.dirname(gpu.__file__), "su2mc")
    print("[driver] su2mc executable", flush=True)
    out = subprocess.run([exe, "--Lt", "8", "--Lx", "8", "--Ly", "8", "--beta", "3", "--init", "1", "--N", "20",
                          "--N_term", "10", "--seed", "5", "--save", "0", "--outdir", str(tmp_path)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "Plaquette" in out.stdout and "accept 1.000000" in out.stdout
    assert (tmp_path / "Check_8_8_8_B_3.00_per_1.txt").exists()
    # progress lines in the reference's format (kernel.cu:172-176,212-216): N_term / Nblock_term and
    # N / N_block sweeps apart, 10 + 20 by default
    progress = [ln for ln in out.stdout.splitlines() if re.fullmatch(r"[0-9.]+ Time -?\d+ Left\(min\)[-0-9.e+]+", ln)]
    assert len(progress) == 30, out.stdout
    out = subprocess.run([exe, "--Lt", "8", "--Lx", "8", "--Ly", "8", "--N", "20", "--N_term", "10", "--N_block", "4",
                          "--Nblock_term", "0"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    progress = [ln for ln in out.stdout.splitlines() if " Left(min)" in ln]
    assert [ln.split()[0] for ln in progress] == ["0.250000", "0.500000", "0.750000", "1.000000"]
    bad = subprocess.run([exe, "--Lt", "7"], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "even" in bad.stderr


def test_fp32_links_agree_with_fp64_within_binned_errors(gpu):
    """The boundary + Metropolis + precision configuration (32x32x48, PCC and
    PMC, eps 0.5, 4 hits, beta 4, hot start): FP32 link storage gives the
    FP64 average plaquette within 3 combined binned errors (20 bins)."""
    q = gpu
    for bc in (q.BC_PCC, q.BC_PMC):
        print(f"[driver] FP32 vs FP64 bc={bc}", flush=True)
        P = {}
        for prec in (q.PREC_F64, q.PREC_F32):
            r = q.run(Lt=32, Lx=32, Ly=48, beta=4.0, boundary=bc, algo=q.ALGO_METROPOLIS, init=q.START_HOT,
                      epsilon=0.5, upd=4, N_term=200, N=600, seed=4321, precision=prec)
            P[prec] = q.stats.binned_error(r.plaquette, 20)
        (m64, e64), (m32, e32) = P[q.PREC_F64], P[q.PREC_F32]
        assert abs(m32 - m64) < 3 * np.hypot(e64, e32), (bc, m64, e64, m32, e32)


def test_strong_coupling_on_gpu(gpu):
    """beta = 0.5: <P> = u + 4 u^5 with u = I2(beta)/I1(beta) (strong-coupling
    expansion of the 3-D SU(2) plaquette), within 4 binned errors."""
    from scipy.special import iv

    q = gpu
    print("[driver] strong coupling 16^3 beta 0.5", flush=True)
    r = q.run(Lt=16, Lx=16, Ly=16, beta=0.5, boundary=0, algo=q.ALGO_HEATBATH, init=q.START_HOT, N_term=100,
              N=1000, seed=11)
    u = iv(2, 0.5) / iv(1, 0.5)
    m, e = q.stats.binned_error(r.plaquette, 20)
    assert abs(m - (u + 4 * u ** 5)) < 4 * e + 2e-4, (m, e, u + 4 * u ** 5)


def test_missing_checkpoint_is_an_error(gpu, tmp_path):
    """readconf without a Config_ file: SU2_EIO with the file named (the
    reference exit(1)s, functions/lattice_setup.h:87-90)."""
    q = gpu
    from qcd_2d_gpu_amd import _lib as L

    print("[driver] missing checkpoint", flush=True)
    with pytest.raises(L.Su2Error) as e:
        q.run(Lt=8, Lx=8, Ly=8, beta=2.0, readconf=1, N_term=2, N=2, outdir=str(tmp_path))
    assert e.value.code == L.SU2_EIO
    assert "Config_8_8_8_B_2.00_per_1.txt" in str(e.value)


def test_runs_are_deterministic(gpu):
    """Two runs from the same seed give the same bits: no atomics or
    order-dependent sums anywhere in the sweep and the reduction."""
    q = gpu
    print("[driver] determinism 24x16x32 heat-bath + over-relaxation", flush=True)
    runs = []
    for _ in range(2):
        g = q.Lattice(24, 16, 32, 3.5)
        try:
            g.setup_kernel(31)
            g.rand_init2()
            for _ in range(5):
                g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_LINK)
                g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_OVERRELAX)
            runs.append((g.field_ref(), g.rng_state(), g.accept(), g.en_av()))
        finally:
            g.close()
    (f1, s1, a1, o1), (f2, s2, a2, o2) = runs
    assert np.array_equal(f1.view(np.uint64), f2.view(np.uint64))
    assert np.array_equal(s1, s2) and np.array_equal(a1, a2) and o1 == o2


@pytest.mark.slow
def test_c2_long_run_matches_checker(gpu, oracle_mod, tmp_path):
    """The large single-lattice configuration on its own schedule: 128^3,
    periodic, heat-bath + one over-relaxation sweep per sweep, beta 5.0,
    Haar-random hot start from seed 7, FP64, 200 thermalisation + 100
    measured sweeps (kernel.cu:161-236), through su2_run.  Every one of the
    100 energies and plaquettes agrees with the checker's run loop to 1e-12
    relative, the acceptance exactly, and the binned means (10 bins) agree
    far inside their binned errors."""
    from _jobs import close as close_rel, oracle_job

    q = gpu
    print("[driver] C2 128^3 HB+OR beta 5 Haar seed 7, 200 + 100 sweeps", flush=True)
    job = q.Job(128, 128, 128, 5.0, bc=0, algo=q.ALGO_HEATBATH, start=q.START_HAAR, seed=7, n_therm=200,
                n_meas=100, or_sweeps=1)
    r = q.run(Lt=128, Lx=128, Ly=128, beta=5.0, boundary=0, algo=q.ALGO_HEATBATH, init=q.START_HAAR, N=100,
              N_term=200, or_sweeps=1, seed=7, save=1, sconf=1, outdir=str(tmp_path))
    print("[driver]   GPU run done; checker run loop", flush=True)
    E, P, _, _, acc = oracle_job(oracle_mod, job, 0)
    assert close_rel(r.energy, E) and close_rel(r.plaquette, P)
    assert abs(r.accept - acc) <= 1e-12 and abs(acc - 1.0) <= 1e-12
    for gpu_series, cpu_series in ((r.energy, E), (r.plaquette, P)):
        mg, eg = q.stats.binned_error(gpu_series, 10)
        mc, ec = q.stats.binned_error(cpu_series, 10)
        assert eg > 0 and abs(mg - mc) <= 1e-12 * abs(mc) and abs(mg - mc) < 0.01 * eg
    # the thermalised plaquette of the weak-coupling side: 1 - <P> close to 1/beta (3-D SU(2))
    assert 0.75 < r.P_mean < 0.85


def test_fp32_run_states_its_error_against_an_fp64_twin(gpu, tmp_path):
    """su2_run with FP32 links and fp64_twin (the boundary + Metropolis +
    precision configuration, PMC, 200 + 600 sweeps): the FP32 series is the
    one the run gives without the twin; the twin's plaquette and 20-bin
    error are those of the same run with FP64 links (errors to 1e-9: the
    bin means are summed in another order than numpy's); the stated
    difference and combined error follow, are written into Energy_av_ and
    the FP32 result is within 3 combined errors of FP64."""
    q = gpu
    print("[driver] FP32 run with FP64 twin, 32x32x48 PMC Metropolis", flush=True)
    kw = dict(Lt=32, Lx=32, Ly=48, beta=4.0, boundary=q.BC_PMC, algo=q.ALGO_METROPOLIS, init=q.START_HOT,
              epsilon=0.5, upd=4, N_term=200, N=600, seed=4321)
    r = q.run(precision=q.PREC_F32, fp64_twin=True, save=0, outdir=str(tmp_path), **kw)
    alone = q.run(precision=q.PREC_F32, **kw)
    r64 = q.run(precision=q.PREC_F64, **kw)
    assert np.array_equal(r.plaquette.view(np.uint64), alone.plaquette.view(np.uint64))
    m64, e64 = q.stats.binned_error(r64.plaquette, 20)
    m32, e32 = q.stats.binned_error(r.plaquette, 20)
    assert abs(r.P_fp64_twin - m64) <= 1e-12 and abs(r.P_fp64_twin_err - e64) <= 1e-9 * e64
    assert abs(r.P_err_binned - e32) <= 1e-9 * e32
    assert abs(r.P_diff_fp32_fp64 - (m32 - m64)) <= 1e-12
    assert abs(r.P_diff_sigma - np.hypot(e32, e64)) <= 1e-9 * r.P_diff_sigma
    z = r.P_diff_fp32_fp64 / r.P_diff_sigma
    assert abs(z) < 3.0, z
    lines = (tmp_path / "Energy_av_32_32_48_B_4.00_pmc_1.txt").read_text().splitlines()
    assert len(lines) == 5 and lines[4].startswith("FP64_twin_plaquette ")
    f = lines[4].split()
    assert f[3] == "FP32_minus_FP64" and f[5] == "sigma" and f[7] == "diff_over_sigma"
    assert abs(float(f[1]) - m64) < 1e-9 and abs(float(f[4]) - (m32 - m64)) < 1e-9
    assert abs(float(f[-1]) - z) < 1e-3
    # without the twin the file keeps the reference's four lines
    q.run(precision=q.PREC_F32, save=0, outdir=str(tmp_path), nsave="2", **kw)
    assert len((tmp_path / "Energy_av_32_32_48_B_4.00_pmc_2.txt").read_text().splitlines()) == 4
