"""bench.py's host side on the CPU: the one-GPU refusal, the numbers it reads
from the committed rocprofv3 summaries, the CPU-baseline record and the
batched-ensemble workload it draws."""
import argparse
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import qcd_2d_gpu_amd as q  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "4"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_profile_lookup_and_staleness(tmp_path):
    """A figure is read from a summary made from the current kernel sources,
    and dropped (None) when the summary's digest names other sources."""
    digest = bench.kernel_sources_digest()
    (tmp_path / "fresh.txt").write_text(f"# summary\nkernel_sources_sha16: {digest}\nbytes_per_launch: 267708585\n")
    (tmp_path / "stale.txt").write_text("kernel_sources_sha16: 0000000000000000\nbytes_per_launch: 267708585\n")
    (tmp_path / "none.txt").write_text("bytes_per_launch: 267708585\n")
    assert bench.load_profile_value("fresh.txt", "bytes_per_launch", str(tmp_path)) == 267708585.0
    assert bench.load_profile_value("fresh.txt", "no_such_key", str(tmp_path)) is None
    assert bench.load_profile_value("stale.txt", "bytes_per_launch", str(tmp_path)) is None
    assert bench.load_profile_value("none.txt", "bytes_per_launch", str(tmp_path)) is None
    assert bench.load_profile_value("missing.txt", "bytes_per_launch", str(tmp_path)) is None


def test_committed_profile_figures():
    """The summaries bench.py reads carry the digest of the sources they were
    made from; where it is the current one, their figures are in range."""
    for name in ("hb_sweep_128cube_traffic.txt", "sweep_128cube_flops.txt"):
        text = open(os.path.join(ROOT, "profiles", name)).read()
        assert "kernel_sources_sha16: " in text, name
    traffic = bench.load_traffic()
    if traffic is not None:
        assert 2.4e8 < traffic < 4.0e8  # HBM bytes per 128^3 heat-bath launch
        assert traffic / (240.0 * 1048576) >= 1.0  # never below the algorithmic bytes
    hb = bench.load_profile_value("sweep_128cube_flops.txt", "hb_fp64_flops_per_update")
    orf = bench.load_profile_value("sweep_128cube_flops.txt", "or_fp64_flops_per_update")
    if hb is not None:
        assert orf is not None and 200 < orf < hb < 1000
        r = {"hb_ms": 0.05, "or_ms": 0.04}
        f = bench.fp64_rate(r, 1048576)
        assert abs(f["heatbath_tflops"] - hb * 1048576 / 0.05e-3 / 1e12) < 1e-9
        assert f["peak_tflops"] == 78.6 and 0 < f["heatbath_frac"] < 1


def test_cpu_baseline_record():
    args = argparse.Namespace(cpu_threads=2, L=8, cpu_steps=2)
    out = bench.cpu_baseline(args, 8 ** 3)
    assert set(out) >= {"value", "unit", "cores", "kind", "sample"}
    assert out["unit"] == "link updates/s" and out["kind"] == "port" and out["cores"] == 2
    assert out["value"] > 0 and "2 steps" in out["sample"]
    args.cpu_steps = 0
    assert bench.cpu_baseline(args, 8 ** 3) is None


def test_ensemble_workload():
    jobs = bench.ensemble_jobs(q, 64, 20, 80)
    assert len(jobs) == 64
    assert jobs == bench.ensemble_jobs(q, 64, 20, 80)  # same draw every run
    for i, j in enumerate(jobs):
        assert j.Lt == j.Lx and 8 <= j.Lt <= 48 and j.Lt % 2 == 0
        assert 8 <= j.Ly <= 64 and j.Ly % 2 == 0
        assert 2.0 <= j.beta <= 8.0 and j.beta == float(np.float32(j.beta))
        assert j.bc in (q.BC_PERIODIC, q.BC_PCC, q.BC_PMC) and j.algo == q.ALGO_HEATBATH
        assert j.seed == 1000 + i and j.n_therm == 20 and j.n_meas == 80 and j.n_cor == 1
    assert {j.bc for j in jobs} == {0, 1, 2}
    assert sum(j.V for j in jobs) == 1891024


def test_ensemble_cpu_sample_record():
    jobs = bench.ensemble_jobs(q, 64, 20, 80)
    args = argparse.Namespace(cpu_threads=2, cpu_steps=-1)
    out = bench.ensemble_cpu_sample(args, jobs[:1])
    assert out["unit"] == "link updates/s" and out["kind"] == "port" and out["cores"] == 2
    assert out["value"] > 0 and out["jobs_per_s"] > 0 and "first 1 jobs" in out["sample"]


def test_small_configs_are_the_reference_configurations():
    names = [c[0] for c in bench.SMALL_CONFIGS]
    assert names == ["C1", "C3_pcc", "C3_pmc"]
    for name, desc, kw in bench.SMALL_CONFIGS:
        assert all(kw[k] % 2 == 0 for k in ("Lt", "Lx", "Ly"))
        q.make_lattice(kw["Lt"], kw["Lx"], kw["Ly"], kw["beta"])
    c1 = bench.SMALL_CONFIGS[0][2]
    assert (c1["Lt"], c1["Lx"], c1["Ly"], c1["beta"], c1["boundary"], c1["algo"], c1["init"], c1["seed"]) == \
        (16, 16, 16, 3.0, q.BC_PERIODIC, q.ALGO_HEATBATH, q.START_COLD, 1234)
    for _, _, kw in bench.SMALL_CONFIGS[1:]:
        assert (kw["Lt"], kw["Lx"], kw["Ly"], kw["beta"], kw["algo"], kw["epsilon"], kw["upd"]) == \
            (32, 32, 48, 4.0, q.ALGO_METROPOLIS, 0.5, 4)
    assert {kw["boundary"] for _, _, kw in bench.SMALL_CONFIGS[1:]} == {q.BC_PCC, q.BC_PMC}


def test_fp32_ensemble_error_record():
    """The per-job FP32-link error: identical series give zero, a shift of
    1.5 combined binned errors gives 1.5 sigma for that job."""
    rng = np.random.default_rng(4)
    p64 = [0.6 + 0.01 * rng.normal(size=80) for _ in range(5)]
    same = bench.fp32_ensemble_error(p64, [a.copy() for a in p64])
    assert same["jobs"] == 5 and same["max_abs_diff"] == 0.0 and same["within_1_sigma"] == 1.0
    from qcd_2d_gpu_amd import stats

    p32 = [a.copy() for a in p64]
    m, e = stats.binned_error(p64[2], 8)
    p32[2] = p64[2] + 1.5 * np.sqrt(2.0) * e  # the same error on both sides: combined sigma = sqrt(2) e
    out = bench.fp32_ensemble_error(p64, p32)
    assert abs(out["max_abs_diff_over_sigma"] - 1.5) < 1e-9
    assert out["within_1_sigma"] == 0.8 and out["within_2_sigma"] == 1.0
