"""The reference's run loop for one job on the CPU checker, shared by the
engine, server and driver GPU tests.

kernel.cu:161-208 as the engine schedules it (csrc/su2_engine.cpp): the
thermalisation loop reunitarises before sweep i when i % N_norm == 0 and once
after the loop (kernel.cu:163-164,179); the counters are cleared
(kernel.cu:190); each measurement runs n_cor sweeps, reunitarises when
i % N_norm == 0 and measures en_av (kernel.cu:192-208).  A "sweep" is one
update sweep followed by the job's over-relaxation sweeps.
"""
from __future__ import annotations

import numpy as np

REL = 1e-12


def oracle_job(O, job, prec):
    """Runs `job` (a qcd_2d_gpu_amd.Job) on the checker with link storage
    `prec`; returns E, P series, the final field (reference layout), the RNG
    planes (device layout) and the mean acceptance (kernel.cu:246-249,
    PCC-corrected as kernel.cu:226-228)."""
    o = O.Oracle(job.Lt, job.Lx, job.Ly, job.beta, job.epsilon, job.hits, f32=(prec == 1))
    o.setup_kernel(job.seed)
    o.start(job.start, job.bc)
    nev = job.normalize_every or 100

    def sweep():
        o.mc_up(job.bc, job.algo, 0)
        for _ in range(job.or_sweeps):
            o.mc_up(job.bc, job.algo, 1)

    for i in range(1, job.n_therm + 1):
        if i % nev == 0:
            o.normalize()
        sweep()
    o.normalize()
    o.accept[:] = 0
    E, P = [], []
    for i in range(1, job.n_meas + 1):
        for _ in range(job.n_cor):
            sweep()
        if i % nev == 0:
            o.normalize()
        e, p = o.en_av(job.bc == 2)
        E.append(e)
        P.append(p)
    hits = job.hits if job.algo == 0 else 1
    n = job.n_meas * job.n_cor * 3 * (hits + job.or_sweeps)
    acc = o.accept.astype(np.float64).sum() / o.V / n
    if job.bc == 1:
        acc = acc / ((job.Ly - 1) / job.Ly + 1 / job.Ly / 3)
    return np.array(E), np.array(P), o.phi.copy(), o.rng_planes(), acc


def close(a, b, rel=REL):
    """Elementwise |a - b| <= rel |b| (energies and plaquettes: the checker
    sums en_av in another order than the device's two-level sum)."""
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and bool(np.all(np.abs(a - b) <= rel * np.maximum(np.abs(b), 1e-300)))


def check_against_oracle(O, job, prec, E, P, field=None, rng=None, accept=None, tag=""):
    """A job's device results against the checker's run loop: the series to
    1e-12 relative; the final links and RNG states bit for bit; the mean
    acceptance to 1e-12."""
    oE, oP, ophi, orng, oacc = oracle_job(O, job, prec)
    assert close(E, oE), f"{tag}: energies differ from the checker"
    assert close(P, oP), f"{tag}: plaquettes differ from the checker"
    if field is not None:
        assert np.array_equal(np.asarray(field).view(np.uint64), ophi.view(np.uint64)), f"{tag}: links differ"
    if rng is not None:
        assert np.array_equal(np.asarray(rng).reshape(orng.shape), orng), f"{tag}: RNG states differ"
    if accept is not None:
        assert abs(accept - oacc) <= 1e-12 * max(1.0, abs(oacc)), f"{tag}: acceptance {accept} vs {oacc}"
