#!/usr/bin/env python3
"""Benchmark of the SU(2) checkerboard sweep on one MI355X.

Workload (the large single-lattice configuration): Lt = Lx = Ly = 128,
periodic, heat-bath followed by one over-relaxation sweep, beta 5.0, hot
start with Haar-random links from seed 7, FP64 links.  One step = one
heat-bath sweep (6 launches) + one over-relaxation sweep (6 launches) + the
fused energy/plaquette measurement, with the reunitarisation every 100
steps, as in the reference's measurement loop (kernel.cu:192-208); as in the
C driver, each measurement is carried by the next step's first launch
(su2_mc_up_measured) and the last one runs on its own, inside the timing.  Before
anything is timed the lattice is thermalised with the configuration's 200
heat-bath + over-relaxation sweeps (--therm, the thermalisation loop of
kernel.cu:161-179); --warmup steps after that are a timing warm-up only.

Prints ONE JSON line.  value = link updates per second of the whole step
(2 x 3 V updates per step).  roofline: the heat-bath sweep kernel, 240
algorithmic bytes per link update (5 link reads + 1 link write of 32 B and
48 B of RNG state read + written), timed with HIP events around the six
launches of a sweep on the stream they run on.  cpu_baseline: the CPU checker (OpenMP C restatement of
the reference) on the host cores, on a bounded number of steps of the same
lattice.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BYTES_HB = 240.0  # per heat-bath link update, FP64 links
BYTES_OR = 224.0  # per over-relaxation link update, FP64 links
BYTES_METRO = 272.0  # per Metropolis link update (any number of hits), FP64 links
BYTES_EN_AV = 96.0  # per site and measurement, FP64 links
# the same with FP32 link storage (16 B links; the 48 B of RNG state per
# heat-bath and Metropolis update are unchanged)
BYTES_F32 = {"hb": 144.0, "or": 112.0, "metro": 160.0, "en_av": 48.0}


def update_bytes(precision):
    """Algorithmic bytes per link update (per site for en_av) of the launches
    with the given link storage."""
    if precision == 1:
        return dict(BYTES_F32)
    return {"hb": BYTES_HB, "or": BYTES_OR, "metro": BYTES_METRO, "en_av": BYTES_EN_AV}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--therm", type=int, default=200,
                    help="thermalisation sweeps (heat-bath + over-relaxation) before any timing, whatever "
                         "--warmup says")
    ap.add_argument("--L", type=int, default=128, help="Lt = Lx = Ly")
    ap.add_argument("--precision", type=int, default=0, help="0 FP64 links, 1 FP32 links")
    ap.add_argument("--cpu-steps", type=int, default=-1, help="CPU baseline steps (-1: auto, 0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--roofline-sweeps", type=int, default=20)
    ap.add_argument("--large-L", type=int, default=256, help="extent of the DRAM-bound timing (0: skip)")
    ap.add_argument("--latency-floor", type=int, default=1,
                    help="phase-separation probe (launch vs cooperative grid barrier) and small-lattice launches "
                         "(0: skip; profiled runs skip it)")
    ap.add_argument("--shard-shares", type=int, default=1,
                    help="with the --large-L timing, time the launches on one GPU's share of a t-slab "
                         "decomposition over 8 GPUs (0: skip)")
    ap.add_argument("--small-configs", type=int, default=1,
                    help="run the 16^3 and 32x32x48 configurations through the C driver, with the checker "
                         "timed beside them (0, or --cpu-steps 0: skip)")
    ap.add_argument("--batch-jobs", type=int, default=64, help="lattices of the batched ensemble (0: skip)")
    ap.add_argument("--batch-therm", type=int, default=20)
    ap.add_argument("--batch-meas", type=int, default=80)
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched-ensemble comparison at --precision and print its JSON")
    ap.add_argument("--serving-only", action="store_true", help="run only the serving stream and print its JSON")
    ap.add_argument("--ceiling-only", action="store_true",
                    help="run only the streaming-bandwidth ceilings (su2_stream_probe) and print their JSON")
    ap.add_argument("--serve-requests", type=int, default=100, help="served requests (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-burst", type=int, default=200,
                    help="requests submitted at once to measure the server's capacity (0, or --serve-requests 0: "
                         "skip)")
    ap.add_argument("--fp32-meas", type=int, default=1000, help="measurements of the FP32-link check (0: skip)")
    ap.add_argument("--fp32-seeds", type=int, default=16, help="seeds of the FP32-link check per boundary")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="after the run, launch the plain copy kernel 5 times on buffers of this size "
                         "(reference dispatches for calibrating rocprofv3 byte counters)")
    return ap.parse_args()


def gpu_bench(args):
    import qcd_2d_gpu_amd as q
    from qcd_2d_gpu_amd import _lib as L

    lib = q.load_library()
    n = q.device_count()
    if n < 1:
        raise SystemExit("bench.py: no HIP device visible")
    stream = C.c_void_p()
    L.check(lib.su2_stream_create(C.byref(stream)), "stream_create")
    Lt = Lx = Ly = args.L
    g = q.Lattice(Lt, Lx, Ly, 5.0, 0.1, 1, precision=args.precision, stream=stream)
    V = g.V
    g.setup_kernel(7)
    g.haar_init()
    res = q.DeviceBuffer(8 * 2 * max(1, args.steps))
    res.zero(stream)

    # the measurement of a step is held back and carried by the next step's
    # first launch (su2_mc_up_measured, the same bits as en_av before the
    # sweep), as the C driver does; flush() measures the last one on its own
    pending = [None]

    def step(i, res_slot=None):
        if pending[0] is not None:
            s = pending[0]
            g.mc_up_measured(q.BC_PERIODIC, res.value + 16 * s, res.value + 16 * s + 8, q.ALGO_HEATBATH,
                             q.UPDATE_LINK)
            pending[0] = None
        else:
            g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_LINK)
        g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_OVERRELAX)
        if i % 100 == 0:
            g.normalize()
        pending[0] = res_slot

    def flush():
        if pending[0] is not None:
            s = pending[0]
            g.en_av_into(False, res.value + 16 * s, res.value + 16 * s + 8)
            pending[0] = None

    # thermalisation (kernel.cu:161-179): reunitarise before the sweep every
    # 100 sweeps and once after the loop
    for i in range(1, args.therm + 1):
        if i % 100 == 0:
            g.normalize()
        g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_LINK)
        g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_OVERRELAX)
    g.normalize()
    for i in range(1, args.warmup + 1):
        step(i, 0)
    flush()
    g.synchronize()
    res.zero(stream)
    g.synchronize()

    ev0, ev1 = C.c_void_p(), C.c_void_p()
    L.check(lib.su2_event_create(C.byref(ev0)), "event")
    L.check(lib.su2_event_create(C.byref(ev1)), "event")
    t0 = time.perf_counter()
    L.check(lib.su2_event_record(ev0, stream), "record")
    for i in range(1, args.steps + 1):
        step(i, i - 1)
    flush()
    L.check(lib.su2_event_record(ev1, stream), "record")
    ms = C.c_float()
    L.check(lib.su2_event_elapsed_ms(ev0, ev1, C.byref(ms)), "elapsed")
    g.synchronize()
    wall = time.perf_counter() - t0
    total_ms = float(ms.value)
    meas = res.download(np.float64, 2 * args.steps, stream).reshape(args.steps, 2)

    # roofline pass: the six heat-bath launches of a sweep bracketed by one
    # pair of HIP events on the stream they are launched on (per launch =
    # bracket / 6, so the event records add little); the over-relaxation
    # sweeps between them run unbracketed, as in the step
    ns = args.roofline_sweeps
    evs = []
    for _ in range(2 * max(ns, 2) + 2):
        e = C.c_void_p()
        L.check(lib.su2_event_create(C.byref(e)), "event")
        evs.append(e)

    def bracketed(fn, n, ends_itself=False):
        """fn(i) for i < n, each call bracketed by an event pair (fn records
        the closing event itself when ends_itself); returns the bracket
        durations in ms."""
        g.synchronize()
        for i in range(n):
            L.check(lib.su2_event_record(evs[2 * i], stream), "record")
            fn(i)
            if not ends_itself:
                L.check(lib.su2_event_record(evs[2 * i + 1], stream), "record")
        g.synchronize()
        out = []
        for i in range(n):
            L.check(lib.su2_event_elapsed_ms(evs[2 * i], evs[2 * i + 1], C.byref(ms)), "elapsed")
            out.append(float(ms.value))
        return out

    # whole sweeps through the sweep entry point, as in the step (it counts
    # acceptance once per sweep; single launches count in each)
    def sweep_pair(i):
        g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_LINK)
        L.check(lib.su2_event_record(evs[2 * i + 1], stream), "record")
        g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_OVERRELAX)

    hb_ms = [t / 6 for t in bracketed(sweep_pair, ns, ends_itself=True)]

    def or_sweep(i):
        g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_OVERRELAX)

    or_ms = [t / 6 for t in bracketed(or_sweep, ns)]

    # en_av (partial + final; six measurements per bracket) and Metropolis
    # (eps 0.5, 4 hits; one sweep of six launches per bracket)
    def en_av6(i):
        for _ in range(6):
            g.en_av_into(False, res.value, res.value + 8)

    en_ms = [t / 6 for t in bracketed(en_av6, ns)]

    # the measurement as the step makes it, carried by the next heat-bath
    # sweep (su2_mc_up_measured): steps with and without it, alternating,
    # one bracket per step; the difference is the measurement's cost
    def step_pair(i):
        if i % 2 == 0:
            g.mc_up_measured(q.BC_PERIODIC, res.value, res.value + 8, q.ALGO_HEATBATH, q.UPDATE_LINK)
        else:
            g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_LINK)
        g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_OVERRELAX)

    sp_ms = bracketed(step_pair, max(2, 2 * (ns // 2)))
    step_meas_ms, step_plain_ms = float(np.mean(sp_ms[0::2])), float(np.mean(sp_ms[1::2]))
    g.set_params(epsilon=0.5, updates=4)

    def metro_sweep(i):
        g.mc_up_per(q.ALGO_METROPOLIS, q.UPDATE_LINK)

    me_ms = [t / 6 for t in bracketed(metro_sweep, ns)]
    for e in evs + [ev0, ev1]:
        lib.su2_event_destroy(e)
    if args.copy_bytes:
        src, dst = q.DeviceBuffer(args.copy_bytes), q.DeviceBuffer(args.copy_bytes)
        src.zero(stream)
        for _ in range(5):
            L.check(lib.su2_device_copy(dst.ptr, src.ptr, args.copy_bytes, stream), "device_copy")
        g.synchronize()
        src.free()
        dst.free()
    det_ok = float(np.mean(meas[:, 1]))
    g.close()
    res.free()
    lib.su2_stream_destroy(stream)
    return dict(total_ms=total_ms, wall=wall, V=V, hb_ms=float(np.mean(hb_ms)), or_ms=float(np.mean(or_ms)),
                hb_ms_min=float(np.min(hb_ms)), E=float(np.mean(meas[:, 0])), P=det_ok,
                en_ms=float(np.mean(en_ms)), step_meas_ms=step_meas_ms, step_plain_ms=step_plain_ms, me_ms=float(np.mean(me_ms[1:] if len(me_ms) > 1 else me_ms)))


def host_cores():
    """Host cores: os.cpu_count() (nproc) and the cores this process may run
    on (its affinity mask)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "affinity": aff}


def host_threads(args):
    """Threads of the checker's OpenMP loops: every core this process may
    run on, unless --cpu-threads says otherwise."""
    return args.cpu_threads or host_cores()["affinity"]


def cpu_baseline(args, V_gpu):
    from oracle import oracle as O

    threads = host_threads(args)
    O.set_threads(threads)
    L = args.L
    o = O.Oracle(L, L, L, 5.0)
    o.setup_kernel(7)
    o.start(2, 0)  # Haar-random hot start, as the GPU leg
    # one warm-up step, then time a bounded number of steps
    o.time_sweeps(0, 1, 1, 1)
    steps = args.cpu_steps
    if steps < 0:
        t1 = o.time_sweeps(0, 1, 1, 1)
        steps = max(1, min(300, int(15.0 / max(t1, 1e-6))))
    if steps == 0:
        return None
    sec = o.time_sweeps(0, 1, 1, steps)
    ups = 2 * 3 * o.V * steps / sec
    out = {"value": ups, "unit": "link updates/s", "cores": threads, "kind": "port", "host": host_cores(),
           "sample": f"{steps} steps (heat-bath + over-relaxation sweep) of the same {L}^3 lattice, "
                     f"{sec:.1f} s, OpenMP C restatement of the reference"}
    if threads > 1 and args.cpu_steps < 0:
        # the same restatement on one thread, over a step or two
        O.set_threads(1)
        s1 = max(1, min(2, int(steps / threads)))
        sec1 = o.time_sweeps(0, 1, 1, s1)
        out["one_thread"] = {"value": 2 * 3 * o.V * s1 / sec1, "steps": s1, "seconds": sec1}
        O.set_threads(threads)
    return out


def lattice_launch_bench(Lt, Lx, Ly, reps=5, therm=2):
    """Heat-bath and over-relaxation launches of one Lt x Lx x Ly lattice
    (FP64 links, periodic, beta = 5), after `therm` heat-bath +
    over-relaxation sweeps from the Haar-random start (reunitarised every
    100 and at the end, as the thermalisation of kernel.cu:161-179): the six
    launches of a sweep bracketed by one pair of HIP events on the launch
    stream, per launch = bracket / 6."""
    import qcd_2d_gpu_amd as q
    from qcd_2d_gpu_amd import _lib as L

    lib = q.load_library()
    stream = C.c_void_p()
    L.check(lib.su2_stream_create(C.byref(stream)), "stream_create")
    g = q.Lattice(Lt, Lx, Ly, 5.0, precision=q.PREC_F64, stream=stream)
    g.setup_kernel(7)
    g.haar_init()
    for i in range(1, therm + 1):
        if i % 100 == 0:
            g.normalize()
        g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_LINK)
        g.mc_up_per(q.ALGO_HEATBATH, q.UPDATE_OVERRELAX)
    g.normalize()
    out = {"shape": f"{Lt}x{Lx}x{Ly}", "sites": g.V, "thermalisation_sweeps": therm}
    for upd, name, nbytes in ((q.UPDATE_LINK, "heatbath", BYTES_HB), (q.UPDATE_OVERRELAX, "overrelax", BYTES_OR)):
        evs = []
        for _ in range(reps):
            a, b = C.c_void_p(), C.c_void_p()
            L.check(lib.su2_event_create(C.byref(a)), "event")
            L.check(lib.su2_event_create(C.byref(b)), "event")
            L.check(lib.su2_event_record(a, stream), "record")
            g.mc_up_per(q.ALGO_HEATBATH, upd)
            L.check(lib.su2_event_record(b, stream), "record")
            evs.append((a, b))
        g.synchronize()
        ms, v = C.c_float(), []
        for a, b in evs:
            lib.su2_event_elapsed_ms(a, b, C.byref(ms))
            v.append(ms.value * 1e-3 / 6)
            lib.su2_event_destroy(a)
            lib.su2_event_destroy(b)
        t = float(np.mean(v))
        out[name] = {"avg_launch_us": t * 1e6, "algorithmic_gbs": nbytes * g.Vh / t / 1e9,
                     "frac_of_hbm_peak": nbytes * g.Vh / t / 1e9 / HBM_PEAK_GBS}
    g.close()
    lib.su2_stream_destroy(stream)
    return out


def latency_floor_bench():
    """What a phase of a small lattice's sweep costs: su2_phase_probe (one
    phase = every thread writes 8 doubles of its own workgroup's, then reads
    8 of the next workgroup's) separated by back-to-back kernel launches, by
    a grid-wide barrier in one cooperative launch, or by a barrier among the
    workgroups of one XCD only (one group per XCD, side by side), for 8, 32
    and 128 workgroups of 256 (1, 4 and 16 per XCD); beside it the heat-bath
    and over-relaxation launches of 16^3 and 32^3 lattices (2048 and 16384
    updates per launch)."""
    import qcd_2d_gpu_amd as q
    from qcd_2d_gpu_amd import _lib as L

    lib = q.load_library()
    stream = C.c_void_p()
    L.check(lib.su2_stream_create(C.byref(stream)), "stream_create")
    words = 8
    out = {"phase_separation": []}
    for blocks in (8, 32, 128):
        buf = q.DeviceBuffer(blocks * 256 * words * 8)
        buf.zero(stream)
        lu, bu, xu = C.c_double(), C.c_double(), C.c_double()
        L.check(lib.su2_phase_probe(buf.ptr, blocks, words, 200, stream, C.byref(lu), C.byref(bu), C.byref(xu)),
                "phase_probe")
        buf.free()
        out["phase_separation"].append({"workgroups": blocks, "workgroups_per_xcd": blocks // 8,
                                        "launch_us": lu.value, "grid_barrier_us": bu.value,
                                        "xcd_barrier_us": xu.value})
    lib.su2_stream_destroy(stream)
    for n in (16, 32):
        r = lattice_launch_bench(n, n, n, reps=20, therm=20)
        out[f"sweep_launch_{n}cube_us"] = {"heatbath": r["heatbath"]["avg_launch_us"],
                                           "overrelax": r["overrelax"]["avg_launch_us"]}
    return out


def large_lattice_bench(args):
    """L^3 launches (256^3: 1.6 GB of FP64 links, far beyond the 256 MB
    Infinity Cache, so DRAM-bound), after the configuration's 200
    thermalisation sweeps (about 1 s at 256^3)."""
    out = lattice_launch_bench(args.large_L, args.large_L, args.large_L, reps=10, therm=args.therm)
    out["L"] = args.large_L
    return out


# One GPU's share of a lattice slab-decomposed along t over 8 GPUs: the
# per-GPU kernel of the sharded scheme at its real shape (the ghost-plane
# exchange between launches is not part of this build).
SHARD_SHARES = ((16, 128, 128, "128^3 over 8 GPUs"), (32, 256, 256, "256^3 over 8 GPUs"))


def shard_share_bench():
    out = []
    for Lt, Lx, Ly, of in SHARD_SHARES:
        r = lattice_launch_bench(Lt, Lx, Ly)
        r["share_of"] = of
        out.append(r)
    return out


# The reference's small configurations, run end to end through the C host
# driver su2_run: start, thermalisation, measurements with en_av, statistics.
SMALL_CONFIGS = (
    ("C1", "16^3 periodic, heat-bath, beta 3.0, cold start, seed 1234",
     dict(Lt=16, Lx=16, Ly=16, beta=3.0, boundary=0, algo=1, init=1, seed=1234, epsilon=0.1, upd=1)),
    ("C3_pcc", "32x32x48 PCC, Metropolis eps 0.5, 4 hits, beta 4.0, hot start, seed 4321",
     dict(Lt=32, Lx=32, Ly=48, beta=4.0, boundary=1, algo=0, init=0, seed=4321, epsilon=0.5, upd=4)),
    ("C3_pmc", "32x32x48 PMC, Metropolis eps 0.5, 4 hits, beta 4.0, hot start, seed 4321",
     dict(Lt=32, Lx=32, Ly=48, beta=4.0, boundary=2, algo=0, init=0, seed=4321, epsilon=0.5, upd=4)),
)


def small_config_bench(args):
    """Each small configuration: 100 thermalisation + 100 measured sweeps
    through su2_run on the GPU (its own clock around the two loops, so
    normalisation, en_av, progress synchronisations and copies are inside),
    and the checker's 200 sweeps of the same lattice, start and update on
    the host cores."""
    import qcd_2d_gpu_amd as q
    from oracle import oracle as O

    threads = host_threads(args)
    O.set_threads(threads)
    n_t = n_m = 100
    out = {}
    for name, desc, kw in SMALL_CONFIGS:
        q.run(N=n_m, N_term=n_t, **kw)  # first run: allocations and code paths warm
        r = q.run(N=n_m, N_term=n_t, **kw)
        gsec = r.seconds_therm + r.seconds_meas
        V = kw["Lt"] * kw["Lx"] * kw["Ly"]
        o = O.Oracle(kw["Lt"], kw["Lx"], kw["Ly"], kw["beta"], kw["epsilon"], kw["upd"])
        o.setup_kernel(kw["seed"])
        o.start(kw["init"], kw["boundary"])
        o.time_sweeps(kw["boundary"], kw["algo"], 0, 2)
        csec = o.time_sweeps(kw["boundary"], kw["algo"], 0, n_t + n_m)
        gpu = r.link_updates / gsec
        cpu = 3 * V * (n_t + n_m) / csec
        out[name] = {"workload": desc, "sweeps": n_t + n_m, "gpu_link_updates_per_s": gpu, "gpu_seconds": gsec,
                     "cpu_link_updates_per_s": cpu, "cpu_seconds": csec, "cpu_cores": threads,
                     "gpu_over_cpu": gpu / cpu, "P_mean": r.P_mean, "P_err": r.P_err, "accept": r.accept}
    return out


def ensemble_jobs(q, n, therm, meas):
    """The batched-ensemble workload: Lt = Lx uniform over the even values
    8..48, Ly uniform over the even values 8..64, beta uniform in [2, 8]
    (rounded to float, as the reference stores it), boundary uniform over
    periodic / PCC / PMC, heat-bath, hot start, seeds 1000 + i."""
    rng = np.random.default_rng(20260101)
    jobs = []
    for i in range(n):
        L = int(rng.choice(np.arange(8, 49, 2)))
        Ly = int(rng.choice(np.arange(8, 65, 2)))
        beta = float(np.float32(rng.uniform(2.0, 8.0)))
        jobs.append(q.Job(L, L, Ly, beta, bc=int(rng.integers(0, 3)), algo=q.ALGO_HEATBATH, start=q.START_HOT,
                          seed=1000 + i, n_therm=therm, n_meas=meas, n_cor=1))
    return jobs


def batched_bench(args, precision):
    """All jobs in one engine (one persistent launch per direction x parity,
    hipGraph-captured ticks, autotuned shape) against the same jobs run one
    by one with the plain single-lattice kernels through the native driver
    su2_run.  Results must be identical."""
    import qcd_2d_gpu_amd as q

    lib = q.load_library()
    jobs = ensemble_jobs(q, args.batch_jobs, args.batch_therm, args.batch_meas)
    need = sum(lib.su2_job_bytes(C.byref(j.c()), precision) for j in jobs)
    eng = q.Engine(precision=precision, max_jobs=len(jobs), budget_bytes=int(need * 1.05) + (16 << 20),
                   use_graph=True)
    slots = [eng.add(j) for j in jobs]
    eng.sync()
    block, grid = eng.autotune()
    t0 = time.perf_counter()
    eng.run()
    t_batch = time.perf_counter() - t0
    series = {s: eng.read(s) for s in slots}
    eng.close()
    t_single = 0.0
    same = True
    for s, j in zip(slots, jobs):
        r = q.run(Lt=j.Lt, Lx=j.Lx, Ly=j.Ly, beta=j.beta, boundary=j.bc, algo=j.algo, init=q.START_HOT,
                  N=j.n_meas, N_term=j.n_therm, seed=j.seed, precision=precision, save=1, sconf=1)
        t_single += r.seconds_therm + r.seconds_meas
        same &= bool(np.array_equal(r.energy.view(np.uint64), series[s]["E"].view(np.uint64)) and
                     np.array_equal(r.plaquette.view(np.uint64), series[s]["P"].view(np.uint64)))
    ups = sum(j.link_updates for j in jobs)
    out = {"jobs": len(jobs), "sites": int(sum(j.V for j in jobs)), "sweeps_per_job": args.batch_therm + args.batch_meas,
           "link_updates": ups, "batched_s": t_batch, "one_by_one_s": t_single,
           "batched_jobs_per_s": len(jobs) / t_batch, "one_by_one_jobs_per_s": len(jobs) / t_single,
           "batched_link_updates_per_s": ups / t_batch, "one_by_one_link_updates_per_s": ups / t_single,
           "speedup": t_single / t_batch, "autotuned_block": block, "autotuned_grid": grid,
           "results_identical": same, "precision": "FP64" if precision == 0 else "FP32 links"}
    if precision == 0 and args.cpu_steps != 0:
        out["cpu_sample"] = ensemble_cpu_sample(args, jobs[:8])
    # per-job plaquette series, for the FP32-link error against FP64 (not printed)
    out["_plaquette"] = [series[s]["P"] for s in slots]
    return out


def fp32_ensemble_error(p64, p32, nbins=8):
    """Each job of the ensemble run with FP32 links against the same job with
    FP64 links (same seed): the difference of the mean plaquettes in units
    of the combined binned error (nbins bins of the measured sweeps)."""
    from qcd_2d_gpu_amd import stats

    z, d = [], []
    for a, b in zip(p64, p32):
        m64, e64 = stats.binned_error(a, nbins)
        m32, e32 = stats.binned_error(b, nbins)
        sig = float(np.hypot(e64, e32))
        d.append(m32 - m64)
        z.append((m32 - m64) / sig if sig > 0 else 0.0)
    za = np.abs(np.array(z))
    return {"jobs": len(z), "bins": nbins, "max_abs_diff": float(np.max(np.abs(d))),
            "max_abs_diff_over_sigma": float(za.max()), "within_1_sigma": float(np.mean(za <= 1.0)),
            "within_2_sigma": float(np.mean(za <= 2.0)), "mean_diff_over_sigma": float(np.mean(z))}


def ensemble_cpu_sample(args, jobs):
    """The checker on the host cores, job after job, over the sweeps of each
    job (start included, measurements left out): a bounded sample of the
    ensemble's CPU rate."""
    from oracle import oracle as O

    threads = host_threads(args)
    O.set_threads(threads)
    sec, ups = 0.0, 0
    for j in jobs:
        o = O.Oracle(j.Lt, j.Lx, j.Ly, j.beta)
        t0 = time.perf_counter()
        o.setup_kernel(j.seed)
        o.start(0, j.bc)
        o.time_sweeps(j.bc, 1, 0, j.n_therm + j.n_meas)
        sec += time.perf_counter() - t0
        ups += j.link_updates
    return {"value": ups / sec, "unit": "link updates/s", "cores": threads, "kind": "port",
            "jobs_per_s": len(jobs) / sec,
            "sample": f"the first {len(jobs)} jobs of the ensemble, {ups} link updates, {sec:.2f} s"}


def serving_jobs(q, n, rng):
    """Run requests of the serving workload: Lt = Lx in {16,24,32,48,64},
    Ly in {16,32,64}, beta uniform in [2, 8] (float), random boundary,
    heat-bath, hot start, 50 thermalisation + 200 measured sweeps."""
    return [q.Job(int(L), int(L), int(Ly), float(np.float32(b)), bc=int(bc), algo=q.ALGO_HEATBATH,
                  start=q.START_HOT, seed=7000 + i, n_therm=50, n_meas=200)
            for i, (L, Ly, b, bc) in enumerate(zip(rng.choice([16, 24, 32, 48, 64], n), rng.choice([16, 32, 64], n),
                                                   rng.uniform(2.0, 8.0, n), rng.integers(0, 3, n)))]


def serving_bench(args, burst=False, use_graph=False):
    """The serving workload into the continuous-batching server (200 GB
    budget): a Poisson stream at --serve-rate, or with burst=True all
    --serve-burst requests at once (the server's capacity).  The server
    enqueues each tick's batched launches on its stream (use_graph=False):
    replaying them as a captured hipGraph made a lone job's tick 4-11 us
    slower on MI355X (38-58 us per tick at 16^3-64^3 without), and the served
    p50 about 8 % higher; use_graph=True measures that form."""
    import threading

    import qcd_2d_gpu_amd as q

    n = args.serve_burst if burst else args.serve_requests
    rng = np.random.default_rng(555)
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, size=n))
    if burst:
        arrivals[:] = 0.0
    jobs = serving_jobs(q, n, rng)
    srv = q.Server(precision=q.PREC_F64, max_jobs=256, budget_bytes=200 * 10 ** 9, use_graph=use_graph)
    # warm the server's kernels and the graphs of every tick plan the stream
    # meets (thermalisation, its last sweep, measuring sweeps, the
    # reunitarisation every 100 sweeps) with one small request of the same
    # schedule, as a server is warmed before it takes traffic
    srv.submit(q.Job(16, 16, 16, 3.0, algo=q.ALGO_HEATBATH, start=q.START_HOT, seed=1, n_therm=50, n_meas=200))
    srv.drain()
    srv.poll()
    base = srv.now()

    def feed():
        for t, j in zip(arrivals, jobs):
            d = base + t - srv.now()
            if d > 0:
                time.sleep(d)
            srv.submit(j)

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    srv.drain()
    res = [r for r in srv.poll() if r.status == 0]
    srv.close()
    lat = np.array([r.t_done - r.t_submit for r in res])
    t_first = min(r.t_submit for r in res)
    t_last = max(r.t_done for r in res)
    ups = sum(r.link_updates for r in res)
    out = {"requests": n, "completed": len(res),
           "jobs_per_s": len(res) / (t_last - t_first), "link_updates_per_s": ups / (t_last - t_first),
           "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
           "mean_s": float(lat.mean()), "window_s": t_last - t_first}
    if burst:
        out["arrival"] = "all at once"
    else:
        out["offered_rate_per_s"] = args.serve_rate
    out["ticks"] = "replayed hipGraphs" if use_graph else "batched launches enqueued on the server's stream"
    return out


def fp32_check(args):
    """The boundary + Metropolis + precision configuration: Lt = Lx = 32,
    Ly = 48, PCC and PMC, Metropolis eps 0.5, 4 hits, beta 4.0, hot start, in
    FP64 and with FP32 link storage; plaquette difference in units of the
    combined binned error.  Seed 4321 is the configuration's own run; with
    --fp32-seeds k, seeds 4321 .. 4321 + k - 1 run too (all in one batched
    engine per precision), and their differences are summarised: how many
    lie within 1 and 2 combined errors, their chi^2, and the mean difference
    over the seeds with its error, a test for a bias of FP32 links that is
    sqrt(k) times more sensitive than one run.  As a control, the same
    statistic between FP64 runs of two different seeds (pairs of seeds)
    shows its spread when only the Markov chain differs."""
    import qcd_2d_gpu_amd as q

    seeds = [4321 + k for k in range(max(1, args.fp32_seeds))]
    out = {}
    runs = {}
    for prec in (q.PREC_F64, q.PREC_F32):
        eng = q.Engine(precision=prec, max_jobs=2 * len(seeds), budget_bytes=2 << 30, use_graph=True)
        slots = {(bc, sd): eng.add(q.Job(32, 32, 48, 4.0, bc=bc, algo=q.ALGO_METROPOLIS, start=q.START_HOT,
                                         seed=sd, n_therm=200, n_meas=args.fp32_meas, epsilon=0.5, hits=4))
                 for bc in (q.BC_PCC, q.BC_PMC) for sd in seeds}
        eng.run()
        for key, s in slots.items():
            runs[(prec,) + key] = eng.read(s)
        eng.close()
    for bc, name in ((q.BC_PCC, "pcc"), (q.BC_PMC, "pmc")):
        rows = []
        for sd in seeds:
            P64, P32 = runs[(0, bc, sd)]["P"], runs[(1, bc, sd)]["P"]
            m64, e64 = q.stats.binned_error(P64, 20)
            m32, e32 = q.stats.binned_error(P32, 20)
            sig = float(np.hypot(e64, e32))
            rows.append({"seed": sd, "P_fp64": m64, "err_fp64": e64, "P_fp32": m32, "err_fp32": e32,
                         "diff": m32 - m64, "diff_over_sigma": (m32 - m64) / sig if sig else 0.0,
                         "accept_fp64": runs[(0, bc, sd)]["accept"], "accept_fp32": runs[(1, bc, sd)]["accept"]})
        out[name] = dict(rows[0])
        if len(rows) > 1:
            z = np.array([r["diff_over_sigma"] for r in rows])
            d = np.array([r["diff"] for r in rows])
            sg = np.array([np.hypot(r["err_fp64"], r["err_fp32"]) for r in rows])
            md, me = float(d.mean()), float(np.sqrt(np.sum(sg ** 2)) / len(rows))
            # control: two FP64 runs of different seeds, the spread of the
            # statistic when only the chain differs
            ctl = []
            for a, b in zip(rows[0::2], rows[1::2]):
                sg2 = np.hypot(a["err_fp64"], b["err_fp64"])
                ctl.append(float((b["P_fp64"] - a["P_fp64"]) / sg2) if sg2 else 0.0)
            out[name]["seeds"] = {"n": len(rows), "diff_over_sigma": [float(x) for x in z],
                                  "within_1_sigma": int(np.sum(np.abs(z) <= 1.0)),
                                  "within_2_sigma": int(np.sum(np.abs(z) <= 2.0)),
                                  "chi2": float(np.sum(z ** 2)),
                                  "mean_diff": md, "mean_diff_err": me,
                                  "mean_diff_over_err": md / me if me else 0.0,
                                  "control_fp64_seed_pairs_diff_over_sigma": ctl}
    return out


NCU = 256  # compute units of an MI355X


def stream_ceiling(args):
    """Measured streaming ceilings that the sweep's HBM rate is compared
    with: su2_stream_probe (reads source arrays and writes their XOR, 16-B
    items, nothing else to do) over unroll x nontemporal x grid, best rate
    per (regime, reads).  'dram': 512 MiB per array, far beyond the 256 MB
    Infinity Cache; 'cache': 40 MiB per array (at most 200 MiB in all),
    resident in it.  reads = 1 is a copy; reads = 4 is close to the heat-bath
    sweep's 3.3 : 1 read : write byte ratio.  Rate = (reads + 1) x bytes per
    array / launch time; the launch time is one HIP-event bracket around 10
    back-to-back launches / 10, after one warm-up launch, so the dispatch
    gap is amortised as in a sweep of back-to-back launches."""
    import qcd_2d_gpu_amd as q
    from qcd_2d_gpu_amd import _lib as L

    lib = q.load_library()
    stream = C.c_void_p()
    L.check(lib.su2_stream_create(C.byref(stream)), "stream_create")
    a, b = C.c_void_p(), C.c_void_p()
    L.check(lib.su2_event_create(C.byref(a)), "event")
    L.check(lib.su2_event_create(C.byref(b)), "event")
    ms = C.c_float()
    out = {}
    for regime, region in (("dram", 512 << 20), ("cache", 40 << 20)):
        for reads in (1, 4):
            src, dst = q.DeviceBuffer(reads * region), q.DeviceBuffer(region)
            src.zero(stream)
            dst.zero(stream)
            rows = []
            for unroll in (1, 2, 4, 8):
                for nt in (0, 1):
                    for bpc in (0, 2, 4, 8, 16, 32):
                        blocks = bpc * NCU
                        L.check(lib.su2_stream_probe(dst.ptr, src.ptr, region, reads, unroll, nt, blocks, stream),
                                "stream_probe")
                        L.check(lib.su2_event_record(a, stream), "record")
                        for _ in range(10):
                            L.check(lib.su2_stream_probe(dst.ptr, src.ptr, region, reads, unroll, nt, blocks,
                                                         stream), "stream_probe")
                        L.check(lib.su2_event_record(b, stream), "record")
                        L.check(lib.su2_event_elapsed_ms(a, b, C.byref(ms)), "elapsed")
                        sec = ms.value * 1e-3 / 10
                        rows.append({"unroll": unroll, "nontemporal": nt, "blocks_per_cu": bpc,
                                     "gbs": (reads + 1) * region / sec / 1e9})
            src.free()
            dst.free()
            best = max(rows, key=lambda r: r["gbs"])
            out[f"{regime}_reads{reads}"] = {"bytes_per_array": region, "best_gbs": best["gbs"], "best": best,
                                             "frac_of_hbm_peak": best["gbs"] / HBM_PEAK_GBS,
                                             "worst_gbs": min(r["gbs"] for r in rows)}
    lib.su2_event_destroy(a)
    lib.su2_event_destroy(b)
    lib.su2_stream_destroy(stream)
    return out


# The sources the sweep, en_av and copy kernels are compiled from.  A
# rocprofv3 summary under profiles/ records their digest on its line
# "kernel_sources_sha16: ..."; a figure read from a summary whose digest
# differs from the current sources' is stale and not reported.
KERNEL_SOURCES = ("su2_sweep.hip", "su2_sweep_site.h", "su2_update.h", "su2_quat.h", "su2_fmath.h",
                  "su2_mrg32k3a.h", "su2_internal.h", "su2_misc.hip", "su2_observe.h")


def kernel_sources_digest():
    import hashlib

    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "qcd_2d_gpu_amd", "csrc", name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


# The probe shapes that came out best in stream_ceiling's full sweep on
# MI355X boxes (unroll, nontemporal, workgroups per CU; 0 = one item per
# thread): the default run measures only these and keeps the best.
CEILING_SHAPES = {("dram", 1): ((1, 1, 0), (2, 1, 0)), ("dram", 4): ((4, 1, 0), (2, 1, 0)),
                  ("cache", 1): ((1, 0, 4), (1, 0, 0), (8, 1, 16)), ("cache", 4): ((1, 0, 0), (2, 0, 8), (1, 0, 2))}


def stream_ceiling_quick():
    """The streaming ceilings of stream_ceiling at their best shapes only:
    {regime_readsN: GB/s}; 'cache' (40 MiB arrays, Infinity-Cache resident)
    bounds the 128^3 sweep (235 MB working set), 'dram' (512 MiB arrays) the
    256^3 one."""
    import qcd_2d_gpu_amd as q
    from qcd_2d_gpu_amd import _lib as L

    lib = q.load_library()
    stream = C.c_void_p()
    L.check(lib.su2_stream_create(C.byref(stream)), "stream_create")
    a, b = C.c_void_p(), C.c_void_p()
    L.check(lib.su2_event_create(C.byref(a)), "event")
    L.check(lib.su2_event_create(C.byref(b)), "event")
    ms = C.c_float()
    out = {}
    for (regime, reads), shapes in CEILING_SHAPES.items():
        region = (512 << 20) if regime == "dram" else (40 << 20)
        src, dst = q.DeviceBuffer(reads * region), q.DeviceBuffer(region)
        src.zero(stream)
        best = 0.0
        for unroll, nt, bpc in shapes:
            for _ in range(2):
                L.check(lib.su2_event_record(a, stream), "record")
                for _ in range(10):
                    L.check(lib.su2_stream_probe(dst.ptr, src.ptr, region, reads, unroll, nt, bpc * NCU, stream),
                            "stream_probe")
                L.check(lib.su2_event_record(b, stream), "record")
                L.check(lib.su2_event_elapsed_ms(a, b, C.byref(ms)), "elapsed")
                best = max(best, (reads + 1) * region / (ms.value * 1e-3 / 10) / 1e9)
        src.free()
        dst.free()
        out[f"{regime}_reads{reads}"] = best
    lib.su2_event_destroy(a)
    lib.su2_event_destroy(b)
    lib.su2_stream_destroy(stream)
    return out


def load_profile_value(fname, key, profiles_dir=None):
    """The number on the line "key: N" of a committed rocprofv3 summary
    under profiles/, or None when the summary is missing, lacks the key, or
    was made from other kernel sources than the current ones."""
    p = os.path.join(profiles_dir or os.path.join(ROOT, "profiles"), fname)
    if not os.path.exists(p):
        return None
    vals = {}
    with open(p) as f:
        for line in f:
            if ":" in line and not line.startswith(" "):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    if vals.get("kernel_sources_sha16") != kernel_sources_digest() or key not in vals:
        return None
    return float(vals[key])


def load_traffic():
    """HBM bytes per heat-bath launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated): the line
    "bytes_per_launch: N" of profiles/hb_sweep_128cube_traffic.txt, if it
    was made from the current kernel sources."""
    return load_profile_value("hb_sweep_128cube_traffic.txt", "bytes_per_launch")


def fp64_rate(r, updates):
    """FP64 flop rate of the heat-bath and over-relaxation launches: flops
    per update counted by rocprofv3 (SQ_INSTS_VALU_FLOPS_FP64 + _TRANS,
    profiles/sweep_128cube_flops.txt) x updates / the launch time measured
    here, against the 78.6 TFLOP/s FP64 vector peak."""
    hb = load_profile_value("sweep_128cube_flops.txt", "hb_fp64_flops_per_update")
    orf = load_profile_value("sweep_128cube_flops.txt", "or_fp64_flops_per_update")
    if hb is None or orf is None:
        return None
    hb_tf = hb * updates / (r["hb_ms"] * 1e-3) / 1e12
    or_tf = orf * updates / (r["or_ms"] * 1e-3) / 1e12
    return {"heatbath_tflops": hb_tf, "overrelax_tflops": or_tf, "peak_tflops": FP64_PEAK_TFLOPS,
            "heatbath_frac": hb_tf / FP64_PEAK_TFLOPS, "flops_per_update_heatbath": hb,
            "flops_per_update_overrelax": orf}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the single-lattice sweep has no multi-GPU path here", file=sys.stderr)
        sys.exit(2)
    if args.batched_only:
        r = batched_bench(args, args.precision)
        r.pop("_plaquette")
        print(json.dumps(r))
        return
    if args.ceiling_only:
        print(json.dumps({"stream_ceiling": stream_ceiling(args)}))
        return
    if args.serving_only:
        out = {"serving": serving_bench(args)}
        if args.serve_burst:
            out["serving_burst"] = serving_bench(args, burst=True)
        print(json.dumps(out))
        return
    r = gpu_bench(args)
    V = r["V"]
    ms_step = r["total_ms"] / args.steps
    ups = 2 * 3 * V / (ms_step * 1e-3)
    per_launch_updates = V / 2
    ub = update_bytes(args.precision)
    # the committed traffic and flop summaries are of the 128^3 FP64 workload
    headline = args.precision == 0 and args.L == 128
    traffic = load_traffic() if headline else None
    hb_gbs = ub["hb"] * per_launch_updates / (r["hb_ms"] * 1e-3) / 1e9
    or_gbs = ub["or"] * per_launch_updates / (r["or_ms"] * 1e-3) / 1e9
    step_bytes = 3 * V * (ub["hb"] + ub["or"]) + ub["en_av"] * V
    extra = {}

    def section(name, fn):
        """One optional part of the line; a failure there is recorded under
        its key instead of losing the headline measurement."""
        try:
            extra[name] = fn()
        except Exception as e:  # noqa: BLE001 - reported in the JSON line
            extra[name] = {"error": f"{type(e).__name__}: {e}"}
            print(f"bench.py: {name} failed: {e}", file=sys.stderr)

    section("stream_ceiling", stream_ceiling_quick)
    ceil = extra.get("stream_ceiling") if "error" not in extra.get("stream_ceiling", {"error": 1}) else None
    if args.large_L:
        section("large_lattice_fp64", lambda: large_lattice_bench(args))
        if ceil and "heatbath" in extra["large_lattice_fp64"]:
            for k in ("heatbath", "overrelax"):
                d = extra["large_lattice_fp64"][k]
                d["frac_of_stream_ceiling"] = d["algorithmic_gbs"] / ceil["dram_reads4"]
    if args.large_L and args.shard_shares:
        section("shard_share_fp64", shard_share_bench)
    if args.latency_floor:
        section("latency_floor", latency_floor_bench)
    if args.small_configs and args.cpu_steps != 0:
        section("small_configs", lambda: small_config_bench(args))
    if args.batch_jobs:
        section("batched_ensemble_fp64", lambda: batched_bench(args, 0))
        section("batched_ensemble_fp32", lambda: batched_bench(args, 1))
        p64 = extra["batched_ensemble_fp64"].pop("_plaquette", None)
        p32 = extra["batched_ensemble_fp32"].pop("_plaquette", None)
        if args.batch_meas >= 8 and p64 is not None and p32 is not None:
            extra["batched_ensemble_fp32"]["plaquette_vs_fp64"] = fp32_ensemble_error(p64, p32)
    if args.serve_requests:
        section("serving", lambda: serving_bench(args))
        section("serving_graph_replay", lambda: serving_bench(args, use_graph=True))
    if args.serve_requests and args.serve_burst:
        section("serving_burst", lambda: serving_bench(args, burst=True))
    if args.fp32_meas:
        section("fp32_links_vs_fp64", lambda: fp32_check(args))
    cpu = None
    if args.cpu_steps != 0:
        cpu = cpu_baseline(args, V)
    prec = "f64" if args.precision == 0 else "f64 (FP32 link storage)"
    out = {
        "metric": "link updates/s + HBM GB/s (FP64 128\u00b3); \u03b2-scan jobs/s with p50/p99 latency",
        "value_is": f"link updates/s of the {args.L}^3 step (heat-bath + over-relaxation sweep + en_av); "
                    "HBM GB/s in roofline, jobs/s and p50/p99 in batched_ensemble_* and serving*",
        "value": ups,
        "unit": "link updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec,
        "data": "synthetic (Haar-random hot start, MRG32k3a seed 7)",
        "config": {"workload": f"SU(2) 2+1D Wilson action, Lt=Lx=Ly={args.L}, periodic, beta=5.0, "
                               "heat-bath + 1 over-relaxation sweep per step + energy/plaquette measurement",
                   "Lt": args.L, "Lx": args.L, "Ly": args.L, "beta": 5.0, "boundary": "periodic",
                   "start": "Haar-random SU(2) links, seed 7",
                   "links": "FP64" if args.precision == 0 else "FP32", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": hb_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": hb_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "heat-bath sweep launch (one direction x parity)",
                     "bytes_per_update": ub["hb"], "updates_per_launch": per_launch_updates,
                     "avg_launch_us": r["hb_ms"] * 1e3,
                     "traffic_gbs": (traffic / (r["hb_ms"] * 1e-3) / 1e9) if traffic else None,
                     "traffic_source": "profiles/hb_sweep_128cube_traffic.txt" if traffic else
                     ("none: no rocprofv3 summary made from the current kernel sources "
                      f"({kernel_sources_digest()})" if headline else "none: the summary is of the 128^3 FP64 run"),
                     "stream_ceiling_gbs": ceil["cache_reads4"] if ceil else None,
                     "frac_of_stream_ceiling": hb_gbs / ceil["cache_reads4"] if ceil else None,
                     "stream_ceiling_is": "su2_stream_probe, 4 arrays in and 1 out of 40 MiB each (resident in "
                                          "the Infinity Cache like the 235 MB of the 128^3 lattice), measured in "
                                          "this run"},
        "overrelax_launch": {"avg_launch_us": r["or_ms"] * 1e3, "achieved_gbs": or_gbs,
                             "updates_per_s_kernel": per_launch_updates / (r["or_ms"] * 1e-3)},
        "fp64_rate": fp64_rate(r, per_launch_updates) if headline else None,
        "metropolis_launch": {"hits": 4, "epsilon": 0.5, "avg_launch_us": r["me_ms"] * 1e3,
                              "achieved_gbs": ub["metro"] * per_launch_updates / (r["me_ms"] * 1e-3) / 1e9,
                              "frac_of_hbm_peak": ub["metro"] * per_launch_updates / (r["me_ms"] * 1e-3) / 1e9
                              / HBM_PEAK_GBS},
        "measurement_in_sweep": {"avg_us": (r["step_meas_ms"] - r["step_plain_ms"]) * 1e3,
                                 "step_with_us": r["step_meas_ms"] * 1e3, "step_without_us": r["step_plain_ms"] * 1e3,
                                 "is": "the step's measurement as it runs there, carried by the next heat-bath "
                                       "sweep: a step with it minus a step without (HIP events, alternating)"},
        "en_av_launch": {"avg_us": r["en_ms"] * 1e3,
                         "achieved_gbs": ub["en_av"] * V / (r["en_ms"] * 1e-3) / 1e9,
                         "frac_of_hbm_peak": ub["en_av"] * V / (r["en_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "step_algorithmic_gbs": step_bytes / (ms_step * 1e-3) / 1e9,
        "heatbath_updates_per_s_kernel": per_launch_updates / (r["hb_ms"] * 1e-3),
        "observables": {"energy": r["E"], "plaquette": r["P"]},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
