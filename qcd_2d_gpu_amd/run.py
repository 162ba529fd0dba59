"""Python entry to the host driver su2_run (kernel.cu:49-268)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


@dataclass
class RunResult:
    energy: np.ndarray
    plaquette: np.ndarray
    E_mean: float
    E_err: float
    P_mean: float
    P_err: float
    accept: float
    det_mean: float
    seconds_therm: float
    seconds_meas: float
    link_updates: int
    Nterms: int
    # FP32-link runs with fp64_twin=True: the FP64 twin's plaquette, the
    # 20-bin errors, the difference and its combined error (else 0)
    P_fp64_twin: float = 0.0
    P_fp64_twin_err: float = 0.0
    P_err_binned: float = 0.0
    P_diff_fp32_fp64: float = 0.0
    P_diff_sigma: float = 0.0


def run(Lt: int = 16, Lx: int = 16, Ly: int = 16, beta: float = 3.0, boundary: int = L.BC_PERIODIC,
        algo: int = L.ALGO_HEATBATH, init: int = L.START_COLD, N: int = 100, N_term: int = 100, N_cor: int = 1,
        Nsave: int = 1, epsilon: float = 0.1, upd: int = 1, W: int = 80, or_sweeps: int = 0,
        precision: int = L.PREC_F64, normalize_every: int = 100, seed: int = 1234, readconf: int = 0,
        sconf: int = 1, save: int = 1, outdir: str = ".", nsave: str = "1", nconfr: str = "1",
        nconfs: str = "1", verbose: bool = False, device: int = 0, N_block: int = 20,
        Nblock_term: int = 10, fp64_twin: bool = False) -> RunResult:
    """Run one simulation with the reference's parameters (kernel.cu:10-46
    as keyword arguments).  sconf/save follow the reference: 0 = write."""
    lib = L.load()
    cfg = L.Su2RunConfig()
    lib.su2_run_config_default(C.byref(cfg))
    for k, v in dict(Lt=Lt, Lx=Lx, Ly=Ly, beta=beta, boundary=boundary, algo=algo, init=init, N=N, N_term=N_term,
                     N_cor=N_cor, Nsave=Nsave, epsilon=epsilon, upd=upd, W=W, or_sweeps=or_sweeps,
                     precision=precision, normalize_every=normalize_every, seed=seed, readconf=readconf,
                     sconf=sconf, save=save, verbose=int(verbose), device=device, N_block=N_block,
                     Nblock_term=Nblock_term, fp64_twin=int(fp64_twin)).items():
        setattr(cfg, k, v)
    keep = [s.encode() for s in (outdir, nsave, nconfr, nconfs)]
    cfg.outdir, cfg.nsave, cfg.nconfr, cfg.nconfs = keep
    en = np.zeros(N, dtype=np.float64)
    pl = np.zeros(N, dtype=np.float64)
    res = L.Su2RunResult()
    res.energy = en.ctypes.data_as(C.POINTER(C.c_double))
    res.plaquette = pl.ctypes.data_as(C.POINTER(C.c_double))
    L.check(lib.su2_run(C.byref(cfg), C.byref(res)), "su2_run")
    return RunResult(en, pl, res.E_mean, res.E_err, res.P_mean, res.P_err, res.accept, res.det_mean,
                     res.seconds_therm, res.seconds_meas, int(res.link_updates), res.Nterms, res.P_fp64_twin,
                     res.P_fp64_twin_err, res.P_err_binned, res.P_diff_fp32_fp64, res.P_diff_sigma)
