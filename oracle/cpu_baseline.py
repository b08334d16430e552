"""CPU baseline: the reference's parallel photon loop restated (test and
CPU-baseline infrastructure only; bench.py's cpu_baseline leg).

Runs oracle/walk.py -- the reference's per-step structure: five numpy draws
per step, numpy scalar math, np.append trajectory growth -- as P worker
processes that reproduce parallelize.py:14-38: every rank draws the N photon
wavelengths from its own stream (mc3D:1519), rank 0's draw is split with
np.array_split into P parts (the scatter), each rank walks its part, and the
answers are gathered and flattened in rank order (mc3D:1618, 1845).
mpi4py is not installed, so the ranks are multiprocessing workers; each
rank's RandomState is seeded with seed + rank (the reference never seeds).
"""
from __future__ import annotations

import multiprocessing as mp
import os
import time

import numpy as np

from oracle import walk


def _rank(args):
    (rank, nranks, seed, n_photon, wvl0, fwhm, ssa, g, ext_cff_mss, tau_tot, rho_snw,
     theta0, lb, R, trajectory) = args
    rs = np.random.RandomState(seed + rank)
    all_wvls = walk.draw_wavelengths(rs, n_photon, wvl0, fwhm)
    if rank == 0:
        share = np.array_split(all_wvls, nranks)[0]
    else:
        # the scatter hands rank r part r of rank 0's draw
        rs0 = np.random.RandomState(seed)
        share = np.array_split(walk.draw_wavelengths(rs0, n_photon, wvl0, fwhm), nranks)[rank]
    n = share.size
    t0 = time.perf_counter()
    ice = dict(ssa=np.full(n, ssa), ext=np.full(n, ext_cff_mss), g=np.full(n, g))
    imp = dict(ssa=np.zeros(n), ext=np.zeros(n))
    par = walk.photon_params(share, ice, imp, 0.0, tau_tot, rho_snw)
    recs = walk.walk_rank(walk.RefStream(rs), par, tau_tot, rho_snw, theta0, lb, R,
                          trajectory=trajectory)
    t1 = time.perf_counter()
    return rank, recs, t1 - t0


def run(n_photon=10_000, processes=None, seed=0, wvl0=1.3, fwhm=1e-12, ssa=0.999, g=0.89,
        r_eff=100.0, rho_snw=300.0, rho_ice=917.0, depth_m=1.0, theta0=60.0, lb=True, R=0.5,
        trajectory=True):
    """Walk ``n_photon`` photons on ``processes`` ranks.  Returns dict with the
    gathered records, walk-phase wall time (max over ranks), end-to-end time."""
    P = processes or min(16, len(os.sched_getaffinity(0)))
    ext_cff_mss = walk.sphere_ext_cff_mss(r_eff, rho_ice)
    tau_tot = depth_m * ext_cff_mss * rho_snw
    args = [(r, P, seed, n_photon, wvl0, fwhm, ssa, g, ext_cff_mss, tau_tot, rho_snw, theta0,
             lb, R, trajectory) for r in range(P)]
    t0 = time.perf_counter()
    ctx = mp.get_context('fork')
    with ctx.Pool(P) as pool:
        parts = pool.map(_rank, args)
    t1 = time.perf_counter()
    parts.sort(key=lambda x: x[0])
    records = [r for _, recs, _ in parts for r in recs]
    walk_s = max(t for _, _, t in parts)
    steps = int(sum(r[4] + 1 for r in records))
    return dict(records=records, processes=P, walk_s=walk_s, total_s=t1 - t0,
                photons=len(records), steps=steps,
                photons_per_s=len(records) / walk_s, steps_per_s=steps / walk_s)


def run_c(n_photon=200_000, threads=None, seed=0, wvl=1.3, ssa=0.999, g=0.89, r_eff=100.0,
          rho_snw=300.0, rho_ice=917.0, depth_m=1.0, theta0=60.0, lb=True, R=0.5):
    """The same case walked by the C restatement of monte_carlo3D (oracle/mcoracle.c,
    FP64, numpy's MT19937 stream per rank): ``threads`` ranks of n_photon / threads
    photons each, one host thread per rank (the GIL is released inside C).  A
    compiled single-core-per-rank CPU line beside the reference-structured one."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import coracle
    T = threads or min(16, len(os.sched_getaffinity(0)))
    ext_cff_mss = walk.sphere_ext_cff_mss(r_eff, rho_ice)
    tau_tot = depth_m * ext_cff_mss * rho_snw
    per = -(-int(n_photon) // T)

    def par(n):
        return dict(wvl=np.full(n, wvl), g=np.full(n, g), ssa_ice=np.full(n, ssa), ssa_imp=np.zeros(n),
                    ext_cff_mss=np.full(n, ext_cff_mss), p_ext_imp=np.zeros(n),
                    snow_depth=np.full(n, tau_tot / (ext_cff_mss * rho_snw)))

    def one(r):
        mt = coracle.mt_from_numpy(np.random.RandomState(seed + r))
        t0 = time.perf_counter()
        out = coracle.walk(par(per), tau_tot, rho_snw, theta0, lb, R, mt=mt)
        return time.perf_counter() - t0, int(out['steps'])

    with ThreadPoolExecutor(max_workers=T) as ex:
        parts = list(ex.map(one, range(T)))
    walk_s = max(t for t, _ in parts)
    steps = sum(k for _, k in parts)
    return dict(threads=T, photons=per * T, steps=steps, walk_s=walk_s,
                photons_per_s=per * T / walk_s, steps_per_s=steps / walk_s)


def cpu_model():
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def run_c_polar(directory, p, tau_tot, n_photon=200_000, threads=None, seed=0, rho_snw=300.0,
                rho_ice=917.0):
    """An aspherical case (``p`` as monte_carlompi_amd.workloads.SNOW_POLAR,
    phase-matrix tables under ``directory``) walked by the C restatement
    (oracle/mcoracle.c with its phase-matrix rejection draws and Stokes update,
    FP64): ``threads`` ranks of n_photon / threads photons, one host thread per
    rank.  Only the walk is timed; the tables are read once."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import coracle
    T = threads or min(16, len(os.sched_getaffinity(0)))
    per = -(-int(n_photon) // T)
    zero = lambda w: dict(ssa=np.zeros(w.shape), ext=np.zeros(w.shape))
    _, par1, case, bands = coracle.polar_run_rank(
        seed, 1, p['wvl'], 1e-12, p['r_eff'], directory, p['stokes0'], zero, 0.0, tau_tot, rho_snw,
        rho_ice, p['theta0'], p['lambertian_bottom'], p['r_lambertian'])
    par = {k: np.full(per, v[0]) for k, v in par1.items()}
    keys = np.full(per, par1['wvl'][0])

    def one(r):
        rs = np.random.RandomState(seed + r)
        walk.draw_wavelengths(rs, per, p['wvl'], 1e-12)
        mt = coracle.mt_from_numpy(rs)
        t0 = time.perf_counter()
        out = coracle.walk(par, tau_tot, rho_snw, p['theta0'], p['lambertian_bottom'], p['r_lambertian'],
                           mt=mt, polar=(case, bands, keys, p['stokes0']))
        return time.perf_counter() - t0, int(out['steps'])

    with ThreadPoolExecutor(max_workers=T) as ex:
        parts = list(ex.map(one, range(T)))
    walk_s = max(t for t, _ in parts)
    steps = sum(k for _, k in parts)
    return dict(threads=T, photons=per * T, steps=steps, walk_s=walk_s,
                photons_per_s=per * T / walk_s, steps_per_s=steps / walk_s)
