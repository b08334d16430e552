"""Batched sweeps: many (wavelength, r_eff, theta0, depth) cases with uneven
photon counts in one persistent launch.

All cases go to the device as one case table + band table; the kernel's work
queue hands out chunks of (case, photon range), so a block keeps walking
photons of whatever case is next instead of waiting at case boundaries, and
each chunk's FP64 BRF histogram lives in LDS until the chunk completes.  The
launch (tally clears + walk + record scatter) is captured once in a hipGraph
and replayed for repeated sweeps of the same shape.  ``autotune`` times a
grid of launch shapes on a capped copy of the cases and keeps the fastest;
results do not depend on the shape because each photon's Philox stream is
keyed by (seed, case, photon).
"""
from __future__ import annotations

import time

import numpy as np

from .gpu import GpuWalker, brf_from_tallies, fractions_from_tallies

DEFAULT_CANDIDATES = [
    (256, 4, 2048, 8), (256, 4, 8192, 8), (256, 6, 4096, 16), (512, 2, 4096, 8),
    (512, 3, 8192, 16), (1024, 1, 8192, 16), (256, 4, 4096, 32), (256, 5, 16384, 8),
]


def autotune(walker: GpuWalker, candidates=None, photons_cap=20_000, repeats=3, seed=1):
    """Pick the fastest launch shape for the cases currently set."""
    cands = list(candidates or DEFAULT_CANDIDATES)
    best, times = walker.autotune(cands, photons_cap=photons_cap, repeats=repeats, seed=seed)
    return best, dict(zip(cands, times.tolist()))


def run(walker: GpuWalker, cases, seed=0, use_graph=True, tune=True, records=False,
        candidates=None, photons_cap=20_000):
    """Walk all cases; returns dict(tallies, counts, brf, fractions, seconds,
    kernel_ms, launch, tune_times, steps, photons)."""
    walker.set_cases(cases)
    tune_times = None
    if tune:
        _, tune_times = autotune(walker, candidates, photons_cap=photons_cap)
    t0 = time.perf_counter()
    walker.walk(mode='native', seed=seed, records=records, use_graph=use_graph)
    t1 = time.perf_counter()
    st = walker.stats()
    t = walker.tallies()
    return dict(tallies=t, counts=walker.counts(), brf=brf_from_tallies(t),
                fractions=fractions_from_tallies(t), seconds=t1 - t0, kernel_ms=st['kernel_ms'],
                launch=walker.get_launch(), tune_times=tune_times, steps=st['steps'],
                photons=st['photons'])
