"""Batched sweeps: many (wavelength, r_eff, theta0, depth) cases with uneven
photon counts in one persistent launch.

All cases go to the device as one case table + band table; the kernel's work
queue hands out chunks of (case, photon range), so a block keeps walking
photons of whatever case is next instead of waiting at case boundaries, and
each chunk's FP64 BRF histogram lives in LDS until the chunk completes.  The
launch (tally clears + walk + record scatter) is captured once in a hipGraph
and replayed for repeated sweeps of the same shape.  ``autotune`` times a
grid of launch shapes on a sampled subset of the cases and keeps the fastest.
Each photon's Philox stream is keyed by (seed, case, photon), so with the
cooperative tail mode off (coop_live 0) results do not depend on the shape;
with it on they agree statistically.
"""
from __future__ import annotations

import time

import numpy as np

from .gpu import GpuWalker, brf_from_tallies, fractions_from_tallies

DEFAULT_CANDIDATES = [
    (256, 6, 8192, 24), (256, 6, 4096, 24), (256, 6, 16384, 24), (256, 6, 8192, 16),
    (256, 6, 4096, 16), (512, 3, 8192, 24), (256, 5, 8192, 24), (512, 3, 16384, 24),
]


def autotune(walker: GpuWalker, candidates=None, sample_cases=1024, repeats=3, seed=1, cases=None,
             finalists=3):
    """Pick the fastest launch shape in two stages: every candidate is timed on
    a seeded random subset of `sample_cases` cases at their full photon counts
    (capping photons per case changes how long walks and chunk tails balance),
    then the `finalists` best are timed once more on the full case set."""
    cands = list(candidates or DEFAULT_CANDIDATES)
    cases = list(cases if cases is not None else walker.cases)
    if len(cases) > sample_cases:
        rng = np.random.default_rng(seed)
        subset = [cases[k] for k in sorted(rng.choice(len(cases), sample_cases, replace=False))]
    else:
        subset = cases
    walker.set_cases(subset)
    best, times = walker.autotune(cands, photons_cap=0, repeats=repeats, seed=seed)
    walker.set_cases(cases)
    result = dict(zip(cands, times.tolist()))
    if len(subset) < len(cases) and finalists > 1:
        order = np.argsort(times)[:finalists]
        fin = [cands[k] for k in order]
        best, ftimes = walker.autotune(fin, photons_cap=0, repeats=1, seed=seed)
        result.update({('full',) + tuple(c): t for c, t in zip(fin, ftimes.tolist())})
    walker.set_launch(*best)
    return best, result


def run(walker: GpuWalker, cases, seed=0, use_graph=True, tune=True, records=False,
        candidates=None, sample_cases=1024):
    """Walk all cases; returns dict(tallies, counts, brf, fractions, seconds,
    kernel_ms, launch, tune_times, steps, photons)."""
    walker.set_cases(cases)
    tune_times = None
    if tune:
        _, tune_times = autotune(walker, candidates, sample_cases=sample_cases, cases=cases)
    t0 = time.perf_counter()
    walker.walk(mode='native', seed=seed, records=records, use_graph=use_graph)
    t1 = time.perf_counter()
    st = walker.stats()
    t = walker.tallies()
    return dict(tallies=t, counts=walker.counts(), brf=brf_from_tallies(t),
                fractions=fractions_from_tallies(t), seconds=t1 - t0, kernel_ms=st['kernel_ms'],
                launch=walker.get_launch(), tune_times=tune_times, steps=st['steps'],
                photons=st['photons'])
