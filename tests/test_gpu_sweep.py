"""Batched multi-case launch: per-case results equal single-case launches
(integer counts exactly), independent of the autotuned launch shape, stable
over hipGraph replays; config-3-shaped sweep statistics against the C oracle."""
import numpy as np
import pytest

from monte_carlompi_amd import sweep
from monte_carlompi_amd.workloads import sweep_cases
from oracle import coracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def walker():
    from monte_carlompi_amd.gpu import GpuWalker
    w = GpuWalker(0)
    yield w
    w.close()


def small_sweep(scale=0.02):
    cases, meta = sweep_cases(seed=3, n_reff=4, n_wvl=2, n_theta=3, photon_scale=scale)
    return cases, meta


def test_batched_counts_are_consistent(walker):
    cases, _ = small_sweep()
    walker.set_cases(cases)
    walker.walk(mode='native', seed=21, records=True)
    cnt = walker.counts()
    rec = walker.records()
    n = np.array([c.n_photons for c in cases])
    assert np.all(cnt[:, 0] == n)
    assert np.all(cnt[:, 1:6].sum(axis=1) == cnt[:, 0])
    # per-case counts equal the records of that case's photon range
    starts = np.concatenate([[0], np.cumsum(n)])
    for k in range(len(cases)):
        sl = slice(starts[k], starts[k + 1])
        assert np.all(np.bincount(rec['condition'][sl], minlength=6)[1:6] == cnt[k, 1:6])
        assert int(np.sum(rec['n_scat'][sl] + 1)) == int(cnt[k, 8])
        assert rec['band'][sl].max() < len(cases[k].bands)


def test_record_columns_pinned_and_pageable_agree(walker):
    """Every record column read into pinned buffers (one DMA per column)
    equals the same read into ordinary numpy arrays (staged copies), and the
    wvn / snow_depth columns are the photon's band values: 1 / wvl and
    tau_tot / ext_cff of its case (mc3D:1489-1490, 1612)."""
    cases, _ = small_sweep()
    walker.set_cases(cases)
    walker.walk(mode='native', seed=33, records=True)
    a = walker.records()
    buf = walker.record_buffers()
    b = walker.records(out=buf)
    for k, _ in walker.RECORD_DTYPES:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    n = np.array([c.n_photons for c in cases])
    starts = np.concatenate([[0], np.cumsum(n)])
    for k, c in enumerate(cases):
        sl = slice(starts[k], starts[k + 1])
        wv = np.array([1.0 / x.wvl for x in c.bands])
        dp = np.array([c.tau_tot / x.ext_cff for x in c.bands])
        np.testing.assert_array_equal(a['wvn'][sl], wv[a['band'][sl]])
        np.testing.assert_array_equal(a['snow_depth'][sl], dp[a['band'][sl]])
    # a second walk into the same pinned buffers overwrites every entry
    walker.walk(mode='native', seed=34, records=True)
    c2 = walker.records(out=buf)
    d2 = walker.records()
    for k, _ in walker.RECORD_DTYPES:
        np.testing.assert_array_equal(c2[k], d2[k], err_msg=k)


def test_first_case_of_batch_equals_lone_launch(walker):
    """Philox streams are keyed by (case, photon): case 0 of a batch and the
    same case launched alone walk identical photons (tail mode off, so that
    the walks do not depend on scheduling)."""
    cases, _ = small_sweep()
    walker.set_launch(coop_live=0)
    walker.set_cases(cases)
    walker.walk(mode='native', seed=5, records=True)
    rb = walker.records()
    n0 = cases[0].n_photons
    walker.set_cases([cases[0]])
    walker.walk(mode='native', seed=5, records=True)
    r1 = walker.records()
    walker.set_launch()
    for f in ('condition', 'n_scat', 'theta_n', 'phi_n', 'path_length', 'band'):
        np.testing.assert_array_equal(rb[f][:n0], r1[f], err_msg=f)


def test_autotune_and_graph_replays_are_invariant(walker):
    """Tail mode off: counts exact and tallies to FP64 summation order over
    autotuned / fixed shapes and graph replays."""
    cases, _ = small_sweep()
    cands = [tuple(c)[:4] + (0,) for c in sweep.DEFAULT_CANDIDATES]
    res = sweep.run(walker, cases, seed=9, use_graph=True, tune=True, sample_cases=8, candidates=cands)
    assert res['launch'] in cands
    ref_counts = res['counts'].copy()
    ref_t = res['tallies'].copy()
    for shape in [(64, 8, 64, 1, 0), (1024, 1, 16384, 64, 0)]:
        walker.set_launch(*shape)
        for _ in range(3):                       # graph replays with host work between
            walker.walk(mode='native', seed=9, records=False, use_graph=True)
            np.testing.assert_array_equal(walker.counts(), ref_counts)
            np.testing.assert_allclose(walker.tallies(), ref_t, rtol=1e-12, atol=1e-300)
            _ = np.linalg.norm(np.random.rand(100, 100))
        walker.walk(mode='native', seed=9, records=False, use_graph=False)
        np.testing.assert_array_equal(walker.counts(), ref_counts)
    walker.set_launch()


def test_sweep_cases_statistics_against_oracle(walker):
    """A handful of config-3 cases (thin and thick, different omega, g,
    theta0, multi-band): reflectance / transmittance / absorption of each case
    within 3 sigma of the C oracle."""
    cases, meta = sweep_cases(seed=11, n_reff=3, n_wvl=2, n_theta=2)
    pick = [0, 5, 9, 14, 19, 23]
    cases = [cases[k] for k in pick]
    for c in cases:
        c.n_photons = 200_000
    walker.set_cases(cases)
    walker.walk(mode='native', seed=77, records=False)
    cnt = walker.counts().astype(float)
    for k, c in enumerate(cases):
        b = c.bands[0]
        par = lambda n: dict(wvl=np.full(n, b.wvl), g=np.full(n, b.g), ssa_ice=np.full(n, b.ssa_ice),
                             ssa_imp=np.zeros(n), ext_cff_mss=np.full(n, b.ext_cff / 300.0),
                             p_ext_imp=np.zeros(n), snow_depth=np.full(n, c.tau_tot / b.ext_cff))
        cpu = coracle.run_ranks(1000 + k, 4, 25_000, par, c.tau_tot, 300.0, c.theta0_deg, False, 0.0)
        m = cpu['condition'].size
        for cond in (1, 2, 3, 4):
            pg, pc = cnt[k, cond] / c.n_photons, np.mean(cpu['condition'] == cond)
            s = np.sqrt(pg * (1 - pg) / c.n_photons + pc * (1 - pc) / m)
            assert abs(pg - pc) <= 3 * s + 1e-9, (k, cond, pg, pc, meta['tau'][pick[k]])


def test_empty_and_tiny_cases_in_a_batch(walker):
    """Cases with no photons, one photon and a partial last chunk, in one launch
    with records: the chunk-to-case search of the walk and of the record
    ordering skips the empty cases, every photon is recorded once in its own
    case's id range, and per-case counts match the records."""
    from monte_carlompi_amd.gpu import Band, Case
    band = lambda: [Band(wvl=1.0, g=0.85, ssa_ice=0.99, ext_cff=5000.0)]
    sizes = [0, 1, 0, 5000, 0, 0, 2049, 64, 0]
    walker.set_cases([Case(tau_tot=3.0 + k, theta0_deg=20.0, n_photons=n, bands=band())
                      for k, n in enumerate(sizes)])
    try:
        for shape in [(256, 4, 2048, 8, 16), (1024, 1, 2048, 16, 32), (64, 2, 64, 1, 0)]:
            walker.set_launch(*shape)
            walker.walk(mode='native', seed=3, records=True)
            r = walker.records()
            c = walker.counts()
            assert r['condition'].size == sum(sizes)
            assert np.all(r['condition'] >= 1)
            off = 0
            for k, n in enumerate(sizes):
                assert int(c[k, 0]) == n, (shape, k)
                seg = r['condition'][off:off + n]
                for cond in range(1, 6):
                    assert int(c[k, cond]) == int(np.sum(seg == cond)), (shape, k, cond)
                assert int(c[k, 8]) == int(np.sum(r['n_scat'][off:off + n] + 1)), (shape, k)
                off += n
    finally:
        walker.set_launch()
