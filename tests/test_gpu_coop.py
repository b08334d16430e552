"""The cooperative tail mode (mcs_launch.coop_live): once a block's queue is
empty, waves with few live photons walk each photon with 64 / L lanes, one
step per lane, through a quaternion frame scan.  Launches with about one
chunk per wave put most photon-steps into that mode; their results must match
the CPU oracle and the one-lane walk at 3 sigma / chi^2, and the known
answers of van de Hulst."""
import numpy as np
import pytest
from scipy import stats

from oracle import coracle
from tests.gpu_cases import flat_params
from tests.test_gpu_native import EXT100, TAU_1M, brf_counts, gpu_case, two_sample_z

pytestmark = pytest.mark.gpu

U_STEPS, U_COOP = 8, 9


@pytest.fixture(scope='module')
def walker():
    from monte_carlompi_amd.gpu import GpuWalker
    w = GpuWalker(0)
    yield w
    w.close()


def coop_runs(walker, case_fn, seeds, coop_live, records=True):
    """One 64-photon chunk per wave (64-thread blocks, one per CU)."""
    n = 64 * walker.cus
    walker.set_cases([case_fn(n)])
    walker.set_launch(64, 1, 64, 1, coop_live)
    recs, counts = [], np.zeros(16, np.uint64)
    for s in seeds:
        walker.walk(mode='native', seed=s, records=records)
        counts += walker.counts()[0]
        if records:
            recs.append(walker.records())
    walker.set_launch()
    out = {k: np.concatenate([r[k] for r in recs]) for k in recs[0]} if records else {}
    return out, counts


def test_coop_dominated_config_case_matches_oracle(walker):
    """1 m snow, 60 deg, omega 0.999, g 0.89, Lambertian bottom 0.5."""
    case = lambda n: gpu_case(n, TAU_1M, 60.0, True, 0.5)
    print('coop config case: gpu', flush=True)
    g, cnt = coop_runs(walker, case, range(100, 112), 32)
    frac = float(cnt[U_COOP]) / float(cnt[U_STEPS])
    print('steps walked cooperatively: %.3f' % frac, flush=True)
    assert frac > 0.5
    print('coop config case: cpu oracle', flush=True)
    cpu = coracle.run_ranks(300, 8, 25_000,
                            lambda n: flat_params(n, 1.3, 0.89, 0.999, EXT100, TAU_1M, 300.0),
                            TAU_1M, 300.0, 60.0, True, 0.5)
    n_g, n_c = g['condition'].size, cpu['condition'].size
    for cond in (1, 2, 4):
        z = two_sample_z(int(np.sum(g['condition'] == cond)), n_g, int(np.sum(cpu['condition'] == cond)), n_c)
        assert abs(z) < 3.0, (cond, z)
    hg = brf_counts(g['theta_n'][g['condition'] == 1], g['phi_n'][g['condition'] == 1])
    hc = brf_counts(cpu['theta_n'][cpu['condition'] == 1], cpu['phi_n'][cpu['condition'] == 1])
    keep = (hg + hc) >= 10
    A, B = n_g, n_c
    chi2 = np.sum((np.sqrt(B / A) * hg[keep] - np.sqrt(A / B) * hc[keep]) ** 2 / (hg[keep] + hc[keep]))
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-3, chi2
    # step counts and path lengths: distributions, not only means
    bins = np.unique(np.round(np.geomspace(1, 2e4, 40)).astype(int))
    cg = np.histogram(g['n_scat'] + 1, bins)[0]
    cc = np.histogram(cpu['n_scat'] + 1, bins)[0]
    keep = (cg + cc) >= 10
    chi2 = np.sum((np.sqrt(B / A) * cg[keep] - np.sqrt(A / B) * cc[keep]) ** 2 / (cg[keep] + cc[keep]))
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-3, chi2
    refl_g = g['path_length'][g['condition'] == 1]
    refl_c = cpu['path_length'][cpu['condition'] == 1]
    assert stats.ks_2samp(refl_g, refl_c).pvalue > 1e-3
    assert int(cnt[U_STEPS]) == int(np.sum(g['n_scat'] + 1))


def test_coop_and_one_lane_walks_agree(walker):
    """Same launch shape, tail mode on (32) and off (0): exit fractions, BRF
    and step counts agree at 3 sigma / chi^2."""
    case = lambda n: gpu_case(n, 40.0, 30.0, False, 0.0, g=0.85, ssa=0.9995)
    a, ca = coop_runs(walker, case, range(10), 32)
    b, cb = coop_runs(walker, case, range(10, 20), 0)
    assert cb[U_COOP] == 0 and ca[U_COOP] > 0.5 * ca[U_STEPS]
    n_a, n_b = a['condition'].size, b['condition'].size
    for cond in (1, 2, 4):
        z = two_sample_z(int(np.sum(a['condition'] == cond)), n_a, int(np.sum(b['condition'] == cond)), n_b)
        assert abs(z) < 3.0, (cond, z)
    ha = brf_counts(a['theta_n'][a['condition'] == 1], a['phi_n'][a['condition'] == 1])
    hb = brf_counts(b['theta_n'][b['condition'] == 1], b['phi_n'][b['condition'] == 1])
    keep = (ha + hb) >= 10
    chi2 = np.sum((np.sqrt(n_b / n_a) * ha[keep] - np.sqrt(n_a / n_b) * hb[keep]) ** 2 / (ha[keep] + hb[keep]))
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-3, chi2
    assert stats.ks_2samp(a['n_scat'], b['n_scat']).pvalue > 1e-3
    # transmitted photons: exit cosines
    ta, tb = a['theta_n'][a['condition'] == 2], b['theta_n'][b['condition'] == 2]
    assert stats.ks_2samp(ta, tb).pvalue > 1e-3


def test_coop_van_de_hulst_known_answer(walker):
    """tau 2, omega 0.9, g 0.75, mu0 1, free bottom: R 0.09739, T 0.66096."""
    case = lambda n: gpu_case(n, 2.0, 0.0, False, 0.0, wvl=0.5, g=0.75, ssa=0.9)
    _, c = coop_runs(walker, case, range(40), 32, records=False)
    n = float(c[0])
    R, T = c[1] / n, (c[2] + c[3]) / n
    assert c[U_COOP] > 0.2 * c[U_STEPS]
    assert abs(R - 0.09739) < 3 * np.sqrt(0.09739 * (1 - 0.09739) / n) + 2e-5
    assert abs(T - 0.66096) < 3 * np.sqrt(0.66096 * (1 - 0.66096) / n) + 2e-5


def test_coop_step_cap(walker):
    """A photon capped inside a cooperative segment stops at exactly max_steps."""
    from monte_carlompi_amd._lib import McsError
    case = gpu_case(64 * walker.cus, 1e4, 0.0, False, 0.0, ssa=1.0)
    walker.set_cases([case])
    walker.set_launch(64, 1, 64, 1, 32)
    try:
        with pytest.raises(McsError, match='STEP_CAP'):
            walker.walk(mode='native', seed=9, records=True, max_steps=300)
        r = walker.records()
        c = walker.counts()[0]
        assert np.all(r['n_scat'] <= 299)
        assert int(np.sum(r['condition'] == 0)) == int(c[6]) > 0
        assert np.all(r['n_scat'][r['condition'] == 0] == 299)
        assert c[U_COOP] > 0
    finally:
        walker.set_launch()


U_MOVED = 10


def handover_runs(walker, case_fn, seeds, coop_live, records=True):
    """One 1024-photon chunk per 16-wave block: the queue is empty at once, and
    waves that finish early take over photons of their siblings' cooperative
    tails."""
    n = 1024 * walker.cus
    walker.set_cases([case_fn(n)])
    walker.set_launch(1024, 1, 1024, 16, coop_live)
    recs, counts = [], np.zeros(16, np.uint64)
    try:
        for s in seeds:
            walker.walk(mode='native', seed=s, records=records)
            counts += walker.counts()[0]
            if records:
                recs.append(walker.records())
    finally:
        walker.set_launch()
    out = {k: np.concatenate([r[k] for r in recs]) for k in recs[0]} if records else {}
    return out, counts


def test_hand_over_between_waves_keeps_the_law(walker):
    """Photons handed from one wave's cooperative tail to an idle sibling keep
    their walk: every photon is recorded once, and exit fractions, BRF and step
    counts agree with the one-lane walk at 3 sigma / chi^2."""
    case = lambda n: gpu_case(n, 40.0, 30.0, False, 0.0, g=0.85, ssa=0.9995)
    a, ca = handover_runs(walker, case, range(3), 32)
    b, cb = handover_runs(walker, case, range(3, 6), 0)
    print('handed over: %d of %d photons' % (int(ca[U_MOVED]), int(ca[0])), flush=True)
    assert ca[U_MOVED] > 0 and cb[U_MOVED] == 0
    assert int(ca[U_STEPS]) == int(np.sum(a['n_scat'] + 1))
    assert np.all(a['condition'] >= 1)   # no photon lost or left unwritten
    n_a, n_b = a['condition'].size, b['condition'].size
    for cond in (1, 2, 4):
        z = two_sample_z(int(np.sum(a['condition'] == cond)), n_a, int(np.sum(b['condition'] == cond)), n_b)
        assert abs(z) < 3.0, (cond, z)
    ha = brf_counts(a['theta_n'][a['condition'] == 1], a['phi_n'][a['condition'] == 1])
    hb = brf_counts(b['theta_n'][b['condition'] == 1], b['phi_n'][b['condition'] == 1])
    keep = (ha + hb) >= 10
    chi2 = np.sum((np.sqrt(n_b / n_a) * ha[keep] - np.sqrt(n_a / n_b) * hb[keep]) ** 2 / (ha[keep] + hb[keep]))
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-3, chi2
    assert stats.ks_2samp(a['n_scat'], b['n_scat']).pvalue > 1e-3
    assert stats.ks_2samp(a['path_length'][a['condition'] == 1],
                          b['path_length'][b['condition'] == 1]).pvalue > 1e-3


def test_hand_over_with_lambertian_bottom_and_step_cap(walker):
    """Handed-over photons that bounce off a Lambertian bottom or reach
    max_steps stop exactly there, in their own chunk's records."""
    from monte_carlompi_amd._lib import McsError
    case = lambda n: gpu_case(n, 30.0, 0.0, True, 0.9, g=0.85, ssa=1.0)
    walker.set_cases([case(1024 * walker.cus)])
    walker.set_launch(1024, 1, 1024, 16, 32)
    try:
        with pytest.raises(McsError, match='STEP_CAP'):
            walker.walk(mode='native', seed=21, records=True, max_steps=600)
        r = walker.records()
        c = walker.counts()[0]
        assert c[U_MOVED] > 0
        assert np.all(r['n_scat'] <= 599)
        assert int(np.sum(r['condition'] == 0)) == int(c[6]) > 0
        assert np.all(r['n_scat'][r['condition'] == 0] == 599)
        assert int(c[U_STEPS]) == int(np.sum(r['n_scat'] + 1))
        assert set(np.unique(r['condition']).tolist()) <= {0, 1, 2, 3}   # no absorption at omega 1
    finally:
        walker.set_launch()


def test_coop_isotropic_scattering(walker):
    """g = 0 bands (isotropic, mc3D:790-800's g == 0 branch) in the cooperative
    mode: exit fractions, BRF and step counts agree with the one-lane walk, and
    the one-lane walk with the C oracle.  (A forward-peaked phase function in
    place of the isotropic one moves the exit fractions by tens of sigma.)"""
    tau = 20.0
    case = lambda n: gpu_case(n, tau, 30.0, False, 0.0, g=0.0, ssa=0.999)
    a, ca = coop_runs(walker, case, range(6), 32)
    b, cb = coop_runs(walker, case, range(6, 12), 0)
    assert ca[U_COOP] > 0.3 * ca[U_STEPS] and cb[U_COOP] == 0
    n_a, n_b = a['condition'].size, b['condition'].size
    for cond in (1, 2, 4):
        z = two_sample_z(int(np.sum(a['condition'] == cond)), n_a, int(np.sum(b['condition'] == cond)), n_b)
        assert abs(z) < 3.0, (cond, z)
    ha = brf_counts(a['theta_n'][a['condition'] == 1], a['phi_n'][a['condition'] == 1])
    hb = brf_counts(b['theta_n'][b['condition'] == 1], b['phi_n'][b['condition'] == 1])
    keep = (ha + hb) >= 10
    chi2 = np.sum((np.sqrt(n_b / n_a) * ha[keep] - np.sqrt(n_a / n_b) * hb[keep]) ** 2 / (ha[keep] + hb[keep]))
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-3, chi2
    assert stats.ks_2samp(a['n_scat'], b['n_scat']).pvalue > 1e-3
    cpu = coracle.run_ranks(77, 8, 5_000, lambda n: flat_params(n, 1.3, 0.0, 0.999, EXT100, tau, 300.0),
                            tau, 300.0, 30.0, False, 0.0)
    n_c = cpu['condition'].size
    for cond in (1, 2, 4):
        z = two_sample_z(int(np.sum(b['condition'] == cond)), n_b, int(np.sum(cpu['condition'] == cond)), n_c)
        assert abs(z) < 3.0, (cond, z)


def test_hand_over_across_cases(walker):
    """Four cases in one launch (thick and thin layers, an isotropic band,
    impurity absorption, a two-band Lambertian-bottom case), 16-wave blocks
    with about one chunk each: photons handed between waves find their own
    case's chunk among slots of other cases.  Per case every photon is
    recorded once, and the condition fractions agree with the one-lane walk
    (3.5 sigma: 15 comparisons)."""
    from monte_carlompi_amd.gpu import Band, Case
    n = 256 * walker.cus // 2
    cases = [
        Case(tau_tot=30.0, theta0_deg=0.0, n_photons=n,
             bands=[Band(wvl=1.0, g=0.85, ssa_ice=0.9995, ext_cff=5000.0)]),
        Case(tau_tot=15.0, theta0_deg=45.0, n_photons=n,
             bands=[Band(wvl=1.1, g=0.0, ssa_ice=0.999, ext_cff=5000.0)]),
        Case(tau_tot=40.0, theta0_deg=60.0, n_photons=n,
             bands=[Band(wvl=1.2, g=0.89, ssa_ice=0.9999, ext_cff=5000.0, ssa_imp=0.5, p_ext_imp=0.002)]),
        Case(tau_tot=20.0, theta0_deg=30.0, n_photons=n, lambertian_bottom=True, r_lambertian=0.7,
             bands=[Band(wvl=1.3, g=0.8, ssa_ice=0.999, ext_cff=5000.0, weight=0.5),
                    Band(wvl=1.31, g=0.75, ssa_ice=0.998, ext_cff=5000.0, weight=0.5)]),
    ]
    walker.set_cases(cases)
    runs = {}
    try:
        for coop in (32, 0):
            walker.set_launch(1024, 1, 2048, 16, coop)
            recs, cnt = [], 0
            for s in range(3):
                walker.walk(mode='native', seed=300 + 10 * coop + s, records=True)
                recs.append(walker.records())
                cnt = cnt + walker.counts()
            runs[coop] = ({k: np.concatenate([r[k] for r in recs]) for k in recs[0]}, cnt)
    finally:
        walker.set_launch()
    (a, ca), (b, cb) = runs[32], runs[0]
    print('handed over per case:', ca[:, U_MOVED], flush=True)
    assert int(ca[:, U_MOVED].sum()) > 0 and int(cb[:, U_MOVED].sum()) == 0
    assert np.all(ca[:, 0] == 3 * n) and np.all(cb[:, 0] == 3 * n)
    per_run = 4 * n
    for k in range(4):
        sel = np.concatenate([np.arange(r * per_run + k * n, r * per_run + (k + 1) * n) for r in range(3)])
        ka, kb = a['condition'][sel], b['condition'][sel]
        assert np.all(ka >= 1)
        for cond in (1, 2, 4, 5) if k == 2 else (1, 2, 4):
            z = two_sample_z(int(np.sum(ka == cond)), ka.size, int(np.sum(kb == cond)), kb.size)
            assert abs(z) < 3.5, (k, cond, z)
        assert int(ca[k, U_STEPS]) == int(np.sum(a['n_scat'][sel] + 1))


@pytest.mark.parametrize('shape', [(1024, 1, 977, 8, 32), (640, 2, 977, 8, 32), (256, 2, 300, 4, 16)])
def test_small_chunks_reach_the_cooperative_tail(walker, shape):
    """Small chunks (a walk raises them to the block's threads) can fill all
    four tally slots with walking chunks; such a block can acquire nothing,
    so it must learn from the global counter that the queue ran dry -- else
    its long walks finish one lane each (5 ms instead of 0.8 ms for the
    1e6-photon case).  Checked through the cooperative photon-steps (count
    slot 9) and the law."""
    from monte_carlompi_amd.workloads import snow_1m_case
    walker.set_cases([snow_1m_case(400_000)])
    walker.set_launch(*shape)
    try:
        walker.walk(mode='native', seed=17, records=True)
        c = walker.counts()[0]
        assert int(c[0]) == 400_000
        assert float(c[U_COOP]) / float(c[U_STEPS]) > 0.1, c
        r = walker.records()
        walker.set_launch(1024, 1, 2048, 16, 0)
        walker.walk(mode='native', seed=18, records=True)
        r0 = walker.records()
        for cond in (1, 2, 4):
            z = two_sample_z(np.sum(r['condition'] == cond), r['condition'].size,
                             np.sum(r0['condition'] == cond), r0['condition'].size)
            assert abs(z) < 4, (cond, z)
    finally:
        walker.set_launch()


def test_chunks_under_the_block_do_not_starve_it(walker):
    """A launch whose chunks are much smaller than its blocks (489 photons
    for 1024 threads) walks with chunks of the block's threads: the
    1e6-photon case takes about as long as at the benchmark's chunk (with
    489-photon chunks its blocks waited for tally slots: ~8 x longer), and
    the records keep the law."""
    from monte_carlompi_amd.workloads import snow_1m_case
    walker.set_cases([snow_1m_case(1_000_000)])
    try:
        t = {}
        recs = {}
        for shape in [(1024, 1, 1954, 16, 32), (1024, 1, 489, 16, 32)]:
            walker.set_launch(*shape)
            for k in range(4):
                walker.walk(mode='native', seed=40 + k, records=True)
            t[shape[2]] = walker.stats()['kernel_ms']
            recs[shape[2]] = walker.records()['condition']
            assert int(walker.counts()[0][0]) == 1_000_000
        print('kernel ms by chunk', t, flush=True)
        assert t[489] < 2.5 * t[1954], t
        a, b = recs[489], recs[1954]
        for cond in (1, 2, 4):
            z = two_sample_z(int(np.sum(a == cond)), a.size, int(np.sum(b == cond)), b.size)
            assert abs(z) < 4, (cond, z)
    finally:
        walker.set_launch()
