"""Native mode on the GPU (FP32, Philox4x32-10) against the CPU oracle and
closed forms: exit fractions within 3 sigma (two-sample binomial), the
18 x 36 BRF histogram by chi^2 and per-bin sigma, known answers, and the
bookkeeping invariants of the tallies."""
import numpy as np
import pytest
from scipy import stats

from oracle import coracle, walk
from tests.gpu_cases import case_from_params, flat_params

pytestmark = pytest.mark.gpu

EXT100 = walk.sphere_ext_cff_mss(100.0, 917.0)
TAU_1M = 1.0 * EXT100 * 300.0            # 1 m of snow, r_eff 100 um, Q_ext 2


@pytest.fixture(scope='module')
def walker():
    from monte_carlompi_amd.gpu import GpuWalker
    w = GpuWalker(0)
    yield w
    w.close()


def gpu_case(n, tau, theta0, lb, R, wvl=1.3, g=0.89, ssa=0.999):
    par = flat_params(n, wvl, g, ssa, EXT100, tau, 300.0)
    return case_from_params(par, tau, 300.0, theta0, lb, R)[0]


def two_sample_z(k1, n1, k2, n2):
    p = (k1 + k2) / (n1 + n2)
    s = np.sqrt(p * (1 - p) * (1 / n1 + 1 / n2))
    return (k1 / n1 - k2 / n2) / s if s > 0 else 0.0


def brf_counts(theta, phi):
    jt = np.clip((theta * (18 / (np.pi / 2))).astype(int), 0, 17)
    kp = np.clip((phi * (36 / (2 * np.pi))).astype(int), 0, 35)
    return np.bincount(jt * 36 + kp, minlength=648)


def test_config_case_matches_cpu_oracle(walker):
    """1 m snow, theta0 60 deg, omega 0.999, g 0.89, Lambertian bottom R 0.5:
    GPU 1e6 photons vs C oracle 2e5 photons (numpy MT19937 streams)."""
    n_gpu = 1_000_000
    walker.set_cases([gpu_case(n_gpu, TAU_1M, 60.0, True, 0.5)])
    print('native 1e6 gpu', flush=True)
    walker.walk(mode='native', seed=2024, records=True)
    g = walker.records()
    cnt = walker.counts()[0]
    print('native 2e5 cpu oracle', flush=True)
    cpu = coracle.run_ranks(100, 8, 25_000,
                            lambda n: flat_params(n, 1.3, 0.89, 0.999, EXT100, TAU_1M, 300.0),
                            TAU_1M, 300.0, 60.0, True, 0.5)
    n_cpu = cpu['condition'].size
    for cond in (1, 2, 3, 4):
        kg = int(np.sum(g['condition'] == cond))
        kc = int(np.sum(cpu['condition'] == cond))
        assert kg == int(cnt[cond])
        z = two_sample_z(kg, n_gpu, kc, n_cpu)
        assert abs(z) < 3.0, (cond, kg / n_gpu, kc / n_cpu, z)
    # BRF histogram of reflected photons: chi^2 two-sample + per-bin sigma
    hg = brf_counts(g['theta_n'][g['condition'] == 1], g['phi_n'][g['condition'] == 1])
    hc = brf_counts(cpu['theta_n'][cpu['condition'] == 1], cpu['phi_n'][cpu['condition'] == 1])
    A, B = n_gpu, n_cpu
    keep = (hg + hc) >= 10
    chi2 = np.sum((np.sqrt(B / A) * hg[keep] - np.sqrt(A / B) * hc[keep]) ** 2 / (hg[keep] + hc[keep]))
    p = stats.chi2.sf(chi2, keep.sum() - 1)
    assert p > 1e-3, (chi2, keep.sum(), p)
    pg, pc = hg / A, hc / B
    sig = np.sqrt(pg * (1 - pg) / A + pc * (1 - pc) / B)
    zb = np.where(sig > 0, (pg - pc) / np.where(sig > 0, sig, 1), 0)
    assert np.mean(np.abs(zb) > 3) < 0.01 and np.max(np.abs(zb)) < 4.5
    # mean step count and path length agree
    sg, sc = g['n_scat'].astype(float), cpu['n_scat'].astype(float)
    zs = (sg.mean() - sc.mean()) / np.sqrt(sg.var() / sg.size + sc.var() / sc.size)
    assert abs(zs) < 3.0
    assert int(cnt[8]) == int(np.sum(g['n_scat'] + 1))


def test_van_de_hulst_known_answer(walker):
    """tau 2, omega 0.9, g 0.75, mu0 1, free bottom: R = 0.09739,
    T_total = 0.66096 (van de Hulst 1980 / Wang et al. 1995 Table 1, the values
    quoted in mc3D:1849-1852); direct part exp(-2)."""
    n = 2_000_000
    walker.set_cases([gpu_case(n, 2.0, 0.0, False, 0.0, wvl=0.5, g=0.75, ssa=0.9)])
    walker.walk(mode='native', seed=77, records=False)
    c = walker.counts()[0].astype(float)
    R, T = c[1] / n, (c[2] + c[3]) / n
    Td = c[3] / n
    assert abs(R - 0.09739) < 3 * np.sqrt(0.09739 * (1 - 0.09739) / n) + 2e-5
    assert abs(T - 0.66096) < 3 * np.sqrt(0.66096 * (1 - 0.66096) / n) + 2e-5
    assert abs(Td - np.exp(-2.0)) < 3 * np.sqrt(np.exp(-2.0) * (1 - np.exp(-2.0)) / n)


def test_direct_transmittance_with_lambertian_bottom(walker):
    """Condition 3 needs an unscattered first flight through the slab and no
    bounce: P = exp(-tau / mu0) (1 - R)."""
    n = 1_000_000
    tau, th, R = 0.7, 40.0, 0.3
    walker.set_cases([gpu_case(n, tau, th, True, R, g=0.8, ssa=0.95)])
    walker.walk(mode='native', seed=5, records=False)
    c = walker.counts()[0].astype(float)
    p = np.exp(-tau / np.cos(np.radians(th))) * (1 - R)
    assert abs(c[3] / n - p) < 3 * np.sqrt(p * (1 - p) / n)


def test_limits_omega_zero_and_one(walker):
    from monte_carlompi_amd.gpu import Band, Case
    n = 200_000
    # omega = 0, free bottom: every photon ends at step 1 (condition 3 or 4)
    walker.set_cases([gpu_case(n, 1.0, 30.0, False, 0.0, ssa=0.0)])
    walker.walk(mode='native', seed=3, records=True)
    r = walker.records()
    assert set(np.unique(r['condition'])) <= {3, 4}
    assert np.all(r['n_scat'] == 0)
    p3 = np.exp(-1.0 / np.cos(np.radians(30.0)))
    assert abs(np.mean(r['condition'] == 3) - p3) < 3 * np.sqrt(p3 * (1 - p3) / n)
    # omega = 1 over a fully reflecting bottom: everything comes back out
    walker.set_cases([gpu_case(n, 3.0, 30.0, True, 1.0, ssa=1.0)])
    walker.walk(mode='native', seed=4, records=True)
    r = walker.records()
    assert np.all(r['condition'] == 1)


def test_tally_bookkeeping(walker):
    """Tallies equal sums over the records; counts add up."""
    from monte_carlompi_amd import gpu
    n = 300_000
    walker.set_cases([gpu_case(n, 5.0, 50.0, True, 0.4, g=0.85, ssa=0.99)])
    walker.walk(mode='native', seed=11, records=True)
    r = walker.records()
    t = walker.tallies()[0]
    u = walker.counts()[0]
    wvn = r['wvn']
    assert u[0] == n and sum(int(u[c]) for c in range(1, 6)) == n
    np.testing.assert_allclose(t[gpu.T_QDOWN], wvn.sum(), rtol=1e-12)
    for c in range(1, 6):
        np.testing.assert_allclose(t[gpu.T_QCOND + c - 1], wvn[r['condition'] == c].sum(), rtol=1e-12)
    refl = r['condition'] == 1
    h = brf_counts(r['theta_n'][refl], r['phi_n'][refl]) * wvn[0]
    np.testing.assert_allclose(t[:648], h, rtol=1e-12)
    assert int(u[8]) == int(np.sum(r['n_scat'] + 1))
    assert np.all(r['theta_n'][refl] < np.pi / 2) and np.all(r['phi_n'] >= 0) and np.all(r['phi_n'] < 2 * np.pi)


def test_records_independent_of_launch_shape_and_graph(walker):
    """With the cooperative tail mode off (coop_live 0) every photon's walk is
    a function of (seed, case, photon) alone: records are bit-identical over
    launch shapes and graph launches.  (With it on, which photons finish
    cooperatively depends on scheduling; the statistics are the same.)"""
    n = 100_000
    walker.set_cases([gpu_case(n, 20.0, 30.0, True, 0.5, g=0.87, ssa=0.995)])
    ref = None
    # 4, 2, 4, 4, 5 and 6 waves per SIMD: every register-budget instance
    shapes = [(256, 4, 2048, 8, 0), (64, 8, 64, 1, 0), (512, 2, 8192, 32, 0), (1024, 1, 4096, 64, 0),
              (256, 5, 2048, 8, 0), (256, 6, 1024, 8, 0)]
    for shape in shapes:
        walker.set_launch(*shape)
        for graph in (False, True):
            walker.walk(mode='native', seed=99, records=True, use_graph=graph)
            r = walker.records()
            u = walker.counts()[0]
            if ref is None:
                ref, uref = r, u
            else:
                for f in ('condition', 'n_scat', 'theta_n', 'phi_n', 'path_length'):
                    np.testing.assert_array_equal(r[f], ref[f], err_msg='%s %s' % (shape, f))
                np.testing.assert_array_equal(u, uref)
    walker.set_launch()


def test_impurity_two_species_against_oracle(walker):
    """Impurity share of extinction P_ext_imp with its own albedo: exit,
    ice-absorption (4) and impurity-absorption (5) fractions within 3 sigma of
    the C oracle (mc3D:1374-1383, 1461-1466)."""
    from monte_carlompi_amd.gpu import Band, Case
    n = 1_000_000
    imp_cnc, ext_imp = 2e-4, 12000.0
    ext_mix = EXT100 * (1 - imp_cnc) + ext_imp * imp_cnc
    p_imp = imp_cnc * ext_imp / (imp_cnc * ext_imp + (1 - imp_cnc) * EXT100)
    tau = 1.0 * ext_mix * 300.0 * 0.01     # 1 cm of dirty snow
    band = Band(wvl=1.0, g=0.86, ssa_ice=0.9995, ext_cff=ext_mix * 300.0, ssa_imp=0.3, p_ext_imp=p_imp)
    walker.set_cases([Case(tau_tot=tau, theta0_deg=30.0, n_photons=n, bands=[band],
                           lambertian_bottom=True, r_lambertian=0.4)])
    walker.walk(mode='native', seed=8, records=False)
    c = walker.counts()[0].astype(float)
    par = lambda k: dict(wvl=np.full(k, 1.0), g=np.full(k, 0.86), ssa_ice=np.full(k, 0.9995),
                         ssa_imp=np.full(k, 0.3), ext_cff_mss=np.full(k, ext_mix),
                         p_ext_imp=np.full(k, p_imp), snow_depth=np.full(k, 0.01))
    cpu = coracle.run_ranks(300, 8, 25_000, par, tau, 300.0, 30.0, True, 0.4)
    m = cpu['condition'].size
    assert c[5] > 0.05 * n
    for cond in (1, 2, 3, 4, 5):
        z = two_sample_z(c[cond], n, np.sum(cpu['condition'] == cond), m)
        assert abs(z) < 3.0, (cond, c[cond] / n, np.mean(cpu['condition'] == cond), z)


def test_lambertian_surface_mode(walker):
    """A Lambertian reflector of reflectance R: albedo R, absorption 1 - R at
    the surface, exit cosines with pdf 2 mu (E[mu] = 2/3), n_scat = 1."""
    from monte_carlompi_amd.gpu import Band, Case
    n, R = 1_000_000, 0.37
    walker.set_cases([Case(tau_tot=5.0, theta0_deg=20.0, n_photons=n, lambertian_surface=True,
                           r_lambertian=R, bands=[Band(wvl=1.0, g=0.85, ssa_ice=0.99, ext_cff=5000.0)])])
    walker.walk(mode='native', seed=12, records=True)
    r = walker.records()
    refl = r['condition'] == 1
    assert set(np.unique(r['condition'])) <= {1, 4}
    assert abs(refl.mean() - R) < 3 * np.sqrt(R * (1 - R) / n)
    mu = np.cos(r['theta_n'][refl])
    assert abs(mu.mean() - 2 / 3) < 3 * np.sqrt((0.5 - 4 / 9) / refl.sum())
    assert np.all(r['n_scat'][refl] == 1) and np.all(r['n_scat'][~refl] == 0)
    assert np.all(np.abs(r['path_length'][refl]) < 1e-6)


def test_reference_known_answer_through_drop_in(tmp_path, monkeypatch):
    """The reference's test() (mc3D:1849-1866): tau 2, omega 0.9, g 0.75,
    normal incidence, set through MonteCarlo(test=True) overrides; with the
    free bottom its docstring values R = 0.09739, T = 0.66096 hold."""
    from monte_carlompi_amd.monte_carlo3D import MonteCarlo
    monkeypatch.setattr('sys.argv', ['prog'])
    mc = MonteCarlo(tau_tot=2.0, imp_cnc=0, output_dir=str(tmp_path), optics_dir=str(tmp_path / 'none'),
                    ssa_ice=0.5, g=0.5, seed=21)
    mc.ssa_ice = 0.9
    mc.g = 0.75
    n = 1_000_000
    mc.run(n, 0.5, 1e-12, 100, test=True, Lambertian_bottom=False, write=False)
    rec = mc.records
    R = np.mean(rec['condition'] == 1)
    T = np.mean((rec['condition'] == 2) | (rec['condition'] == 3))
    assert abs(R - 0.09739) < 3 * np.sqrt(0.09739 * 0.90261 / n) + 2e-5
    assert abs(T - 0.66096) < 3 * np.sqrt(0.66096 * 0.33904 / n) + 2e-5
    assert abs(mc.calculate_albedo(rec) - R) < 1e-12
    mc.close()



def test_native_band_sampling_follows_the_weights(walker):
    """Native mode draws each photon's wavelength bin from its case's band
    weights (an alias table over the cdf column): band frequencies in the
    records match the weights (chi^2), a zero-weight band is never drawn, and
    two cases with different band sets keep their own distributions."""
    from monte_carlompi_amd.gpu import Band, Case
    wa = np.array([3.0, 0.0, 1.0, 7.0, 2.0, 0.5, 4.0, 9.0, 1.5, 2.0])
    wb = np.array([0.2, 0.8])
    mk = lambda w, wvl0: [Band(wvl=wvl0 + 0.01 * j, g=0.85, ssa_ice=0.99, ext_cff=5000.0, weight=float(x))
                          for j, x in enumerate(w)]
    n = 400_000
    walker.set_cases([Case(tau_tot=2.0, theta0_deg=30.0, n_photons=n, bands=mk(wa, 1.0)),
                      Case(tau_tot=2.0, theta0_deg=30.0, n_photons=n, bands=mk(wb, 1.5))])
    walker.walk(mode='native', seed=5, records=True)
    r = walker.records()
    for k, w in enumerate((wa, wb)):
        b = r['band'][k * n:(k + 1) * n].astype(int)
        cnt = np.bincount(b, minlength=w.size)
        assert cnt.size == w.size
        assert np.all(cnt[w == 0] == 0)
        p = w / w.sum()
        keep = p > 0
        chi2 = np.sum((cnt[keep] - n * p[keep]) ** 2 / (n * p[keep]))
        assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-3, (k, chi2, cnt)
