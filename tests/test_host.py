"""Host side of the drop-in (no GPU): record file format and naming, config
reading, Parallel order, optics interpolation, BRF/albedo estimators."""
import os

import numpy as np
import pytest

from monte_carlompi_amd import optics, post_processing
from monte_carlompi_amd.monte_carlo3D import MonteCarlo
from monte_carlompi_amd.parallelize import Parallel
from oracle import walk
from tests.golden_cases import load

META, CASES = load()


def test_record_file_format_roundtrip(tmp_path):
    g = [c for c in CASES if c['name'] == 'thin_lamb'][0]
    rec = dict(g['records'])
    path = tmp_path / 'out.txt'
    MonteCarlo.write_records(str(path), rec)
    lines = path.read_text().splitlines()
    assert lines[0] == 'condition wvn[um^-1] theta_n phi_n n_scat path_length[m], snow_depth[m]'
    assert len(lines) == 1 + g['n']
    f = lines[1].split()
    assert int(f[0]) == rec['condition'][0] and int(f[4]) == rec['n_scat'][0]
    assert float(f[2]) == rec['theta_n'][0] and repr(float(rec['theta_n'][0])) == f[2]
    df = post_processing.read_output(str(path))
    np.testing.assert_array_equal(df['theta_n'].to_numpy(), rec['theta_n'])
    np.testing.assert_array_equal(df['path_length[m],'].to_numpy(), rec['path_length'])
    assert list(df.columns)[5] == 'path_length[m],'   # the reference header's stray comma


def test_output_naming_matches_reference(tmp_path):
    for c in CASES:
        mc = MonteCarlo(output_dir=str(tmp_path / c['name']))
        mc.theta_0 = (np.pi * c['theta0']) / 180.
        mc.snow_effective_radius = c['rds']
        p1 = mc.setup_output(c['n'], c['wvl0'], c['fwhm'])
        assert os.path.basename(p1) == c['file_name']
        open(p1, 'w').close()
        p2 = mc.setup_output(c['n'], c['wvl0'], c['fwhm'])
        assert os.path.basename(p2) == c['file_name'][:-4] + '_1.txt'


def test_config_ini_and_flags(tmp_path, monkeypatch):
    (tmp_path / 'config.ini').write_text('[model parameters]\ntau_tot=12.5\nimp_cnc=0\nrho_snw=250\n'
                                         'rho_ice=917\n[data]\noutput_dir=o\noptics_dir=x\nfi_imp=f.nc\n'
                                         '[gpu]\nmode=replay\nssa_ice=0.95\n')
    monkeypatch.chdir(tmp_path)
    monkeypatch.setattr('sys.argv', ['prog', '--rho_snw', '310'])
    mc = MonteCarlo(imp_cnc=1e-6)
    assert mc.tau_tot == 12.5 and mc.rho_snw == 310.0 and mc.imp_cnc == 1e-6
    assert mc.mode == 'replay' and mc._param_ssa_ice == 0.95 and mc.fi_imp == 'f.nc'


def test_parallel_split_order():
    data = np.arange(10)
    parts = [Parallel(data, size=3, rank=r).working_set for r in range(3)]
    np.testing.assert_array_equal(np.concatenate(parts), data)
    assert [p.size for p in parts] == [4, 3, 3]
    assert Parallel(data).answer_and_reduce([1, 2], MonteCarlo.flatten_list) == [1, 2]


@pytest.mark.parametrize('case', [c for c in CASES if c['name'] in ('spectral_imp', 'one_photon')],
                         ids=lambda c: c['name'])
def test_spectral_lookup_matches_reference_interpolation(case):
    t = optics.SpectralTable(case['ice']['wvl'], {k: case['ice'][k] for k in ('ss_alb', 'ext_cff_mss', 'asm_prm')})
    for w in np.round(np.linspace(1.0, 1.6, 61), 2):
        for k in ('ss_alb', 'ext_cff_mss', 'asm_prm'):
            want = float(walk.interp_two_nearest(case['ice']['wvl'], case['ice'][k], w))
            assert t.lookup(k, w) == want
    # outside the grid: nearest-value fallback, as the reference's ValueError path
    assert t.lookup('ss_alb', 0.1) == float(walk.interp_two_nearest(case['ice']['wvl'], case['ice']['ss_alb'], 0.1))


def test_netcdf_reader(tmp_path):
    from scipy.io import netcdf_file
    case = [c for c in CASES if c['name'] == 'spectral_imp'][0]
    p = tmp_path / 'ice_wrn_0250.nc'
    f = netcdf_file(str(p), 'w')
    f.createDimension('wvl', case['ice']['wvl'].size)
    for k, v in case['ice'].items():
        var = f.createVariable(k, 'd', ('wvl',))
        var[:] = v
    f.close()
    t = optics.SpectralTable.from_netcdf(str(p))
    np.testing.assert_array_equal(t.columns['asm_prm'], case['ice']['asm_prm'])


def test_brf_and_albedo_estimators():
    assert post_processing.calculate_bins(1., 175.) == 137
    g = [c for c in CASES if c['name'] == 'vdh_free'][0]['records']
    mid, b = post_processing.brf(g, 90)
    refl = g['condition'] == 1
    h, e = np.histogram(g['theta_n'][refl], bins=90, range=(0, np.pi / 2), weights=g['wvn'][refl])
    m = e[:-1] + np.diff(e) / 2
    np.testing.assert_allclose(b, h / (g['wvn'].sum() * np.sin(m) * np.cos(m) / np.sum(np.sin(m) * np.cos(m))))
    # the BRF integrates back to the albedo: sum_k BRF_k w_k = albedo
    w = np.sin(mid) * np.cos(mid) / np.sum(np.sin(mid) * np.cos(mid))
    assert abs(np.sum(b * w) - post_processing.albedo(g)) < 1e-12
    assert post_processing.albedo(g) == walk.calculate_albedo(
        list(zip(g['condition'], g['wvn'])))


def test_unsupported_reference_modes_raise(tmp_path):
    mc = MonteCarlo(optics_dir=str(tmp_path))
    with pytest.raises(NotImplementedError):
        mc.run(10, 1.3, 0.085, 100, debug=True)
    # aspherical grains need the Yang-layout tables (the reference reads them too)
    with pytest.raises(FileNotFoundError):
        mc.run(10, 1.3, 0.085, 100, shape='droxtal')
    with pytest.raises(ValueError):
        mc.run(10, 15.5, 0.085, 100, shape='droxtal')     # between the two table bands


def test_aspherical_output_naming(tmp_path):
    """output_dir/<shape dir>/<roughness dir>/..._I<IQUV>.txt or ..._HG.txt
    (mc3D:105-134) and the 10-column header with Q_n, U_n, V_n (mc3D:1637-1646)."""
    import numpy as np
    mc = MonteCarlo(output_dir=str(tmp_path / 'out'))
    mc.theta_0 = np.pi / 6
    mc.snow_effective_radius = 100.0000125
    mc.shape, mc.roughness = 'hollow bullet rosette', 'moderately rough'
    mc.initial_stokes_params = np.array([1, 0, 1, 0])
    p = mc.setup_output(300, 1.3, 1e-12)
    assert p.endswith('out/hollow_bullet_rosette/Rough003/1.3_1e-12_100.0000125_300_29.999999999999996_I1010.txt')
    mc.HG = True
    assert mc.setup_output(300, 1.3, 1e-12).endswith('_HG.txt')
    rec = dict(condition=np.array([1, 4]), wvn=np.array([1 / 1.3, 1 / 1.3]), theta_n=np.array([0.5, 1.0]),
               phi_n=np.array([1.0, 2.0]), n_scat=np.array([3, 0]), path_length=np.array([0.1, 0.2]),
               snow_depth=np.array([1.0, 1.0]), Q_n=np.array([0.25, 0.0]), U_n=np.array([-0.5, 0.0]),
               V_n=np.array([0.125, 0.0]))
    MonteCarlo.write_records(p, rec)
    lines = open(p).read().splitlines()
    assert lines[0].endswith('snow_depth[m], Q_n, U_n, V_n')
    assert lines[1].split()[-3:] == ['0.25', '-0.5', '0.125']


def test_balanced_chunk_gives_every_block_the_same_number_of_chunks():
    from monte_carlompi_amd.gpu import balanced_chunk
    for n, bpc, cus, per in [(1_000_000, 1, 256, 2), (1_000_000, 2, 256, 2), (12345, 1, 256, 2), (10_000_000, 1, 256, 4)]:
        c = balanced_chunk(n, bpc, cus, per)
        chunks = -(-n // c)
        assert chunks <= per * bpc * cus                    # no block gets an extra chunk
        assert c == 64 or -(-n // (c - 1)) > per * bpc * cus  # and the chunk is the smallest that does it
    assert balanced_chunk(1_000_000, 1, 256) == 1954


def test_bench_cli_parses_and_refuses_several_gpus():
    import subprocess, sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, 'bench.py'), '--help'], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and '--launch' in r.stdout
    r = subprocess.run([sys.executable, os.path.join(root, 'bench.py'), '--gpus', '2'], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and 'one GPU' in (r.stdout + r.stderr)


def test_pack_polar_layout():
    """Polar cases -> mcs_polar_case rows and 7 x n tables (grid, P11..P44);
    HG cases get table -1."""
    import numpy as np
    from monte_carlompi_amd.gpu import Band, Case, PolarTables, pack_polar
    x = np.linspace(-1, 1, 4)
    b = [Band(wvl=1.0, g=0.8, ssa_ice=0.99, ext_cff=1e4), Band(wvl=1.1, g=0.8, ssa_ice=0.99, ext_cff=1e4)]
    Y0, Y1 = np.arange(24.).reshape(6, 4) + 1, np.arange(24.).reshape(6, 4) + 100
    cases = [Case(tau_tot=1, theta0_deg=0, n_photons=3, bands=b[:1]),
             Case(tau_tot=1, theta0_deg=0, n_photons=3, bands=b, polar=PolarTables(x, [Y0, Y1], 1, (1, 0.5, 0, 0)))]
    parr, tabs, n = pack_polar(cases)
    assert n == 4 and tabs.shape == (2, 7, 4)
    assert parr[0].table == -1 and parr[1].table == 0 and parr[1].cutoff == 1
    assert list(parr[1].stokes0) == [1.0, 0.5, 0.0, 0.0]
    np.testing.assert_array_equal(tabs[1, 0], x)
    np.testing.assert_array_equal(tabs[1, 1:], Y1)
    assert pack_polar(cases[:1]) is None


def test_spectral_albedo_table(tmp_path):
    """spectral_albedo (pp:250-312) over record files named as run() names
    them: grouped by radius then nominal wavelength, actual wavelength
    mean(1 / wvn), albedo and non-absorbed fraction -- against numbers worked
    out by hand for small files (wvn weights 1/1.0 and 1/1.1 etc.)."""
    def rec(wvls, conds):
        wvls = np.asarray(wvls, float)
        n = wvls.size
        return dict(condition=np.asarray(conds), wvn=1.0 / wvls, theta_n=np.full(n, 0.3), phi_n=np.full(n, 1.0),
                    n_scat=np.arange(n), path_length=np.full(n, 0.01), snow_depth=np.full(n, 1.0))
    files = {
        # (wvl0, rds) -> (wavelengths, conditions)
        (1.0, 100.0): ([1.0, 1.0, 1.1, 0.9], [1, 4, 2, 1]),
        (1.3, 100.0): ([1.3, 1.3], [1, 3]),
        (1.0, 50.0): ([1.0, 1.0, 1.0, 1.0, 1.0], [1, 1, 1, 4, 5]),
    }
    names = []
    for (wvl0, rds), (wv, cd) in files.items():
        name = '%s_0.01_%s_%d_59.99999999999999_HG.txt' % (wvl0, rds, len(wv))
        MonteCarlo.write_records(str(tmp_path / name), rec(wv, cd))
        names.append(name)
    rows = post_processing.spectral_albedo(post_processing.group_files(names, str(tmp_path)))
    assert [(r[0], r[1]) for r in rows] == [(50.0, 1.0), (100.0, 1.0), (100.0, 1.3)]
    # r = 50: five photons at 1.0 um, three reflected; nothing leaves through the bottom
    assert rows[0][2:] == (1.0, 0.6, 0.6)
    # r = 100, 1.0 um: wvn 1, 1, 1/1.1, 1/0.9; reflected 1 + 1/0.9; bottom diffuse 1/1.1
    q = 1 + 1 + 1 / 1.1 + 1 / 0.9
    np.testing.assert_allclose(rows[1][2], (1.0 + 1.0 + 1.1 + 0.9) / 4, rtol=1e-15)
    np.testing.assert_allclose(rows[1][3], (1 + 1 / 0.9) / q, rtol=1e-14)
    np.testing.assert_allclose(rows[1][4], (1 + 1 / 0.9 + 1 / 1.1) / q, rtol=1e-14)
    # r = 100, 1.3 um: one reflected, one direct transmission (condition 3 counts as non-absorbed)
    np.testing.assert_allclose(rows[2][2:], (1.3, 0.5, 1.0), rtol=1e-15)


def test_wavelength_bins_equal_numpy_unique():
    """run() bins the photons' wavelengths by integer hundredths; the result
    equals np.unique(return_inverse=True) for draws on the 0.01 um grid (as
    mc3D:1519-1520 rounds them), for off-grid and non-finite values (which go
    through np.unique), a single value and no photons."""
    from monte_carlompi_amd.monte_carlo3D import wavelength_bins
    rng = np.random.default_rng(4)
    inputs = [np.around(rng.normal(1.3, 0.085 / 2.355, 100000), 2),
              np.around(rng.normal(0.5, 0.2, 50000), 2),              # includes negative draws
              np.around(rng.normal(1.3, 1e-12, 1000), 2),
              rng.uniform(0.3, 2.5, 1000),                            # off the grid
              np.array([1.3, np.nan, 1.31]), np.array([0.01]), np.array([], dtype=float)]
    for w in inputs:
        u, inv = wavelength_bins(w)
        u0, inv0 = np.unique(w, return_inverse=True)
        np.testing.assert_array_equal(u, u0)
        np.testing.assert_array_equal(inv, inv0.ravel())


class _CaptureWalker:
    """Stands in for GpuWalker in run(): keeps the case and the photon bands it
    is given and returns empty records (no device work)."""
    def set_cases(self, cases):
        self.cases = cases

    def walk(self, mode, seed, records, photon_band=None, replay=None):
        self.photon_band = np.array(photon_band)

    def records(self):
        n = self.cases[0].n_photons
        return {k: np.zeros(n, t) for k, t in (('condition', np.int8), ('band', np.uint8), ('n_scat', np.int64),
                                              ('theta_n', float), ('phi_n', float), ('path_length', float),
                                              ('wvn', float), ('snow_depth', float))}

    def tallies(self):
        return np.zeros((1, 656))

    def stats(self):
        return {}


def test_run_bins_photon_wavelengths_and_optics_per_bin(tmp_path, monkeypatch):
    """run() evaluates the optics once per 0.01 um wavelength bin: the case's
    bands, photon weights and per-photon band indices, tau_tot from the snow
    depth and the per-photon snow_depth equal the per-photon evaluation of
    mc3D:1515-1612 (np.unique over the photons' wavelengths, a lookup in the
    ice table for every photon)."""
    from scipy.io import netcdf_file
    monkeypatch.setattr('sys.argv', ['x'])
    ice = [c for c in CASES if c['name'] == 'spectral_imp'][0]['ice']
    path = optics.ice_file(str(tmp_path), 100)
    os.makedirs(os.path.dirname(path))
    f = netcdf_file(path, 'w')
    f.createDimension('wvl', ice['wvl'].size)
    for k, v in ice.items():
        f.createVariable(k, 'd', ('wvl',))[:] = v
    f.close()
    mc = MonteCarlo(output_dir=str(tmp_path), optics_dir=str(tmp_path), depth_m=0.4, seed=5)
    cap = _CaptureWalker()
    mc._walker = cap
    n, wvl0, fwhm = 20000, 1.3, 0.085
    np.random.seed(11)
    mc.run(n, wvl0, fwhm, 100, theta_0=30.0, write=False)
    np.random.seed(11)
    wvls = np.around(np.random.normal(loc=wvl0, scale=fwhm / 2.355, size=n), decimals=2)
    uniq, first, band = np.unique(wvls, return_index=True, return_inverse=True)
    t = optics.SpectralTable.from_netcdf(path)
    ssa_p, ext_p, g_p = (np.array([t.lookup(k, w) for w in wvls]) for k in ('ss_alb', 'ext_cff_mss', 'asm_prm'))
    ssa, ext, g = ssa_p[first], ext_p[first], g_p[first]
    assert np.unique(ext).size > 5                         # the optics do vary over the bins
    case = cap.cases[0]
    assert case.n_photons == n and len(case.bands) == uniq.size > 5
    np.testing.assert_array_equal(cap.photon_band, band)
    np.testing.assert_array_equal([b.wvl for b in case.bands], uniq)
    np.testing.assert_array_equal([b.weight for b in case.bands], np.bincount(band))
    np.testing.assert_array_equal([b.ext_cff for b in case.bands], ext * mc.rho_snw)
    np.testing.assert_array_equal([b.ssa_ice for b in case.bands], ssa)
    np.testing.assert_array_equal([b.g for b in case.bands], g)
    tau = float(0.4 * ext_p[0] * mc.rho_snw)
    assert case.tau_tot == tau
    np.testing.assert_array_equal(mc.snow_depth, tau / (ext_p * mc.rho_snw))
