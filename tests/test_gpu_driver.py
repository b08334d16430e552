"""The drop-in path end to end on the GPU: MonteCarlo.run as the reference's
driver calls it, in replay mode (records equal to the reference's golden
records) and native mode (record file, tallies and CSV agree)."""
import os

import numpy as np
import pytest

from monte_carlompi_amd import post_processing
from monte_carlompi_amd.monte_carlo3D import MonteCarlo
from tests.golden_cases import load
from tests.test_oracle_golden import META, run_case

pytestmark = pytest.mark.gpu
_, CASES = load()


def write_nc(path, table):
    from scipy.io import netcdf_file
    f = netcdf_file(path, 'w')
    f.createDimension('wvl', table['wvl'].size)
    for key, arr in table.items():
        v = f.createVariable(key, 'd', ('wvl',))
        v[:] = arr
    f.close()


@pytest.mark.parametrize('case', [c for c in CASES if c['name'] in ('spectral_imp', 'thin_lamb', 'cfg1_semi_inf',
                                                                   'lamb_surface')],
                         ids=lambda c: c['name'])
def test_run_replay_reproduces_reference_file(tmp_path, monkeypatch, case):
    print('driver replay', case['name'], flush=True)
    monkeypatch.setattr('sys.argv', ['prog'])
    optics_dir = tmp_path / 'optics'
    os.makedirs(optics_dir / 'mie' / 'snicar')
    write_nc(str(optics_dir / 'mie' / 'snicar' / ('ice_wrn_%04d.nc' % case['rds'])), case['ice'])
    write_nc(str(optics_dir / 'mie' / 'snicar' / 'imp.nc'), case['imp'])
    _, _, values, offsets = run_case(case)
    mc = MonteCarlo(tau_tot=case['tau'], imp_cnc=case['imp_cnc'], rho_snw=META['rho_snw'],
                    output_dir=str(tmp_path / 'out'), optics_dir=str(optics_dir), fi_imp='imp.nc')
    np.random.seed(case['seed'])
    path = mc.run(case['n'], case['wvl0'], case['fwhm'], case['rds'], theta_0=case['theta0'],
                  Lambertian_bottom=case['lb'], Lambertian_reflectance=case['R'],
                  Lambertian_surface=case.get('ls', False), replay=(values, offsets))
    mc.close()
    assert os.path.basename(path) == case['file_name']
    df = post_processing.read_output(path)
    ref = case['records']
    np.testing.assert_array_equal(df['condition'].to_numpy(), ref['condition'])
    np.testing.assert_array_equal(df['n_scat'].to_numpy(), ref['n_scat'])
    for col, key in (('wvn[um^-1]', 'wvn'), ('theta_n', 'theta_n'), ('phi_n', 'phi_n'),
                     ('path_length[m],', 'path_length'), ('snow_depth[m]', 'snow_depth')):
        np.testing.assert_allclose(df[col].to_numpy(), ref[key], rtol=1e-5, atol=1e-12, err_msg=col)


def test_run_native_file_matches_tallies(tmp_path, monkeypatch):
    monkeypatch.setattr('sys.argv', ['prog'])
    mc = MonteCarlo(output_dir=str(tmp_path / 'out'), optics_dir=str(tmp_path / 'none'),
                    ssa_ice=0.999, g=0.89, depth_m=1.0, seed=3)
    np.random.seed(0)
    path = mc.run(200_000, 1.3, 0.085, 100., theta_0=15., Lambertian_bottom=True,
                  Lambertian_reflectance=0.5)
    df = post_processing.read_output(path)
    assert len(df) == 200_000
    assert abs(df['snow_depth[m]'].to_numpy() - 1.0).max() < 1e-12
    # BRF over 18 zenith bins from the CSV equals the phi marginal of the GPU tallies
    m1, b_csv = post_processing.brf(df, 18)
    m2, b_tal = post_processing.brf_1d_from_tallies(mc.tallies)
    np.testing.assert_allclose(b_csv, b_tal, rtol=1e-6)
    assert abs(post_processing.albedo(df) - mc.calculate_albedo(mc.records)) < 1e-12
    wv = df['wvn[um^-1]'].to_numpy()
    assert len(np.unique(wv)) > 10          # a spread of wavelengths, FWHM 0.085 um
    mc.close()


@pytest.mark.parametrize('name', ['pm_spectral_u', 'pm_surface', 'asph_hg'])
def test_run_aspherical_replay_reproduces_reference_file(tmp_path, monkeypatch, name):
    """MonteCarlo.run(shape=..., stokes_params=...) on Yang-layout tables:
    the file name (shape / roughness directories, r_eff, I<IQUV> or HG
    appendix) and the 7- or 10-column records of the reference's own run."""
    from tests import golden_polar
    from tests.test_oracle_polar import run_case as polar_case
    _, pcases = golden_polar.load()
    case = [c for c in pcases if c['name'] == name][0]
    print('driver aspherical replay', name, flush=True)
    monkeypatch.setattr('sys.argv', ['prog'])
    optics_dir = tmp_path / 'optics'
    os.makedirs(optics_dir / 'mie' / 'snicar')
    n = 480
    grid = (0.205 + 0.01 * np.arange(n)) * 1e-6
    write_nc(str(optics_dir / 'mie' / 'snicar' / 'imp.nc'),
             dict(wvl=grid, ss_alb=np.full(n, 0.3), ext_cff_mss=np.full(n, 12000.0)))
    _, _, values, offsets, _, _, _ = polar_case(case, str(optics_dir))
    mc = MonteCarlo(tau_tot=case['tau'], imp_cnc=case['imp_cnc'], rho_snw=META['rho_snw'],
                    output_dir=str(tmp_path / 'out'), optics_dir=str(optics_dir), fi_imp='imp.nc',
                    HG=case['HG'])
    np.random.seed(case['seed'])
    path = mc.run(case['n'], case['wvl0'], case['fwhm'], case['rds'], theta_0=case['theta0'],
                  stokes_params=np.array(case['stokes']), shape=case['shape'], roughness=case['rough'],
                  Lambertian_bottom=case['lb'], Lambertian_reflectance=case['R'],
                  Lambertian_surface=case.get('ls', False), replay=(values, offsets))
    mc.close()
    assert os.path.basename(path) == case['file_name']
    assert os.path.relpath(os.path.dirname(path), str(tmp_path / 'out')) == case['subdir']
    df = post_processing.read_output(path)
    ref = case['records']
    np.testing.assert_array_equal(df['condition'].to_numpy(), ref['condition'])
    np.testing.assert_array_equal(df['n_scat'].to_numpy(), ref['n_scat'])
    cols = [('wvn[um^-1]', 'wvn'), ('theta_n', 'theta_n'), ('phi_n', 'phi_n'),
            ('path_length[m],', 'path_length'), ('snow_depth[m]' + ('' if case['HG'] else ','), 'snow_depth')]
    if not case['HG']:
        cols += [('Q_n,', 'Q_n'), ('U_n,', 'U_n'), ('V_n', 'V_n')]
    for col, key in cols:
        np.testing.assert_allclose(df[col].to_numpy(), ref[key], rtol=1e-5, atol=1e-9, err_msg=key)
