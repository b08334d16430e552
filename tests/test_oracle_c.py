"""The C oracle: MT19937 stream == numpy RandomState, and records equal to the
Python oracle (hence to the reference's golden records)."""
import numpy as np
import pytest

from oracle import coracle, walk
from tests.golden_cases import load
from tests.test_oracle_golden import META, oracle_ssp_fn, run_case

_, CASES = load()


def test_mt19937_matches_numpy_draw_for_draw():
    rs = np.random.RandomState(12345)
    rs.normal(size=7)               # leave the state mid-block
    m = coracle.mt_from_numpy(rs)
    got = coracle.mt_doubles(m, 5000)
    want = rs.random_sample(5000)
    np.testing.assert_array_equal(got, want)
    coracle.mt_to_numpy(m, rs)
    # state handed back continues identically
    rs2 = np.random.RandomState(12345); rs2.normal(size=7); rs2.random_sample(5000)
    np.testing.assert_array_equal(rs.random_sample(10), rs2.random_sample(10))


@pytest.mark.parametrize('case', CASES, ids=[c['name'] for c in CASES])
def test_c_oracle_matches_reference_records(case):
    out, par = coracle.run_rank(case['seed'], case['n'], case['wvl0'], case['fwhm'],
                                oracle_ssp_fn(case), case['imp_cnc'], case['tau'],
                                META['rho_snw'], case['theta0'], case['lb'], case['R'],
                                record=True, lambertian_surface=case.get('ls', False))
    ref = case['records']
    np.testing.assert_array_equal(out['condition'], ref['condition'])
    np.testing.assert_array_equal(out['n_scat'], ref['n_scat'])
    rtol = 2e-5 if case['lowprec_columns'] else 1e-12
    for f in ('theta_n', 'phi_n', 'path_length'):
        np.testing.assert_allclose(out[f], ref[f], rtol=rtol, atol=1e-13, err_msg=f)
    # the recorded replay stream equals the Python oracle's
    _, _, values, offsets = run_case(case)
    np.testing.assert_array_equal(out['stream'][1], offsets)
    np.testing.assert_array_equal(out['stream'][0], values)


def test_c_oracle_replay_consumes_slices_exactly():
    case = [c for c in CASES if c['name'] == 'nadir_bounce'][0]
    records, par, values, offsets = run_case(case)
    out = coracle.walk(par, case['tau'], META['rho_snw'], case['theta0'], case['lb'],
                       case['R'], replay=(values, offsets))
    ref = walk.records_to_arrays(records)
    np.testing.assert_array_equal(out['condition'], ref['condition'])
    np.testing.assert_array_equal(out['n_scat'], ref['n_scat'])
    # a truncated slice is reported, not read past
    bad = offsets.copy(); bad[1:] -= 1; bad[0] = 0
    bad = np.concatenate([[0], np.cumsum(np.diff(offsets) - 0)])
    short = (values, offsets.copy())
    short[1][-1] -= 1
    with pytest.raises(RuntimeError, match='replay'):
        coracle.walk(par, case['tau'], META['rho_snw'], case['theta0'], case['lb'],
                     case['R'], replay=short)


def _polar_cases():
    from tests import golden_polar
    meta, cases = golden_polar.load()
    return meta, [c for c in cases if not c['HG']]


PMETA, PCASES = _polar_cases()


@pytest.mark.parametrize('case', PCASES, ids=[c['name'] for c in PCASES])
def test_c_oracle_polar_matches_reference_records(case, tmp_path):
    """The aspherical walk in C (phase-matrix rejection draws, Stokes update)
    follows the reference's records: the same conditions and scattering
    counts, the same uniform stream draw for draw as the Python oracle, and
    angles, paths and Stokes vectors to libm-vs-numpy rounding."""
    from tests import golden_polar
    from tests.test_oracle_polar import run_case as py_polar_case
    d = golden_polar.table_dir(str(tmp_path), case)
    imp_fn = lambda w: dict(ssa=np.full(w.shape, 0.3), ext=np.full(w.shape, 12000.0))
    out, par, pc, bands = coracle.polar_run_rank(
        case['seed'], case['n'], case['wvl0'], case['fwhm'], case['rds'], d, case['stokes'], imp_fn,
        case['imp_cnc'], case['tau'], PMETA['rho_snw'], PMETA['rho_ice'], case['theta0'], case['lb'],
        case['R'], lambertian_surface=case.get('ls', False), record=True)
    ref = case['records']
    np.testing.assert_array_equal(out['condition'], ref['condition'])
    np.testing.assert_array_equal(out['n_scat'], ref['n_scat'])
    rtol = 2e-5 if case['lowprec_columns'] else 1e-12
    for f in ('theta_n', 'phi_n', 'path_length'):
        np.testing.assert_allclose(out[f], ref[f], rtol=rtol, atol=1e-13, err_msg=f)
    for f in ('Q_n', 'U_n', 'V_n'):
        np.testing.assert_allclose(out[f], ref[f], rtol=0, atol=1e-13, err_msg=f)
    _, _, values, offsets, _, _, _ = py_polar_case(case, str(tmp_path / 'py'))
    np.testing.assert_array_equal(out['stream'][1], offsets)
    np.testing.assert_array_equal(out['stream'][0], values)
    # replaying that stream walks the same photons and ends on its last draw
    rep = coracle.walk(par, case['tau'], PMETA['rho_snw'], case['theta0'], case['lb'], case['R'],
                       replay=(values, offsets), lambertian_surface=case.get('ls', False),
                       polar=(pc, bands, 1. / out['wvn'], case['stokes']))
    for f in ('condition', 'n_scat', 'theta_n', 'phi_n', 'path_length', 'Q_n', 'U_n', 'V_n'):
        np.testing.assert_array_equal(rep[f], out[f], err_msg=f)


def test_c_oracle_polar_catches_a_dropped_stokes_term(tmp_path):
    """A planted mistake in the restated tables (P12 zeroed: the phase-matrix
    coupling of I and Q dropped) changes the Stokes vectors the walk ends
    with, so the record comparison above pins that term."""
    from tests import golden_polar
    case = [c for c in PCASES if c['name'] == 'pm_linear_q'][0]
    d = golden_polar.table_dir(str(tmp_path), case)
    imp_fn = lambda w: dict(ssa=np.full(w.shape, 0.3), ext=np.full(w.shape, 12000.0))
    out, par, pc, bands = coracle.polar_run_rank(
        case['seed'], case['n'], case['wvl0'], case['fwhm'], case['rds'], d, case['stokes'], imp_fn,
        case['imp_cnc'], case['tau'], PMETA['rho_snw'], PMETA['rho_ice'], case['theta0'], case['lb'],
        case['R'], record=True)
    broken = {k: dict(b, Y=dict(b['Y'], P12=np.zeros_like(b['Y']['P12']))) for k, b in bands.items()}
    rs = np.random.RandomState(case['seed'])
    walk.draw_wavelengths(rs, case['n'], case['wvl0'], case['fwhm'])
    bad = coracle.walk(par, case['tau'], PMETA['rho_snw'], case['theta0'], case['lb'], case['R'],
                       mt=coracle.mt_from_numpy(rs), polar=(pc, broken, 1. / out['wvn'], case['stokes']),
                       allow_unbound_phi=True)
    assert np.abs(bad['Q_n'] - case['records']['Q_n']).max() > 1e-3


def test_polar_law_check_rejects_a_dropped_p12(tmp_path):
    """The at-scale law comparison of the GPU tests (exit fractions,
    reflected zenith angles, mean final Q, U, V, mean scattering count at 4
    sigma) has the power to see a wrong phase matrix: 5e4 C-oracle photons of
    the Q-polarised case pass against 5e4 more of their own law and fail
    against 5e4 walked with P12 zeroed (the I-Q coupling dropped)."""
    from concurrent.futures import ThreadPoolExecutor
    from tests import golden_polar
    from tests.test_gpu_polar import assert_same_law, assert_same_mean, c_oracle_law
    case = [c for c in PCASES if c['name'] == 'pm_linear_q'][0]
    d = golden_polar.table_dir(str(tmp_path), case)
    imp_fn = lambda w: dict(ssa=np.full(w.shape, 0.3), ext=np.full(w.shape, 12000.0))
    _, par1, pc, bands = coracle.polar_run_rank(
        case['seed'], 1, case['wvl0'], 1e-12, case['rds'], d, case['stokes'], imp_fn, case['imp_cnc'],
        case['tau'], PMETA['rho_snw'], PMETA['rho_ice'], case['theta0'], case['lb'], case['R'])
    broken = {k: dict(b, Y=dict(b['Y'], P12=np.zeros_like(b['Y']['P12']))) for k, b in bands.items()}
    per = 6250
    par = {k: np.full(per, v[0]) for k, v in par1.items()}
    keys = par['wvl']

    def sample(tabs, seed0):
        def one(r):
            mt = coracle.mt_from_numpy(np.random.RandomState(seed0 + r))
            return coracle.walk(par, case['tau'], PMETA['rho_snw'], case['theta0'], case['lb'], case['R'], mt=mt,
                                polar=(pc, tabs, keys, case['stokes']), allow_unbound_phi=True)
        with ThreadPoolExecutor(max_workers=8) as ex:
            parts = list(ex.map(one, range(8)))
        out = {k: np.concatenate([p[k] for p in parts]) for k in parts[0] if k != 'steps'}
        law = c_oracle_law(out)
        return law, dict(condition=law['condition'], theta_n=law['theta_n'], n_scat=law['n_scat']), \
            {k: law[k] for k in ('Q_n', 'U_n', 'V_n')}

    a, _, _ = sample(bands, 1000)
    _, rb, sb = sample(bands, 2000)
    assert_same_law(a, rb, sb, polarised=True)
    assert_same_mean(a['n_scat'], rb['n_scat'], 'n_scat')
    _, rx, sx = sample(broken, 2000)
    with pytest.raises(AssertionError):
        assert_same_law(a, rx, sx, polarised=True)
