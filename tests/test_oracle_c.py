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
