"""The Python oracle reproduces, bit for bit, the records the reference's own
MonteCarlo.run wrote for seeded streams (tests/golden/make_golden.py)."""
import numpy as np
import pytest

from oracle import walk
from tests.golden_cases import FIELDS, load

META, CASES = load()


def oracle_ssp_fn(case):
    def fn(wvls):
        ice = walk.lookup_ssps(case['ice'], wvls)
        imp = walk.lookup_ssps(case['imp'], wvls)
        return (dict(ssa=ice['ss_alb'], ext=ice['ext_cff_mss'], g=ice['asm_prm']),
                dict(ssa=imp['ss_alb'], ext=imp['ext_cff_mss']))
    return fn


def run_case(case, trajectory=False):
    return walk.replay_run(case['seed'], case['n'], case['wvl0'], case['fwhm'],
                           oracle_ssp_fn(case), case['imp_cnc'], case['tau'],
                           META['rho_snw'], case['theta0'], case['lb'], case['R'],
                           trajectory=trajectory, lambertian_surface=case.get('ls', False))


@pytest.mark.parametrize('case', CASES, ids=[c['name'] for c in CASES])
def test_oracle_matches_reference_records(case):
    records, par, values, offsets = run_case(case, trajectory=(case['n'] <= 250))
    got = walk.records_to_arrays(records)
    ref = case['records']
    lowprec = set(case['lowprec_columns'])
    for j, f in enumerate(FIELDS):
        if j in lowprec:
            np.testing.assert_allclose(got[f], ref[f], rtol=2e-5, err_msg=f)
        else:
            np.testing.assert_array_equal(got[f], ref[f], err_msg=f)
    # every photon's replay slice = initial triple + its walk draws
    assert offsets[-1] == values.size and np.all(np.diff(offsets) >= 3)


def test_golden_cover_every_exit_condition():
    seen = np.zeros(6, int)
    for c in CASES:
        seen += np.bincount(c['records']['condition'], minlength=6)
    assert np.all(seen[1:] > 0), seen


def test_replay_stream_reproduces_records():
    """Walking each photon from its CSR slice alone gives the same records and
    consumes the slice exactly (the contract the GPU replay walk follows)."""
    case = [c for c in CASES if c['name'] == 'thin_lamb'][0]
    records, par, values, offsets = run_case(case)
    theta_0 = (np.pi * case['theta0']) / 180.
    for k in range(case['n']):
        src = walk.ArrayStream(values[offsets[k]:offsets[k + 1]])
        first = (-np.log(src.rand1())[0], src.rand1()[0], src.rand1()[0])
        rec = walk.walk_photon(src, first, par['wvl'][k], par['g'][k], par['ssa_ice'][k],
                               par['ssa_imp'][k], par['ext_cff_mss'][k], par['p_ext_imp'][k],
                               par['snow_depth'][k], case['tau'], META['rho_snw'], theta_0,
                               case['lb'], case['R'], trajectory=False)
        assert rec == records[k]
        assert src.used == src.values.size
