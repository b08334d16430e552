"""Replay mode on the GPU: the FP64 walk consuming the reference's own
uniform-deviate stream gives, photon by photon, the reference's condition and
n_scat exactly and theta_n / phi_n / path_length within 1e-5 relative
(observed: ~1e-15)."""
import numpy as np
import pytest

from oracle import coracle, walk
from tests.golden_cases import load
from tests.gpu_cases import case_from_params
from tests.test_oracle_golden import META, oracle_ssp_fn, run_case

pytestmark = pytest.mark.gpu
_, CASES = load()
RTOL = 1e-5


@pytest.fixture(scope='module')
def walker():
    from monte_carlompi_amd.gpu import GpuWalker
    w = GpuWalker(0)
    yield w
    w.close()


def check_records(got, ref, lowprec=False):
    np.testing.assert_array_equal(got['condition'], ref['condition'])
    np.testing.assert_array_equal(got['n_scat'], ref['n_scat'])
    rtol = 2e-5 if lowprec else RTOL
    for f in ('theta_n', 'phi_n', 'path_length', 'wvn', 'snow_depth'):
        np.testing.assert_allclose(got[f], ref[f], rtol=rtol, atol=1e-12, err_msg=f)


@pytest.mark.parametrize('case', CASES, ids=[c['name'] for c in CASES])
def test_replay_matches_reference_golden(walker, case):
    print('replay golden', case['name'], flush=True)
    records, par, values, offsets = run_case(case)
    gcase, pband = case_from_params(par, case['tau'], META['rho_snw'], case['theta0'],
                                    case['lb'], case['R'], case.get('ls', False))
    walker.set_cases([gcase])
    walker.walk(mode='replay', replay=(values, offsets), photon_band=pband)
    got = walker.records()
    check_records(got, case['records'], lowprec=bool(case['lowprec_columns']))
    st = walker.stats()
    assert st['photons'] == case['n']
    assert st['steps'] == int(np.sum(case['records']['n_scat'] + 1))


def test_replay_config_case_100k(walker):
    """100,000 photons of the 1 m / 60 deg / omega 0.999 / g 0.89 case:
    stream and records from the C oracle (numpy MT19937, seed 7)."""
    ext = walk.sphere_ext_cff_mss(100.0, 917.0)
    tau = 1.0 * ext * 300.0

    def ssp(wvls):
        n = wvls.size
        return (dict(ssa=np.full(n, 0.999), ext=np.full(n, ext), g=np.full(n, 0.89)),
                dict(ssa=np.zeros(n), ext=np.zeros(n)))
    print('replay 1e5 photons: oracle', flush=True)
    out, par = coracle.run_rank(7, 100_000, 1.3, 1e-12, ssp, 0.0, tau, 300.0, 60.0, True, 0.5,
                                record=True)
    gcase, pband = case_from_params(par, tau, 300.0, 60.0, True, 0.5)
    walker.set_cases([gcase])
    print('replay 1e5 photons: gpu', flush=True)
    walker.walk(mode='replay', replay=out['stream'], photon_band=pband)
    check_records(walker.records(), out)
    assert walker.stats()['steps'] == out['steps']


def test_replay_rejects_short_slice(walker):
    from monte_carlompi_amd._lib import McsError
    case = [c for c in CASES if c['name'] == 'thin_lamb'][0]
    records, par, values, offsets = run_case(case)
    gcase, pband = case_from_params(par, case['tau'], META['rho_snw'], case['theta0'],
                                    case['lb'], case['R'])
    walker.set_cases([gcase])
    short = offsets.copy()
    short[-1] -= 1      # last photon's slice one deviate short
    with pytest.raises(McsError, match='REPLAY'):
        walker.walk(mode='replay', replay=(values, short), photon_band=pband)
    # the context stays usable
    walker.walk(mode='replay', replay=(values, offsets), photon_band=pband)
    check_records(walker.records(), walk.records_to_arrays(records))


def test_replay_config_case_1e6_streamed(walker):
    """The 1e6-photon config case (1 m, 60 deg, omega 0.999, g 0.89, Lambertian
    bottom 0.5) in replay mode, in ten chunks of 1e5 photons so the recorded
    deviates stay under 1 GB at a time.  One MT19937 state (seed 11) runs on
    through all ten oracle chunks; every photon's record must match."""
    from oracle import coracle as C
    ext = walk.sphere_ext_cff_mss(100.0, 917.0)
    tau = 1.0 * ext * 300.0
    n, k = 1_000_000, 100_000
    rs = np.random.RandomState(11)
    wvls = walk.draw_wavelengths(rs, n, 1.3, 1e-12)
    mt = C.mt_from_numpy(rs)
    par_all = walk.photon_params(wvls, dict(ssa=0.999, ext=ext, g=0.89), dict(ssa=0.0, ext=0.0),
                                 0.0, tau, 300.0)
    total_steps = 0
    for s in range(0, n, k):
        print('replay 1e6 chunk', s // k, flush=True)
        par = {key: v[s:s + k] for key, v in par_all.items()}
        out = C.walk(par, tau, 300.0, 60.0, True, 0.5, mt=mt, record=True)
        out['wvn'] = 1. / par['wvl']
        out['snow_depth'] = par['snow_depth']
        gcase, pband = case_from_params(par, tau, 300.0, 60.0, True, 0.5)
        walker.set_cases([gcase])
        walker.walk(mode='replay', replay=out['stream'], photon_band=pband)
        check_records(walker.records(), out)
        total_steps += out['steps']
    assert total_steps > 1.5e8
