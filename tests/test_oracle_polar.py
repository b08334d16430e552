"""The aspherical oracle (oracle/polar.py + walk.walk_photon(polar=...))
reproduces, bit for bit, the records and final Stokes vectors the reference's
own MonteCarlo.run wrote for seeded streams on synthetic Yang-layout tables;
the product's table reader agrees with the oracle's restatement."""
import numpy as np
import pytest

from oracle import polar, walk
from tests import golden_polar
from tests.golden_polar import FIELDS, STOKES

META, CASES = golden_polar.load()
IMP = dict(ssa=0.3, ext=12000.0)


def run_case(case, root):
    d = golden_polar.table_dir(root, case)
    imp_fn = lambda w: dict(ssa=np.full(w.shape, IMP['ssa']), ext=np.full(w.shape, IMP['ext']))
    return polar.polar_run(case['seed'], case['n'], case['wvl0'], case['fwhm'], case['rds'], d, case['HG'],
                           case['stokes'], imp_fn, case['imp_cnc'], case['tau'], META['rho_snw'],
                           META['rho_ice'], case['theta0'], case['lb'], case['R'],
                           lambertian_surface=case.get('ls', False))


@pytest.mark.parametrize('case', CASES, ids=[c['name'] for c in CASES])
def test_polar_oracle_matches_reference_records(case, tmp_path):
    records, par, values, offsets, reff, tabs, bands = run_case(case, str(tmp_path))
    assert reff == case['snow_effective_radius']
    ref = case['records']
    cols = list(zip(*records))
    lowprec = set(case['lowprec_columns'])
    names = FIELDS + (STOKES if not case['HG'] else ())
    for j, f in enumerate(names):
        got = np.asarray(cols[j], dtype=np.float64)
        if j in lowprec:
            np.testing.assert_allclose(got, ref[f], rtol=2e-5, err_msg=f)
        else:
            np.testing.assert_array_equal(got, ref[f], err_msg=f)
    assert offsets[-1] == values.size and np.all(np.diff(offsets) >= 3)


def test_polar_golden_cover_conditions_and_polarisation():
    seen = np.zeros(6, int)
    for c in CASES:
        seen += np.bincount(c['records']['condition'], minlength=6)
    assert np.all(seen[1:] > 0), seen
    pm = [c for c in CASES if not c['HG']]
    assert any(np.any(np.abs(c['records']['V_n']) > 1e-3) for c in pm)
    assert any(c.get('ls') for c in pm) and any(c['lb'] and c['R'] > 0 for c in pm)


def test_product_reader_matches_oracle(tmp_path):
    """monte_carlompi_amd.optics.aspherical_ssps / PhaseMatrix against the
    oracle's restatement of mc3D:219-496, 819-883 on the same files."""
    from monte_carlompi_amd import optics
    case = [c for c in CASES if c['name'] == 'pm_spectral_u'][0]
    d = golden_polar.table_dir(str(tmp_path), case)
    wvls = np.around(np.random.RandomState(5).normal(1.3, 0.2 / 2.355, 300), 2)
    o = polar.aspherical_ssps(d, 1.3, 60, wvls, False, 917.0)
    p = optics.aspherical_ssps(wvls, 60, 1.3, str(tmp_path), 'droxtal', 'smooth', False)
    np.testing.assert_array_equal(p['wvls'], o[0])
    np.testing.assert_array_equal(p['ssa'], o[1])
    np.testing.assert_array_equal(p['ext_cff_mss'], o[2])
    np.testing.assert_array_equal(p['g'], o[3])
    assert p['snow_effective_radius'] == o[4]
    np.testing.assert_array_equal(p['pm'].theta_deg, o[5])
    tabs, bands = polar.polar_tables(o[5], o[6], o[0], 1.3)
    assert p['pm'].cutoff(1.3, p['wvls']) == tabs['cut']
    for key, b in bands.items():
        x, Y = p['pm'].interp_arrays(key)
        np.testing.assert_array_equal(x, tabs['x'])
        for j, e in enumerate(optics.PM_ELEMENTS):
            np.testing.assert_array_equal(Y[j], b['Y'][e])


def test_interp_is_scipy_interp1d():
    """np.interp (what the oracle calls) equals the reference's
    interp1d(cos[::-1], P[::-1]) at random points and on every grid point."""
    from scipy import interpolate
    from monte_carlompi_amd import workloads
    th = workloads.yang_theta_grid()
    el = workloads.synthetic_phase_matrix(th, 1.3, 100.0)
    c = np.cos((th * np.pi) / 180.)
    f = interpolate.interp1d(c[::-1], el['P12'][::-1])
    rs = np.random.RandomState(3)
    q = np.concatenate([rs.uniform(-1, 1, 2000), c])
    for v in q:
        assert np.interp(v, c[::-1], el['P12'][::-1]) == f(v)


def test_stokes_update_closed_forms():
    """Pins of the Stokes update that do not come from the reference run:
    rotations by +a then -a cancel and keep the degree of linear
    polarisation; with a diagonal matrix P11 = P22 = P33 = P44 the update is a
    pure frame rotation, so the result has I = 1 and the same degree."""
    s = np.array([1.0, 0.3, -0.2, 0.1])
    back = polar.rotate(-0.7, polar.rotate(0.7, s))
    np.testing.assert_allclose(back, s, atol=1e-15)
    # a pure rotation keeps the degree of linear polarisation
    r = polar.rotate(1.1, s)
    assert abs(np.hypot(r[1], r[2]) - np.hypot(s[1], s[2])) < 1e-15
    x = np.array([-1.0, 0.0, 1.0])
    ones = np.ones(3)
    pol = dict(x=x, Y={'P11': 2 * ones, 'P12': 0 * ones, 'P22': 2 * ones, 'P33': 2 * ones,
                       'P43': 0 * ones, 'P44': 2 * ones})
    # theta = 0, phi = 0: den = 0, so cos_i = 0 and the return rotation is by
    # -pi/2 (double angle -pi): Q and U change sign, I and V stay
    out = polar.stokes_update(pol, s, 1.0, 1.0, 0.5, 0.5)
    np.testing.assert_allclose(out, [1.0, -0.3, 0.2, 0.1], atol=1e-15)
    out = polar.stokes_update(pol, s, 0.3, np.cos(0.4), -0.2, 0.6)
    assert out[0] == 1.0 and np.hypot(out[1], out[2]) == pytest.approx(np.hypot(s[1], s[2]), abs=1e-12)


def single_scatter_expectations(recs, tabs, bands, key, stokes0):
    """Closed form of one scattering at normal incidence (muz_0 = -1, so the
    scattering plane is the exit direction's meridian plane and the return
    rotation is by pi): a photon reflected after one scattering at cos theta
    = -cos(theta_n) has Q_n = (P12 + c P22) / (P11 + c P12) with c =
    cos(2 phi_n) Q_0, and |U_n|, |V_n| = |sin 2 phi_n| Q_0 (P33, P43) /
    (P11 + c P12).  Returns (got, expected) arrays for such photons."""
    cols = list(zip(*recs))
    cond = np.asarray(cols[0]); ns = np.asarray(cols[4])
    th = np.asarray(cols[2], float); ph = np.asarray(cols[3], float)
    q = np.asarray(cols[7], float); u = np.asarray(cols[8], float); v = np.asarray(cols[9], float)
    m = (cond == 1) & (ns == 1)
    mu = -np.cos(th[m])
    Y = bands[key]['Y']
    P = {e: np.interp(mu, tabs['x'], Y[e]) for e in polar.ELEMENTS}
    c = np.cos(2 * ph[m]) * stokes0[1]
    s = np.abs(np.sin(2 * ph[m])) * stokes0[1]
    den = P['P11'] + c * P['P12']
    return (q[m], np.abs(u[m]), np.abs(v[m])), ((P['P12'] + c * P['P22']) / den, np.abs(s * P['P33'] / den),
                                                 np.abs(s * P['P43'] / den))


@pytest.mark.parametrize('stokes0', [(1, 0, 0, 0), (1, 1, 0, 0)])
def test_polar_oracle_single_scattering_closed_form(tmp_path, stokes0):
    """Independent of the reference's run: singly reflected photons of a thin
    layer under normal incidence carry the Stokes vector the phase matrix
    gives in closed form (unpolarised: Q_n = P12 / P11, U_n = V_n = 0)."""
    case = dict(CASES[0], n=8000, seed=31, fwhm=1e-12, theta0=0.0, lb=False, R=0.0, tau=0.5,
                stokes=list(stokes0))
    recs, par, values, offsets, reff, tabs, bands = run_case(case, str(tmp_path))
    key = par['wvl'][0]
    got, exp = single_scatter_expectations(recs, tabs, bands, key, stokes0)
    assert got[0].size > 30
    for g, e in zip(got, exp):
        np.testing.assert_allclose(g, e, rtol=1e-9, atol=1e-12)
