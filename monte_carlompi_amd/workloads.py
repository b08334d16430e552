"""Named workloads of the photon walk (sizes as the benchmarks quote them).

``snow_1m``: the single case of the reference's driver settings as the
benchmark names it -- monochromatic 1.30 um, ice spheres r_eff = 100 um,
rho_snw = 300, rho_ice = 917 kg/m^3, Q_ext = 2, omega = 0.999, g = 0.89,
1.0 m of snow (tau = 4907.3, effectively semi-infinite), theta0 = 60 deg,
phi0 = 0, Lambertian bottom R = 0.5 (monte_carlo3D-run.py:9-12).
"""
from __future__ import annotations

import os

import numpy as np

from .gpu import Band, Case

RHO_SNW = 300.0
RHO_ICE = 917.0


def sphere_ext_cff_mss(r_eff_um, rho_ice=RHO_ICE, q_ext=2.0):
    """3 Q_ext / (4 rho_ice r_eff) [m^2/kg]: extinction per mass of ice spheres."""
    return 3.0 * q_ext / (4.0 * rho_ice * (r_eff_um * 1e-6))


SNOW_1M = dict(wvl=1.30, r_eff=100.0, q_ext=2.0, ssa=0.999, g=0.89, depth_m=1.0,
               theta0=60.0, lambertian_bottom=True, r_lambertian=0.5)


def snow_1m_case(n_photons: int) -> Case:
    p = SNOW_1M
    ext_cff = sphere_ext_cff_mss(p['r_eff'], RHO_ICE, p['q_ext']) * RHO_SNW
    return Case(tau_tot=p['depth_m'] * ext_cff, theta0_deg=p['theta0'], n_photons=int(n_photons),
                bands=[Band(wvl=p['wvl'], g=p['g'], ssa_ice=p['ssa'], ext_cff=ext_cff)],
                lambertian_bottom=p['lambertian_bottom'], r_lambertian=p['r_lambertian'])


def sweep_cases(seed: int = 0, n_reff: int = 64, n_wvl: int = 16, n_theta: int = 8,
                depths=(0.05, 1.0), photons_lo: float = 1e4, photons_hi: float = 1e6,
                photon_scale: float = 1.0):
    """The batched sweep: r_eff log-spaced 30..1500 um x n_wvl Gaussian
    wavelength distributions (centre U(0.9, 1.6) um, FWHM U(0.01, 0.05) um,
    rounded to 0.01 um bins as mc3D:1519-1520 rounds photon wavelengths) x
    theta0 linspace(0, 70) deg x depths; photons per case log-uniform in
    [photons_lo, photons_hi]; omega = 1 - U(1e-5, 5e-2), g = U(0.85, 0.90) per
    case (the SSP tables are absent, so optics are parametric: Q_ext = 2).
    Returns (cases, meta) with meta arrays per case."""
    rng = np.random.Generator(np.random.PCG64(seed))
    reffs = np.logspace(np.log10(30.0), np.log10(1500.0), n_reff)
    centres = rng.uniform(0.9, 1.6, n_wvl)
    fwhms = rng.uniform(0.01, 0.05, n_wvl)
    thetas = np.linspace(0.0, 70.0, n_theta)
    cases, meta = [], dict(r_eff=[], wvl0=[], fwhm=[], theta0=[], depth=[], omega=[], g=[], tau=[], n=[])
    for r in reffs:
        ext_cff = sphere_ext_cff_mss(r) * RHO_SNW
        for c0, fw in zip(centres, fwhms):
            bands = spectral_bands(c0, fw)
            for th in thetas:
                for d in depths:
                    n = int(round(10 ** rng.uniform(np.log10(photons_lo), np.log10(photons_hi)) * photon_scale))
                    om = 1.0 - rng.uniform(1e-5, 5e-2)
                    g = rng.uniform(0.85, 0.90)
                    bl = [Band(wvl=w, g=g, ssa_ice=om, ext_cff=ext_cff, weight=wt) for w, wt in bands]
                    cases.append(Case(tau_tot=d * ext_cff, theta0_deg=float(th), n_photons=max(n, 1),
                                      bands=bl, lambertian_bottom=False, r_lambertian=0.0))
                    for k, v in (('r_eff', r), ('wvl0', c0), ('fwhm', fw), ('theta0', th), ('depth', d),
                                 ('omega', om), ('g', g), ('tau', d * ext_cff), ('n', max(n, 1))):
                        meta[k].append(v)
    return cases, {k: np.asarray(v) for k, v in meta.items()}


def spectral_bands(wvl0: float, fwhm: float, max_bands: int = 64, tail: float = 1e-6):
    """Wavelength bins of photons drawn from N(wvl0, fwhm / 2.355) and rounded
    to 0.01 um (mc3D:1516-1520): list of (wavelength, probability)."""
    from math import erf, sqrt
    sigma = fwhm / 2.355
    lo = np.round(wvl0 - 6 * sigma, 2)
    hi = np.round(wvl0 + 6 * sigma, 2)
    centres = np.round(np.arange(lo, hi + 0.005, 0.01), 2)
    cdf = lambda x: 0.5 * (1 + erf((x - wvl0) / (sigma * sqrt(2))))
    out = [(float(c), cdf(c + 0.005) - cdf(c - 0.005)) for c in centres]
    out = [(c, p) for c, p in out if p > tail]
    tot = sum(p for _, p in out)
    out = [(c, p / tot) for c, p in out]
    if len(out) > max_bands:
        raise ValueError('distribution spans more than %d bins' % max_bands)
    return out


# ---------------------------------------------------------------- aspherical
# The Yang et al. (2013) ice tables are not distributed with the reference;
# this writes synthetic tables in their layout so that the aspherical
# (phase-matrix + Stokes) walk can run: an ice-like P11 (forward peak, 22 deg
# halo, weak backscatter) and smooth, bounded ratios P12/P11 .. P44/P11.
YANG_WVLS = np.round(np.arange(0.30, 2.0001, 0.05), 2)
YANG_WVLS_FAR_IR = np.round(np.arange(16.5, 30.001, 0.5), 2)
YANG_REFF = (50.0, 100.0, 200.0)


def yang_theta_grid():
    """207 angles [deg]: 0.1 deg steps to 2, 0.5 to 10, 1 deg to 180."""
    return np.concatenate([np.arange(0.0, 2.0, 0.1), np.arange(2.0, 10.0, 0.5), np.arange(10.0, 180.5, 1.0)])


def _hg(mu, g):
    return (1 - g * g) / (1 + g * g - 2 * g * mu) ** 1.5


def synthetic_phase_matrix(theta_deg, wvl, reff):
    """The six elements (P11 normalised to (1/2) int P11 dmu = 1; the others
    as P_ij itself, not divided by P11) of one synthetic particle."""
    th = np.deg2rad(theta_deg)
    mu = np.cos(th)
    g1 = 0.86 + 0.04 * np.sin(1.7 * wvl) + 0.01 * np.log(reff / 100.0)
    halo = np.exp(-((theta_deg - 22.0) / 1.5) ** 2)
    p11 = 0.82 * _hg(mu, g1) + 0.14 * _hg(mu, -0.25) + 0.6 * halo
    norm = 0.5 * np.trapezoid(p11[::-1], mu[::-1])
    p11 = p11 / norm
    s2 = np.sin(th) ** 2
    dol = 0.25 + 0.1 * np.cos(wvl)
    el = {'P11': p11,
          'P12': -dol * s2 / (1 + mu * mu) * p11,
          'P22': (0.92 + 0.08 * mu) * p11,
          'P33': (0.85 * mu + 0.1) * p11,
          'P43': 0.12 * np.sin(th) ** 3 * p11,
          'P44': (0.8 * mu + 0.15) * p11}
    return el


def write_synthetic_yang(optics_dir, shape='droxtal', roughness='smooth', wvls=YANG_WVLS, reffs=YANG_REFF):
    """Write isca.dat and P11..P44.dat under
    ``optics_dir/ice_optics/<band>/<shape>/<roughness>`` (band 0.2-15.25 or,
    for far-infrared ``wvls``, 16.4-99.0): rows ordered by wavelength then
    size; the projected area and volume give r_eff = 3/4 V / A exactly
    ``reffs``.  Returns the directory."""
    from . import optics
    band = '16.4-99.0' if float(np.min(wvls)) >= 16.4 else '0.2-15.25'
    d = os.path.join(optics_dir, 'ice_optics', band, optics.SHAPE_DIRS[shape],
                     optics.ROUGHNESS_DIRS[roughness])
    theta = yang_theta_grid()
    rows = dict(wvl=[], max_dim=[], volume=[], area=[], q_ext=[], ssa=[], g=[])
    els = {e: [] for e in optics.PM_ELEMENTS}
    for w in wvls:
        for r in reffs:
            area = 2.0 * r * r
            rows['wvl'].append(w)
            rows['max_dim'].append(2.6 * r)
            rows['volume'].append(area * r / 0.75)
            rows['area'].append(area)
            rows['q_ext'].append(2.0 + 0.02 * np.sin(3.0 * w))
            ab = 2e-4 * (1.0 + 8.0 * w * w) * (r / 100.0)      # stronger absorption in the infrared
            rows['ssa'].append(1.0 - ab if ab < 0.05 else 0.95 - 0.2 * ab / (1.0 + ab))
            el = synthetic_phase_matrix(theta, w, r)
            mu = np.cos(np.deg2rad(theta))
            rows['g'].append(float(0.5 * np.trapezoid((mu * el['P11'])[::-1], mu[::-1])))
            for e in optics.PM_ELEMENTS:
                els[e].append(el[e])
    optics.write_yang_tables(d, *[np.array(rows[k]) for k in ('wvl', 'max_dim', 'volume', 'area', 'q_ext',
                                                                'ssa', 'g')],
                             theta_deg=theta, elements={e: np.array(v) for e, v in els.items()})
    return d


SNOW_POLAR = dict(wvl=1.30, r_eff=100.0, depth_m=1.0, theta0=60.0, lambertian_bottom=True,
                  r_lambertian=0.5, shape='droxtal', roughness='smooth', stokes0=(1.0, 1.0, 0.0, 0.0))


def snow_polar_case(n_photons: int, optics_dir: str) -> Case:
    """The snow_1m geometry (1 m of snow at 300 kg/m^3, theta0 60 deg,
    Lambertian bottom R 0.5) of aspherical droxtals (synthetic Yang-layout
    tables written under ``optics_dir``, r_eff 100 um, 1.30 um) lit by
    linearly polarised light (Stokes (1, 1, 0, 0)): the phase-matrix walk."""
    from . import optics
    from .gpu import PolarTables
    p = SNOW_POLAR
    write_synthetic_yang(optics_dir, shape=p['shape'], roughness=p['roughness'])
    a = optics.aspherical_ssps(np.array([p['wvl']]), p['r_eff'], p['wvl'], optics_dir, p['shape'],
                               p['roughness'], False, RHO_ICE)
    key = float(a['wvls'][0])
    ext_cff = float(a['ext_cff_mss'][0]) * RHO_SNW
    x, Y = a['pm'].interp_arrays(key)
    return Case(tau_tot=p['depth_m'] * ext_cff, theta0_deg=p['theta0'], n_photons=int(n_photons),
                bands=[Band(wvl=key, g=float(a['g'][0]), ssa_ice=float(a['ssa'][0]), ext_cff=ext_cff)],
                lambertian_bottom=p['lambertian_bottom'], r_lambertian=p['r_lambertian'],
                polar=PolarTables(x=x, Y=[Y], cutoff=a['pm'].cutoff(p['wvl'], a['wvls']),
                                  stokes0=p['stokes0']))
