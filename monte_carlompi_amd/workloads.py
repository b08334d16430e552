"""Named workloads of the photon walk (sizes as the benchmarks quote them).

``snow_1m``: the single case of the reference's driver settings as the
benchmark names it -- monochromatic 1.30 um, ice spheres r_eff = 100 um,
rho_snw = 300, rho_ice = 917 kg/m^3, Q_ext = 2, omega = 0.999, g = 0.89,
1.0 m of snow (tau = 4907.3, effectively semi-infinite), theta0 = 60 deg,
phi0 = 0, Lambertian bottom R = 0.5 (monte_carlo3D-run.py:9-12).
"""
from __future__ import annotations

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
