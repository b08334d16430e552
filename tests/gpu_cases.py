"""Build GPU cases from per-photon oracle parameters (test helper)."""
import numpy as np

from monte_carlompi_amd.gpu import Band, Case


def case_from_params(par, tau_tot, rho_snw, theta0, lb, R, ls=False):
    """One GPU case whose bands are the unique photon wavelengths; returns
    (case, per-photon band index)."""
    wvl = par['wvl']
    uniq, inv = np.unique(wvl, return_inverse=True)
    bands = []
    for j, w in enumerate(uniq):
        k = np.nonzero(inv == j)[0][0]
        bands.append(Band(wvl=float(w), g=float(par['g'][k]), ssa_ice=float(par['ssa_ice'][k]),
                          ext_cff=float(par['ext_cff_mss'][k] * rho_snw),
                          ssa_imp=float(par['ssa_imp'][k]), p_ext_imp=float(par['p_ext_imp'][k]),
                          weight=float(np.sum(inv == j))))
    case = Case(tau_tot=tau_tot, theta0_deg=theta0, n_photons=int(wvl.size), bands=bands,
                lambertian_bottom=bool(lb), r_lambertian=float(R), lambertian_surface=bool(ls))
    return case, inv.astype(np.uint8)


def flat_params(n, wvl, g, ssa, ext_cff_mss, tau_tot, rho_snw):
    return dict(wvl=np.full(n, wvl), g=np.full(n, g), ssa_ice=np.full(n, ssa),
                ssa_imp=np.zeros(n), ext_cff_mss=np.full(n, ext_cff_mss),
                p_ext_imp=np.zeros(n), snow_depth=np.full(n, tau_tot / (ext_cff_mss * rho_snw)))
