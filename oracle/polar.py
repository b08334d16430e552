"""Python/numpy oracle of the aspherical-ice walk (TEST INFRASTRUCTURE ONLY).

Restates, for non-spherical particles (``monte_carloMPI/monte_carlo3D.py``,
abbreviated mc3D):

* ``get_aspherical_SSPs``             mc3D:170-496  (``aspherical_ssps``)
* ``interpolate_phase_matrix``        mc3D:819-883  (``polar_tables``)
* ``populate_pdfs`` aspherical branch mc3D:898-1004, 1010-1025 (``draw_step``)
* ``rotate_stokes_vector``            mc3D:1090-1109 (``rotate``)
* Stokes update of the step loop      mc3D:1283-1342 (``stokes_update``)

The walk itself is ``walk.walk_photon`` with ``polar=`` given.  Arithmetic is
numpy float64 in the reference's operation order, with ``math.cos`` /
``math.sin`` where the reference calls them (the rejection test) and numpy's
elsewhere, so that for one deviate stream the records and the final Stokes
vectors equal the reference's bit for bit (tests/golden/golden_polar.*).

Interpolation: the reference's ``interp1d(cos_theta[::-1], P[::-1])`` on a
float64 1-D table delegates to ``np.interp`` (scipy 1.15,
``interp1d._call_linear_np``), which is what is called here.
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import walk

TWO_PIE = 2 * np.pi          # mc3D:26
FOUR_PIE = 4 * np.pi         # mc3D:27
ELEMENTS = ('P11', 'P12', 'P22', 'P33', 'P43', 'P44')


# --------------------------------------------------------------------------
# tables
# --------------------------------------------------------------------------
def aspherical_ssps(directory, wvl0, rds_snw, wvls, HG, rho_ice):
    """mc3D:219-496 on the files of ``directory`` (isca.dat, P11..P44.dat).

    Returns (wvls replaced by the nearest table wavelengths, ssa, ext_cff_mss,
    g, r_eff, theta_deg or None, {key: {element: array}} or None)."""
    rows = [line.split() for line in open(os.path.join(directory, 'isca.dat')).readlines()]
    wvl_in = np.array([float(r[0]) for r in rows])
    vol_in = np.array([float(r[2]) for r in rows])
    G_in = np.array([float(r[3]) for r in rows])
    Q_in = np.array([float(r[4]) for r in rows])
    ssa_in = np.array([float(r[5]) for r in rows])
    asm_in = np.array([float(r[6]) for r in rows])
    wvl0_idxs = np.where(wvl_in == wvl0)
    RE = (3. / 4.) * (np.array(vol_in) / np.array(G_in))
    idx_RE = np.argsort(np.absolute(rds_snw - RE[wvl0_idxs]))[0]
    valid = np.where(RE == RE[wvl0_idxs][idx_RE])
    reff = RE[wvl0_idxs][idx_RE]
    wvl_in, vol_in, G_in, Q_in, ssa_in, asm_in = (a[valid] for a in (wvl_in, vol_in, G_in, Q_in,
                                                                        ssa_in, asm_in))
    lines = {}
    line0 = None
    if not HG:
        for e in ELEMENTS:
            with open(os.path.join(directory, e + '.dat')) as fh:
                l0 = fh.readline().split()
                lines[e] = np.array(fh.readlines())[valid]
            if e == 'P11':
                line0 = l0
    out_w = np.empty(wvls.shape)
    ssa = np.empty(wvls.shape)
    ext = np.empty(wvls.shape)
    g = np.empty(wvls.shape)
    P = {}
    last = None
    for wvl in np.sort(wvls):
        if wvl == last:
            continue
        last = wvl
        ws = np.where(wvls == wvl)
        idx = np.argsort(np.absolute(wvl - wvl_in))
        key = wvl_in[idx[:2]][0]
        out_w[ws] = key
        ssa[ws] = ssa_in[idx[0]]
        ext[ws] = (1e6 * G_in[idx[0]] * Q_in[idx[0]]) / (rho_ice * vol_in[idx[0]])
        g[ws] = asm_in[idx[0]]
        if not HG:
            p11 = np.empty(len(line0))
            for j in range(len(line0)):
                p11[j] = float(lines['P11'][idx[0]].split()[j])
            P[key] = {'P11': p11}
            for e in ELEMENTS[1:]:
                v = np.empty(len(line0))
                for j in range(len(line0)):
                    v[j] = float(lines[e][idx[0]].split()[j])
                P[key][e] = v * p11
    theta = np.array([float(t) for t in line0]) if line0 is not None else None
    return out_w, ssa, ext, g, reff, theta, (P if not HG else None)


def polar_tables(theta_deg, P, wvls, wvl0):
    """mc3D:819-883: the interpolation grid (ascending cos theta), the
    rejection split ``cutoff_idx`` on P11 at wvl0 (or the last sorted
    wavelength), and per wavelength the six elements on the grid and the area
    maxima of populate_pdfs (mc3D:926-941).  Returns (case dict, {key: band dict})."""
    cos_t = np.cos((theta_deg * np.pi) / 180.)
    sorted_w = np.sort(wvls)
    wvl0_exists = bool(np.any(sorted_w == wvl0))
    wvl = wvl0 if wvl0_exists else sorted_w[-1]
    P11 = P[wvl]['P11']
    cmax = cos_t[0]
    cmin = cos_t[-1]
    pct = []
    n = len(cos_t)
    for i, c in enumerate(cos_t):
        if i > 0 and i < n - 1:
            a1 = (cmax - c) * P11[:i + 1].max()
            a2 = (c - cmin) * P11[i:].max()
            pct.append(a1 / (a1 + a2))
        else:
            pct.append(1e5)
    cut = int(np.argsort(np.absolute(0.5 - np.array(pct)))[0])
    case = dict(cmax=cmax, cmin=cmin, ccut=cos_t[cut], cut=cut, x=cos_t[::-1])
    bands = {}
    for key in np.unique(wvls):
        el = P[key]
        bands[key] = dict(Y={e: el[e][::-1] for e in ELEMENTS},
                          A1=np.absolute(el['P11'][:cut + 1]).max(), B1=np.absolute(el['P12'][:cut + 1]).max(),
                          A2=np.absolute(el['P11'][cut:]).max(), B2=np.absolute(el['P12'][cut:]).max())
    return case, bands


# --------------------------------------------------------------------------
# sampling and the Stokes update
# --------------------------------------------------------------------------
def draw_step(src, pol, stokes):
    """The per-step draws of ``populate_pdfs`` for an aspherical particle, in
    order (mc3D:898-1023): area, cos theta, phi, then the rejection loop
    (r3; after a rejection: area, cos theta, 2 phi, r3), then -ln u, u_ssa,
    u_ext.  The target density is I S11 + S12 (Q cos 2phi + U sin 2phi) with
    the photon's current Stokes vector.  Returns (cos theta, tau, phi, ssa, ext)."""
    I = stokes[0]
    Q = stokes[1]
    U = stokes[2]
    if Q == 0:
        two_beta = np.arctan(U / 1e-15)
    else:
        two_beta = np.arctan(U / Q)
    x, Y = pol['x'], pol['Y']
    cmax, cmin, ccut = pol['cmax'], pol['cmin'], pol['ccut']
    A = np.absolute(I) * pol['A1']
    B = pol['B1']
    C = np.absolute(Q * np.cos(two_beta) + U * np.sin(two_beta))
    max_val1 = A + B * C
    A = np.absolute(I) * pol['A2']
    B = pol['B2']
    max_val2 = A + B * C
    area1 = (cmax - ccut) * max_val1
    area2 = (ccut - cmin) * max_val2
    percent_area1 = area1 / (area1 + area2)
    area_rand = src.rand1()
    cos_theta_rand = src.rand1()
    phi_rand = src.rand1() * TWO_PIE
    two_phi = (2 * phi_rand)[0]
    area_rand_j = area_rand[0]
    cos_theta = cos_theta_rand[0]
    if area_rand_j < percent_area1:
        area = 1
        cos_theta = cos_theta * (cmax - ccut) + ccut
    else:
        area = 2
        cos_theta = -cos_theta * (ccut - cmin) + ccut
    r3 = 1
    phase_func_val = 0
    k = 0
    while r3 > phase_func_val:
        if k > 0:
            area_rand_j = src.rand()
            if area_rand_j < percent_area1:
                area = 1
                cos_theta = src.rand() * (cmax - ccut) + ccut
            else:
                area = 2
                cos_theta = -src.rand() * (ccut - cmin) + ccut
            two_phi = src.rand() * FOUR_PIE
        r3 = 1.000001 * src.rand()
        if area == 1:
            r3 = r3 * max_val1
        elif area == 2:
            r3 = r3 * max_val2
        S11 = np.interp(cos_theta, x, Y['P11'])
        S12 = np.interp(cos_theta, x, Y['P12'])
        phase_func_val = I * S11 + S12 * (Q * math.cos(two_phi) + U * math.sin(two_phi))
        k += 1
    phi = two_phi / 2.
    tau = -np.log(src.rand1())
    ssa = src.rand1()
    ext = src.rand1()
    return cos_theta, tau[0], phi, ssa[0], ext[0]


def rotate(angle, stokes):
    """mc3D:1090-1109: the Stokes vector in a frame rotated by ``angle``."""
    I_old, Q_old, U_old, V_old = stokes[0], stokes[1], stokes[2], stokes[3]
    double_angle = 2 * angle
    c = np.cos(double_angle)
    s = np.sin(double_angle)
    return (I_old, Q_old * c + U_old * s, -Q_old * s + U_old * c, V_old)


def stokes_update(pol, stokes, costheta, cosphi, muz_n, muz_0):
    """mc3D:1287-1342: rotate into the scattering plane (by arccos(cos phi),
    which folds phi into [0, pi]), scatter by the phase matrix at cos theta,
    rotate back to the new meridian plane, normalise to I = 1.  The sign of
    the return rotation is +1 only for phi_sca >= pi, as in the reference."""
    x, Y = pol['x'], pol['Y']
    cos2theta = costheta ** 2
    phi_sca = np.arccos(cosphi)
    I_sp, Q_sp, U_sp, V_sp = rotate(phi_sca, stokes)
    P11 = np.interp(costheta, x, Y['P11'])
    P12 = np.interp(costheta, x, Y['P12'])
    P22 = np.interp(costheta, x, Y['P22'])
    P33 = np.interp(costheta, x, Y['P33'])
    P43 = np.interp(costheta, x, Y['P43'])
    P44 = np.interp(costheta, x, Y['P44'])
    I_sca = I_sp * P11 + Q_sp * P12
    Q_sca = I_sp * P12 + Q_sp * P22
    U_sca = U_sp * P33 - V_sp * P43
    V_sca = U_sp * P43 + V_sp * P44
    den = np.sqrt((1 - cos2theta) * (1 - muz_n ** 2))
    if den == 0:
        cos_i = 0
    else:
        num = muz_n * costheta - muz_0
        sign = 1 if phi_sca >= np.pi else -1
        cos_i = num / (sign * den)
        if cos_i < -1:
            cos_i = -1
        elif cos_i > 1:
            cos_i = 1
    i_2 = np.arccos(cos_i)
    I_m, Q_m, U_m, V_m = rotate(-i_2, (I_sca, Q_sca, U_sca, V_sca))
    return np.array([I_m, Q_m, U_m, V_m]) / I_m


# --------------------------------------------------------------------------
# run(): one rank
# --------------------------------------------------------------------------
def polar_run(seed, n_photon, wvl0, half_width, rds_snw, directory, HG, stokes0, imp_fn, imp_cnc,
              tau_tot, rho_snw, rho_ice, theta_0_deg, lambertian_bottom, r_lambertian,
              lambertian_surface=False):
    """Single-rank run() of an aspherical case with ``RandomState(seed)``
    (mc3D:1492-1619).  ``imp_fn(wvls) -> dict(ssa, ext)`` gives the impurity
    optics.  Returns (records (10-tuples with Q, U, V unless HG), params,
    CSR replay stream (values, offsets), reff, case tables, band tables)."""
    rs = np.random.RandomState(seed)
    wvls = walk.draw_wavelengths(rs, n_photon, wvl0, half_width)
    wvls, ssa, ext, g, reff, theta, P = aspherical_ssps(directory, wvl0, rds_snw, wvls, HG, rho_ice)
    imp = imp_fn(wvls)
    par = walk.photon_params(wvls, dict(ssa=ssa, ext=ext, g=g), imp, imp_cnc, tau_tot, rho_snw)
    case = bands = None
    if not HG:
        case, bands = polar_tables(theta, P, wvls, wvl0)
    src = walk.RefStream(rs, record=True)
    theta_0 = (np.pi * theta_0_deg) / 180.
    tau0, ssa0, ext0 = walk.initial_triples(src, n_photon)
    starts = []
    records = []
    for k in range(n_photon):
        starts.append(len(src.log))
        pol = None
        if not HG:
            b = bands[wvls[k]]
            pol = dict(case, Y=b['Y'], A1=b['A1'], B1=b['B1'], A2=b['A2'], B2=b['B2'],
                       stokes0=np.asarray(stokes0))
        records.append(walk.walk_photon(src, (tau0[k], ssa0[k], ext0[k]), par['wvl'][k], par['g'][k],
                                        par['ssa_ice'][k], par['ssa_imp'][k], par['ext_cff_mss'][k],
                                        par['p_ext_imp'][k], par['snow_depth'][k], tau_tot, rho_snw,
                                        theta_0, lambertian_bottom, r_lambertian, trajectory=False,
                                        lambertian_surface=lambertian_surface, polar=pol))
    log = np.asarray(src.log, dtype=np.float64)
    init = log[:3 * n_photon].reshape(n_photon, 3)
    bounds = starts + [log.size]
    slices = [np.concatenate([init[k], log[bounds[k]:bounds[k + 1]]]) for k in range(n_photon)]
    offsets = np.zeros(n_photon + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([s.size for s in slices])
    values = np.concatenate(slices) if slices else np.zeros(0)
    return records, par, values, offsets, reff, case, bands
