"""Single-scattering properties (SSPs) per photon wavelength.

Two sources:
* SNICAR-style NetCDF tables (``optics_dir/mie/snicar/ice_wrn_%04d.nc`` and
  the impurity file), read with scipy.io.netcdf_file and interpolated
  linearly between the two grid wavelengths nearest to each photon
  wavelength, falling back to the nearest value when the query lies outside
  those two points -- the behaviour of get_optical_properties /
  get_impurity_optics (monte_carloMPI/monte_carlo3D.py:498-777);
* parametric ice spheres (the archived tables are absent here):
  ext_cff_mss = 3 Q_ext / (4 rho_ice r_eff), with omega and g given.

Lookups are done once per unique wavelength; the walk consumes the result as
one band row per wavelength.
"""
from __future__ import annotations

import os
import sys
from dataclasses import dataclass

import numpy as np


@dataclass
class SpectralTable:
    wvl_m: np.ndarray                  # grid wavelengths [m]
    columns: dict                      # name -> values on the grid

    @classmethod
    def from_netcdf(cls, path, names=('ss_alb', 'ext_cff_mss', 'asm_prm')):
        from scipy.io import netcdf_file
        with netcdf_file(path, 'r', mmap=False) as f:
            wvl = np.array(f.variables['wvl'].data, dtype=np.float64)
            cols = {k: np.array(f.variables[k].data, dtype=np.float64)
                    for k in names if k in f.variables}
        return cls(wvl, cols)

    def lookup(self, name, wvl_um):
        """Value of column `name` at photon wavelength `wvl_um` [um]."""
        x = wvl_um * 1e-6
        order = np.argsort(np.absolute(x - self.wvl_m))
        xs = self.wvl_m[order[:2]]
        ys = self.columns[name][order[:2]]
        if xs[0] > xs[1]:
            xs, ys = xs[::-1], ys[::-1]
        if xs[0] <= x <= xs[1] and xs[1] > xs[0]:
            slope = (ys[1] - ys[0]) / (xs[1] - xs[0])
            return float(slope * (x - xs[0]) + ys[0])
        sys.stderr.write('error: exception raised while interpolating %s, using nearest value '
                         'instead\n' % name)
        return float(self.columns[name][order[0]])


def sphere_ext_cff_mss(r_eff_um, rho_ice=917.0, q_ext=2.0):
    """3 Q_ext / (4 rho_ice r_eff) [m^2/kg]."""
    return 3.0 * q_ext / (4.0 * rho_ice * (r_eff_um * 1e-6))


def ice_file(optics_dir, rds_snw):
    return os.path.join(optics_dir, 'mie', 'snicar', 'ice_wrn_%04d.nc' % rds_snw)


def impurity_file(optics_dir, fi_imp):
    return os.path.join(optics_dir, 'mie', 'snicar', fi_imp)


def ice_ssps(unique_wvls, rds_snw, optics_dir=None, parametric=None):
    """(ssa, ext_cff_mss, g) arrays over ``unique_wvls``.  Uses the NetCDF
    table when it exists, else ``parametric`` = dict(ssa, g, q_ext, rho_ice)."""
    path = ice_file(optics_dir, rds_snw) if optics_dir else None
    if path and os.path.exists(path):
        t = SpectralTable.from_netcdf(path)
        return (np.array([t.lookup('ss_alb', w) for w in unique_wvls]),
                np.array([t.lookup('ext_cff_mss', w) for w in unique_wvls]),
                np.array([t.lookup('asm_prm', w) for w in unique_wvls]))
    if parametric is None:
        raise FileNotFoundError('%s not found and no parametric optics given' % path)
    n = len(unique_wvls)
    ext = sphere_ext_cff_mss(rds_snw, parametric.get('rho_ice', 917.0), parametric.get('q_ext', 2.0))
    return (np.full(n, float(parametric['ssa'])), np.full(n, ext), np.full(n, float(parametric['g'])))


def impurity_ssps(unique_wvls, optics_dir=None, fi_imp=None, parametric=None):
    """(ssa_imp, ext_cff_mss_imp) over ``unique_wvls`` (NetCDF if present)."""
    path = impurity_file(optics_dir, fi_imp) if (optics_dir and fi_imp) else None
    if path and os.path.exists(path):
        t = SpectralTable.from_netcdf(path, names=('ss_alb', 'ext_cff_mss'))
        return (np.array([t.lookup('ss_alb', w) for w in unique_wvls]),
                np.array([t.lookup('ext_cff_mss', w) for w in unique_wvls]))
    p = parametric or {}
    n = len(unique_wvls)
    return np.full(n, float(p.get('ssa', 0.0))), np.full(n, float(p.get('ext', 0.0)))


# --------------------------------------------------------------------------
# aspherical ice: Yang et al. (2013) single-scattering tables
# --------------------------------------------------------------------------
# directory names of mc3D:185-211
SHAPE_DIRS = {'solid hexagonal column': 'solid_column', 'hexagonal plate': 'plate',
              'hollow hexagonal column': 'hollow_column', 'droxtal': 'droxtal',
              'hollow bullet rosette': 'hollow_bullet_rosette',
              'solid bullet rosette': 'solid_bullet_rosette',
              '8-element column aggregate': 'column_8elements',
              '5-element plate aggregate': 'plate_5elements',
              '10-element plate aggregate': 'plate_10elements'}
ROUGHNESS_DIRS = {'smooth': 'Rough000', 'moderately rough': 'Rough003', 'severely rough': 'Rough050'}
PM_ELEMENTS = ('P11', 'P12', 'P22', 'P33', 'P43', 'P44')


def aspherical_dir(optics_dir, shape, roughness, wvl0):
    """``optics_dir/ice_optics/<band>/<shape>/<roughness>`` (mc3D:176-217,
    1525-1545): band 0.2-15.25 um or 16.4-99.0 um by the centre wavelength."""
    if 0.2 <= wvl0 <= 15.25:
        band = '0.2-15.25'
    elif 16.4 <= wvl0 <= 99.0:
        band = '16.4-99.0'
    else:
        raise ValueError('wvl0 = %r um lies outside the aspherical tables (0.2-15.25, 16.4-99.0 um)' % wvl0)
    if shape not in SHAPE_DIRS:
        raise ValueError('unknown shape %r' % (shape,))
    if roughness not in ROUGHNESS_DIRS:
        raise ValueError('unknown roughness %r' % (roughness,))
    return os.path.join(optics_dir, 'ice_optics', band, SHAPE_DIRS[shape], ROUGHNESS_DIRS[roughness])


@dataclass
class PhaseMatrix:
    """Scattering phase matrix of one particle size at several wavelengths,
    as the reference holds it (mc3D:267-496): ``theta_deg`` the angle grid of
    the files (first line), ``P[key][e]`` element e of wavelength ``key`` on
    that grid, P12..P44 already multiplied by P11 (the files hold them
    divided by P11)."""
    theta_deg: np.ndarray
    P: dict

    def interp_arrays(self, key):
        """(x, Y): the ascending cos(theta) grid and the six elements on it --
        the arrays interp1d(cos_theta[::-1], P[::-1]) holds (mc3D:842-857)."""
        x = np.cos((self.theta_deg * np.pi) / 180.)[::-1]
        return x, np.stack([self.P[key][e][::-1] for e in PM_ELEMENTS])

    def cutoff(self, wvl0, keys):
        """Index (file order) of cos_theta_cutoff, the split of the rejection
        sampler's two areas where area 1 = (cos_max - cos) max P11[:i+1] is
        nearest half of the total (mc3D:858-883): on P11 at the centre
        wavelength when it is a table wavelength, else at the largest key."""
        keys = np.sort(np.asarray(keys))
        wvl = wvl0 if np.any(keys == wvl0) else keys[-1]
        P11 = self.P[wvl]['P11']
        cos_t = np.cos((self.theta_deg * np.pi) / 180.)
        cmax, cmin = cos_t[0], cos_t[-1]
        n = len(cos_t)
        pct = []
        for i, c in enumerate(cos_t):
            if 0 < i < n - 1:
                a1 = (cmax - c) * P11[:i + 1].max()
                a2 = (c - cmin) * P11[i:].max()
                pct.append(a1 / (a1 + a2))
            else:
                pct.append(1e5)
        return int(np.argsort(np.absolute(0.5 - np.array(pct)))[0])


def read_isca(path):
    """Columns of isca.dat: wavelength [um], maximum dimension [um], volume
    [um^3], projected area [um^2], Q_ext, omega, g (mc3D:219-245)."""
    rows = [[float(v) for v in line.split()[:7]] for line in open(path) if line.strip()]
    a = np.array(rows, dtype=np.float64).reshape(-1, 7)
    return {k: a[:, j] for j, k in enumerate(('wvl', 'max_dim', 'volume', 'area', 'q_ext', 'ssa', 'g'))}


def aspherical_ssps(wvls, rds_snw, wvl0, optics_dir, shape, roughness, HG, rho_ice=917.0):
    """``get_aspherical_SSPs`` (mc3D:170-496) over per-photon wavelengths.

    Picks the particle size whose effective radius 3/4 V / A at ``wvl0`` is
    nearest ``rds_snw`` (``wvl0`` must be a table wavelength, as in the
    reference), then for every photon the nearest table wavelength, without
    interpolation.  Returns a dict: ``wvls`` (the photons' wavelengths
    replaced by those table wavelengths, as mc3D:1535-1536 does), per-photon
    ``ssa``, ``ext_cff_mss`` = 1e6 A Q_ext / (rho_ice V), ``g``, the chosen
    ``snow_effective_radius`` and ``pm`` (a PhaseMatrix, or None with HG)."""
    d = aspherical_dir(optics_dir, shape, roughness, wvl0)
    t = read_isca(os.path.join(d, 'isca.dat'))
    wvl0_idxs = np.where(t['wvl'] == wvl0)
    if wvl0_idxs[0].size == 0:
        raise ValueError('wvl0 = %r um is not a wavelength of %s (the reference requires it)'
                         % (wvl0, os.path.join(d, 'isca.dat')))
    RE = (3. / 4.) * (t['volume'] / t['area'])
    idx_RE = np.argsort(np.absolute(rds_snw - RE[wvl0_idxs]))[0]
    valid = np.where(RE == RE[wvl0_idxs][idx_RE])
    reff = RE[wvl0_idxs][idx_RE]
    t = {k: v[valid] for k, v in t.items()}
    pm_rows = None
    theta = None
    if not HG:
        pm_rows = {}
        for e in PM_ELEMENTS:
            with open(os.path.join(d, e + '.dat')) as fh:
                line0 = fh.readline().split()
                lines = np.array(fh.readlines())[valid]
            th = np.array([float(v) for v in line0])
            if theta is None:
                theta = th
            elif th.shape != theta.shape or np.any(th != theta):
                raise ValueError('%s.dat: its angle grid differs from P11.dat (one grid is supported)' % e)
            pm_rows[e] = lines
    wvls = np.asarray(wvls, dtype=np.float64)
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
        idx = np.argsort(np.absolute(wvl - t['wvl']))[0]
        key = t['wvl'][idx]
        out_w[ws] = key
        ssa[ws] = t['ssa'][idx]
        ext[ws] = (1e6 * t['area'][idx] * t['q_ext'][idx]) / (rho_ice * t['volume'][idx])
        g[ws] = t['g'][idx]
        if pm_rows is not None:
            p11 = np.array([float(v) for v in pm_rows['P11'][idx].split()])
            el = {'P11': p11}
            for e in PM_ELEMENTS[1:]:
                el[e] = np.array([float(v) for v in pm_rows[e][idx].split()]) * p11
            P[key] = el
    pm = PhaseMatrix(theta, P) if pm_rows is not None else None
    return dict(wvls=out_w, ssa=ssa, ext_cff_mss=ext, g=g, snow_effective_radius=reff, pm=pm)


def write_yang_tables(directory, wvl, max_dim, volume, area, q_ext, ssa, g, theta_deg, elements):
    """Write tables in the layout the reader above (and the reference) parses:
    ``isca.dat`` rows (wvl, max_dim, volume, area, Q_ext, omega, g) and
    ``P11.dat`` .. ``P44.dat`` with the angle grid on the first line and one
    row per isca.dat row (P11 itself, the others divided by P11).
    ``elements[e]`` is an array rows x angles."""
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, 'isca.dat'), 'w') as fh:
        for row in zip(wvl, max_dim, volume, area, q_ext, ssa, g):
            fh.write(' '.join('%.6e' % v for v in row) + '\n')
    for e in PM_ELEMENTS:
        with open(os.path.join(directory, e + '.dat'), 'w') as fh:
            fh.write(' '.join('%.4f' % v for v in theta_deg) + '\n')
            vals = elements[e] if e == 'P11' else elements[e] / elements['P11']
            for row in vals:
                fh.write(' '.join('%.6e' % v for v in row) + '\n')
