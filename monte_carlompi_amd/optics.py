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
