"""GPU-backed drop-in for monte_carloMPI/monte_carlo3D.py's ``MonteCarlo``.

Same construction (``MonteCarlo(**model_kwargs)`` over config.ini defaults
and argparse flags, mc3D:43-94, 1778-1843), same ``run()`` signature
(mc3D:1492-1496), same output file naming (mc3D:96-143) and per-photon
record ``condition wvn theta_n phi_n n_scat path_length snow_depth``
(mc3D:1621-1648).  What changes is the region mc3D:1521 / 1594-1619: the
Parallel scatter, the per-rank photon loop over ``monte_carlo3D`` and the
gather are one call of the HIP walk (libmcsnow.so through ctypes).

Modes (keyword ``mode`` of ``run``, default from ``--mode`` / config):
* ``native``: FP32 walk with a Philox4x32-10 stream per photon;
* ``replay``: FP64 walk consuming a given uniform-deviate stream in the
  reference's draw order (``replay=(values, offsets)``), which reproduces the
  reference's records photon for photon.

Scope: spheres / Henyey-Greenstein and aspherical grains (Yang et al. 2013
tables: the HG phase function with HG=True, else the scattering phase matrix
with Stokes vectors), with the Lambertian bottom and the Lambertian surface
modes.  Debug trajectory plots and phase-function plots of the reference
raise NotImplementedError.
"""
from __future__ import annotations

import argparse
import configparser
import os
import sys

import numpy as np

from . import optics
from .gpu import Band, Case, GpuWalker, PolarTables
from .parallelize import Parallel

TWO_PIE = 2 * np.pi


def wavelength_bins(wvls):
    """(uniq, inverse) of the photons' wavelengths, equal to
    ``np.unique(wvls, return_inverse=True)``.  Wavelengths on the 0.01 um grid
    (the ``np.around(..., 2)`` draws of mc3D:1519-1520) are binned by their
    integer hundredths in O(n); any other input goes through np.unique."""
    w = np.asarray(wvls, dtype=np.float64).ravel()
    if w.size:
        m = w * 100.0
        np.rint(m, out=m)
        lo, hi = m.min(), m.max()
        if lo == hi and np.all(w == lo / 100.0):           # one bin (a narrow distribution)
            return w[:1].copy(), np.zeros(w.size, dtype=np.int64)
        if np.isfinite(lo) and np.isfinite(hi) and hi - lo < 65536 and np.array_equal(m / 100.0, w):
            k = (m - lo).astype(np.int64)
            present = np.flatnonzero(np.bincount(k))
            lut = np.zeros(int(hi - lo) + 1, dtype=np.int64)
            lut[present] = np.arange(present.size)
            return (present + lo) / 100.0, lut[k]
    return np.unique(w, return_inverse=True)
_DEFAULTS = {'tau_tot': 1000000.0, 'imp_cnc': 0.0, 'rho_snw': 300.0, 'rho_ice': 917.0,
             'flg_crt': 1, 'flg_3D': 999, 'output_dir': 'monte_carlo_results',
             'optics_dir': 'inputdata', 'fi_imp': 'mie_sot_ChC90_dns_1317.nc'}


class MonteCarlo(object):
    def __init__(self, **model_kwargs):
        """Valid model_kwargs: those of the reference (tau_tot, imp_cnc,
        rho_snw, rho_ice, rsensor, hsensor, flg_crt, flg_3D, output_dir,
        optics_dir, fi_imp, HG, phase_functions) plus mode, seed, device,
        depth_m (snow depth [m]; overrides tau_tot), ssa_ice, g, q_ext
        (parametric optics when the NetCDF tables are absent)."""
        args = self.get_model_args()
        d = dict(vars(args))
        for k, v in model_kwargs.items():
            d[k] = v
        for k in ('tau_tot', 'imp_cnc', 'rho_snw', 'rho_ice', 'rsensor', 'hsensor', 'flg_crt',
                  'flg_3D', 'output_dir', 'optics_dir', 'fi_imp', 'HG', 'phase_functions', 'mode',
                  'seed', 'device', 'depth_m', 'ssa_ice', 'g', 'q_ext'):
            setattr(self, k if k not in ('ssa_ice', 'g') else '_param_' + k, d.get(k))
        self.shape = 'sphere'                      # set by run() (mc3D:1509-1510)
        self.roughness = 'smooth'
        self.initial_stokes_params = np.array([1, 0, 0, 0])
        self._walker = None

    # ------------------------------------------------------------------ config
    def get_model_args(self):
        """config.ini (CWD) then command-line flags (mc3D:1778-1843), plus the
        GPU keys of the ``[gpu]`` section: mode, seed, device, depth_m,
        ssa_ice, g, q_ext."""
        config = configparser.ConfigParser()
        config.read('config.ini')

        def get(section, key, conv, default):
            try:
                return conv(config.get(section, key))
            except (configparser.Error, ValueError):
                return default

        mp, po, da, gp = 'model parameters', 'plot options', 'data', 'gpu'
        p = argparse.ArgumentParser(description='MI355X photon random walk in snow')
        p.add_argument('--tau_tot', type=float, default=get(mp, 'tau_tot', float, _DEFAULTS['tau_tot']),
                       help='snow optical depth')
        p.add_argument('--imp_cnc', type=float, default=get(mp, 'imp_cnc', float, _DEFAULTS['imp_cnc']),
                       help='mass concentration of impurity [mIMP/(mIMP+mICE)]')
        p.add_argument('--rho_snw', type=float, default=get(mp, 'rho_snw', float, _DEFAULTS['rho_snw']),
                       help='snow density (kg/m3)')
        p.add_argument('--rho_ice', type=float, default=get(mp, 'rho_ice', float, _DEFAULTS['rho_ice']),
                       help='ice density (kg/m3)')
        p.add_argument('--rsensor', type=float, default=None, help='sensor radius [m]')
        p.add_argument('--hsensor', type=float, default=None, help='sensor height above snow [m]')
        p.add_argument('--flg_crt', type=int, default=get(po, 'flg_crt', int, _DEFAULTS['flg_crt']))
        p.add_argument('--flg_3D', type=int, default=get(po, 'flg_3D', int, _DEFAULTS['flg_3D']))
        p.add_argument('--output_dir', type=str, default=get(da, 'output_dir', str, _DEFAULTS['output_dir']))
        p.add_argument('--optics_dir', type=str, default=get(da, 'optics_dir', str, _DEFAULTS['optics_dir']))
        p.add_argument('--fi_imp', type=str, default=get(da, 'fi_imp', str, _DEFAULTS['fi_imp']))
        p.add_argument('--HG', action='store_true', help='Henyey-Greenstein phase function')
        p.add_argument('--phase_functions', action='store_true', help='plot phase functions')
        p.add_argument('--mode', type=str, default=get(gp, 'mode', str, 'native'),
                       choices=('native', 'replay'))
        p.add_argument('--seed', type=int, default=get(gp, 'seed', int, None))
        p.add_argument('--device', type=int, default=get(gp, 'device', int, 0))
        p.add_argument('--depth_m', type=float, default=get(gp, 'depth_m', float, None),
                       help='snow depth [m]; sets tau_tot = depth * ext_cff_mss * rho_snw')
        p.add_argument('--ssa_ice', type=float, default=get(gp, 'ssa_ice', float, None))
        p.add_argument('--g', type=float, default=get(gp, 'g', float, None))
        p.add_argument('--q_ext', type=float, default=get(gp, 'q_ext', float, 2.0))
        args, _unknown = p.parse_known_args()
        return args

    # ------------------------------------------------------------------ output
    def setup_output(self, n_photon, wvl0, half_width):
        """Output path ``output_dir/sphere/WVL_FWHM_REFF_N_THETA_HG[_i].txt``
        (aspherical grains: ``output_dir/<shape>/<roughness>/..._HG`` or
        ``..._I<IQUV>``), never overwriting (mc3D:96-143)."""
        output_dir = self.output_dir
        if not os.path.isdir(output_dir):
            os.mkdir(output_dir)
        if self.shape == 'sphere':
            save_dir = os.path.join(output_dir, 'sphere')
        else:
            save_dir = os.path.join(output_dir, optics.SHAPE_DIRS[self.shape],
                                    optics.ROUGHNESS_DIRS[self.roughness])
        os.makedirs(save_dir, exist_ok=True)
        theta0_deg = np.rad2deg(self.theta_0)
        if self.shape == 'sphere' or self.HG:
            appendix = 'HG'
        else:
            I, Q, U, V = self.initial_stokes_params[:4]
            appendix = 'I%d%d%d%d' % (I, Q, U, V)
        base = '%s_%s_%s_%s_%s_%s' % (wvl0, half_width, self.snow_effective_radius, n_photon,
                                      theta0_deg, appendix)
        output_file = os.path.join(save_dir, base + '.txt')
        i = 0
        while os.path.isfile(output_file):
            i += 1
            output_file = os.path.join(save_dir, '%s_%d.txt' % (base, i))
        return output_file

    HEADER = 'condition wvn[um^-1] theta_n phi_n n_scat path_length[m], snow_depth[m]\n'
    HEADER_STOKES = ('condition wvn[um^-1] theta_n phi_n n_scat path_length[m], snow_depth[m], Q_n, '
                     'U_n, V_n\n')

    @staticmethod
    def write_records(path, rec):
        """Header and rows of mc3D:1626-1646, written by the library
        (mcs_write_records_text: host threads, floats as Python's shortest
        round-trip repr -- the text ``%r`` produced for numpy float64 before
        numpy 2; numpy >= 2 would print ``np.float64(...)``, which
        post_processing.py cannot parse)."""
        import ctypes
        from ._lib import McsTextRecords, check, lib
        cols = dict(condition=np.ascontiguousarray(rec['condition'], dtype=np.int8),
                    n_scat=np.ascontiguousarray(rec['n_scat'], dtype=np.int64))
        for k in ('wvn', 'theta_n', 'phi_n', 'path_length', 'snow_depth'):
            cols[k] = np.ascontiguousarray(rec[k], dtype=np.float64)
        stokes = 'Q_n' in rec
        if stokes:
            for k, c in (('Q_n', 'q'), ('U_n', 'u'), ('V_n', 'v')):
                cols[c] = np.ascontiguousarray(rec[k], dtype=np.float64)
        n = cols['condition'].size
        if any(a.size != n for a in cols.values()):
            raise ValueError('record columns differ in length')
        R = McsTextRecords(*[cols[k].ctypes.data if k in cols else None for k in
                             ('condition', 'wvn', 'theta_n', 'phi_n', 'n_scat', 'path_length', 'snow_depth',
                              'q', 'u', 'v')])
        header = MonteCarlo.HEADER_STOKES if stokes else MonteCarlo.HEADER
        code = lib().mcs_write_records_text(os.fsencode(path), header.encode(), ctypes.byref(R), n)
        if code != 0:
            raise OSError('cannot write record file %s' % path)

    # ------------------------------------------------------------------ optics
    def get_optical_properties(self, wvls, rds_snw):
        """Per-photon (ssa_ice, ext_cff_mss_ice, g) (mc3D:498-658)."""
        self.snow_effective_radius = rds_snw
        uniq, inv = np.unique(wvls, return_inverse=True)
        param = None
        if self._param_ssa_ice is not None or self._param_g is not None:
            param = dict(ssa=self._param_ssa_ice if self._param_ssa_ice is not None else 0.999,
                         g=self._param_g if self._param_g is not None else 0.89,
                         q_ext=self.q_ext or 2.0, rho_ice=self.rho_ice)
        ssa, ext, g = optics.ice_ssps(uniq, rds_snw, self.optics_dir, param)
        return ssa[inv], ext[inv], g[inv]

    def get_impurity_optics(self, wvls):
        """Per-photon (ssa_imp, ext_cff_mss_imp) (mc3D:660-777).  Without the
        impurity table the impurity has no extinction (allowed for imp_cnc = 0)."""
        uniq, inv = np.unique(wvls, return_inverse=True)
        ssa, ext = optics.impurity_ssps(uniq, self.optics_dir, self.fi_imp)
        if not os.path.exists(optics.impurity_file(self.optics_dir, self.fi_imp)) and self.imp_cnc:
            raise FileNotFoundError('impurity optics %s needed for imp_cnc = %r'
                                    % (optics.impurity_file(self.optics_dir, self.fi_imp), self.imp_cnc))
        return ssa[inv], ext[inv]

    # ------------------------------------------------------------------ run
    def walker(self):
        if self._walker is None:
            self._walker = GpuWalker(self.device or 0)
        return self._walker

    def run(self, n_photon, wvl0, half_width, rds_snw, theta_0=0., stokes_params=np.array([1, 0, 0, 0]),
            shape='sphere', roughness='smooth', test=False, debug=False, Lambertian_surface=False,
            Lambertian_bottom=True, Lambertian_reflectance=1., replay=None, write=True):
        """Walk n_photon packets and write the per-photon record file.
        Mirrors mc3D:1492-1648; returns the output path (None if write=False).
        ``replay=(values, offsets)``: CSR uniform deviates of every photon in
        the reference's order (mode 'replay')."""
        if debug or self.phase_functions:
            raise NotImplementedError('debug trajectory and phase-function plots are not built')
        self.Lambertian_bottom = Lambertian_bottom
        self.R_Lambertian = Lambertian_reflectance
        self.theta_0 = (np.pi * theta_0) / 180.            # mc3D:1508
        self.shape = shape
        self.roughness = roughness
        self.wvl0 = wvl0
        self.initial_stokes_params = stokes_params

        scale = half_width / 2.355                         # mc3D:1516
        wvls = np.random.normal(loc=wvl0, scale=scale, size=(n_photon))   # mc3D:1519-1520
        np.around(wvls, decimals=2, out=wvls)
        par_wvls = Parallel(wvls)
        wvls = par_wvls.working_set

        # The reference evaluates the optics per photon (mc3D:1522-1612); they
        # depend on the wavelength alone, so here they are evaluated once per
        # wavelength bin (uniq) and the photons carry their bin (band).
        pm = None
        if shape == 'sphere':
            uniq, band = wavelength_bins(wvls)
            self._check_bins(uniq)
            ssa_ice, ext_ice, g = self.get_optical_properties(uniq, rds_snw)
        else:                                              # mc3D:1525-1545
            a = optics.aspherical_ssps(wvls, rds_snw, wvl0, self.optics_dir, shape, roughness,
                                       bool(self.HG), self.rho_ice)
            wvls = a['wvls']                               # nearest table wavelengths
            par_wvls.working_set = wvls
            uniq, first, band = np.unique(wvls, return_index=True, return_inverse=True)
            self._check_bins(uniq)
            ssa_ice, ext_ice, g = a['ssa'][first], a['ext_cff_mss'][first], a['g'][first]
            self.snow_effective_radius = a['snow_effective_radius']
            pm = a['pm']
        ssa_imp, ext_imp = self.get_impurity_optics(uniq)
        if test:                                           # mc3D:1553-1573
            for name in ('ssa_ice', 'ext_cff_mss_ice', 'g', 'ssa_imp', 'ext_cff_mss_imp'):
                if name in self.__dict__:
                    val = self.__dict__[name] * np.ones(uniq.size)
                    if name == 'ssa_ice': ssa_ice = val
                    elif name == 'ext_cff_mss_ice': ext_ice = val
                    elif name == 'g': g = val
                    elif name == 'ssa_imp': ssa_imp = val
                    else: ext_imp = val
        imp_cnc = self.imp_cnc
        ext_cff_mss = ext_ice * (1 - imp_cnc) + ext_imp * imp_cnc           # mc3D:1577
        with np.errstate(invalid='ignore', divide='ignore'):
            p_ext_imp = (imp_cnc * ext_imp) / (imp_cnc * ext_imp + (1 - imp_cnc) * ext_ice)
        p_ext_imp = np.nan_to_num(p_ext_imp)
        tau_tot = self.tau_tot
        if self.depth_m is not None:                       # the first photon's extinction
            tau_tot = float(self.depth_m * ext_cff_mss[band[0]] * self.rho_snw)
        self.tau_tot_used = tau_tot
        self.snow_depth = (tau_tot / (ext_cff_mss * self.rho_snw))[band]  # per photon, mc3D:1612

        weight = np.bincount(band, minlength=uniq.size)
        bands = [Band(wvl=float(uniq[j]), g=float(g[j]), ssa_ice=float(ssa_ice[j]),
                      ext_cff=float(ext_cff_mss[j] * self.rho_snw),
                      ssa_imp=float(ssa_imp[j]), p_ext_imp=float(p_ext_imp[j]),
                      weight=float(weight[j])) for j in range(uniq.size)]
        polar = None
        if pm is not None:                                 # mc3D:819-883, 1132
            x = pm.interp_arrays(float(uniq[0]))[0]
            polar = PolarTables(x=x, Y=[pm.interp_arrays(float(w))[1] for w in uniq],
                                cutoff=pm.cutoff(wvl0, uniq), stokes0=tuple(float(v) for v in stokes_params))
        case = Case(tau_tot=tau_tot, theta0_deg=theta_0, n_photons=n_photon, bands=bands,
                    lambertian_bottom=Lambertian_bottom, r_lambertian=Lambertian_reflectance,
                    lambertian_surface=Lambertian_surface, polar=polar)
        w = self.walker()
        w.set_cases([case])
        mode = 'replay' if replay is not None else (self.mode or 'native')
        if mode == 'replay' and replay is None:
            raise ValueError('replay mode needs replay=(values, offsets)')
        seed = self.seed if self.seed is not None else int(np.random.randint(0, 2**31 - 1))
        w.walk(mode=mode, seed=seed, records=True, photon_band=band.astype(np.uint8), replay=replay)
        rec = w.records()
        if polar is not None:
            rec.update(w.stokes())
        self.records = rec
        self.tallies = w.tallies()[0]
        self.stats = w.stats()
        all_answers = par_wvls.answer_and_reduce(rec, lambda parts: parts[0])
        if all_answers is None or not write:
            return None
        output_file = self.setup_output(n_photon, wvl0, half_width)
        self.write_records(output_file, all_answers)
        print('%s' % output_file)
        return output_file

    @staticmethod
    def _check_bins(uniq):
        if uniq.size > 256:
            raise ValueError('%d distinct photon wavelengths: at most 256 bins of 0.01 um per run'
                             % uniq.size)

    def calculate_albedo(self, answers):
        """Directional-hemispherical reflectance sum(wvn | cond 1) / sum(wvn)
        (mc3D:1659-1671); accepts a list of record tuples or a record dict."""
        if isinstance(answers, dict):
            wvn = answers['wvn']
            return float(wvn[answers['condition'] == 1].sum() / wvn.sum())
        q_down = sum(a[1] for a in answers)
        q_up = sum(a[1] for a in answers if a[0] == 1)
        return q_up / q_down

    @classmethod
    def flatten_list(klass, l):
        return [item for sublist in l for item in sublist]

    def close(self):
        if self._walker is not None:
            self._walker.close()
            self._walker = None
