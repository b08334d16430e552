#!/usr/bin/env python
"""Write golden per-photon records by running the reference's own
``MonteCarlo.run`` (monte_carloMPI/monte_carlo3D.py:1492-1648) in THIS
container.  Nothing here runs on a GPU box; the fixtures it writes
(``golden_records.npz``) are data: inputs (case parameters, SSP tables, seed)
and the records the reference printed.

Why an import is needed: no closed form pins the exact per-photon records
of ``monte_carlo3D`` for a given deviate stream (branch choices on
``muz_0 == +-1``, correction order, phi_n quadrants, the draw order around
``normal()`` / ``initial_pdfs``), and the reference ships no tests or
fixtures.

How the reference is made to run here (nothing of it is copied):
* ``mpi4py`` is not installed: a one-rank ``COMM_WORLD`` (scatter returns
  part 0, gather returns ``[answer]``) is written to a temp dir;
* the SSP NetCDF tarball is absent: small synthetic ``ice_wrn_%04d.nc`` and
  impurity files are written with scipy.io.netcdf_file;
* the global numpy stream is seeded with ``np.random.seed(seed)`` before
  each ``run()``; the reference never seeds it itself.

Usage:  python tests/golden/make_golden.py [--polar]   (needs /root/reference)
"""
import json
import os
import re
import sys
import tempfile

import numpy as np

REF = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))

MPI_STUB = '''
class _Comm(object):
    size = 1
    rank = 0
    def scatter(self, data, root=0):
        return data[0]
    def gather(self, data, root=0):
        return [data]
class MPI(object):
    COMM_WORLD = _Comm()
'''

CONFIG_INI = '''[model parameters]
tau_tot=1000000.0
imp_cnc=0e-9
rho_snw = 300
rho_ice = 917
[plot options]
flg_crt=1
flg_3D=999
[data]
output_dir=out
optics_dir=optics
fi_imp=imp.nc
'''

# spectral grid in metres: 0.205 .. 4.995 um every 0.01 um, so a photon
# wavelength rounded to 0.01 um sits half way between two grid points
GRID_M = (0.205 + 0.01 * np.arange(480)) * 1e-6


def flat_table(ssa, ext, g):
    n = GRID_M.size
    return dict(wvl=GRID_M, ss_alb=np.full(n, ssa), ext_cff_mss=np.full(n, ext),
                asm_prm=np.full(n, g))


def spectral_table():
    lam = GRID_M * 1e6
    return dict(wvl=GRID_M,
                ss_alb=1.0 - 1e-3 * (1.0 + 40.0 * (lam - 1.2) ** 2),
                ext_cff_mss=16.0 + 0.5 * np.sin(3.0 * lam),
                asm_prm=0.86 + 0.03 * np.cos(2.0 * lam))


def imp_table(ssa, ext):
    n = GRID_M.size
    return dict(wvl=GRID_M, ss_alb=np.full(n, ssa), ext_cff_mss=np.full(n, ext))


EXT_100 = 3 * 2.0 / (4 * 917.0 * 100e-6)      # Q_ext = 2 spheres, r_eff = 100 um

CASES = [
    # name, seed, run kwargs, model kwargs, ice table, impurity table
    dict(name='cfg1_semi_inf', seed=1001, n=120, wvl0=1.3, fwhm=1e-12, rds=100,
         theta0=60.0, lb=True, R=0.5, tau=4907.306, imp_cnc=0.0,
         ice=flat_table(0.999, EXT_100, 0.89), imp=imp_table(0.3, 12000.0)),
    dict(name='vdh_free', seed=1002, n=1500, wvl0=0.5, fwhm=1e-12, rds=100,
         theta0=0.0, lb=False, R=0.0, tau=2.0, imp_cnc=0.0,
         ice=flat_table(0.9, EXT_100, 0.75), imp=imp_table(0.3, 12000.0)),
    dict(name='thin_lamb', seed=1003, n=800, wvl0=1.0, fwhm=1e-12, rds=100,
         theta0=60.0, lb=True, R=0.5, tau=2.0, imp_cnc=0.0,
         ice=flat_table(0.99, EXT_100, 0.8), imp=imp_table(0.3, 12000.0)),
    dict(name='spectral_imp', seed=1004, n=400, wvl0=1.3, fwhm=0.085, rds=250,
         theta0=30.0, lb=True, R=0.8, tau=10.0, imp_cnc=1e-4,
         ice=spectral_table(), imp=imp_table(0.3, 12000.0)),
    dict(name='nadir_bounce', seed=1005, n=250, wvl0=0.9, fwhm=1e-12, rds=100,
         theta0=0.0, lb=True, R=1.0, tau=1.0, imp_cnc=0.0,
         ice=flat_table(0.999, EXT_100, 0.85), imp=imp_table(0.3, 12000.0)),
    dict(name='isotropic', seed=1006, n=400, wvl0=1.1, fwhm=1e-12, rds=100,
         theta0=45.0, lb=False, R=0.0, tau=5.0, imp_cnc=0.0,
         ice=flat_table(0.95, EXT_100, 0.0), imp=imp_table(0.3, 12000.0)),
    dict(name='backward_g', seed=1007, n=400, wvl0=0.7, fwhm=1e-12, rds=100,
         theta0=75.0, lb=True, R=0.3, tau=3.0, imp_cnc=0.0,
         ice=flat_table(0.97, EXT_100, -0.3), imp=imp_table(0.3, 12000.0)),
    dict(name='one_photon', seed=1008, n=1, wvl0=1.3, fwhm=0.085, rds=100,
         theta0=20.0, lb=True, R=0.5, tau=30.0, imp_cnc=0.0,
         ice=spectral_table(), imp=imp_table(0.3, 12000.0)),
    dict(name='lamb_surface', seed=1009, n=500, wvl0=1.0, fwhm=1e-12, rds=100,
         theta0=35.0, lb=False, R=0.6, tau=5.0, imp_cnc=0.0, ls=True,
         ice=flat_table(0.99, EXT_100, 0.85), imp=imp_table(0.3, 12000.0)),
]


def write_nc(path, table):
    from scipy.io import netcdf_file
    f = netcdf_file(path, 'w')
    f.createDimension('wvl', table['wvl'].size)
    for key, arr in table.items():
        v = f.createVariable(key, 'd', ('wvl',))
        v[:] = arr
    f.close()


_NUM = re.compile(r'np\.float64\(([^)]*)\)')
_ARR = re.compile(r'array\(\[([^\]]*)\]\)')


def parse_output(path):
    """Rows of the reference CSV.  Under numpy 2 ``%r`` of a float64 prints
    ``np.float64(x)`` (full precision); with one photon the reference's
    size-1 SSP path (mc3D:519-564) makes some fields 1-element arrays, printed
    as ``array([x])`` with numpy's 8-decimal array repr.  Returns the rows and
    the set of column indices that were printed at that reduced precision."""
    rows = []
    lowprec = set()
    with open(path) as fh:
        header = fh.readline()
        for line in fh:
            fields = line.split(' ')
            for j, f in enumerate(fields):
                if _ARR.search(f):
                    lowprec.add(j)
            line = _ARR.sub(r'\1', _NUM.sub(r'\1', line))
            rows.append(line.split())
    return header, rows, sorted(lowprec)


def main():
    tmp = tempfile.mkdtemp(prefix='mc_golden_')
    os.makedirs(os.path.join(tmp, 'mpi4py'))
    with open(os.path.join(tmp, 'mpi4py', '__init__.py'), 'w') as fh:
        fh.write(MPI_STUB)
    with open(os.path.join(tmp, 'config.ini'), 'w') as fh:
        fh.write(CONFIG_INI)
    os.chdir(tmp)
    sys.argv = [sys.argv[0]]
    sys.path[:0] = [tmp, REF]
    import matplotlib
    matplotlib.use('Agg')
    from monte_carloMPI import monte_carlo3D

    store = {}
    meta = []
    for c in CASES:
        optics = os.path.join(tmp, 'optics_' + c['name'])
        os.makedirs(os.path.join(optics, 'mie', 'snicar'))
        write_nc(os.path.join(optics, 'mie', 'snicar', 'ice_wrn_%04d.nc' % c['rds']), c['ice'])
        write_nc(os.path.join(optics, 'mie', 'snicar', 'imp.nc'), c['imp'])
        outdir = os.path.join(tmp, 'out_' + c['name'])
        np.random.seed(c['seed'])
        mc = monte_carlo3D.MonteCarlo(tau_tot=c['tau'], imp_cnc=c['imp_cnc'],
                                      output_dir=outdir, optics_dir=optics,
                                      fi_imp='imp.nc')
        mc.run(c['n'], c['wvl0'], c['fwhm'], c['rds'], theta_0=c['theta0'],
               Lambertian_bottom=c['lb'], Lambertian_reflectance=c['R'],
               Lambertian_surface=c.get('ls', False))
        files = os.listdir(os.path.join(outdir, 'sphere'))
        assert len(files) == 1, files
        header, rows, lowprec = parse_output(os.path.join(outdir, 'sphere', files[0]))
        arr = np.array(rows, dtype=object)
        p = c['name'] + '/'
        store[p + 'condition'] = arr[:, 0].astype(np.int64)
        for j, key in enumerate(['wvn', 'theta_n', 'phi_n'], start=1):
            store[p + key] = np.array([float(x) for x in arr[:, j]])
        store[p + 'n_scat'] = arr[:, 4].astype(np.int64)
        store[p + 'path_length'] = np.array([float(x) for x in arr[:, 5]])
        store[p + 'snow_depth'] = np.array([float(x) for x in arr[:, 6]])
        for tname in ('ice', 'imp'):
            for key, val in c[tname].items():
                store[p + tname + '_' + key] = np.asarray(val)
        m = {k: v for k, v in c.items() if k not in ('ice', 'imp')}
        m.setdefault('ls', False)
        m['file_name'] = files[0]
        m['header'] = header
        m['lowprec_columns'] = lowprec
        meta.append(m)
        print('%-14s n=%5d  conditions=%s  file=%s' % (
            c['name'], c['n'], np.bincount(store[p + 'condition'], minlength=6).tolist(), files[0]))

    np.savez_compressed(os.path.join(HERE, 'golden_records.npz'), **store)
    with open(os.path.join(HERE, 'golden_cases.json'), 'w') as fh:
        json.dump(dict(rho_snw=300.0, rho_ice=917.0, cases=meta), fh, indent=1)


# ---------------------------------------------------------------- aspherical
# Phase-matrix + Stokes walk (mc3D:145-496, 819-1004, 1287-1342) and the
# aspherical SSPs with the HG phase function (HG=True).  The Yang et al.
# tables are absent: synthetic tables in their layout are written by
# monte_carlompi_amd.workloads.write_synthetic_yang (wavelengths 0.30..2.00 um
# every 0.05 um, r_eff 50 / 100 / 200 um).
POLAR_CASES = [
    dict(name='pm_free', seed=2001, n=300, wvl0=1.3, fwhm=1e-12, rds=100, theta0=30.0, lb=False,
         R=0.0, tau=3.0, imp_cnc=0.0, shape='droxtal', rough='smooth', HG=False, stokes=[1, 0, 0, 0]),
    dict(name='pm_linear_q', seed=2002, n=250, wvl0=1.0, fwhm=1e-12, rds=200, theta0=0.0, lb=True,
         R=0.5, tau=2.0, imp_cnc=0.0, shape='droxtal', rough='smooth', HG=False, stokes=[1, 1, 0, 0]),
    dict(name='pm_spectral_u', seed=2003, n=200, wvl0=1.3, fwhm=0.2, rds=60, theta0=60.0, lb=True,
         R=0.8, tau=4.0, imp_cnc=1e-4, shape='droxtal', rough='smooth', HG=False, stokes=[1, 0, 1, 0]),
    dict(name='pm_circular', seed=2004, n=200, wvl0=0.5, fwhm=1e-12, rds=100, theta0=45.0, lb=False,
         R=0.0, tau=5.0, imp_cnc=0.0, shape='droxtal', rough='smooth', HG=False, stokes=[1, 0, 0, 1]),
    dict(name='pm_surface', seed=2005, n=200, wvl0=1.3, fwhm=1e-12, rds=100, theta0=50.0, lb=False,
         R=0.7, tau=5.0, imp_cnc=0.0, shape='droxtal', rough='smooth', HG=False, stokes=[1, 1, 0, 0],
         ls=True),
    dict(name='pm_one_photon', seed=2006, n=1, wvl0=1.3, fwhm=0.085, rds=100, theta0=20.0, lb=True,
         R=0.5, tau=3.0, imp_cnc=0.0, shape='droxtal', rough='smooth', HG=False, stokes=[1, 0, 0, 0]),
    dict(name='asph_hg', seed=2007, n=400, wvl0=1.3, fwhm=0.085, rds=100, theta0=30.0, lb=True,
         R=0.5, tau=6.0, imp_cnc=0.0, shape='hexagonal plate', rough='severely rough', HG=True,
         stokes=[1, 0, 0, 0]),
    # far infrared (the 16.4-99.0 um tables), another shape and roughness
    dict(name='pm_far_ir', seed=2008, n=250, wvl0=20.0, fwhm=1.0, rds=200, theta0=40.0, lb=True,
         R=0.6, tau=3.0, imp_cnc=0.0, shape='solid bullet rosette', rough='moderately rough', HG=False,
         stokes=[1, 0, 1, 0], far_ir=True),
    dict(name='asph_hg_column', seed=2009, n=300, wvl0=0.5, fwhm=0.05, rds=50, theta0=0.0, lb=False,
         R=0.0, tau=4.0, imp_cnc=0.0, shape='hollow hexagonal column', rough='smooth', HG=True,
         stokes=[1, 0, 0, 0]),
]
_INT = re.compile(r'np\.int64\(([^)]*)\)')


def polar_main():
    """Run POLAR_CASES through the reference; write golden_polar.npz/json
    (records plus Q_n, U_n, V_n, the table directory is regenerated from
    the same synthetic writer at test time)."""
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from monte_carlompi_amd import workloads
    tmp = tempfile.mkdtemp(prefix='mc_golden_pm_')
    os.makedirs(os.path.join(tmp, 'mpi4py'))
    with open(os.path.join(tmp, 'mpi4py', '__init__.py'), 'w') as fh:
        fh.write(MPI_STUB)
    with open(os.path.join(tmp, 'config.ini'), 'w') as fh:
        fh.write(CONFIG_INI)
    os.chdir(tmp)
    sys.argv = [sys.argv[0]]
    sys.path[:0] = [tmp, REF]
    import matplotlib
    matplotlib.use('Agg')
    from monte_carloMPI import monte_carlo3D
    store = {}
    meta = []
    for c in POLAR_CASES:
        optics = os.path.join(tmp, 'optics_' + c['name'])
        os.makedirs(os.path.join(optics, 'mie', 'snicar'))
        write_nc(os.path.join(optics, 'mie', 'snicar', 'imp.nc'), imp_table(0.3, 12000.0))
        workloads.write_synthetic_yang(optics, shape=c['shape'], roughness=c['rough'],
                                       wvls=workloads.YANG_WVLS_FAR_IR if c.get('far_ir') else workloads.YANG_WVLS)
        outdir = os.path.join(tmp, 'out_' + c['name'])
        np.random.seed(c['seed'])
        mc = monte_carlo3D.MonteCarlo(tau_tot=c['tau'], imp_cnc=c['imp_cnc'], output_dir=outdir,
                                      optics_dir=optics, fi_imp='imp.nc', HG=c['HG'])
        mc.run(c['n'], c['wvl0'], c['fwhm'], c['rds'], theta_0=c['theta0'],
               stokes_params=np.array(c['stokes']), shape=c['shape'], roughness=c['rough'],
               Lambertian_bottom=c['lb'], Lambertian_reflectance=c['R'],
               Lambertian_surface=c.get('ls', False))
        sub = os.path.join(outdir, mc.shape_dir, mc.roughness_dir)
        files = os.listdir(sub)
        assert len(files) == 1, files
        header, rows, lowprec = parse_output(os.path.join(sub, files[0]))
        rows = [[_INT.sub(r'\1', f) for f in r] for r in rows]
        arr = np.array(rows, dtype=object)
        p = c['name'] + '/'
        store[p + 'condition'] = arr[:, 0].astype(np.int64)
        for j, key in enumerate(['wvn', 'theta_n', 'phi_n'], start=1):
            store[p + key] = np.array([float(x) for x in arr[:, j]])
        store[p + 'n_scat'] = arr[:, 4].astype(np.int64)
        store[p + 'path_length'] = np.array([float(x) for x in arr[:, 5]])
        store[p + 'snow_depth'] = np.array([float(x) for x in arr[:, 6]])
        if not c['HG']:
            for j, key in enumerate(['Q_n', 'U_n', 'V_n'], start=7):
                store[p + key] = np.array([float(x) for x in arr[:, j]])
        m = dict(c)
        m.setdefault('ls', False)
        m['file_name'] = files[0]
        m['subdir'] = os.path.join(mc.shape_dir, mc.roughness_dir)
        m['header'] = header
        m['lowprec_columns'] = lowprec
        m['snow_effective_radius'] = float(mc.snow_effective_radius)
        meta.append(m)
        print('%-14s n=%5d  conditions=%s  file=%s' % (
            c['name'], c['n'], np.bincount(store[p + 'condition'], minlength=6).tolist(), files[0]))
    np.savez_compressed(os.path.join(HERE, 'golden_polar.npz'), **store)
    with open(os.path.join(HERE, 'golden_polar.json'), 'w') as fh:
        json.dump(dict(rho_snw=300.0, rho_ice=917.0, cases=meta), fh, indent=1)


if __name__ == '__main__':
    if '--polar' in sys.argv:
        polar_main()
    else:
        main()
