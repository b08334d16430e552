"""ctypes wrapper of oracle/mcoracle.c (TEST INFRASTRUCTURE ONLY).

Runs the float64 C restatement of the walk from a numpy legacy RandomState
(continuing the reference's stream exactly) or from replay slices, and can
record the CSR replay stream that the GPU replay walk consumes.
"""
from __future__ import annotations

import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class MTState(ctypes.Structure):
    _fields_ = [('key', ctypes.c_uint32 * 624), ('pos', ctypes.c_int32)]


class CaseC(ctypes.Structure):
    _fields_ = [('tau_tot', ctypes.c_double), ('rho_snw', ctypes.c_double),
                ('mux0', ctypes.c_double), ('muz0', ctypes.c_double),
                ('lambertian_bottom', ctypes.c_int32), ('lambertian_surface', ctypes.c_int32),
                ('r_lambertian', ctypes.c_double), ('max_steps', ctypes.c_int64)]


class PmC(ctypes.Structure):
    _fields_ = [('n', ctypes.c_int32), ('pad', ctypes.c_int32), ('x', ctypes.c_void_p),
                ('P', ctypes.c_void_p), ('A1', ctypes.c_double), ('B1', ctypes.c_double),
                ('A2', ctypes.c_double), ('B2', ctypes.c_double)]


class PolarC(ctypes.Structure):
    _fields_ = [('cmax', ctypes.c_double), ('ccut', ctypes.c_double), ('cmin', ctypes.c_double),
                ('stokes0', ctypes.c_double * 4), ('pm', ctypes.c_void_p), ('q', ctypes.c_void_p),
                ('u', ctypes.c_void_p), ('v', ctypes.c_void_p)]


ELEMENTS = ('P11', 'P12', 'P22', 'P33', 'P43', 'P44')


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'libmcoracle.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle/libmcoracle.so is not built: run `make -C oracle` '
                               '(or __graft_entry__.build())')
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        L.mco_walk.argtypes = [ctypes.POINTER(CaseC), ctypes.c_int64, P, P, P, P, P,
                               ctypes.POINTER(MTState), P, P, P, ctypes.c_int64, P,
                               P, P, P, P, P, P, P]
        L.mco_walk.restype = ctypes.c_int
        L.mco_mt_doubles.argtypes = [ctypes.POINTER(MTState), P, ctypes.c_int64]
        L.mco_mt_doubles.restype = ctypes.c_int
        _LIB = L
    return _LIB


def mt_from_numpy(rs: np.random.RandomState) -> MTState:
    st = rs.get_state()
    m = MTState()
    np.ctypeslib.as_array(m.key)[:] = st[1]
    m.pos = int(st[2])
    return m


def mt_to_numpy(m: MTState, rs: np.random.RandomState):
    st = rs.get_state()
    rs.set_state((st[0], np.ctypeslib.as_array(m.key).copy(), int(m.pos), st[3], st[4]))


def mt_doubles(m: MTState, n: int) -> np.ndarray:
    out = np.empty(n)
    lib().mco_mt_doubles(ctypes.byref(m), out.ctypes.data, n)
    return out


def _p(a):
    return None if a is None else a.ctypes.data


ERRORS = {-1: 'bad arguments', -2: 'record capacity exceeded', -3: 'replay slice mismatch',
          -4: 'phi_n unbound (reference raises UnboundLocalError)', -5: 'step cap reached'}


def walk(par, tau_tot, rho_snw, theta_0_deg, lambertian_bottom, r_lambertian,
         mt: MTState | None = None, replay=None, record=False, max_steps=0,
         allow_unbound_phi=False, lambertian_surface=False, polar=None):
    """Walk len(par['g']) photons.  ``replay`` = (values, offsets) CSR slices.
    ``polar`` = (case, bands, keys, stokes0) walks aspherical grains: ``case``
    and ``bands`` as oracle.polar.polar_tables returns them, ``keys`` each
    photon's wavelength (its key in ``bands``).
    Returns dict of record arrays (+ 'steps', 'Q_n', 'U_n', 'V_n' when
    ``polar``, and 'stream' = (values, offsets) when ``record``)."""
    theta_0 = (np.pi * theta_0_deg) / 180.
    c = CaseC(float(tau_tot), float(rho_snw), float(np.sin(theta_0)), float(-np.cos(theta_0)),
              int(bool(lambertian_bottom)), int(bool(lambertian_surface)), float(r_lambertian),
              int(max_steps))
    n = int(par['g'].size)
    arrs = [np.ascontiguousarray(par[k], dtype=np.float64)
            for k in ('g', 'ssa_ice', 'ssa_imp', 'ext_cff_mss', 'p_ext_imp')]
    cond = np.empty(n, np.int8)
    nscat = np.empty(n, np.int64)
    theta = np.empty(n)
    phi = np.empty(n)
    path = np.empty(n)
    steps = ctypes.c_int64(0)
    rep_v = rep_o = None
    if replay is not None:
        rep_v = np.ascontiguousarray(replay[0], dtype=np.float64)
        rep_o = np.ascontiguousarray(replay[1], dtype=np.int64)
        assert rep_o.size == n + 1
    pol = None
    if polar is not None:
        pcase, bands, keys, stokes0 = polar
        x = np.ascontiguousarray(pcase['x'], dtype=np.float64)
        keep = [x]
        tabs = {}
        for key, b in bands.items():
            P = np.ascontiguousarray(np.stack([b['Y'][e] for e in ELEMENTS]), dtype=np.float64)
            keep.append(P)
            tabs[key] = PmC(x.size, 0, x.ctypes.data, P.ctypes.data,
                            b['A1'], b['B1'], b['A2'], b['B2'])
        pms = list(tabs.values())
        arr = (PmC * len(pms))(*pms)
        slot = {key: j for j, key in enumerate(tabs)}
        ptrs = (ctypes.c_void_p * max(n, 1))()
        base = ctypes.addressof(arr)
        for k, key in enumerate(np.asarray(keys)):
            ptrs[k] = base + slot[key] * ctypes.sizeof(PmC)
        qo, uo, vo = np.empty(n), np.empty(n), np.empty(n)
        pol = PolarC(pcase['cmax'], pcase['ccut'], pcase['cmin'],
                     (ctypes.c_double * 4)(*[float(v) for v in stokes0]),
                     ctypes.addressof(ptrs), qo.ctypes.data, uo.ctypes.data, vo.ctypes.data)
    cap = 0
    rec = rec_off = None
    saved = None
    if record:
        cap = 3 * n + 5 * 64 * n + 1024
        if mt is not None:
            saved = MTState()
            ctypes.memmove(ctypes.byref(saved), ctypes.byref(mt), ctypes.sizeof(MTState))
    while True:
        if record:
            rec = np.empty(cap)
            rec_off = np.empty(n + 1, np.int64)
        ret = lib().mco_walk(ctypes.byref(c), n, *[a.ctypes.data for a in arrs],
                             None if mt is None else ctypes.byref(mt), _p(rep_v), _p(rep_o),
                             _p(rec), cap, _p(rec_off), cond.ctypes.data, nscat.ctypes.data,
                             theta.ctypes.data, phi.ctypes.data, path.ctypes.data,
                             ctypes.byref(steps), None if pol is None else ctypes.byref(pol))
        if ret == -2 and record:
            cap = int(rec_off[n]) + 16
            if saved is not None:
                ctypes.memmove(ctypes.byref(mt), ctypes.byref(saved), ctypes.sizeof(MTState))
            continue
        break
    if ret != 0 and not (ret == -4 and allow_unbound_phi):
        raise RuntimeError('mco_walk: %s (%d)' % (ERRORS.get(ret, '?'), ret))
    out = dict(condition=cond.astype(np.int64), theta_n=theta, phi_n=phi,
               n_scat=nscat, path_length=path, steps=int(steps.value))
    if pol is not None:
        out.update(Q_n=qo, U_n=uo, V_n=vo)
    if record:
        out['stream'] = (rec[:rec_off[n]].copy(), rec_off)
    return out


def run_rank(seed, n_photon, wvl0, half_width, ssp_fn, imp_cnc, tau_tot, rho_snw,
             theta_0_deg, lambertian_bottom, r_lambertian, record=False, max_steps=0,
             lambertian_surface=False):
    """Single-rank run() (mc3D:1492-1619) with ``RandomState(seed)``: the normal
    wavelength draw in numpy, then initial triples + walks in C continuing the
    same MT19937 stream.  Returns (records dict incl. wvn/snow_depth, params)."""
    from oracle import walk as pywalk
    rs = np.random.RandomState(seed)
    wvls = pywalk.draw_wavelengths(rs, n_photon, wvl0, half_width)
    ice, imp = ssp_fn(wvls)
    par = pywalk.photon_params(wvls, ice, imp, imp_cnc, tau_tot, rho_snw)
    mt = mt_from_numpy(rs)
    out = walk(par, tau_tot, rho_snw, theta_0_deg, lambertian_bottom, r_lambertian,
               mt=mt, record=record, max_steps=max_steps, lambertian_surface=lambertian_surface)
    out['wvn'] = 1. / wvls
    out['snow_depth'] = par['snow_depth']
    return out, par


def polar_run_rank(seed, n_photon, wvl0, half_width, rds_snw, directory, stokes0, imp_fn, imp_cnc,
                   tau_tot, rho_snw, rho_ice, theta_0_deg, lambertian_bottom, r_lambertian,
                   lambertian_surface=False, record=False, max_steps=0):
    """Single-rank run() of an aspherical case (HG = False) with
    ``RandomState(seed)``: wavelengths, SSPs and the phase-matrix tables as in
    oracle.polar.polar_run, then the initial triples and walks in C continuing
    the same MT19937 stream.  Returns (records dict incl. wvn/snow_depth and
    Q_n/U_n/V_n, params, case tables, band tables)."""
    from oracle import polar as pypolar
    from oracle import walk as pywalk
    rs = np.random.RandomState(seed)
    wvls = pywalk.draw_wavelengths(rs, n_photon, wvl0, half_width)
    wvls, ssa, ext, g, reff, theta, P = pypolar.aspherical_ssps(directory, wvl0, rds_snw, wvls, False,
                                                                rho_ice)
    par = pywalk.photon_params(wvls, dict(ssa=ssa, ext=ext, g=g), imp_fn(wvls), imp_cnc, tau_tot, rho_snw)
    case, bands = pypolar.polar_tables(theta, P, wvls, wvl0)
    mt = mt_from_numpy(rs)
    out = walk(par, tau_tot, rho_snw, theta_0_deg, lambertian_bottom, r_lambertian, mt=mt,
               record=record, max_steps=max_steps, lambertian_surface=lambertian_surface,
               polar=(case, bands, wvls, stokes0))
    out['wvn'] = 1. / wvls
    out['snow_depth'] = par['snow_depth']
    return out, par, case, bands


def polar_run_ranks(seed, n_ranks, n_per_rank, wvl0, half_width, rds_snw, directory, stokes0, imp_fn,
                    imp_cnc, tau_tot, rho_snw, rho_ice, theta_0_deg, lambertian_bottom, r_lambertian,
                    lambertian_surface=False, threads=None):
    """Independent aspherical ranks r = 0..n_ranks-1, each polar_run_rank with
    ``RandomState(seed + r)`` and ``n_per_rank`` photons, on host threads (the
    GIL is released inside C).  Returns (records incl. Q_n, U_n, V_n and the
    per-photon parameters, both concatenated rank-major; rank 0's case
    tables; every rank's band tables by wavelength).  Used for the native
    polarised walk's statistics."""
    def one(r):
        return polar_run_rank(seed + r, n_per_rank, wvl0, half_width, rds_snw, directory, stokes0, imp_fn,
                              imp_cnc, tau_tot, rho_snw, rho_ice, theta_0_deg, lambertian_bottom,
                              r_lambertian, lambertian_surface=lambertian_surface)
    T = threads or min(n_ranks, 16, len(os.sched_getaffinity(0)))
    with ThreadPoolExecutor(max_workers=T) as ex:
        parts = list(ex.map(one, range(n_ranks)))
    recs = [p[0] for p in parts]
    out = {k: np.concatenate([q[k] for q in recs]) for k in recs[0] if k != 'steps'}
    out['steps'] = sum(q['steps'] for q in recs)
    par = {k: np.concatenate([p[1][k] for p in parts]) for k in parts[0][1]}
    bands = {}
    for p in parts:
        bands.update(p[3])
    return out, par, parts[0][2], bands


def run_ranks(seed, n_ranks, n_per_rank, par_fn, tau_tot, rho_snw, theta_0_deg,
              lambertian_bottom, r_lambertian, threads=None):
    """Independent ranks r = 0..n_ranks-1 with ``RandomState(seed + r)`` walking
    ``n_per_rank`` photons each (threads release the GIL inside C).
    ``par_fn(n) -> params`` gives each rank's per-photon arrays.  Used for
    native-mode statistics; records are concatenated rank-major."""
    def one(r):
        rs = np.random.RandomState(seed + r)
        mt = mt_from_numpy(rs)
        return walk(par_fn(n_per_rank), tau_tot, rho_snw, theta_0_deg,
                    lambertian_bottom, r_lambertian, mt=mt)
    with ThreadPoolExecutor(max_workers=threads or min(n_ranks, os.cpu_count() or 1)) as ex:
        parts = list(ex.map(one, range(n_ranks)))
    out = {k: np.concatenate([p[k] for p in parts]) for k in parts[0] if k != 'steps'}
    out['steps'] = sum(p['steps'] for p in parts)
    return out
