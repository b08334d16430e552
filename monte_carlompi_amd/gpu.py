"""High-level wrapper of the C ABI (include/mcsnow.h) for the photon walk.

A ``GpuWalker`` owns one library context (one device, one stream).  Cases and
their wavelength bands are uploaded with ``set_cases``; ``walk`` runs every
photon of every case in one persistent launch; records (photon order),
per-case FP64 tallies and integer counts are read back afterwards.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (BRF_BINS, MCS_NATIVE, MCS_REPLAY, PHI_BINS, TALLY_F64, TALLY_U64,
                   THETA_BINS, McsBand, McsCase, McsLaunch, McsPolarCase, McsRecords,
                   McsStats, McsWalkArgs, check, lib)

# tally columns (see include/mcsnow.h)
T_QDOWN = BRF_BINS
T_QCOND = BRF_BINS + 1          # +cond-1 for condition 1..5
T_QCAP = BRF_BINS + 6
T_QPATH = BRF_BINS + 7
U_N, U_CAP, U_AXIS, U_STEPS, U_COOP, U_MOVED = 0, 6, 7, 8, 9, 10


@dataclass
class Band:
    """One wavelength bin of a case (the per-photon SSPs of run())."""
    wvl: float                   # wavelength [um]
    g: float
    ssa_ice: float
    ext_cff: float               # ext_cff_mss * rho_snw [1/m]
    ssa_imp: float = 0.0
    p_ext_imp: float = 0.0
    weight: float = 1.0          # photon fraction (GPU wavelength sampling)


@dataclass
class PolarTables:
    """Scattering phase matrices of an aspherical-grain case (one per band, in
    band order): ``x`` the ascending cos(theta) grid, ``Y[j]`` the (6, n)
    array P11, P12, P22, P33, P43, P44 of band j on it (P12.. multiplied by
    P11), ``cutoff`` the rejection split as a file-order grid index
    (mc3D:858-883), ``stokes0`` the initial Stokes vector."""
    x: np.ndarray
    Y: list
    cutoff: int
    stokes0: tuple = (1.0, 0.0, 0.0, 0.0)


@dataclass
class Case:
    """A snowpack/illumination configuration walked by ``n_photons`` photons.
    ``polar``: phase matrices for aspherical grains (None: Henyey-Greenstein)."""
    tau_tot: float
    theta0_deg: float
    n_photons: int
    bands: list = field(default_factory=list)
    lambertian_bottom: bool = False
    r_lambertian: float = 0.0
    lambertian_surface: bool = False
    polar: PolarTables = None

    def direction(self):
        th = (np.pi * self.theta0_deg) / 180.           # mc3D:1508
        return float(np.sin(th)), float(-np.cos(th))    # mc3D:1119-1121


def pack_cases(cases):
    """Case list -> (mcs_case array, mcs_band array)."""
    n_bands = sum(len(c.bands) for c in cases)
    carr = (McsCase * len(cases))()
    barr = (McsBand * max(n_bands, 1))()
    b = 0
    for k, c in enumerate(cases):
        if not c.bands:
            raise ValueError('case %d has no bands' % k)
        mux0, muz0 = c.direction()
        carr[k] = McsCase(float(c.tau_tot), mux0, muz0, float(c.r_lambertian),
                          int(bool(c.lambertian_bottom)), b, len(c.bands),
                          int(bool(c.lambertian_surface)), int(c.n_photons))
        w = np.array([x.weight for x in c.bands], dtype=np.float64)
        cdf = np.cumsum(w) / w.sum()
        cdf[-1] = 1.0
        for j, x in enumerate(c.bands):
            barr[b] = McsBand(1.0 / x.wvl, x.g, x.ssa_ice, x.ssa_imp, x.ext_cff, x.p_ext_imp,
                              c.tau_tot / x.ext_cff, float(cdf[j]))
            b += 1
    return carr, barr


def pack_polar(cases):
    """Case list -> (mcs_polar_case array, tables (n_tables, 7, n), n) or None
    when no case is polar."""
    pol = [c.polar for c in cases if c.polar is not None]
    if not pol:
        return None
    n = int(np.asarray(pol[0].x).size)
    parr = (McsPolarCase * len(cases))()
    tabs = []
    for k, c in enumerate(cases):
        if c.polar is None:
            parr[k] = McsPolarCase(-1, 0, (ctypes.c_double * 4)(1.0, 0.0, 0.0, 0.0))
            continue
        x = np.asarray(c.polar.x, dtype=np.float64)
        if x.size != n:
            raise ValueError('case %d: polar grids must all have %d points' % (k, n))
        if len(c.polar.Y) != len(c.bands):
            raise ValueError('case %d: %d phase matrices for %d bands' % (k, len(c.polar.Y), len(c.bands)))
        parr[k] = McsPolarCase(len(tabs), int(c.polar.cutoff),
                               (ctypes.c_double * 4)(*[float(v) for v in c.polar.stokes0]))
        for Y in c.polar.Y:
            Y = np.asarray(Y, dtype=np.float64).reshape(6, n)
            tabs.append(np.concatenate([x[None, :], Y], axis=0))
    return parr, np.ascontiguousarray(np.stack(tabs)), n


class _HeldArrays(dict):
    """A dict of numpy views that keeps the tensors owning their memory."""
    def __init__(self):
        super().__init__()
        self.hold = []


class GpuWalker:
    """One library context on ``device``."""

    def __init__(self, device: int = 0):
        self._ctx = ctypes.c_void_p()
        check(lib().mcs_create(device, ctypes.byref(self._ctx)))
        self.cases = []
        self.n_photons = 0
        self._bands_of_case = []

    # ------------------------------------------------------------------
    def close(self):
        if self._ctx:
            lib().mcs_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code):
        check(code, self._ctx)

    @property
    def cus(self) -> int:
        v = ctypes.c_int32()
        self._check(lib().mcs_device_cus(self._ctx, ctypes.byref(v)))
        return v.value

    # ------------------------------------------------------------------
    def set_cases(self, cases):
        carr, barr = pack_cases(cases)
        self._check(lib().mcs_set_cases(self._ctx, carr, len(cases), barr, len(barr)))
        pp = pack_polar(cases)
        self.polar = pp is not None
        if pp is not None:
            parr, tabs, n = pp
            self._check(lib().mcs_set_polar(self._ctx, parr, len(cases), tabs.ctypes.data,
                                            tabs.shape[0], n))
        self.cases = list(cases)
        self.n_photons = int(sum(c.n_photons for c in cases))

    def set_launch(self, block=256, blocks_per_cu=4, chunk=2048, refill_min=8,
                   coop_live=_lib.DEFAULT_COOP_LIVE):
        L = McsLaunch(block, blocks_per_cu, chunk, refill_min, coop_live)
        self._check(lib().mcs_set_launch(self._ctx, ctypes.byref(L)))

    def get_launch(self):
        L = McsLaunch()
        self._check(lib().mcs_get_launch(self._ctx, ctypes.byref(L)))
        return L.as_tuple()

    def walk(self, mode='native', seed=0, records=True, photon_band=None, replay=None,
             max_steps=0, use_graph=False, async_=False):
        """Walk every photon of the cases set.  ``photon_band``: per-photon
        band index (uint8, photon order); ``replay``: (values, offsets) CSR."""
        a = McsWalkArgs()
        a.mode = MCS_REPLAY if mode == 'replay' else MCS_NATIVE
        a.want_records = int(bool(records))
        a.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        a.max_steps = int(max_steps)
        a.async_ = int(bool(async_))
        a.use_graph = int(bool(use_graph))
        keep = []
        if photon_band is not None:
            pb = np.ascontiguousarray(photon_band, dtype=np.uint8)
            if pb.size != self.n_photons:
                raise ValueError('photon_band has %d entries, %d photons' % (pb.size, self.n_photons))
            a.photon_band = pb.ctypes.data
            keep.append(pb)
        if replay is not None:
            vals = np.ascontiguousarray(replay[0], dtype=np.float64)
            offs = np.ascontiguousarray(replay[1], dtype=np.int64)
            if offs.size != self.n_photons + 1 or offs[-1] > vals.size:
                raise ValueError('replay offsets do not match the photon count / values')
            a.replay_dev = vals.ctypes.data
            a.replay_off = offs.ctypes.data
            keep += [vals, offs]
        self._check(lib().mcs_walk(self._ctx, ctypes.byref(a)))
        self._keep = keep
        return self

    def walk_noalloc(self, tallies, counts, records=None, stokes=None, seed=0, max_steps=0, async_=False):
        """The native walk into caller-owned device memory (mcs_walk_noalloc:
        no device allocation, no host copies).  ``tallies`` (n_cases x 656
        float64) and ``counts`` (n_cases x 16 uint64) are device addresses;
        ``records`` an optional dict of device addresses n_scat, cond_band
        (uint32) and theta_n, phi_n, path_length (float32), photon order;
        ``stokes`` an optional device address of 3 x n float32 (Q, U, V
        planes, aspherical cases).  Any object with ``data_ptr()`` (a torch
        tensor on the walker's device) or an int address is accepted."""
        def addr(x):
            if x is None:
                return None
            return int(x.data_ptr()) if hasattr(x, 'data_ptr') else int(x)
        a = McsWalkArgs()
        a.mode = MCS_NATIVE
        a.want_records = int(records is not None)
        a.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        a.max_steps = int(max_steps)
        a.async_ = int(bool(async_))
        rec = records or {}
        o = _lib.McsDevOut(addr(tallies), addr(counts), addr(rec.get('n_scat')), addr(rec.get('cond_band')),
                           addr(rec.get('theta_n')), addr(rec.get('phi_n')), addr(rec.get('path_length')),
                           addr(stokes))
        self._check(lib().mcs_walk_noalloc(self._ctx, ctypes.byref(a), ctypes.byref(o)))
        return self

    def sync(self):
        self._check(lib().mcs_sync(self._ctx))

    # ------------------------------------------------------------------
    RECORD_DTYPES = (('condition', np.int8), ('band', np.uint8), ('n_scat', np.int64), ('theta_n', np.float64),
                     ('phi_n', np.float64), ('path_length', np.float64), ('wvn', np.float64),
                     ('snow_depth', np.float64))

    def record_buffers(self, pinned=True):
        """Empty record arrays for ``records(out=...)``, one per column of
        RECORD_DTYPES, n_photons long.  ``pinned``: in page-locked host memory
        (torch pin_memory), which the library fills by DMA straight from the
        device; keep the returned dict alive as long as its arrays are used."""
        n = self.n_photons
        if not pinned:
            return {k: np.empty(n, t) for k, t in self.RECORD_DTYPES}
        import torch
        of = {np.int8: torch.int8, np.uint8: torch.uint8, np.int64: torch.int64, np.float64: torch.float64}
        out = _HeldArrays()
        for k, t in self.RECORD_DTYPES:
            ten = torch.empty(n, dtype=of[t], pin_memory=True)
            out.hold.append(ten)
            out[k] = ten.numpy()
        return out

    def records(self, out=None):
        """Per-photon records in photon order: the CSV columns of mc3D:1626-1636
        plus the band index.  ``out``: a dict of arrays of those dtypes and
        the photon count to fill (a consumer that reads walk after walk can
        keep its buffers, best from record_buffers(); new arrays otherwise)."""
        n = self.n_photons
        if out is None:
            out = {k: np.empty(n, t) for k, t in self.RECORD_DTYPES}
        else:
            for k, t in self.RECORD_DTYPES:
                a = out[k]
                if a.dtype != t or a.shape != (n,) or not a.flags.c_contiguous:
                    raise ValueError('out[%r] must be a contiguous %s array of %d entries' % (k, np.dtype(t), n))
        R = McsRecords(*[out[k].ctypes.data for k in
                         ('condition', 'band', 'n_scat', 'theta_n', 'phi_n', 'path_length', 'wvn', 'snow_depth')])
        self._check(lib().mcs_read_records(self._ctx, ctypes.byref(R), n))
        return out

    def stokes(self):
        """Final Stokes vector (Q_n, U_n, V_n; I = 1) of every photon of the
        last walk, photon order (polar cases; HG cases read zeros)."""
        n = self.n_photons
        q, u, v = np.empty(n), np.empty(n), np.empty(n)
        self._check(lib().mcs_read_stokes(self._ctx, q.ctypes.data, u.ctypes.data, v.ctypes.data, n))
        return dict(Q_n=q, U_n=u, V_n=v)

    def tallies(self):
        t = np.empty((len(self.cases), TALLY_F64))
        self._check(lib().mcs_read_tallies(self._ctx, t.ctypes.data, len(self.cases)))
        return t

    def counts(self):
        u = np.empty((len(self.cases), TALLY_U64), np.uint64)
        self._check(lib().mcs_read_counts(self._ctx, u.ctypes.data, len(self.cases)))
        return u

    def stats(self):
        s = McsStats()
        self._check(lib().mcs_last_stats(self._ctx, ctypes.byref(s)))
        return dict(photons=s.photons, steps=s.steps, capped=s.capped,
                    phi_unbound=s.phi_unbound, kernel_ms=s.kernel_ms)

    def autotune(self, candidates, photons_cap=0, repeats=3, seed=1):
        arr = (McsLaunch * len(candidates))(*[McsLaunch.of(c) for c in candidates])
        times = np.zeros(len(candidates))
        best = McsLaunch()
        self._check(lib().mcs_autotune(self._ctx, arr, len(candidates), int(photons_cap),
                                       int(repeats), int(seed), times.ctypes.data,
                                       ctypes.byref(best)))
        return best.as_tuple(), times

    def philox(self, counters, key):
        ctr = np.ascontiguousarray(counters, dtype=np.uint32).reshape(-1, 4)
        out = np.empty_like(ctr)
        self._check(lib().mcs_philox(self._ctx, ctr.ctypes.data, int(key[0]), int(key[1]),
                                     out.ctypes.data, ctr.shape[0]))
        return out


def balanced_chunk(n_photons, blocks_per_cu, cus, per_block=2):
    """Chunk size that splits n_photons into per_block chunks for every block
    of a blocks_per_cu x cus grid (the queue then ends with no block a whole
    chunk behind the others).  A block keeps four chunks in flight, so chunks
    should hold at least twice the block's lanes; the caller picks per_block
    accordingly."""
    return max(64, -(-int(n_photons) // (int(per_block) * int(blocks_per_cu) * int(cus))))


def brf_from_tallies(t):
    """18 x 36 BRF grid of each case from its tallies (the 2-D form of
    post_processing.brf, pp:73-81): BRF_jk = H_jk / (Q_down w_jk) with
    w_jk = sin m_j cos m_j / (36 sum_j sin m_j cos m_j) at bin midpoints m_j.
    Summing H over phi reproduces pp's 1-D estimator with 18 bins."""
    t = np.atleast_2d(t)
    edges = np.linspace(0.0, np.pi / 2, THETA_BINS + 1)
    mid = (np.diff(edges) / 2.) + edges[:-1]
    w = np.sin(mid) * np.cos(mid)
    w = w / np.sum(w)
    H = t[:, :BRF_BINS].reshape(-1, THETA_BINS, PHI_BINS)
    q = t[:, T_QDOWN][:, None, None]
    return H / (q * (w[None, :, None] / PHI_BINS))


def fractions_from_tallies(t):
    """albedo (cond 1), diffuse T (2), direct T (3), absorbed (4+5), as
    wvn-weighted fractions of Q_down (pp:125-128, 269-275)."""
    t = np.atleast_2d(t)
    q = t[:, T_QDOWN]
    return dict(albedo=t[:, T_QCOND] / q, t_diffuse=t[:, T_QCOND + 1] / q,
                t_direct=t[:, T_QCOND + 2] / q,
                absorbed=(t[:, T_QCOND + 3] + t[:, T_QCOND + 4]) / q)
