"""Python/numpy oracle of the photon random walk (TEST INFRASTRUCTURE ONLY).

Restates, for the sphere / Henyey-Greenstein path of the reference
(``monte_carloMPI/monte_carlo3D.py``, abbreviated mc3D below):

* ``MonteCarlo.run`` photon section   mc3D:1492-1619 (``run_rank``)
* ``initial_pdfs``                    mc3D:1027-1044 (``initial_triples``)
* ``populate_pdfs`` HG branch         mc3D:885-921, 1010-1025 (``_draw_step``)
* ``Henyey_Greenstein2``              mc3D:790-800 (``hg2``)
* ``monte_carlo3D``                   mc3D:1111-1490 (``walk_photon``)
* ``get_optical_properties`` interp   mc3D:498-658 (``interp_two_nearest``)
* ``calculate_albedo``                mc3D:1659-1671

Arithmetic is numpy float64 in the reference's operation order, so that for
one deviate stream the records equal the reference's bit for bit (this is
what the golden fixtures under tests/golden check).  The trajectory arrays
(``x_tau`` ... ``z_crt``) are kept with ``np.append`` like the reference when
``trajectory=True`` so that the CPU baseline timed on this module carries the
reference's per-step cost; only ``z_tau`` reaches an output.

Parity pins: golden records (bit-exact), closed forms (HG mean = g, chi^2 vs
the HG pdf, exponential free path), van de Hulst (1980) R/T for tau=2,
omega=0.9, g=0.75, mu0=1, direct transmittance exp(-tau/mu0).
"""
from __future__ import annotations

import numpy as np

TWO_PIE = 2 * np.pi          # mc3D:26


# --------------------------------------------------------------------------
# deviate sources
# --------------------------------------------------------------------------
class RefStream:
    """Uniform deviates from a numpy legacy ``RandomState`` in the reference's
    call pattern (``np.random.rand(1)``, ``np.random.rand()``,
    ``np.random.uniform(0, pi/2)``).

    With ``record=True`` every raw uniform is appended to ``log`` so that a
    replay stream (what the GPU replay walk consumes) can be cut from it.
    """

    def __init__(self, rs: np.random.RandomState, record: bool = False):
        self.rs = rs
        self.record = record
        self.log: list[float] = []

    def rand1(self):
        a = self.rs.random_sample(1)
        if self.record:
            self.log.append(float(a[0]))
        return a

    def rand(self) -> float:
        v = self.rs.random_sample()
        if self.record:
            self.log.append(v)
        return v

    def uniform_half_pi(self) -> float:
        # legacy uniform(low, high) = low + (high - low) * next_double
        v = self.rs.random_sample()
        if self.record:
            self.log.append(v)
        return 0.0 + (np.pi / 2) * v


class ArrayStream:
    """Replays a photon's recorded uniforms (one CSR slice) in order."""

    def __init__(self, values):
        self.values = np.asarray(values, dtype=np.float64)
        self.used = 0

    def _next(self) -> float:
        if self.used >= self.values.size:
            raise IndexError('replay stream exhausted after %d deviates' % self.used)
        v = self.values[self.used]
        self.used += 1
        return float(v)

    def rand1(self):
        return np.array([self._next()])

    def rand(self) -> float:
        return self._next()

    def uniform_half_pi(self) -> float:
        return 0.0 + (np.pi / 2) * self._next()


# --------------------------------------------------------------------------
# scattering phase function sampling
# --------------------------------------------------------------------------
def hg2(g, costheta_p):
    """Inverse CDF of the Henyey-Greenstein phase function (mc3D:790-800).

    ``g`` is a float64 scalar, ``costheta_p`` the uniform deviate (a 1-element
    array in the walk, as in mc3D:915-916).  ``g == 0`` gives ``1 - 2u``.
    """
    if g == 0:
        return 1 - 2 * costheta_p
    return (1. / (2. * g)) * (1 + g ** 2 - ((1 - g ** 2) / (1 - g + 2 * g * costheta_p)) ** 2)


def hg_pdf(g, mu):
    """HG phase function p(mu) normalised over mu in [-1, 1] (mc3D:779-788 / 2)."""
    return 0.5 * (1 - g * g) / (1 + g * g - 2 * g * mu) ** 1.5


def _draw_step(src, g_k):
    """The five per-step draws of ``populate_pdfs`` (HG branch), in order:
    u_hg -> HG cos(theta), u_phi -> 2 pi u, u_tau -> -ln u, u_ssa, u_ext
    (mc3D:915-916, 921, 1014, 1020, 1023)."""
    g = np.array([g_k])
    p = hg2(g[0], src.rand1())
    phi = src.rand1() * TWO_PIE
    tau = -np.log(src.rand1())
    ssa = src.rand1()
    ext = src.rand1()
    return p[0], tau[0], phi[0], ssa[0], ext[0]


def initial_triples(src, n):
    """``initial_pdfs`` (mc3D:1027-1044): for each photon in order draw
    tau0 = -ln u, ssa0 = u, ext0 = u.  Returns three float64 arrays."""
    tau = np.empty((n, 1))
    ssa = np.empty((n, 1))
    ext = np.empty((n, 1))
    for k in range(n):
        tau[k, :] = -np.log(src.rand1())
        ssa[k, :] = src.rand1()
        ext[k, :] = src.rand1()
    return tau[:, 0], ssa[:, 0], ext[:, 0]


# --------------------------------------------------------------------------
# one photon packet
# --------------------------------------------------------------------------
class PhotonUnboundPhi(UnboundLocalError):
    """The reference leaves ``phi_n`` unbound (UnboundLocalError, mc3D:1474-1485)
    when the final direction has an exactly-zero x or y component."""


def walk_photon(src, first, wvl, g_k, ssa_ice_k, ssa_imp_k, ext_cff_mss_k,
                p_ext_imp_k, snow_depth_k, tau_tot, rho_snw, theta_0,
                lambertian_bottom, r_lambertian, trajectory=True, lambertian_surface=False,
                polar=None):
    """Walk one photon packet (mc3D:1111-1490, sphere/HG; with
    ``lambertian_surface`` the layer is a Lambertian reflector, mc3D:1228-1238,
    1385-1387).

    ``first`` is the photon's initial triple (tau0, ssa0, ext0) from
    ``initial_triples``.  Returns the 7 CSV fields
    ``(condition, wvn, theta_n, phi_n, n_scat, path_length, snow_depth)``.

    ``polar`` (an aspherical particle with its phase matrix, from
    ``oracle.polar``): the angles come from the phase-matrix rejection
    sampler and the Stokes vector is carried (mc3D:898-1004, 1283-1342); the
    record then ends with its final Q, U, V (mc3D:1637-1646).
    """
    if polar is not None:
        from . import polar as _polar
        stokes = polar['stokes0']
    # initial state: mc3D:1114-1132
    x_tau = np.array([0]); y_tau = np.array([0]); z_tau = np.array([0])
    mux_0 = np.sin(theta_0)
    muy_0 = 0
    muz_0 = -np.cos(theta_0)
    muz2_0 = muz_0 ** 2
    x_crt = np.array([0]); y_crt = np.array([0]); z_crt = np.array([0])
    z_list = [0.0]
    path_length = 0

    tau_r, ssa_r, ext_r = first
    p_r = phi_r = None
    bottom_reflection = False
    condition = 0
    i = 0
    while condition == 0:                                   # mc3D:1212
        i += 1
        if i >= 2:                                          # i_rand > i_max: mc3D:1215-1224
            if polar is not None:
                p_r, tau_r, phi_r, ssa_r, ext_r = _polar.draw_step(src, polar, stokes)
            else:
                p_r, tau_r, phi_r, ssa_r, ext_r = _draw_step(src, g_k)
        dtau_current = tau_r                                # mc3D:1227
        if lambertian_surface and i == 1:                   # mc3D:1228-1229
            dtau_current = 0

        if i == 1:                                          # mc3D:1232-1237
            costheta = 1; sintheta = 0; cosphi = 0; sinphi = 0
        elif lambertian_surface or bottom_reflection:       # mc3D:1238-1250
            mux_0 = 0.; muy_0 = 0.; muz_0 = 1.
            while True:
                theta_rand = src.uniform_half_pi()
                r1 = src.rand()
                if r1 < 2 * np.sin(theta_rand) * np.cos(theta_rand):
                    costheta = np.cos(theta_rand)
                    sintheta = np.sqrt(1 - costheta ** 2)
                    break
        else:                                               # mc3D:1251-1253
            costheta = p_r
            sintheta = np.sqrt(1 - costheta ** 2)

        if i > 1:                                           # mc3D:1255-1285
            cosphi = np.cos(phi_r)
            sinphi = np.sin(phi_r)
            if muz_0 == 1:
                mux_n = sintheta * cosphi
                muy_n = sintheta * sinphi
                muz_n = costheta
            elif muz_0 == -1:
                mux_n = sintheta * cosphi
                muy_n = -sintheta * sinphi
                muz_n = -costheta
            else:
                mux_n = ((sintheta * (mux_0 * muz_0 * cosphi - muy_0 * sinphi)) /
                         (np.sqrt(1 - muz2_0)) + mux_0 * costheta)
                muy_n = ((sintheta * (muy_0 * muz_0 * cosphi + mux_0 * sinphi)) /
                         (np.sqrt(1 - muz2_0)) + muy_0 * costheta)
                muz_n = (-np.sqrt(1 - muz2_0) * sintheta * cosphi + muz_0 * costheta)
            if bottom_reflection:                           # mc3D:1283-1285
                if polar is not None:
                    stokes = np.array([1, 0, 0, 0])
                bottom_reflection = False
            elif polar is not None:                         # mc3D:1287-1342
                stokes = _polar.stokes_update(polar, stokes, costheta, cosphi, muz_n, muz_0)
        else:                                               # mc3D:1344-1347
            mux_n = mux_0; muy_n = muy_0; muz_n = muz_0

        # move: mc3D:1349-1362
        z_prev = z_list[i - 1]
        z_i = z_prev + dtau_current * muz_n
        z_list.append(z_i)
        ext_cff = ext_cff_mss_k * rho_snw
        if trajectory:
            x_tau = np.append(x_tau, x_tau[i - 1] + dtau_current * mux_n)
            y_tau = np.append(y_tau, y_tau[i - 1] + dtau_current * muy_n)
            z_tau = np.append(z_tau, z_tau[i - 1] + dtau_current * muz_n)
            multiplier = dtau_current / ext_cff
            x_crt = np.append(x_crt, x_crt[i - 1] + multiplier * mux_n)
            y_crt = np.append(y_crt, y_crt[i - 1] + multiplier * muy_n)
            z_crt = np.append(z_crt, z_crt[i - 1] + multiplier * muz_n)

        if i > 1:                                           # mc3D:1364-1369
            mux_0 = mux_n; muy_0 = muy_n; muz_0 = muz_n
            muz2_0 = muz_0 ** 2

        path_length += dtau_current / ext_cff               # mc3D:1371-1372

        if ext_r > p_ext_imp_k:                             # mc3D:1374-1383
            ext_state = 1
            ssa_event = ssa_ice_k
        else:
            ext_state = 2
            ssa_event = ssa_imp_k
        if lambertian_surface:                              # mc3D:1385-1387
            ssa_event = r_lambertian

        if z_i > 0:                                         # mc3D:1390-1397
            condition = 1
            correction = -((z_i * dtau_current) / ((z_i - z_prev) * ext_cff))
            path_length += correction
        elif z_i < -tau_tot:                                # mc3D:1399-1459
            correction = -(((z_i + tau_tot) * dtau_current) / ((z_i - z_prev) * ext_cff))
            path_length += correction
            dtau_correction = -(((z_i + tau_tot) / (z_prev - z_i)) * dtau_current)
            z_list[i] = z_i - (muz_n * dtau_correction)
            if trajectory:
                x_tau[i] = x_tau[i] - (mux_n * dtau_correction)
                y_tau[i] = y_tau[i] - (muy_n * dtau_correction)
                z_tau[i] = z_tau[i] - (muz_n * dtau_correction)
                multiplier = dtau_correction / ext_cff
                x_crt[i] = x_crt[i] - (mux_n * multiplier)
                y_crt[i] = y_crt[i] - (muy_n * multiplier)
                z_crt[i] = z_crt[i] - (muz_n * multiplier)
            exit_condition = 3 if i == 1 else 2
            if lambertian_bottom:
                reflectance_rand = src.rand()
                if reflectance_rand <= r_lambertian:
                    bottom_reflection = True
                else:
                    condition = exit_condition
            else:
                condition = exit_condition
        elif ssa_r >= ssa_event:                            # mc3D:1461-1466
            condition = 4 if ext_state == 1 else 5

    # epilogue: mc3D:1468-1490
    wvn = 1. / wvl
    theta_n = np.arccos(muz_0)
    if i == 1:
        phi_n = 0.
    elif mux_0 > 0 and muy_0 > 0:
        phi_n = np.arctan(muy_0 / mux_0)
    elif mux_0 < 0 and muy_0 > 0:
        phi_n = np.arctan(muy_0 / mux_0) + np.pi
    elif mux_0 < 0 and muy_0 < 0:
        phi_n = np.arctan(muy_0 / mux_0) + np.pi
    elif mux_0 > 0 and muy_0 < 0:
        phi_n = np.arctan(muy_0 / mux_0) + TWO_PIE
    else:
        raise PhotonUnboundPhi("local variable 'phi_n' referenced before assignment")
    n_scat = i - 1
    if polar is not None:
        return (condition, wvn, theta_n, phi_n, n_scat, path_length, snow_depth_k,
                stokes[1], stokes[2], stokes[3])
    return (condition, wvn, theta_n, phi_n, n_scat, path_length, snow_depth_k)


# --------------------------------------------------------------------------
# single-scattering properties
# --------------------------------------------------------------------------
def interp_two_nearest(grid_m, values, wvl_um):
    """Two-nearest-point linear interpolation of ``get_optical_properties`` /
    ``get_impurity_optics`` (mc3D:519-577, 579-655, 670-711): the grid is in
    metres, the query in microns; a ``ValueError`` from ``interp1d``
    (query outside the two nearest points) falls back to the nearest value."""
    from scipy import interpolate
    wvl = wvl_um * 1e-6
    idx = np.argsort(np.absolute(wvl - grid_m))
    near_w = grid_m[idx[:2]]
    near_v = values[idx[:2]]
    try:
        if near_w[0] < near_w[1]:
            return interpolate.interp1d(near_w, near_v)(wvl)
        return interpolate.interp1d(near_w[::-1], near_v[::-1])(wvl)
    except ValueError:
        return near_v[0]


def lookup_ssps(table, wvls):
    """Per-photon (ssa, ext_cff_mss, g) for ``wvls`` from a spectral table
    ``{'wvl': metres, 'ss_alb', 'ext_cff_mss', 'asm_prm'}`` (mc3D:566-655)."""
    out = {}
    for key in ('ss_alb', 'ext_cff_mss', 'asm_prm'):
        if key not in table:
            continue
        arr = np.empty(wvls.shape)
        for w in np.unique(wvls):
            arr[wvls == w] = interp_two_nearest(table['wvl'], table[key], w)
        out[key] = arr
    return out


def sphere_ext_cff_mss(r_eff_um, rho_ice, q_ext=2.0):
    """Mass extinction cross-section of ice spheres, 3 Q_ext / (4 rho_ice r_eff)
    [m^2/kg] (geometric-optics limit; the SNICAR tables tend to Q_ext = 2)."""
    return 3.0 * q_ext / (4.0 * rho_ice * (r_eff_um * 1e-6))


# --------------------------------------------------------------------------
# run(): photon-side section for one rank
# --------------------------------------------------------------------------
def draw_wavelengths(rs, n_photon, wvl0, half_width):
    """mc3D:1515-1520: Gaussian photon wavelengths rounded to 0.01 um."""
    scale = half_width / 2.355
    return np.around(rs.normal(loc=wvl0, scale=scale, size=(n_photon)), decimals=2)


def photon_params(wvls, ice, imp, imp_cnc, tau_tot, rho_snw):
    """Per-photon parameter arrays of run() (mc3D:1577-1612).

    ``ice``/``imp`` map to callables ``f(wvls) -> dict`` or to dicts of
    per-photon arrays with keys ssa, ext, (g)."""
    ssa_ice = np.asarray(ice['ssa'], dtype=np.float64) * np.ones(wvls.shape)
    ext_ice = np.asarray(ice['ext'], dtype=np.float64) * np.ones(wvls.shape)
    g = np.asarray(ice['g'], dtype=np.float64) * np.ones(wvls.shape)
    ssa_imp = np.asarray(imp['ssa'], dtype=np.float64) * np.ones(wvls.shape)
    ext_imp = np.asarray(imp['ext'], dtype=np.float64) * np.ones(wvls.shape)
    ext_cff_mss = ext_ice * (1 - imp_cnc) + ext_imp * imp_cnc
    p_ext_imp = (imp_cnc * ext_imp) / (imp_cnc * ext_imp + (1 - imp_cnc) * ext_ice)
    snow_depth = tau_tot / (ext_cff_mss * rho_snw)
    return dict(wvl=wvls, g=g, ssa_ice=ssa_ice, ssa_imp=ssa_imp,
                ext_cff_mss=ext_cff_mss, p_ext_imp=p_ext_imp, snow_depth=snow_depth)


def walk_rank(src, par, tau_tot, rho_snw, theta_0_deg, lambertian_bottom,
              r_lambertian, trajectory=True, photon_hook=None, lambertian_surface=False):
    """initial_pdfs + the photon loop of run() (mc3D:1594-1616) for one rank's
    working set.  ``photon_hook(k, start_index)`` is called before each walk
    (used to cut per-photon replay slices out of a recorded stream)."""
    theta_0 = (np.pi * theta_0_deg) / 180.
    n = par['wvl'].size
    tau0, ssa0, ext0 = initial_triples(src, n)
    out = []
    for k in range(n):
        if photon_hook is not None:
            photon_hook(k)
        out.append(walk_photon(src, (tau0[k], ssa0[k], ext0[k]), par['wvl'][k],
                               par['g'][k], par['ssa_ice'][k], par['ssa_imp'][k],
                               par['ext_cff_mss'][k], par['p_ext_imp'][k],
                               par['snow_depth'][k], tau_tot, rho_snw, theta_0,
                               lambertian_bottom, r_lambertian, trajectory, lambertian_surface))
    return out


def records_to_arrays(records):
    """List of 7-tuples -> dict of numpy arrays (CSV column order)."""
    cols = list(zip(*records)) if records else [[]] * 7
    return dict(condition=np.asarray(cols[0], dtype=np.int64),
                wvn=np.asarray(cols[1], dtype=np.float64),
                theta_n=np.asarray(cols[2], dtype=np.float64),
                phi_n=np.asarray(cols[3], dtype=np.float64),
                n_scat=np.asarray(cols[4], dtype=np.int64),
                path_length=np.asarray(cols[5], dtype=np.float64),
                snow_depth=np.asarray(cols[6], dtype=np.float64))


def replay_run(seed, n_photon, wvl0, half_width, ssp_fn, imp_cnc, tau_tot,
               rho_snw, theta_0_deg, lambertian_bottom, r_lambertian,
               trajectory=False, lambertian_surface=False):
    """Single-rank run() with ``RandomState(seed)``: returns (records,
    per-photon wavelengths, params, CSR replay stream (values, offsets)).

    ``ssp_fn(wvls) -> (ice_dict, imp_dict)`` supplies the optics.  The replay
    slice of photon k is its initial triple followed by every uniform its walk
    drew, which is exactly what the GPU replay walk consumes."""
    rs = np.random.RandomState(seed)
    wvls = draw_wavelengths(rs, n_photon, wvl0, half_width)
    ice, imp = ssp_fn(wvls)
    par = photon_params(wvls, ice, imp, imp_cnc, tau_tot, rho_snw)
    src = RefStream(rs, record=True)
    starts = []
    records = walk_rank(src, par, tau_tot, rho_snw, theta_0_deg,
                        lambertian_bottom, r_lambertian, trajectory,
                        photon_hook=lambda k: starts.append(len(src.log)),
                        lambertian_surface=lambertian_surface)
    log = np.asarray(src.log, dtype=np.float64)
    n = n_photon
    init = log[:3 * n].reshape(n, 3)
    bounds = starts + [log.size]
    slices = []
    for k in range(n):
        slices.append(np.concatenate([init[k], log[bounds[k]:bounds[k + 1]]]))
    offsets = np.zeros(n + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([s.size for s in slices])
    values = np.concatenate(slices) if slices else np.zeros(0)
    return records, par, values, offsets


def calculate_albedo(records):
    """mc3D:1659-1671: sum(wvn | condition 1) / sum(wvn)."""
    q_down = 0
    q_up = 0
    for r in records:
        q_down += r[1]
        if r[0] == 1:
            q_up += r[1]
    return q_up / q_down
