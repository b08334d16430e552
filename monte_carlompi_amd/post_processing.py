"""BRF and albedo from per-photon record files or from GPU tallies.

The estimators of post_processing.py (pp):
* ``brf``: zenith histogram of reflected photons (condition 1) weighted by
  wvn over [0, pi/2) with ``calculate_bins`` bins; BRF_k = H_k / (Q_down w_k),
  w_k = sin m_k cos m_k / sum sin m cos m at bin midpoints m (pp:46-100);
* ``albedo``: sum(wvn | condition 1) / sum(wvn) (pp:125-128);
* ``non_absorbed``: adds conditions 2 and 3 (pp:269-275);
* ``spectral_albedo``: albedo and non-absorbed fraction of every record file
  grouped by snow effective radius and nominal wavelength, with the actual
  wavelength mean(1 / wvn) of each file (pp:250-312, files grouped as in
  MonteCarloData, pp:19-43);
* ``calculate_bins``: int(pi d_dome / (4 active_area)) (pp:435-444).
Plotting is not reproduced; the CLI prints the numbers.

    python -m monte_carlompi_amd.post_processing --output_dir D FILE [FILE ...]
    python -m monte_carlompi_amd.post_processing --spectral --output_dir D FILE [FILE ...]
"""
from __future__ import annotations

import argparse
import os

import numpy as np


def read_output(path):
    """A record file as a DataFrame (the reference reads it with
    pd.read_csv(delim_whitespace=True), pp:38)."""
    import pandas as pd
    return pd.read_csv(path, sep=r'\s+', float_precision='round_trip')


def parse_name(name):
    """(wvl0, half_width, rds_snw) from WVL_FWHM_REFF_... (pp:24, 31-33)."""
    parts = os.path.basename(name).split('_')
    return float(parts[0]), float(parts[1]), float(parts[2])


def calculate_bins(active_area=1., d_dome=175.):
    return int((np.pi * d_dome) / (4. * active_area))


def _cols(data):
    if isinstance(data, dict):
        return np.asarray(data['condition']), np.asarray(data['wvn']), np.asarray(data['theta_n'])
    return (data['condition'].to_numpy(), data['wvn[um^-1]'].to_numpy(), data['theta_n'].to_numpy())


def brf(data, n_bins=None, hist_range=(0., np.pi / 2)):
    """(midpoints [rad], BRF per bin) of pp:46-100 for a record DataFrame or
    dict of arrays."""
    if n_bins is None:
        n_bins = calculate_bins(1., 175.)
    cond, wvn, theta = _cols(data)
    q_down = wvn.sum()
    refl = cond == 1
    h = np.histogram(theta[refl], bins=n_bins, range=hist_range, weights=wvn[refl])
    midpoints = (np.diff(h[1]) / 2.) + h[1][:-1]
    w = np.sin(midpoints) * np.cos(midpoints) / np.sum(np.sin(midpoints) * np.cos(midpoints))
    return midpoints, h[0] / (q_down * w)


def albedo(data):
    cond, wvn, _ = _cols(data)
    return float(wvn[cond == 1].sum() / wvn.sum())


def non_absorbed(data):
    cond, wvn, _ = _cols(data)
    return float(wvn[(cond == 1) | (cond == 2) | (cond == 3)].sum() / wvn.sum())


def group_files(file_list, output_dir='.'):
    """{rds_snw: {wvl0: data}} from record file names WVL_FWHM_REFF_...
    read under ``output_dir`` (MonteCarloData, pp:19-43: the radius is the
    third field and the nominal wavelength the first; a later file with the
    same pair replaces an earlier one)."""
    out = {}
    for name in file_list:
        base = os.path.basename(name)
        wvl0, _, rds = parse_name(base)
        path = os.path.join(output_dir, base)
        out.setdefault(rds, {})[wvl0] = read_output(path)
    return out


def spectral_albedo(data_dict):
    """Rows (rds_snw, wvl0, actual_wvl0, albedo, non_absorbed) in order of
    radius then nominal wavelength, for {rds_snw: {wvl0: data}} (a record
    DataFrame or a dict of arrays per file) -- the numbers pp:250-312 plots:
    actual_wvl0 = mean(1 / wvn), albedo = sum(wvn | cond 1) / sum(wvn),
    non_absorbed adds conditions 2 and 3."""
    rows = []
    for rds in sorted(data_dict):
        for wvl0 in sorted(data_dict[rds]):
            data = data_dict[rds][wvl0]
            cond, wvn, _ = _cols(data)
            actual = float(np.mean(1.0 / wvn))
            rows.append((rds, wvl0, actual, albedo(data), non_absorbed(data)))
    return rows


def brf_1d_from_tallies(tally_row):
    """Zenith BRF (18 bins of 5 deg) from one case's GPU tallies: the phi
    marginal of the 18 x 36 histogram in pp's estimator."""
    from .gpu import BRF_BINS, PHI_BINS, T_QDOWN, THETA_BINS
    H = np.asarray(tally_row[:BRF_BINS]).reshape(THETA_BINS, PHI_BINS).sum(axis=1)
    edges = np.linspace(0., np.pi / 2, THETA_BINS + 1)
    mid = (np.diff(edges) / 2.) + edges[:-1]
    w = np.sin(mid) * np.cos(mid) / np.sum(np.sin(mid) * np.cos(mid))
    return mid, H / (tally_row[T_QDOWN] * w)


def main():
    ap = argparse.ArgumentParser(description='BRF / albedo of monte_carlo3D record files')
    ap.add_argument('--output_dir', type=str, default='.')
    ap.add_argument('--active_area', type=float, default=1.)
    ap.add_argument('--d_dome', type=float, default=175.)
    ap.add_argument('--spectral', action='store_true',
                    help='albedo per snow radius and wavelength over all files (spectral_albedo)')
    ap.add_argument('file_list', nargs=argparse.REMAINDER)
    a = ap.parse_args()
    if a.spectral:
        for rds, wvl0, actual, alb, non_abs in spectral_albedo(group_files(a.file_list, a.output_dir)):
            print('albedo(%s, %s) = %r   non-absorbed %r   (nominal wavelength %s um)'
                  % (actual, rds, alb, non_abs, wvl0))
        return
    n_bins = calculate_bins(a.active_area, a.d_dome)
    print('BRF computing over %d bins' % n_bins)
    for name in a.file_list:
        path = name if os.path.exists(name) else os.path.join(a.output_dir, os.path.basename(name))
        data = read_output(path)
        wvl0, fwhm, rds = parse_name(path)
        mid, b = brf(data, n_bins)
        print('%s: wvl0 %s um, FWHM %s um, r_eff %s um, albedo %.6f, non-absorbed %.6f'
              % (os.path.basename(path), wvl0, fwhm, rds, albedo(data), non_absorbed(data)))
        for m, v in zip(np.rad2deg(mid[::-1]), b[::-1]):
            print('  elevation %7.3f deg  BRF %.5f' % (90 - m, v))


if __name__ == '__main__':
    main()
