"""Loader for the aspherical golden records (tests/golden/make_golden.py --polar)."""
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
FIELDS = ('condition', 'wvn', 'theta_n', 'phi_n', 'n_scat', 'path_length', 'snow_depth')
STOKES = ('Q_n', 'U_n', 'V_n')


def load():
    with open(os.path.join(HERE, 'golden_polar.json')) as fh:
        meta = json.load(fh)
    data = np.load(os.path.join(HERE, 'golden_polar.npz'), allow_pickle=False)
    cases = []
    for m in meta['cases']:
        p = m['name'] + '/'
        c = dict(m)
        c['records'] = {f: data[p + f] for f in FIELDS + STOKES if p + f in data.files}
        cases.append(c)
    return meta, cases


def table_dir(root, case):
    """Write the synthetic Yang-layout tables the case was run on; returns
    the directory of its shape and roughness."""
    from monte_carlompi_amd import workloads
    wvls = workloads.YANG_WVLS_FAR_IR if case.get('far_ir') else workloads.YANG_WVLS
    return workloads.write_synthetic_yang(root, shape=case['shape'], roughness=case['rough'], wvls=wvls)
