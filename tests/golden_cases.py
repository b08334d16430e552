"""Loader for the golden records written by tests/golden/make_golden.py."""
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
FIELDS = ('condition', 'wvn', 'theta_n', 'phi_n', 'n_scat', 'path_length', 'snow_depth')


def load():
    with open(os.path.join(HERE, 'golden_cases.json')) as fh:
        meta = json.load(fh)
    data = np.load(os.path.join(HERE, 'golden_records.npz'), allow_pickle=False)
    cases = []
    for m in meta['cases']:
        p = m['name'] + '/'
        c = dict(m)
        c['records'] = {f: data[p + f] for f in FIELDS}
        c['ice'] = {k[len(p) + 4:]: data[k] for k in data.files if k.startswith(p + 'ice_')}
        c['imp'] = {k[len(p) + 4:]: data[k] for k in data.files if k.startswith(p + 'imp_')}
        cases.append(c)
    return meta, cases
