"""rocprofv3 helpers: a copy kernel of known size for counter calibration and
summaries of rocprofv3 CSV output (kernel stats, FETCH_SIZE / WRITE_SIZE).

    python -m monte_carlompi_amd.profiling copy [--mib 256]
        three device copies of MIB MiB (run under rocprofv3 --pmc to read what
        FETCH_SIZE / WRITE_SIZE report for a known number of bytes)
    python -m monte_carlompi_amd.profiling stats KERNEL_STATS_CSV
    python -m monte_carlompi_amd.profiling traffic --kernel NAME --fetch F.csv
        --write W.csv --cal-fetch CF.csv --cal-write CW.csv [--mib 256] [--json OUT]
"""
from __future__ import annotations

import argparse
import collections
import csv
import json


def copy_kernel(mib: int = 256):
    import torch
    n = mib * 1024 * 1024 // 4
    x = torch.ones(n, device='cuda')
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print('copied %d bytes x3' % (n * 4))


def kernel_stats(path):
    rows = list(csv.DictReader(open(path)))
    return [dict(name=r['Name'], calls=int(r['Calls']), avg_ns=float(r['AverageNs']),
                 min_ns=float(r['MinNs']), max_ns=float(r['MaxNs']), pct=float(r['Percentage']))
            for r in rows]


def counter_means(path, counter):
    acc = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r['Counter_Name'] == counter:
            acc[r['Kernel_Name']].append(float(r['Counter_Value']))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def _pick(means, substr):
    hits = [(k, v) for k, v in means.items() if substr in k]
    if not hits:
        raise KeyError(substr)
    return max(hits, key=lambda kv: kv[1][0])


def traffic(kernel, fetch, write, cal_fetch, cal_write, mib=256):
    """HBM bytes per dispatch of `kernel`: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) each scaled by known_bytes / counter of the calibration copy."""
    known_kib = mib * 1024.0
    cf = _pick(counter_means(cal_fetch, 'FETCH_SIZE'), 'copyBuffer')[1][0]
    cw = _pick(counter_means(cal_write, 'WRITE_SIZE'), 'copyBuffer')[1][0]
    f_scale, w_scale = known_kib / cf, known_kib / cw
    kf = _pick(counter_means(fetch, 'FETCH_SIZE'), kernel)
    kw = _pick(counter_means(write, 'WRITE_SIZE'), kernel)
    fetch_b = kf[1][0] * f_scale * 1024
    write_b = kw[1][0] * w_scale * 1024
    return dict(kernel=kf[0], dispatches=kf[1][1], fetch_kib_raw=kf[1][0], write_kib_raw=kw[1][0],
                fetch_scale=f_scale, write_scale=w_scale, fetch_bytes=fetch_b, write_bytes=write_b,
                hbm_bytes_per_launch=fetch_b + write_b)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest='cmd', required=True)
    c = sub.add_parser('copy'); c.add_argument('--mib', type=int, default=256)
    s = sub.add_parser('stats'); s.add_argument('csv')
    t = sub.add_parser('traffic')
    for k in ('kernel', 'fetch', 'write', 'cal_fetch', 'cal_write'):
        t.add_argument('--' + k.replace('_', '-'), dest=k, required=True)
    t.add_argument('--mib', type=int, default=256)
    t.add_argument('--json')
    a = ap.parse_args()
    if a.cmd == 'copy':
        copy_kernel(a.mib)
    elif a.cmd == 'stats':
        for r in kernel_stats(a.csv):
            print('%-60s calls %4d  avg %12.1f ns  min %12.1f  max %12.1f  %6.2f%%' % (
                r['name'][:60], r['calls'], r['avg_ns'], r['min_ns'], r['max_ns'], r['pct']))
    else:
        res = traffic(a.kernel, a.fetch, a.write, a.cal_fetch, a.cal_write, a.mib)
        print(json.dumps(res, indent=1))
        if a.json:
            with open(a.json, 'w') as fh:
                json.dump(res, fh, indent=1)


if __name__ == '__main__':
    main()
