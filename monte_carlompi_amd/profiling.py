"""rocprofv3 helpers: a copy kernel of known size for counter calibration and
summaries of rocprofv3 CSV output (kernel stats, FETCH_SIZE / WRITE_SIZE).

    python -m monte_carlompi_amd.profiling copy [--mib 256]
        three device copies of MIB MiB (run under rocprofv3 --pmc to read what
        FETCH_SIZE / WRITE_SIZE report for a known number of bytes)
    python -m monte_carlompi_amd.profiling sweep [--repeats 2]
        the batched sweep (16,384 cases) alone, to profile it
    python -m monte_carlompi_amd.profiling mlp [--repeats 3]
        the surrogate network's kernels at their benchmark shapes: serving 1e6
        queries (features, 4 x 256 SiLU GEMMs, output transform) and batch-4096
        Adam steps (random weights and data: timing and counters only)
    python -m monte_carlompi_amd.profiling valu COUNTERS_CSV [--kernel k_walk] [--json OUT]
        [--steps PHOTON_STEPS_PER_LAUNCH] [--launch SHAPE]
        VALU issue occupancy, dual issue and lane utilisation from SQ_INSTS_VALU,
        SQ_ACTIVE_INST_VALU(2), SQ_THREAD_CYCLES_VALU, GRBM_GUI_ACTIVE (medians
        over dispatches to OUT)
    python -m monte_carlompi_amd.profiling stats KERNEL_STATS_CSV
    python -m monte_carlompi_amd.profiling traffic --kernel NAME --fetch F.csv
        --write W.csv --cal-fetch CF.csv --cal-write CW.csv [--mib 256] [--json OUT]
"""
from __future__ import annotations

import argparse
import collections
import csv
import json

import numpy as np


def copy_kernel(mib: int = 256):
    import torch
    n = mib * 1024 * 1024 // 4
    x = torch.ones(n, device='cuda')
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print('copied %d bytes x3' % (n * 4))


def run_sweep(repeats=2, launch=(256, 6, 8192, 24)):
    """The batched sweep (16,384 cases) `repeats` times with a fixed launch shape
    (a workload to profile on its own)."""
    from .gpu import GpuWalker
    from .workloads import sweep_cases
    cases, _ = sweep_cases(seed=0)
    with GpuWalker(0) as w:
        w.set_cases(cases)
        w.set_launch(*launch)
        for k in range(repeats):
            w.walk(mode='native', seed=1 + k, records=False)
            st = w.stats()
            print('sweep %d: %.2f ms, %d photons, %d photon-steps' % (k, st['kernel_ms'], st['photons'],
                                                                      st['steps']), flush=True)


def run_polar(repeats=3, launch=(256, 3, 1302, 16, 16), photons=1_000_000):
    """The aspherical-grain workload of bench.py (1e6 photons, droxtal phase
    matrices, polarised light) `repeats` times (a workload to profile on its
    own)."""
    import tempfile
    from .gpu import GpuWalker
    from .workloads import snow_polar_case
    case = snow_polar_case(photons, tempfile.mkdtemp(prefix='mc_polar_'))
    with GpuWalker(0) as w:
        w.set_cases([case])
        w.set_launch(*launch)
        for k in range(repeats):
            w.walk(mode='native', seed=1 + k, records=True)
            st = w.stats()
            print('polar %d: %.2f ms, %d photons, %d photon-steps' % (k, st['kernel_ms'], st['photons'],
                                                                      st['steps']), flush=True)


def run_mlp(repeats=3, queries=1_000_000, batch=4096, steps=100):
    """The surrogate network's kernels at the bench shapes (a workload to
    profile on its own): serving `queries` raw inputs on the device, then
    `steps` Adam steps of batch `batch` on a resident random dataset."""
    import torch
    from .mlp import Mlp
    from ._lib import MCS_FEATURES_SNOW_BRF
    from . import surrogate as S
    net = Mlp(S.N_FEAT, 256, 4, S.N_OUT, max_batch=65536)
    net.init_params(seed=1)
    # small output scale: the exp of the transform stays finite on random weights
    net.set_serving(MCS_FEATURES_SNOW_BRF, np.full(S.N_OUT, 1e-3), np.zeros(S.N_OUT), np.zeros(S.N_OUT),
                    S.FEAT_MU, S.FEAT_SD)
    q = torch.from_numpy(S.sample_inputs(queries, 1).astype(np.float32)).cuda()
    y = torch.empty((queries, S.N_OUT), dtype=torch.float32, device='cuda')
    torch.cuda.synchronize()
    for k in range(repeats):
        net.serve(q, queries, y)
        print('serve %d: %.3f ms for %d queries' % (k, net.last_ms(), queries), flush=True)
    rng = np.random.default_rng(2)
    n_data = 65536
    net.set_dataset(rng.normal(size=(n_data, S.N_FEAT)).astype(np.float32),
                    rng.normal(size=(n_data, S.N_OUT)).astype(np.float32))
    w = np.full(S.N_OUT, 1.0 / S.N_OUT, np.float32)
    import time
    for k in range(repeats):
        idx = rng.integers(0, n_data, steps * batch).astype(np.int32)
        t0 = time.perf_counter()
        net.train_steps(idx, batch, w, lr=1e-4)
        print('train %d: %d steps in %.3f s' % (k, steps, time.perf_counter() - t0), flush=True)
    net.close()


def run_snow1m(repeats=6, launch=(1024, 1, 1954, 16, 32), photons=1_000_000, json_out=None):
    """The headline workload (bench.py: snow_1m, 1e6 photons, records on) with
    a fixed launch shape `repeats` times; prints and optionally writes the
    photon-steps per launch (for the per-step instruction counts)."""
    from .gpu import GpuWalker
    from .workloads import snow_1m_case
    steps = []
    with GpuWalker(0) as w:
        w.set_cases([snow_1m_case(photons)])
        w.set_launch(*launch)
        for k in range(repeats):
            w.walk(mode='native', seed=k, records=True)
            st = w.stats()
            steps.append(st['steps'])
            print('snow1m %d: %.4f ms, %d photon-steps' % (k, st['kernel_ms'], st['steps']), flush=True)
    if json_out:
        with open(json_out, 'w') as fh:
            json.dump(dict(photon_steps=float(np.mean(steps)), launch=list(launch), photons=photons), fh)


def valu_metrics(path, kernel, simds=1024, n_xcd=8, clock_hz=2.4e9):
    """VALU metrics of `kernel` per dispatch from one --pmc pass.  rocprofv3
    writes one row per counter per dispatch, summed over all XCDs, so the
    kernel's cycles are GRBM_GUI_ACTIVE / n_xcd.  Reported:
      issue_busy   quad-cycles in which a SIMD issued >= 1 VALU instruction,
                   (SQ_ACTIVE_INST_VALU2 + SQ_INSTS_VALU - 2 SQ_ACTIVE_INST_VALU2)
                   / (cycles / 4 * SIMDs)
      dual_issue   quad-cycles with two VALU issues, SQ_ACTIVE_INST_VALU2 / (cycles / 4 * SIMDs)
      VALUUtilization  rocprofv3's formula: 100 SQ_THREAD_CYCLES_VALU / (SQ_ACTIVE_INST_VALU 64)
                   (share of the 64 lanes active per VALU instruction)."""
    vals = collections.defaultdict(dict)
    for r in csv.DictReader(open(path)):
        if kernel in r['Kernel_Name']:
            vals[r['Dispatch_Id']][r['Counter_Name']] = float(r['Counter_Value'])
    out = []
    for d, g in vals.items():
        res = dict(dispatch=int(d), **g)
        if g.get('GRBM_GUI_ACTIVE'):
            cycles = g['GRBM_GUI_ACTIVE'] / n_xcd
            quads = cycles / 4 * simds
            res['kernel_s'] = cycles / clock_hz
            if 'SQ_INSTS_VALU' in g and 'SQ_ACTIVE_INST_VALU2' in g:
                dual = g['SQ_ACTIVE_INST_VALU2']
                res['issue_busy'] = (dual + g['SQ_INSTS_VALU'] - 2 * dual) / quads
                res['dual_issue'] = dual / quads
        if g.get('SQ_ACTIVE_INST_VALU'):
            res['VALUUtilization'] = 100 * g.get('SQ_THREAD_CYCLES_VALU', 0) / (g['SQ_ACTIVE_INST_VALU'] * 64)
        out.append(res)
    return out


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


def kernel_table(stats_csv, fetch, write, mfma, cal_fetch, cal_write, mib=256, top=12):
    """Per kernel of a kernel-trace run: calls and average time, HBM bytes per
    dispatch (FETCH_SIZE / WRITE_SIZE scaled by a copy of known size), BF16
    MFMA work (SQ_INSTS_VALU_MFMA_MOPS_BF16, 512 flops each) and its rate over
    the average time."""
    known_kib = mib * 1024.0
    f_scale = known_kib / _pick(counter_means(cal_fetch, 'FETCH_SIZE'), 'copyBuffer')[1][0]
    w_scale = known_kib / _pick(counter_means(cal_write, 'WRITE_SIZE'), 'copyBuffer')[1][0]
    cf, cw = counter_means(fetch, 'FETCH_SIZE'), counter_means(write, 'WRITE_SIZE')
    mo = counter_means(mfma, 'SQ_INSTS_VALU_MFMA_MOPS_BF16')
    rows = []
    for r in sorted(kernel_stats(stats_csv), key=lambda r: -r['pct'])[:top]:
        name = r['name']
        def get(m, sc=1.0):
            hit = [v for k, v in m.items() if k == name or k.startswith(name) or name.startswith(k)]
            return hit[0][0] * sc if hit else None
        fb, wb = get(cf, f_scale * 1024), get(cw, w_scale * 1024)
        mops = get(mo)
        tf = mops * 512.0 / (r['avg_ns'] * 1e-9) / 1e12 if mops else None
        rows.append(dict(kernel=name, calls=r['calls'], avg_us=r['avg_ns'] / 1e3, share=r['pct'],
                         fetch_mb=fb / 1e6 if fb is not None else None,
                         write_mb=wb / 1e6 if wb is not None else None,
                         bf16_tflops=tf, hbm_gbs=((fb or 0) + (wb or 0)) / (r['avg_ns'] * 1e-9) / 1e9))
    return rows


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest='cmd', required=True)
    c = sub.add_parser('copy'); c.add_argument('--mib', type=int, default=256)
    sw = sub.add_parser('sweep'); sw.add_argument('--repeats', type=int, default=2)
    po = sub.add_parser('polar'); po.add_argument('--repeats', type=int, default=3)
    po.add_argument('--photons', type=int, default=1_000_000)
    po.add_argument('--launch', default='256,3,1302,16,16')
    ml = sub.add_parser('mlp'); ml.add_argument('--repeats', type=int, default=3)
    sn = sub.add_parser('snow1m'); sn.add_argument('--repeats', type=int, default=6)
    sn.add_argument('--launch', default='1024,1,1954,16,32'); sn.add_argument('--json')
    vm = sub.add_parser('valu'); vm.add_argument('csv'); vm.add_argument('--kernel', default='k_walk')
    vm.add_argument('--json', help='write the per-dispatch medians here')
    vm.add_argument('--steps', type=float, help='photon-steps per launch of the profiled workload')
    vm.add_argument('--launch', help='launch shape of the profiled workload')
    s = sub.add_parser('stats'); s.add_argument('csv')
    t = sub.add_parser('traffic')
    for k in ('kernel', 'fetch', 'write', 'cal_fetch', 'cal_write'):
        t.add_argument('--' + k.replace('_', '-'), dest=k, required=True)
    t.add_argument('--mib', type=int, default=256)
    t.add_argument('--json')
    kt = sub.add_parser('kernels')
    for k in ('stats', 'fetch', 'write', 'mfma', 'cal_fetch', 'cal_write'):
        kt.add_argument('--' + k.replace('_', '-'), dest=k, required=True)
    kt.add_argument('--json')
    a = ap.parse_args()
    if a.cmd == 'copy':
        copy_kernel(a.mib)
    elif a.cmd == 'sweep':
        run_sweep(a.repeats)
    elif a.cmd == 'polar':
        run_polar(a.repeats, launch=tuple(int(v) for v in a.launch.split(',')), photons=a.photons)
    elif a.cmd == 'mlp':
        run_mlp(a.repeats)
    elif a.cmd == 'snow1m':
        run_snow1m(a.repeats, tuple(int(v) for v in a.launch.split(',')), json_out=a.json)
    elif a.cmd == 'valu':
        rows = valu_metrics(a.csv, a.kernel)
        for r in rows:
            print(json.dumps(r))
        if a.json:
            keys = ('issue_busy', 'dual_issue', 'VALUUtilization', 'kernel_s', 'SQ_INSTS_VALU',
                    'SQ_INSTS_VALU_TRANS_F32')
            med = {k: float(np.median([r[k] for r in rows if k in r])) for k in keys
                   if any(k in r for r in rows)}
            med.update(kernel=a.kernel, dispatches=len(rows))
            if a.steps:
                med['photon_steps'] = a.steps
                med['valu_insts_per_step'] = med['SQ_INSTS_VALU'] / a.steps
            if a.launch:
                med['launch'] = [int(v) for v in a.launch.split(',')]
            with open(a.json, 'w') as fh:
                json.dump(med, fh, indent=1)
    elif a.cmd == 'kernels':
        rows = kernel_table(a.stats, a.fetch, a.write, a.mfma, a.cal_fetch, a.cal_write)
        for r in rows:
            print('%-48s %5d calls %10.1f us %6.2f %%  fetch %s MB  write %s MB  %s TFLOP/s  %7.0f GB/s' % (
                r['kernel'][:48], r['calls'], r['avg_us'], r['share'],
                '%.2f' % r['fetch_mb'] if r['fetch_mb'] is not None else '-',
                '%.2f' % r['write_mb'] if r['write_mb'] is not None else '-',
                '%.1f' % r['bf16_tflops'] if r['bf16_tflops'] else '-', r['hbm_gbs']))
        if a.json:
            with open(a.json, 'w') as fh:
                json.dump(rows, fh, indent=1)
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
