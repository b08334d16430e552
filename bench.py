#!/usr/bin/env python
"""Benchmark of the photon random walk on one MI355X.

Workload (one "step"): one pass of the hot path over one batch -- 1,000,000
photon packets of the single snowpack case (1.30 um, r_eff 100 um ice
spheres, Q_ext 2, omega 0.999, g 0.89, 1.0 m of snow at rho 300 kg/m^3,
theta0 60 deg, Lambertian bottom R 0.5), FP32 walk with Philox4x32-10
streams, per-photon records (photon order) and FP64 BRF tallies kept in HBM.
Case tables are resident before the timed region; D2H of records is not
timed.  Prints one JSON line.

The line also carries, as blocks beside the headline: the end-to-end
drop-in path (MonteCarlo.run: walk, record read-back, record file), the
batched 16,384-case sweep, the aspherical (phase-matrix) walk, and the
learned surrogate (trained on walk output, serving 1e6 queries with a 1 %
Monte Carlo audit at 1e6 photons and region-wise fallback).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--autotune]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'photons/sec on 1 GPU; BRF cases/sec at ≤1% albedo err vs 1e6-photon MC'
# (block, blocks_per_cu, chunk, refill_min, coop_live): small chunks spread the
# 1e6 photons over every block; the cooperative tail mode finishes the walks
# still running when the queue is empty, and 16-wave blocks spread those
# photons over the waves of the block that run out of work.  Chunks sized for
# exactly two (or four) per 1024-thread block keep any block from ending the
# queue a whole chunk behind.
def snow_1m_candidates(n_photons, cus):
    from monte_carlompi_amd.gpu import balanced_chunk
    b1 = balanced_chunk(n_photons, 1, cus)
    b4 = balanced_chunk(n_photons, 1, cus, per_block=4)
    return [(1024, 1, b1, 16, 32), (1024, 1, b1, 16, 24), (1024, 1, b1, 8, 32), (1024, 1, 2048, 16, 32),
            (1024, 1, 2048, 16, 24), (512, 2, b1, 16, 32), (1024, 1, b4, 16, 32), (256, 5, 512, 16, 16)]


N_PHOTONS = 1_000_000
# the launch shape of the headline workload, measured best on MI355X
# (1024-thread blocks, one per CU, 1954-photon chunks = two per block,
# refill at 16 idle lanes, cooperative tail at <= 32 live photons per wave),
# pinned so that the committed rocprofv3 profiles describe the timed launch
PINNED_LAUNCH = (1024, 1, 1954, 16, 32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    # the walk kernel settles after ~40 launches (clocks and power state:
    # 0.76 ms averaged over launches 3-23, 0.70 ms from launch 43 on), so the
    # default warm-up covers that before the timed steps
    ap.add_argument('--warmup', type=int, default=50)
    ap.add_argument('--photons', type=int, default=N_PHOTONS)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--cpu-photons', type=int, default=10_000)
    ap.add_argument('--cpu-procs', type=int, default=0)
    ap.add_argument('--no-sweep', action='store_true', help='skip the batched sweep block')
    ap.add_argument('--no-polar', action='store_true', help='skip the aspherical (phase-matrix) block')
    ap.add_argument('--no-surrogate', action='store_true', help='skip the surrogate (configs 4/5) block')
    ap.add_argument('--no-e2e', action='store_true', help='skip the end-to-end MonteCarlo.run block')
    ap.add_argument('--no-replay', action='store_true', help='skip the FP64 replay block')
    ap.add_argument('--no-repro', action='store_true', help='skip the reproducible-mode (tail mode off) timing')
    ap.add_argument('--launch', type=str, default=','.join(str(v) for v in PINNED_LAUNCH),
                    help='launch shape block,blocks_per_cu,chunk,refill_min[,coop_live] of the headline walk')
    ap.add_argument('--autotune', action='store_true',
                    help='pick the headline launch shape with the autotuner instead of --launch')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py measures one GPU: the photon walk is built and measured on a single '
              'MI355X here (multi-GPU sharding is a design note, not a launched path); '
              'run with --gpus 1', file=sys.stderr)
        sys.exit(2)

    import torch
    from monte_carlompi_amd import roofline
    from monte_carlompi_amd.gpu import GpuWalker, fractions_from_tallies
    from monte_carlompi_amd.workloads import SNOW_1M, snow_1m_case

    w = GpuWalker(0)
    w.set_cases([snow_1m_case(args.photons)])
    # the project's autotuner picks the launch shape for this workload (untimed)
    tune_ms = None
    if args.autotune:
        shape, tune_ms = w.autotune(snow_1m_candidates(args.photons, w.cus), photons_cap=0, repeats=9, seed=99)
    else:
        shape = tuple(int(v) for v in args.launch.split(','))
        w.set_launch(*shape)
    for k in range(args.warmup):
        w.walk(mode='native', seed=10_000 + k, records=True)
    torch.cuda.synchronize()
    w.sync()
    kern_ms, steps = [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        w.walk(mode='native', seed=k, records=True)   # returns after the walk is checked
        st = w.stats()
        kern_ms.append(st['kernel_ms'])
        steps.append(st['steps'])
    w.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    cnt = w.counts()[0]
    coop_frac = float(cnt[9]) / max(float(cnt[8]), 1.0)
    handed_over = int(cnt[10])
    fr = fractions_from_tallies(w.tallies())
    # PCIe-inclusive (not `value`): the same walk plus the read-back of every
    # record into host arrays in the reference's record dtypes (int8 condition,
    # int64 n_scat, float64 angles / path / wvn / depth), written in that
    # layout on the device and copied by DMA into pinned host buffers kept
    # from walk to walk, as a streaming consumer keeps them
    rec_buf = w.record_buffers()
    w.records(out=rec_buf)
    t2 = time.perf_counter()
    for k in range(5):
        w.walk(mode='native', seed=500 + k, records=True)
        w.records(out=rec_buf)
    pcie_ms = (time.perf_counter() - t2) * 1e3 / 5
    # the same case in the bit-reproducible mode (tail mode off: every record
    # depends only on seed, case and photon), for reference
    repro_ms = []
    if not args.no_repro:
        w.set_launch(*shape[:4], 0)
        for k in range(3):
            w.walk(mode='native', seed=700 + k, records=True)
            repro_ms.append(w.stats()['kernel_ms'])
        w.set_launch(*shape)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    photons_per_s = args.photons / (ms_per_step * 1e-3)
    avg_kern = sum(kern_ms) / len(kern_ms)
    avg_steps = sum(steps) / len(steps)
    rf = roofline.walk_roofline(avg_steps, avg_kern)

    traffic = profiled('k_walk_native_snow1m_traffic.json', 'hbm_bytes_per_launch')
    issue = profiled('k_walk_native_snow1m_valu.json', 'issue_busy')
    prof_insts = profiled('k_walk_native_snow1m_valu.json', 'SQ_INSTS_VALU')
    prof_steps = profiled('k_walk_native_snow1m_valu.json', 'photon_steps')
    prof_shape = profiled('k_walk_native_snow1m_valu.json', 'launch')
    vi = (roofline.valu_issue(prof_insts / prof_steps, avg_steps, avg_kern)
          if prof_insts and prof_steps else None)

    out = {
        'metric': METRIC, 'value': round(photons_per_s, 1), 'unit': 'photons/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(ms_per_step, 4), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'fp32', 'data': 'synthetic',
        'config': {'workload': 'snow_1m single case, 1e6 photons, native Philox FP32 walk '
                               '(per-photon records + FP64 18x36 BRF tallies in HBM)',
                   'photons': args.photons, 'wvl_um': SNOW_1M['wvl'], 'r_eff_um': SNOW_1M['r_eff'],
                   'omega': SNOW_1M['ssa'], 'g': SNOW_1M['g'], 'depth_m': SNOW_1M['depth_m'],
                   'theta0_deg': SNOW_1M['theta0'], 'parallelism': 'photon-per-lane, 1 GPU'},
        'photon_steps_per_s': round(avg_steps / (ms_per_step * 1e-3), 1),
        'photon_steps_per_launch': round(avg_steps, 1),
        'brf_cases_per_s': round(1e3 / ms_per_step, 3),
        'pcie_inclusive_photons_per_s': round(args.photons / (pcie_ms * 1e-3), 1),
        'kernel_ms_avg': round(avg_kern, 4),
        'launch': list(shape),
        'autotune_ms': ({str(tuple(c)): round(float(t), 4) for c, t in zip(snow_1m_candidates(args.photons, w.cus), tune_ms)}
                        if tune_ms is not None else None),
        'coop_steps_frac': round(coop_frac, 4),
        'reproducible_mode_kernel_ms': round(sorted(repro_ms)[1], 4) if repro_ms else None,
        'photons_handed_over': handed_over,
        'albedo': round(float(fr['albedo'][0]), 6),
        'roofline': {
            'bound': 'valu', 'achieved': round(rf['fp32_tflops'], 4),
            'peak': roofline.PEAK_FP32_TFLOPS, 'unit': 'TFLOP/s',
            'frac': round(rf['fp32_frac'], 5), 'traffic': traffic,
            'valu_issue_busy': issue,
            'valu_issue': ({'achieved': round(vi['achieved'], 1), 'peak': vi['peak'], 'unit': 'wave64 VALU insts/s',
                            'frac': round(vi['frac'], 4),
                            'insts_per_step': round(prof_insts / prof_steps, 4)} if vi else None),
            'profiled_launch': prof_shape,
            'basis': '48 FP32 flops per photon-step x photon-steps per launch / walk-kernel '
                     'HIP-event time (the kernel is bound by VALU issue, not FP32 flops: valu_issue = '
                     'VALU wave instructions per photon-step x photon-steps per launch / the same time, '
                     'against one wave64 instruction per SIMD per 4 cycles, which the SIMDs exceed '
                     'when they dual-issue VALU work of two waves, ~26 % of quad-cycles here); traffic (HBM bytes per '
                     'launch), valu_issue_busy (share of SIMD quad-cycles issuing VALU work) and the '
                     'instructions per step from rocprofv3 runs of the same workload and launch shape, '
                     'profiles/',
        },
    }
    if not args.no_e2e:
        out['end_to_end'] = e2e_block(w, args.photons)
    if not args.no_replay:
        out['replay'] = replay_block(w)
    if not args.no_sweep:
        out['sweep'] = sweep_block(w)
    if not args.no_polar:
        out['polar'] = polar_block(w, args.no_cpu)
    if not args.no_surrogate:
        out['surrogate'] = surrogate_block(w)
    w.close()
    if not args.no_cpu:
        from oracle import cpu_baseline
        cb = cpu_baseline.run(n_photon=args.cpu_photons, processes=args.cpu_procs or None, seed=1)
        out['cpu_baseline'] = {
            'value': round(cb['photons_per_s'], 2), 'unit': 'photons/s', 'cores': cb['processes'],
            'kind': 'port',
            'sample': '%d photons of the same case (the reference CPU configuration), %d ranks '
                      '(array_split scatter / ordered gather) running the oracle restatement of '
                      'monte_carlo3D with its per-step numpy structure; walk phase %.2f s, '
                      '%.3g photon-steps/s; CPU: %s' % (cb['photons'], cb['processes'], cb['walk_s'],
                                                         cb['steps_per_s'], cpu_baseline.cpu_model()),
        }
        out['gpu_over_cpu'] = round(photons_per_s / cb['photons_per_s'], 1)
        # the same 1e6-photon case walked by the compiled C restatement (FP64)
        cc = cpu_baseline.run_c(n_photon=args.photons, threads=args.cpu_procs or None, seed=1)
        out['cpu_baseline']['c_restatement'] = {
            'value': round(cc['photons_per_s'], 1), 'unit': 'photons/s', 'cores': cc['threads'],
            'sample': '%d photons of the same case, %d host threads each walking its share with '
                      'oracle/mcoracle.c (FP64, gcc -O2); walk phase %.2f s, %.3g photon-steps/s'
                      % (cc['photons'], cc['threads'], cc['walk_s'], cc['steps_per_s']),
        }
        out['gpu_over_cpu_c'] = round(photons_per_s / cc['photons_per_s'], 1)
    print(json.dumps(out), flush=True)


def profiled(name, key):
    """A figure of a committed rocprofv3 summary (profiles/NAME), or None."""
    path = os.path.join(ROOT, 'profiles', name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get(key)


def polar_block(w, no_cpu):
    """Aspherical grains: the snow_1m geometry with droxtal phase matrices
    and polarised light, 1e6 photons, native FP32 walk (rejection sampling of
    the phase matrix, Stokes vector per photon); the oracle restatement of the
    reference's walk on a bounded sample beside it."""
    import tempfile
    import numpy as np
    from monte_carlompi_amd.workloads import SNOW_POLAR, snow_polar_case
    d = tempfile.mkdtemp(prefix='mc_polar_')
    n = 1_000_000
    case = snow_polar_case(n, d)
    w.set_cases([case])
    # 512-thread blocks (the polarised kernel's limit; 1024 requests are split),
    # or three of 256 per CU (12 waves: all that 3 waves per SIMD allow; chunks
    # of 1302 make 768 chunks for 1e6 photons, one per block);
    # coop_live is the tail mode's threshold (at most 16 photons per wave)
    cands = [(512, 2, 1024, 16, 16), (256, 3, 1024, 16, 16), (256, 3, 1302, 16, 16), (256, 3, 2048, 16, 16),
             (1024, 1, 2048, 16, 16), (512, 2, 2048, 16, 8)]
    shape, tune = w.autotune(cands, photons_cap=0, repeats=3, seed=7)
    # walk-to-walk spread is ~10 % (the longest walk of each seed sets the
    # launch): median of 15 walks after 10 warm-up walks
    for k in range(10):
        w.walk(mode='native', seed=900 + k, records=True)
    times, steps = [], []
    for k in range(15):
        w.walk(mode='native', seed=k, records=True)
        st = w.stats()
        times.append(st['kernel_ms'])
        steps.append(st['steps'])
    rec = w.records()
    sto = w.stokes()
    ms = float(np.median(times))
    refl = rec['condition'] == 1
    out = {'workload': '1e6 photons, 1 m of snow, droxtal phase matrices (synthetic Yang-layout tables, '
                       'r_eff 100 um, 1.30 um), theta0 60 deg, Lambertian bottom R 0.5, Stokes (1, 1, 0, 0)',
           'photons': n, 'launch': list(shape), 'autotune_ms': {str(c): round(float(t), 3) for c, t in zip(cands, tune)},
           'kernel_ms': round(ms, 3), 'photons_per_s': round(n / (ms * 1e-3), 1),
           'photon_steps_per_s': round(float(np.mean(steps)) / (ms * 1e-3), 1),
           'mean_steps': round(float(np.mean(steps)) / n, 2),
           'albedo': round(float(np.mean(refl)), 5),
           'reflected_mean_QUV': [round(float(np.mean(sto[k][refl])), 5) for k in ('Q_n', 'U_n', 'V_n')]}
    if not no_cpu:
        import time
        from oracle import polar as opolar
        from monte_carlompi_amd.workloads import RHO_ICE, RHO_SNW
        from monte_carlompi_amd import optics
        p = SNOW_POLAR
        m = 3000
        t0 = time.perf_counter()
        opolar.polar_run(1, m, p['wvl'], 1e-12, p['r_eff'], optics.aspherical_dir(d, p['shape'], p['roughness'], p['wvl']),
                         False, p['stokes0'], lambda w_: dict(ssa=np.zeros(w_.shape), ext=np.zeros(w_.shape)),
                         0.0, case.tau_tot, RHO_SNW, RHO_ICE, p['theta0'], True, p['r_lambertian'])
        dt = time.perf_counter() - t0
        out['cpu_baseline'] = {'value': round(m / dt, 2), 'unit': 'photons/s', 'cores': 1, 'kind': 'port',
                               'sample': '%d photons of the same case walked by oracle/polar.py (the reference\'s '
                                         'per-step numpy structure, one core): %.1f s' % (m, dt)}
        out['gpu_over_cpu'] = round(out['photons_per_s'] / out['cpu_baseline']['value'], 1)
        # the whole 1e6-photon case walked by the compiled C restatement (FP64)
        from oracle import cpu_baseline
        cc = cpu_baseline.run_c_polar(optics.aspherical_dir(d, p['shape'], p['roughness'], p['wvl']), p,
                                      case.tau_tot, n_photon=n, seed=1, rho_snw=RHO_SNW, rho_ice=RHO_ICE)
        out['cpu_baseline']['c_restatement'] = {
            'value': round(cc['photons_per_s'], 1), 'unit': 'photons/s', 'cores': cc['threads'],
            'sample': '%d photons of the same case, %d host threads each walking its share with '
                      'oracle/mcoracle.c (phase-matrix rejection draws and Stokes update, FP64, gcc -O2); '
                      'walk phase %.2f s, %.3g photon-steps/s'
                      % (cc['photons'], cc['threads'], cc['walk_s'], cc['steps_per_s'])}
        out['gpu_over_cpu_c'] = round(out['photons_per_s'] / cc['photons_per_s'], 1)
    return out


def replay_block(w, n=100_000):
    """The FP64 replay walk (the parity path): n photons of the headline case
    consuming the reference's uniform-deviate stream, recorded by the C
    restatement (MT19937, oracle/mcoracle.c) outside the timed region; the
    kernel time of the replay launch, with every record's condition and
    n_scat checked against the stream's own walk."""
    import numpy as np
    from oracle import coracle
    from oracle import walk as owalk
    from monte_carlompi_amd.gpu import Band, Case
    from monte_carlompi_amd.workloads import RHO_SNW, SNOW_1M, sphere_ext_cff_mss
    p = SNOW_1M
    ext = sphere_ext_cff_mss(p['r_eff'])
    tau = p['depth_m'] * ext * RHO_SNW
    ssp = lambda wv: (dict(ssa=np.full(wv.shape, p['ssa']), ext=np.full(wv.shape, ext), g=np.full(wv.shape, p['g'])),
                      dict(ssa=np.zeros(wv.shape), ext=np.zeros(wv.shape)))
    out, par = coracle.run_rank(7, n, p['wvl'], 1e-12, ssp, 0.0, tau, RHO_SNW, p['theta0'], True, 0.5, record=True)
    vals, offs = out['stream']
    case = Case(tau_tot=tau, theta0_deg=p['theta0'], n_photons=n, lambertian_bottom=True, r_lambertian=0.5,
                bands=[Band(wvl=float(par['wvl'][0]), g=p['g'], ssa_ice=p['ssa'], ext_cff=ext * RHO_SNW)])
    w.set_cases([case])
    w.set_launch(256, 4, 512, 8, 0)
    pband = np.zeros(n, np.uint8)
    w.walk(mode='replay', replay=(vals, offs), photon_band=pband, records=True)    # warm-up (uploads)
    times = []
    for k in range(3):
        w.walk(mode='replay', replay=(vals, offs), photon_band=pband, records=True)
        times.append(w.stats()['kernel_ms'])
    rec = w.records()
    same = bool(np.array_equal(rec['condition'], out['condition']) and np.array_equal(rec['n_scat'], out['n_scat']))
    ms = float(np.median(times))
    dev = float(offs[-1])
    return {'workload': '%d photons of the headline case, FP64 walk consuming the reference\'s deviate stream '
                        '(recorded by oracle/mcoracle.c, MT19937 seed 7)' % n,
            'photons': n, 'deviates': int(dev), 'kernel_ms': round(ms, 4),
            'photons_per_s': round(n / (ms * 1e-3), 1), 'photon_steps_per_s': round(out['steps'] / (ms * 1e-3), 1),
            'deviate_gbs': round(dev * 8 / (ms * 1e-3) / 1e9, 2),
            'records_equal_to_stream_walk': same}


def e2e_block(w, n):
    """The drop-in path end to end: MonteCarlo.run of the headline case
    (1e6 photons; wavelengths drawn and binned, tables uploaded, walk, record
    read-back, record file written), timed by the wall clock, and the same
    without the file.  numpy's normal draws of the wavelengths and the file
    write (Python's shortest repr of every float, the reference's numpy-1
    text) take most of it."""
    import contextlib
    import io
    import sys
    import tempfile
    import numpy as np
    from monte_carlompi_amd.monte_carlo3D import MonteCarlo
    from monte_carlompi_amd.workloads import SNOW_1M
    p = SNOW_1M
    argv = sys.argv
    sys.argv = [argv[0]]
    try:
        d = tempfile.mkdtemp(prefix='mc_e2e_')
        mc = MonteCarlo(output_dir=d, optics_dir=os.path.join(d, 'none'), ssa_ice=p['ssa'], g=p['g'],
                        depth_m=p['depth_m'], seed=5)
        mc._walker = w
        w.set_launch(*PINNED_LAUNCH)
        kw = dict(theta_0=p['theta0'], Lambertian_bottom=True, Lambertian_reflectance=0.5)
        np.random.seed(1)
        mc.run(n, p['wvl'], 1e-12, p['r_eff'], write=False, **kw)      # warm-up
        t0 = time.perf_counter()
        mc.run(n, p['wvl'], 1e-12, p['r_eff'], write=False, **kw)
        t1 = time.perf_counter()
        with contextlib.redirect_stdout(io.StringIO()):                 # run() prints the file name
            path = mc.run(n, p['wvl'], 1e-12, p['r_eff'], write=True, **kw)
        t2 = time.perf_counter()
        size = os.path.getsize(path)
        os.remove(path)
        mc._walker = None
    finally:
        sys.argv = argv
    return {'workload': 'MonteCarlo.run(1e6, 1.30 um, FWHM 1e-12, r_eff 100 um, theta0 60 deg) on the headline case, '
                        'native walk, records in photon order, record file text as the reference writes it',
            'photons': n, 'run_no_file_s': round(t1 - t0, 4), 'photons_per_s_no_file': round(n / (t1 - t0), 1),
            'run_with_file_s': round(t2 - t1, 4), 'photons_per_s_with_file': round(n / (t2 - t1), 1),
            'file_bytes': size}


def surrogate_block(w):
    """Configs 4/5: the surrogate trained here on walk output (200,000 cases
    x 1e5 photons from batched launches, 200 epochs of batch-4096 Adam on
    the BF16-MFMA kernels), then 1e6 random queries served on the GPU with a
    seeded 1 % audit by the walk at 1e6 photons and region-wise fallback;
    2,000 further queries walked at 1e6 photons judge the served answers.
    The same network and queries are served again with a 0.2 % audit
    (`audit_0p2pct`): a 1 % audit alone caps the served rate at 100x."""
    import numpy as np
    from monte_carlompi_amd import roofline
    from monte_carlompi_amd import surrogate as S
    w.set_launch(256, 6, 8192, 16, 16)
    t0 = time.perf_counter()
    sur, trep = S.train(w, n_cases=200_000, photons=100_000, epochs=200, seed=0)
    t1 = time.perf_counter()
    q = S.sample_inputs(1_000_000, 12345)
    _, rep = S.serve(sur, w, q, audit_frac=0.01, audit_photons=1_000_000, seed=3, validate=2000)
    t2 = time.perf_counter()
    # the same network and queries with a 0.2 % audit (2,000 walks), judged
    # by 2,000 fresh queries
    _, rep2 = S.serve(sur, w, q, audit_frac=0.002, audit_photons=1_000_000, seed=4, validate=2000)
    per_case = S.plain_mc_rate_per_case(w, S.sample_inputs(200, 777), 1_000_000)
    sur.net.close()
    flops = roofline.mlp_flops_per_query() * rep['n_queries'] / rep['predict_s_device'] / 1e12
    out = {'workload': 'MLP 9 features (of r_eff, omega, g, theta0, tau) -> 4 x 256 SiLU -> 651 (648 BRF bins, albedo, '
                       'T_diffuse, T_direct); trained on 200,000 walk-generated cases x 1e5 photons; 1e6 queries '
                       'served, 1 % audit at 1e6 photons, fallback by region; 2,000 extra queries as judges',
           'train': trep, 'train_wall_s': round(t1 - t0, 2), 'serve_wall_s': round(t2 - t1, 2),
           'plain_mc_cases_per_s_per_case': round(per_case, 2),
           'inference_roofline': {'bound': 'mfma', 'achieved': round(flops, 2), 'peak': roofline.PEAK_BF16_TFLOPS,
                                  'unit': 'TFLOP/s', 'frac': round(flops / roofline.PEAK_BF16_TFLOPS, 4)}}
    out.update({k: (round(v, 6) if isinstance(v, float) else v) for k, v in rep.items()})
    out['gain_surrogate_vs_batched_mc'] = round(rep['cases_per_s_surrogate'] / rep['plain_mc_cases_per_s_batched'], 1)
    out['gain_served_vs_batched_mc'] = round(rep['cases_per_s_served'] / rep['plain_mc_cases_per_s_batched'], 2)
    out['gain_served_vs_per_case_mc'] = round(rep['cases_per_s_served'] / per_case, 2)
    out['served_ceiling_vs_batched_mc'] = round(1.0 / (rep['n_audit'] / rep['n_queries']), 1)
    v2 = rep2['validation']
    out['audit_0p2pct'] = {
        'n_audit': rep2['n_audit'], 'audit_fail_frac': round(rep2['audit_fail_frac'], 6),
        'fallback_regions': rep2['fallback_regions'], 'fallback_frac': round(rep2['fallback_frac'], 6),
        'predict_s_device': round(rep2['predict_s_device'], 6), 'audit_s_device': round(rep2['audit_s_device'], 6),
        'fallback_s_device': round(rep2['fallback_s_device'], 6), 'served_s_host': round(rep2['served_s_host'], 4),
        'cases_per_s_served': round(rep2['cases_per_s_served'], 1),
        'cases_per_s_served_host': round(rep2['cases_per_s_served_host'], 1),
        'gain_served_vs_batched_mc': round(rep2['cases_per_s_served'] / rep['plain_mc_cases_per_s_batched'], 2),
        'gain_served_vs_per_case_mc': round(rep2['cases_per_s_served'] / per_case, 2),
        'served_ceiling_vs_batched_mc': round(rep2['n_queries'] / rep2['n_audit'], 1),
        'validation': {k: v2[k] for k in ('cases', 'served_by_walk', 'fail_frac', 'fail_frac_network',
                                          'network_bins_over_bar', 'albedo_err', 'network_bin_rms')}}
    return out


def sweep_block(w):
    """Batched sweep: 64 r_eff x 16 wavelength distributions x 8 theta0 x 2
    depths = 16,384 cases, photons per case log-uniform in [1e4, 1e6]
    (seeded), one persistent launch, autotuned shape, replayed through a
    hipGraph."""
    from monte_carlompi_amd import roofline, sweep
    from monte_carlompi_amd.workloads import sweep_cases
    cases, meta = sweep_cases(seed=0)
    res = sweep.run(w, cases, seed=1, use_graph=True, tune=True, sample_cases=1024)
    times = []
    for k in range(3):                          # graph replays of the captured launch
        w.walk(mode='native', seed=2 + k, records=False, use_graph=True)
        times.append(w.stats()['kernel_ms'])
    st = w.stats()
    ms = sorted(times)[1]
    rf = roofline.walk_roofline(st['steps'], ms)
    return {'workload': '16384 cases (64 r_eff x 16 Gaussian wavelength distributions x 8 theta0 x 2 depths), '
                        'photons/case log-uniform 1e4..1e6, free bottom',
            'cases': len(cases), 'photons': int(st['photons']), 'photon_steps': int(st['steps']),
            'ms_per_sweep': round(ms, 3), 'cases_per_s': round(len(cases) / (ms * 1e-3), 1),
            'photons_per_s': round(st['photons'] / (ms * 1e-3), 1),
            'photon_steps_per_s': round(st['steps'] / (ms * 1e-3), 1),
            'launch': list(res['launch']), 'autotune_ms': {str(k): round(v, 3) for k, v in res['tune_times'].items()},
            'fp32_frac': round(rf['fp32_frac'], 5),
            'valu_issue_busy': profiled('k_walk_native_sweep16384_valu.json', 'issue_busy')}


if __name__ == '__main__':
    main()
