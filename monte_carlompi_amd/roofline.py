"""Algorithmic work per unit of the photon walk and MI355X peaks.

Unit = one photon-step (one iteration of the walk loop of
monte_carloMPI/monte_carlo3D.py:1212-1466); photon-steps = sum(n_scat + 1).
Per step the FP32 native walk needs 38 FP32 VALU operations = 48 flops (an FMA
counts 2): HG inverse CDF 6, azimuth 1, sin(theta) 2, direction rotation 17,
free path scale 1, position / path 4, uniform conversion 4 x 2 + 2, exit /
absorption tests 3.  Besides them a step issues one Philox4x32-10 block (19
32x32->64 multiplies, 18 XORs, the key schedule on the scalar unit) and 6
transcendentals (log, sin, cos, sqrt, rsqrt, rcp): the kernel is bound by VALU
issue, which rocprofv3 measures (profiling.valu_metrics), not by FP32 flops.
Bytes: each photon writes one 24-byte compacted record (id, condition+band,
n_scat, theta_n, phi_n, path_length); replay mode reads 8 bytes per uniform
deviate.
"""

FP32_FLOPS_PER_STEP = 48
RECORD_BYTES_PER_PHOTON = 24
REPLAY_BYTES_PER_DEVIATE = 8

# MI355X (gfx950) published peaks
MI355X_CUS = 256
MI355X_CLOCK_HZ = 2.4e9
PEAK_FP32_TFLOPS = 157.3                      # vector FP32
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0                     # dense MFMA BF16
# VALU issue: one wave64 VALU instruction per SIMD every 4 cycles (4 SIMDs
# per CU) without dual issue -- 256 x 4 x 2.4e9 / 4 = 6.14e11 per second
PEAK_VALU_WAVE_INSTS = MI355X_CUS * 4 * MI355X_CLOCK_HZ / 4


def walk_roofline(steps_per_launch: float, kernel_ms: float):
    """FP32 roofline figures of one walk launch from its photon-steps and time."""
    s = kernel_ms * 1e-3
    flops = steps_per_launch * FP32_FLOPS_PER_STEP / s / 1e12
    return dict(fp32_tflops=flops, fp32_frac=flops / PEAK_FP32_TFLOPS)


def valu_issue(insts_per_step: float, steps_per_launch: float, kernel_ms: float):
    """VALU issue rate of a walk launch: wave64 VALU instructions per
    photon-step (rocprofv3 SQ_INSTS_VALU of the same workload and launch
    shape) x photon-steps per launch / kernel time, against the single-issue
    peak (dual issue can lift the fraction above 1)."""
    rate = insts_per_step * steps_per_launch / (kernel_ms * 1e-3)
    return dict(achieved=rate, peak=PEAK_VALU_WAVE_INSTS, frac=rate / PEAK_VALU_WAVE_INSTS)


def mlp_flops_per_query(n_in_padded=32, hidden=256, layers=4, n_out=651):
    """Dense MFMA flops of one inference of the surrogate network (inputs
    padded to 32 columns, FMA = 2 flops)."""
    macs = n_in_padded * hidden + (layers - 1) * hidden * hidden + hidden * n_out
    return 2.0 * macs
