"""Algorithmic work per unit of the photon walk and MI355X peaks.

Unit = one photon-step (one iteration of the walk loop of
monte_carloMPI/monte_carlo3D.py:1212-1466); photon-steps = sum(n_scat + 1).
Per step the algorithm needs (FP32 native walk):
  * 38 FP32 VALU operations = 48 flops (an FMA counts 2): HG inverse CDF 6,
    azimuth 1, sin(theta) 2, direction rotation 17, free path scale 1,
    position / path 4, uniform conversion 4 x 2 + 2, exit / absorption tests 3;
  * 7 transcendentals: log, sin, cos, 2 sqrt, rsqrt, rcp;
  * ~125 integer operations: 5/4 of a Philox4x32-10 call (~100 ops each:
    20 32x32->64 multiplies, XORs, key schedule) for the 5 uniforms of a step;
  about 170 VALU lane-operations in all.
Bytes: each photon writes one 20-byte record (condition+band, n_scat, theta_n,
phi_n, path_length); replay mode reads 8 bytes per uniform deviate.
"""

FP32_FLOPS_PER_STEP = 48
VALU_LANE_OPS_PER_STEP = 170
RECORD_BYTES_PER_PHOTON = 20
REPLAY_BYTES_PER_DEVIATE = 8

# MI355X (gfx950) published peaks
MI355X_CUS = 256
MI355X_CLOCK_HZ = 2.4e9
PEAK_FP32_TFLOPS = 157.3                      # vector FP32
PEAK_VALU_LANE_OPS = MI355X_CUS * 64 * MI355X_CLOCK_HZ   # 39.3e12 non-packed lane-ops/s
PEAK_HBM_GBS = 8000.0


def walk_roofline(steps_per_launch: float, kernel_ms: float):
    """Roofline figures of one walk launch from its photon-steps and time."""
    s = kernel_ms * 1e-3
    flops = steps_per_launch * FP32_FLOPS_PER_STEP / s / 1e12
    ops = steps_per_launch * VALU_LANE_OPS_PER_STEP / s
    return dict(fp32_tflops=flops, fp32_frac=flops / PEAK_FP32_TFLOPS,
                valu_lane_ops=ops, valu_frac=ops / PEAK_VALU_LANE_OPS)
