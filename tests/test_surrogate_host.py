"""Surrogate host logic (no GPU): feature map, log-target bias correction,
target variances, regions, symmetrisation, audit arithmetic."""
import numpy as np

from monte_carlompi_amd import gpu, surrogate as S


def test_log_target_bias_correction_is_unbiased():
    rng = np.random.default_rng(0)
    for mu in (20.0, 60.0, 400.0):
        c = rng.poisson(mu, 400_000)
        b = c / 100.0
        t = np.zeros((c.size, S.N_OUT)); t[:, 0] = b
        var = np.zeros_like(t); var[:, 0] = np.where(c > 0, b * b / np.maximum(c, 1), 0)
        truth = np.log(mu / 100 + S.OFFSETS[0])
        raw = S.to_log(t)[:, 0].mean()
        cor = S.to_log(t, var)[:, 0].mean()
        assert abs(cor - truth) < 0.25 * abs(raw - truth) + 2e-4, (mu, raw - truth, cor - truth)


def test_from_log_inverts_to_log():
    rng = np.random.default_rng(1)
    t = rng.uniform(0, 2, (10, S.N_OUT))
    np.testing.assert_allclose(S.from_log(S.to_log(t)), t, rtol=1e-12, atol=1e-12)


def test_features_are_finite_and_standardised():
    x = S.features(S.sample_inputs(100_000, seed=3))
    assert np.all(np.isfinite(x))
    assert np.all(np.abs(x.mean(axis=0)) < 0.8) and np.all((x.std(axis=0) > 0.3) & (x.std(axis=0) < 2.0))


def test_target_variance_and_symmetrisation():
    photons = 1000
    t = np.zeros((1, gpu.TALLY_F64))
    t[0, gpu.T_QDOWN] = photons * 1.0          # wvn = 1
    t[0, 0] = 30.0                             # 30 photons in bin (0, 0)
    t[0, 35] = 10.0                            # 10 in its mirror (0, 35)
    t[0, gpu.T_QCOND] = 40.0
    tg = S.targets_from_tallies(t)
    assert tg[0, 0] == tg[0, 35]                                   # mirrored bins pooled
    var = S.target_variance(t, tg, photons)
    np.testing.assert_allclose(var[0, 0], tg[0, 0] ** 2 / 40.0)
    np.testing.assert_allclose(var[0, S.BRF_BINS], 0.04 * 0.96 / photons)


def test_regions_cover_the_input_box():
    r = S.region_of(S.sample_inputs(50_000, seed=5))
    assert r.min() >= 0 and r.max() < 8 * 6 * 3 and len(np.unique(r)) > 100
