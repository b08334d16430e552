"""Surrogate serving at the benchmark's size (configs 4 and 5): a network
trained here on 200,000 walk-generated cases x 1e5 photons (200 epochs of
batch-4096 Adam on the BF16-MFMA kernels) serves 1e6 random queries; a seeded
1 % audit (10,000 queries) is re-walked at 1e6 photons and regions where the
audit misses the bars fall back to the walk.  The served answers -- network or
fallback -- are then judged against 2,000 further queries walked
independently at 1e6 photons (the audit's own queries carry the walk's
answers, so they cannot judge the server).  Bars per case: albedo within 1 %,
and no BRF bin (>= 100 expected photons) off by more than 3 % plus 4.3
standard errors of the judging walk; per bin over the judged cases: no bin's
rms error (walk noise removed) above 3 % with 4-sigma confidence.  The same
network then serves with a 0.2 % audit (2,000 queries), which lifts the
100x cap that a 1 % audit puts on the served rate."""
import numpy as np
import pytest

from monte_carlompi_amd import surrogate as S

pytestmark = pytest.mark.gpu


def test_served_answers_meet_the_bars_at_config5_size():
    from monte_carlompi_amd.gpu import GpuWalker
    w = GpuWalker(0)
    try:
        w.set_launch(256, 6, 8192, 16, 16)
        print('surrogate serve: training (200,000 cases x 1e5 photons)', flush=True)
        sur, trep = S.train(w, n_cases=200_000, photons=100_000, epochs=200, seed=0)
        assert trep['final_loss'] < 0.05, trep
        q = S.sample_inputs(1_000_000, 12345)
        print('surrogate serve: 1e6 queries, 1 % audit at 1e6 photons', flush=True)
        ans, rep = S.serve(sur, w, q, audit_frac=0.01, audit_photons=1_000_000, seed=3, validate=2000)
        print({k: rep[k] for k in ('audit_fail_frac', 'fallback_frac', 'cases_per_s_served',
                                   'plain_mc_cases_per_s_batched')}, rep['validation'], flush=True)
        assert rep['n_queries'] == 1_000_000 and rep['n_audit'] == 10_000
        assert tuple(ans.shape) == (1_000_000, S.N_OUT)
        import torch
        assert bool(torch.isfinite(ans).all())
        v = rep['validation']
        assert v['cases'] == 2000
        assert v['albedo_err']['p95'] <= 0.01, v
        assert v['fail_frac'] <= 0.02, v
        # per bin over the network-served judged answers (judging walks' noise
        # removed): no bin's rms error above 3 % with 4-sigma confidence
        assert v['network_bins_over_bar'] == 0, v
        assert rep['fallback_frac'] <= 0.05, rep
        # what the network gains: the network path alone, and the served rate
        # with the audit and fallback walks, against the plain walk's cases/s
        # (batched, the audit's own launches); a 1 % audit caps the latter at
        # 100x
        gain_net = rep['cases_per_s_surrogate'] / rep['plain_mc_cases_per_s_batched']
        gain_served = rep['cases_per_s_served'] / rep['plain_mc_cases_per_s_batched']
        assert gain_net >= 100.0, gain_net
        assert gain_served >= 50.0, gain_served
        # the same server with a 0.2 % audit (2,000 queries): the audit walks
        # no longer cap the served rate at 100x; the bars hold on 2,000 fresh
        # judging queries
        print('surrogate serve: 1e6 queries, 0.2 % audit', flush=True)
        _, rep2 = S.serve(sur, w, q, audit_frac=0.002, audit_photons=1_000_000, seed=4, validate=2000)
        print({k: rep2[k] for k in ('audit_fail_frac', 'fallback_frac', 'cases_per_s_served')},
              rep2['validation'], flush=True)
        v2 = rep2['validation']
        assert rep2['n_audit'] == 2000 and v2['cases'] == 2000
        assert v2['albedo_err']['p95'] <= 0.01, v2
        assert v2['fail_frac'] <= 0.02, v2
        assert v2['network_bins_over_bar'] == 0, v2
        assert rep2['cases_per_s_served'] / rep['plain_mc_cases_per_s_batched'] >= 100.0, rep2
        sur.net.close()
    finally:
        w.close()
