"""The learned surrogate at small scale: trained only on walk output, it
predicts held-out cases' albedo and BRF far better than the training mean
(what the network gains in accuracy) at >= 100x the cases/s of the plain walk
(what it gains in speed); serving audits and falls back by region."""
import numpy as np
import pytest

from monte_carlompi_amd import surrogate as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def walker():
    from monte_carlompi_amd.gpu import GpuWalker
    w = GpuWalker(0)
    yield w
    w.close()


def test_surrogate_learns_and_beats_plain_mc_rate(walker):
    inp = S.sample_inputs(6144, seed=1)
    print('surrogate: training data', flush=True)
    tg, _, gen_s = S.run_mc(walker, inp, 20_000, seed=11)
    sur = S.Surrogate(max_batch=8192)
    print('surrogate: fit', flush=True)
    sur.fit(inp, tg, epochs=60, batch=1024, lr=3e-3)
    sur.enable_serving()
    assert sur.history[-1] < 0.5 * sur.history[0]
    test_inp = S.sample_inputs(128, seed=2)
    print('surrogate: audit walk', flush=True)
    mc, mct, mc_s = S.run_mc(walker, test_inp, 200_000, seed=13)
    pred = sur.predict(test_inp)
    err = S.case_errors(pred, mc, 200_000)
    base = np.abs(tg[:, S.BRF_BINS].mean() - mc[:, S.BRF_BINS]) / mc[:, S.BRF_BINS]
    assert np.median(err['albedo']) < 0.02, np.median(err['albedo'])
    assert np.median(err['albedo']) < 0.2 * np.median(base)
    # speed: network throughput vs the walk at 200k photons per case
    q = S.sample_inputs(200_000, seed=3)
    sur.predict(q)
    nn_rate = q.shape[0] / (sur.predict_ms() * 1e-3)
    mc_rate = test_inp.shape[0] / mc_s
    assert nn_rate > 100 * mc_rate, (nn_rate, mc_rate)
    # serving: audit 1 % with the walk, fall back where the audit fails
    ans, rep = S.serve(sur, walker, q[:20_000], audit_frac=0.01, audit_photons=200_000, seed=4,
                       fallback_photons=20_000)
    assert rep['n_audit'] == 200 and 0.0 <= rep['fallback_frac'] <= 1.0
    assert tuple(ans.shape) == (20_000, S.N_OUT)
    # the device serving path (features and output transform on the GPU)
    # gives the host path's answers: inputs rounded alike to bf16, outputs
    # through expf instead of float64 exp
    import torch
    q2 = q[:100_000].astype(np.float32)            # the device path reads float32 inputs
    host = sur.predict(q2.astype(np.float64))
    dev = sur.predict_device(torch.from_numpy(q2).cuda()).cpu().numpy()
    rel = np.abs(dev - host) / np.maximum(np.abs(host), 1e-3)
    assert np.mean(np.all(rel < 1e-4, axis=1)) > 0.995, np.quantile(rel, [0.5, 0.999])
    assert rel.max() < 0.05, rel.max()
    sur.net.close()
