"""MLP surrogate kernels: the float64 reference is pinned by finite
differences (CPU); the BF16-MFMA forward and gradients match it within BF16
tolerance (GPU); training lowers the loss."""
import numpy as np
import pytest

from tests import mlp_ref


def small_net(rng, n_in=5, H=64, L=2, n_out=70):
    shapes = [(H, n_in)] + [(H, H)] * (L - 1) + [(n_out, H)]
    return [(rng.normal(0, np.sqrt(2.0 / c), (r, c)), rng.normal(0, 0.1, r)) for r, c in shapes]


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    layers = small_net(rng, H=8, L=2, n_out=5)
    x = rng.normal(size=(16, 5)); t = rng.normal(size=(16, 5)); w = rng.uniform(0.5, 2, 5)
    loss, grads = mlp_ref.loss_and_grads(layers, x, t, w)
    h = 1e-6
    for l, (W, b) in enumerate(layers):
        for idx in [(0, 0), (W.shape[0] - 1, W.shape[1] - 1), (1, 2)]:
            Wp = [(A.copy(), c.copy()) for A, c in layers]
            Wm = [(A.copy(), c.copy()) for A, c in layers]
            Wp[l][0][idx] += h; Wm[l][0][idx] -= h
            fd = (mlp_ref.loss_and_grads(Wp, x, t, w)[0] - mlp_ref.loss_and_grads(Wm, x, t, w)[0]) / (2 * h)
            assert abs(fd - grads[l][0][idx]) < 1e-6 * max(1.0, abs(fd))
        bp = [(A.copy(), c.copy()) for A, c in layers]
        bm = [(A.copy(), c.copy()) for A, c in layers]
        bp[l][1][0] += h; bm[l][1][0] -= h
        fd = (mlp_ref.loss_and_grads(bp, x, t, w)[0] - mlp_ref.loss_and_grads(bm, x, t, w)[0]) / (2 * h)
        assert abs(fd - grads[l][1][0]) < 1e-6 * max(1.0, abs(fd))


def bf16_round(a):
    """Round float array to bfloat16 (RNE), back to float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.gpu
@pytest.mark.parametrize('dims', [(5, 64, 2, 70, 128), (5, 256, 4, 651, 4096), (3, 128, 3, 100, 256),
                                  (4, 64, 2, 37, 192)])
def test_gpu_forward_and_grads_match_reference(dims):
    from monte_carlompi_amd.mlp import Mlp
    n_in, H, L, n_out, batch = dims
    rng = np.random.default_rng(1)
    net = Mlp(n_in, H, L, n_out, max_batch=max(batch, 256))
    layers = small_net(rng, n_in, H, L, n_out)
    net.set_params(Mlp.flatten(layers))
    got = net.get_params()
    np.testing.assert_allclose(got, Mlp.flatten(layers), rtol=1e-7)
    x = rng.normal(size=(batch, n_in)); t = rng.normal(size=(batch, n_out)) * 0.5
    w = rng.uniform(0.5, 1.5, n_out)
    # reference on bf16-rounded weights and inputs (what the GEMMs consume)
    lref = [(bf16_round(W), b) for W, b in layers]
    y_ref, _, _ = mlp_ref.forward(lref, bf16_round(x))
    y = net.forward(x)
    err = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    assert err < 2e-2, err
    g, loss = net.grads(x, t, w)
    loss_ref, gref = mlp_ref.loss_and_grads(lref, bf16_round(x), t, w)
    assert abs(loss - loss_ref) < 2e-2 * loss_ref
    for l, (gW, gb) in enumerate(net.unflatten(g)):
        rW, rb = gref[l]
        eW = np.linalg.norm(gW - rW) / np.linalg.norm(rW)
        eb = np.linalg.norm(gb - rb) / np.linalg.norm(rb)
        assert eW < 3e-2 and eb < 3e-2, (l, eW, eb)
    # cosine similarity of the full gradient
    gr = np.concatenate([np.concatenate([a.ravel(), b]) for a, b in gref])
    assert np.dot(g, gr) / np.linalg.norm(g) / np.linalg.norm(gr) > 0.999
    net.close()


@pytest.mark.gpu
def test_gpu_training_fits_a_smooth_function():
    from monte_carlompi_amd.mlp import Mlp
    rng = np.random.default_rng(2)
    n, n_in, n_out = 8192, 5, 40
    x = rng.uniform(-1, 1, (n, n_in))
    basis = rng.normal(size=(n_in, n_out))
    t = np.tanh(x @ basis) + 0.3 * np.sin(2 * x[:, :1])
    net = Mlp(n_in, 128, 3, n_out, max_batch=1024)
    net.init_params(seed=3)
    net.set_dataset(x, t)
    w = np.full(n_out, 1.0 / n_out)
    losses = []
    for epoch in range(20):
        idx = rng.permutation(n)[: (n // 512) * 512]
        losses.extend(net.train_steps(idx, 512, w, lr=3e-3))
    assert losses[-1] < 0.05 * losses[0], (losses[0], losses[-1])
    y = net.forward(x[:1000])
    mse = np.mean((y - t[:1000]) ** 2)
    assert abs(mse - losses[-1]) < 0.5 * losses[-1] + 1e-3
    net.close()


@pytest.mark.gpu
def test_fused_inference_equals_the_per_layer_kernels(monkeypatch):
    """256-wide networks infer through one fused kernel (activations in LDS);
    its MFMA operands, k order and epilogues are the per-layer GEMMs', so
    forward and serving outputs are bit-identical to theirs, for row counts
    that are not multiples of the fused kernel's 32-row tiles and over
    several passes of max_batch rows."""
    import torch
    from monte_carlompi_amd import surrogate as S
    from monte_carlompi_amd._lib import MCS_FEATURES_SNOW_BRF
    from monte_carlompi_amd.mlp import Mlp
    rng = np.random.default_rng(7)
    shift = rng.normal(size=S.N_OUT) * 0.1
    nets = []
    for layered in ('0', '1'):
        monkeypatch.setenv('MCS_MLP_LAYERED', layered)
        net = Mlp(S.N_FEAT, 256, 4, S.N_OUT, max_batch=4096)
        net.init_params(seed=3)
        net.set_serving(MCS_FEATURES_SNOW_BRF, np.full(S.N_OUT, 1e-2), shift, np.full(S.N_OUT, 1e-3),
                        S.FEAT_MU, S.FEAT_SD)
        nets.append(net)
    try:
        x = rng.normal(size=(9001, S.N_FEAT)).astype(np.float32)
        np.testing.assert_array_equal(nets[0].forward(x), nets[1].forward(x))
        q = torch.from_numpy(S.sample_inputs(10_007, 5).astype(np.float32)).cuda()
        ys = []
        for net in nets:
            y = torch.empty((q.shape[0], S.N_OUT), dtype=torch.float32, device='cuda')
            net.serve(q, q.shape[0], y)
            ys.append(y.cpu().numpy())
        assert np.all(np.isfinite(ys[0]))
        np.testing.assert_array_equal(ys[0], ys[1])
    finally:
        for net in nets:
            net.close()
