"""Float64 numpy reference of the MLP (forward, loss, backward) used to check
the BF16-MFMA kernels.  Pinned by a central finite-difference check."""
import numpy as np


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(layers, x):
    acts, zs = [x], []
    a = x
    for l, (W, b) in enumerate(layers):
        z = a @ W.T + b
        zs.append(z)
        a = z if l == len(layers) - 1 else silu(z)
        acts.append(a)
    return a, acts, zs


def loss_and_grads(layers, x, t, w):
    """loss = sum_rows sum_o w_o (y - t)^2 / rows and its parameter gradients."""
    y, acts, zs = forward(layers, x)
    n = x.shape[0]
    d = y - t
    loss = float(np.sum(w * d * d) / n)
    dz = 2.0 * w * d / n
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        grads[l] = (dz.T @ acts[l], dz.sum(axis=0))
        if l > 0:
            dz = (dz @ W) * dsilu(zs[l - 1])
    return loss, grads
