"""ctypes wrapper of the BF16-MFMA MLP (mcs_mlp_* in include/mcsnow.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import McsAdam, McsMlpDims, check, lib


class Mlp:
    """MLP n_in -> layers x hidden (SiLU) -> n_out on one device."""

    def __init__(self, n_in=5, hidden=256, layers=4, n_out=651, max_batch=65536, device=0):
        self.dims = (n_in, hidden, layers, n_out, max_batch)
        d = McsMlpDims(n_in, hidden, layers, n_out, max_batch, 0)
        self._h = ctypes.c_void_p()
        check(lib().mcs_mlp_create(device, ctypes.byref(d), ctypes.byref(self._h)))
        n = ctypes.c_int64()
        self._check(lib().mcs_mlp_param_count(self._h, ctypes.byref(n)))
        self.n_params = n.value
        self.adam_step = 0

    def _check(self, code):
        check(code, self._h, errfn='mcs_mlp_last_error')

    def close(self):
        if self._h:
            lib().mcs_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameter layout: per layer W [rows x cols] row-major, then b [rows]
    def shapes(self):
        n_in, H, L, n_out, _ = self.dims
        out = []
        for l in range(L + 1):
            rows = n_out if l == L else H
            cols = n_in if l == 0 else H
            out.append((rows, cols))
        return out

    def unflatten(self, flat):
        res, k = [], 0
        for rows, cols in self.shapes():
            W = flat[k:k + rows * cols].reshape(rows, cols); k += rows * cols
            b = flat[k:k + rows]; k += rows
            res.append((W, b))
        return res

    @staticmethod
    def flatten(layers):
        return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers]).astype(np.float32)

    def init_params(self, seed=0, out_scale=0.1):
        rng = np.random.default_rng(seed)
        layers = []
        shp = self.shapes()
        for l, (rows, cols) in enumerate(shp):
            std = np.sqrt(2.0 / cols) * (out_scale if l == len(shp) - 1 else 1.0)
            layers.append((rng.normal(0, std, (rows, cols)), np.zeros(rows)))
        self.set_params(self.flatten(layers))
        return layers

    def set_params(self, flat):
        p = np.ascontiguousarray(flat, dtype=np.float32)
        self._check(lib().mcs_mlp_set_params(self._h, p.ctypes.data, p.size))
        self.adam_step = 0

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        self._check(lib().mcs_mlp_get_params(self._h, p.ctypes.data, p.size))
        return p

    # ---- compute
    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((x.shape[0], self.dims[3]), np.float32)
        self._check(lib().mcs_mlp_forward(self._h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def set_serving(self, feature_map, out_scale, out_shift, out_offset, feat_mu=None, feat_sd=None):
        """Feature map (_lib.MCS_FEATURES_*) and output transform
        exp(scale y + shift) - offset of mcs_mlp_serve."""
        from ._lib import MCS_FEATURES_IDENTITY
        no = self.dims[3]
        arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(no) for a in (out_scale, out_shift, out_offset)]
        fm = fs = None
        if feature_map != MCS_FEATURES_IDENTITY:
            fm = np.ascontiguousarray(feat_mu, dtype=np.float32)
            fs = np.ascontiguousarray(feat_sd, dtype=np.float32)
        self._serving = arrs + [fm, fs]
        self._check(lib().mcs_mlp_set_serving(self._h, int(feature_map), None if fm is None else fm.ctypes.data,
                                              None if fs is None else fs.ctypes.data, arrs[0].ctypes.data,
                                              arrs[1].ctypes.data, arrs[2].ctypes.data))

    def serve(self, x_dev, n, y_dev):
        """mcs_mlp_serve on device memory: ``x_dev`` (n x 5 or n_in float32)
        and ``y_dev`` (n x n_out float32) are device addresses or objects
        with data_ptr()."""
        addr = lambda t: int(t.data_ptr()) if hasattr(t, 'data_ptr') else int(t)
        self._check(lib().mcs_mlp_serve(self._h, addr(x_dev), int(n), addr(y_dev)))

    def last_ms(self):
        v = ctypes.c_double()
        self._check(lib().mcs_mlp_last_ms(self._h, ctypes.byref(v)))
        return v.value

    def grads(self, x, t, w):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        g = np.empty(self.n_params, np.float32)
        loss = ctypes.c_float()
        self._check(lib().mcs_mlp_grads(self._h, x.ctypes.data, t.ctypes.data, w.ctypes.data,
                                        x.shape[0], g.ctypes.data, ctypes.byref(loss)))
        return g, loss.value

    def set_dataset(self, x, t):
        self._x = np.ascontiguousarray(x, dtype=np.float32)
        self._t = np.ascontiguousarray(t, dtype=np.float32)
        self._check(lib().mcs_mlp_set_dataset(self._h, self._x.ctypes.data, self._t.ctypes.data,
                                              self._x.shape[0]))
        self.n_data = self._x.shape[0]

    def train_steps(self, idx, batch, w, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        n_steps = idx.size // batch
        w = np.ascontiguousarray(w, dtype=np.float32)
        a = McsAdam(lr, beta1, beta2, eps, weight_decay, self.adam_step)
        losses = np.empty(n_steps, np.float32)
        self._check(lib().mcs_mlp_train_steps(self._h, idx.ctypes.data, n_steps, batch, w.ctypes.data,
                                              ctypes.byref(a), losses.ctypes.data))
        self.adam_step += n_steps
        return losses
