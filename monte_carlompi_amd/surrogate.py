"""Learned BRF surrogate with the plain Monte Carlo as fallback and auditor.

Inputs per case: (r_eff [um], omega, g, theta0 [deg], tau) -- the snowpack
optical properties the walk consumes.  The network sees eight standardised
functions of those five numbers (log r_eff, log10(1 - omega), g, cos theta0,
log10 tau, and three classic combinations of them, see ``features``).
Outputs (learned in log space so the loss tracks relative error): the
18 x 36 BRF grid (648 bins, phi-symmetrised: the layer is mirror symmetric
about the plane of incidence, so bins k and 35 - k estimate the same value),
albedo, diffuse and direct transmittance -- 651 numbers.

Training data come only from the GPU walk (batched sweeps, Philox streams);
the network is an MLP 8 -> 4 x 256 (SiLU) -> 651 trained by the BF16-MFMA
kernels with FP32 Adam.  Serving evaluates the network for every query,
re-runs a seeded fraction of the queries with the plain walk at high photon
counts, and falls back to the plain walk for every query in an input region
where the audited errors exceed the bars (albedo > 1 %, or BRF bins beyond
3 % once the audit walk's own Monte Carlo noise is removed).

    python -m monte_carlompi_amd.surrogate --help
"""
from __future__ import annotations

import json
import time
from dataclasses import dataclass, field

import numpy as np

from . import gpu as gpumod
from .gpu import BRF_BINS, PHI_BINS, THETA_BINS, Band, Case, GpuWalker
from .workloads import RHO_SNW, sphere_ext_cff_mss

N_OUT = BRF_BINS + 3
RANGES = dict(r_eff=(30.0, 1500.0), eps=(1e-5, 5e-2), g=(0.85, 0.90), theta0=(0.0, 70.0),
              tau=(1.0, 1e4))


def sample_training_inputs(n, seed, thin_frac=0.25, thin_tau=10.0):
    """Training cases: the query distribution of sample_inputs plus an extra
    `thin_frac` share with tau log-uniform in [1, thin_tau] (thin layers
    carry the sharpest dependence on tau and theta0)."""
    base = sample_inputs(n, seed)
    k = int(round(thin_frac * n))
    if k:
        extra = sample_inputs(k, seed + 1)
        rng = np.random.Generator(np.random.PCG64(seed + 2))
        extra[:, 4] = np.exp(rng.uniform(0.0, np.log(thin_tau), k))
        base = np.concatenate([base, extra])
    return base


def sample_inputs(n, seed):
    """n cases: log-uniform r_eff, omega = 1 - U(1e-5, 5e-2), g ~ U(0.85, 0.90),
    theta0 ~ U(0, 70) deg, log-uniform tau in [1, 1e4].  Columns:
    r_eff, omega, g, theta0, tau."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = RANGES['r_eff']
    r = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    om = 1.0 - rng.uniform(*RANGES['eps'], n)
    g = rng.uniform(*RANGES['g'], n)
    th = rng.uniform(*RANGES['theta0'], n)
    tlo, thi = RANGES['tau']
    tau = np.exp(rng.uniform(np.log(tlo), np.log(thi), n))
    return np.stack([r, om, g, th, tau], axis=1)


N_FEAT = 9
# standardisation (f - mu) / sd of the nine features (fixed, roughly the
# query distribution's mean and spread); the device feature map
# (MCS_FEATURES_SNOW_BRF) takes the same constants
FEAT_MU = np.array([5.2, -2.0, 0.875, 0.7, 2.0, 0.05, 0.35, 1.1, 2.15])
FEAT_SD = np.array([1.13, 0.75, 0.0144, 0.25, 1.15, 0.15, 0.17, 1.15, 1.15])


def raw_features(inp):
    """The nine network inputs before standardisation, all functions of
    (r_eff, omega, g, theta0, tau): log r_eff, log10(1 - omega), g,
    cos theta0, log10 tau, plus three classic radiative-transfer
    combinations -- the direct-beam transmittance exp(-tau / mu0), the
    similarity parameter sqrt((1 - omega) / (1 - omega g)), the scaled depth
    log10(tau (1 - omega g)) -- and the slant depth log10(tau / mu0)."""
    r, om, g, th, tau = (np.asarray(inp[:, k], np.float64) for k in range(5))
    mu0 = np.cos(np.radians(th))
    return np.stack([np.log(r), np.log10(np.maximum(1.0 - om, 1e-12)), g, mu0, np.log10(tau),
                     np.exp(-tau / mu0), np.sqrt((1.0 - om) / (1.0 - om * g)),
                     np.log10(tau * (1.0 - om * g)), np.log10(tau / mu0)], axis=1)


def features(inp):
    """Standardised network inputs (raw_features - FEAT_MU) / FEAT_SD."""
    return ((raw_features(inp) - FEAT_MU) / FEAT_SD).astype(np.float32)


# targets are learned in log space so that the loss follows relative error
OFFSETS = np.concatenate([np.full(BRF_BINS, 1e-2), [1e-4, 1e-3, 1e-6]])


def to_log(targets, var=None):
    """log(target + offset), plus var / (2 (target + offset)^2) when the
    targets' Monte Carlo variance is given: the delta-method correction that
    makes the log of a noisy estimate unbiased for the log of its mean to
    second order (E[log x] ~ log E[x] - var / (2 E[x]^2))."""
    x = np.maximum(targets, 0.0) + OFFSETS
    y = np.log(x)
    if var is not None:
        y = y + var / (2.0 * x * x)
    return y


def from_log(y):
    return np.exp(y) - OFFSETS


def mc_cases(inp, photons, wvl=1.0):
    """Plain-walk cases for inputs (free bottom, one wavelength band)."""
    out = []
    for r, om, g, th, tau in inp:
        ext_cff = sphere_ext_cff_mss(r) * RHO_SNW
        out.append(Case(tau_tot=float(tau), theta0_deg=float(th), n_photons=int(photons),
                        bands=[Band(wvl=wvl, g=float(g), ssa_ice=float(om), ext_cff=ext_cff)]))
    return out


def symmetrise_brf(brf):
    """Average the 18 x 36 grid with its mirror phi -> 2 pi - phi."""
    b = brf.reshape(-1, THETA_BINS, PHI_BINS)
    return 0.5 * (b + b[:, :, ::-1])


def targets_from_tallies(t):
    brf = symmetrise_brf(gpumod.brf_from_tallies(t)).reshape(-1, BRF_BINS)
    fr = gpumod.fractions_from_tallies(t)
    return np.concatenate([brf, fr['albedo'][:, None], fr['t_diffuse'][:, None],
                           fr['t_direct'][:, None]], axis=1)


def target_variance(t, targets, photons):
    """Monte Carlo variance of each target of targets_from_tallies for cases
    of `photons` photons of one wavelength: a BRF bin estimated from c
    reflected photons (pooled over the two mirrored bins) has variance
    BRF^2 / c; a fraction p has variance p (1 - p) / photons."""
    H = t[:, :BRF_BINS].reshape(-1, THETA_BINS, PHI_BINS)
    per_photon = t[:, gpumod.T_QDOWN] / photons               # wvn of the case
    c = (H + H[:, :, ::-1]) / per_photon[:, None, None]        # photons in the pooled bins
    c = c.reshape(-1, BRF_BINS)
    var = np.zeros_like(targets)
    b = targets[:, :BRF_BINS]
    var[:, :BRF_BINS] = np.where(c > 0, b * b / np.maximum(c, 1.0), 0.0)
    p = targets[:, BRF_BINS:]
    var[:, BRF_BINS:] = np.clip(p, 0, 1) * (1 - np.clip(p, 0, 1)) / photons
    return var


def run_mc(walker, inp, photons, seed, chunk_cases=None, use_graph=True):
    """Targets, tallies and device seconds of the plain walk for all inputs
    (launches of at most ~2e9 photons)."""
    tg, tl = [], []
    dev_s = 0.0
    if chunk_cases is None:
        chunk_cases = int(max(1, min(20000, 2e9 // max(photons, 1))))
    for s in range(0, inp.shape[0], chunk_cases):
        part = inp[s:s + chunk_cases]
        walker.set_cases(mc_cases(part, photons))
        walker.walk(mode='native', seed=seed + s, records=False, use_graph=use_graph)
        dev_s += walker.stats()['kernel_ms'] * 1e-3
        t = walker.tallies()
        tl.append(t)
        tg.append(targets_from_tallies(t))
    return np.concatenate(tg), np.concatenate(tl), dev_s


@dataclass
class Surrogate:
    hidden: int = 256
    layers: int = 4
    device: int = 0
    max_batch: int = 65536
    y_mean: np.ndarray = None
    y_std: np.ndarray = None
    history: list = field(default_factory=list)

    def __post_init__(self):
        from .mlp import Mlp
        self.net = Mlp(N_FEAT, self.hidden, self.layers, N_OUT, max_batch=self.max_batch, device=self.device)

    def out_weights(self):
        w = np.full(N_OUT, 1.0 / BRF_BINS)
        w[BRF_BINS:] = (1.0, 0.5, 0.5)
        return (w / w.sum()).astype(np.float32)

    def fit(self, inp, targets, epochs=200, batch=4096, lr=2e-3, lr_end=2e-5, seed=0,
            weight_decay=0.0, log_every=0, var=None):
        """Train on walk targets; `var` (target_variance) bias-corrects the
        log transform of noisy targets."""
        x = features(inp)
        lt = to_log(targets, var)
        self.y_mean = lt.mean(axis=0)
        self.y_std = lt.std(axis=0) + 1e-6
        y = ((lt - self.y_mean) / self.y_std).astype(np.float32)
        self.net.init_params(seed=seed)
        self.net.set_dataset(x, y)
        n = x.shape[0]
        rng = np.random.default_rng(seed)
        w = self.out_weights()
        steps_per_epoch = n // batch
        total = epochs * steps_per_epoch
        done = 0
        t0 = time.perf_counter()
        for ep in range(epochs):
            perm = rng.permutation(n)[:steps_per_epoch * batch].astype(np.int32)
            # cosine schedule, updated per chunk of steps
            for s in range(0, steps_per_epoch, 256):
                k = min(256, steps_per_epoch - s)
                frac = done / max(total, 1)
                lr_now = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * frac))
                losses = self.net.train_steps(perm[s * batch:(s + k) * batch], batch, w, lr=lr_now,
                                              weight_decay=weight_decay)
                done += k
            self.history.append(float(losses[-1]))
            if log_every and (ep % log_every == 0 or ep == epochs - 1):
                print('epoch %4d loss %.5f lr %.2e (%.1f s)' % (ep, losses[-1], lr_now,
                                                                time.perf_counter() - t0), flush=True)
        return self.history

    def predict(self, inp):
        """Host path: features and output transform in numpy around
        mcs_mlp_forward (host buffers)."""
        y = self.net.forward(features(inp))
        return from_log(y * self.y_std + self.y_mean)

    def enable_serving(self):
        """Device path: features (MCS_FEATURES_SNOW_BRF) and the output
        transform exp(y_std y + y_mean) - OFFSETS run on the GPU."""
        from ._lib import MCS_FEATURES_SNOW_BRF
        self.net.set_serving(MCS_FEATURES_SNOW_BRF, self.y_std, self.y_mean, OFFSETS, FEAT_MU, FEAT_SD)

    def predict_device(self, q_dev, out=None):
        """Answers for queries already on the device (torch float32 tensor
        n x 5: r_eff, omega, g, theta0, tau) into a device tensor n x 651
        (allocated when ``out`` is None): one mcs_mlp_serve call."""
        import torch
        n = int(q_dev.shape[0])
        if out is None:
            out = torch.empty((n, N_OUT), dtype=torch.float32, device=q_dev.device)
        torch.cuda.synchronize(q_dev.device)       # q_dev written on torch's stream
        self.net.serve(q_dev, n, out)
        return out

    def predict_ms(self):
        return self.net.last_ms()

    def save(self, path):
        np.savez(path, params=self.net.get_params(), y_mean=self.y_mean, y_std=self.y_std,
                 hidden=self.hidden, layers=self.layers)

    def load(self, path):
        d = np.load(path, allow_pickle=False)
        self.y_mean, self.y_std = d['y_mean'], d['y_std']
        self.net.set_params(d['params'])


# ---------------------------------------------------------------- judging
# A BRF bin of a case fails when its relative difference from the 1e6-photon
# walk exceeds the 3 % bar by more than Z_BIN standard errors of that walk's
# own estimate (binomial: 1 / sqrt(photons in the bin)); Z_BIN = 4.3 makes the
# chance that a prediction equal to the true BRF fails any of ~300 judged
# bins through the walk's noise alone about 1 %.  Bins judged: those holding
# at least MIN_BIN_PHOTONS reflected photons (two mirrored bins pooled).
Z_BIN = 4.3
MIN_BIN_PHOTONS = 100


def bin_weights():
    """w_jk of brf_from_tallies per 18 x 36 bin: BRF_jk = H_jk / (Q_down w_jk)."""
    edges = np.linspace(0.0, np.pi / 2, THETA_BINS + 1)
    mid = (np.diff(edges) / 2.) + edges[:-1]
    w = np.sin(mid) * np.cos(mid)
    return np.repeat(w / np.sum(w) / PHI_BINS, PHI_BINS)


def case_errors(pred, mc, photons, min_bin_photons=MIN_BIN_PHOTONS, brf_bar=0.03, z_bin=Z_BIN,
                pred_photons=None):
    """Per-case errors of predictions against Monte Carlo runs of `photons`
    photons (one wavelength).  A BRF bin's relative difference is taken
    against the prediction, rel = (mc - pred) / pred, and its binomial
    standard error from the count the prediction expects there, n = pred x 2
    w_jk x photons (two mirrored bins pooled): both unbiased for an exact
    prediction, unlike scaling by the walk's own noisy count.  Bins expecting
    >= min_bin_photons photons are judged.
      albedo          |A_pred - A_mc| / A_mc
      brf_bins_fail   judged bins with |rel| > brf_bar + z_bin / sqrt(n)
      brf_bin_excess  max over judged bins of |rel| - z_bin / sqrt(n)
      brf_rms_excess  rms of rel over judged bins, walk noise removed in quadrature
      brf_median      median |rel| over judged bins (walk noise included).
    ``pred_photons`` (per case, 0 for network answers): where a prediction is
    itself a walk's answer (a fallback), its own binomial noise is added in
    quadrature, and its albedo gets 3 standard errors of the difference of
    the two walks beside the 1 % -- otherwise two exact walks of 1e6 photons
    would fail each other."""
    alb_err = np.abs(pred[:, BRF_BINS] - mc[:, BRF_BINS]) / np.maximum(mc[:, BRF_BINS], 1e-12)
    pb = pred[:, :BRF_BINS]
    nbin = np.maximum(pb, 0.0) * (2.0 * photons) * bin_weights()[None, :]
    rel = (mc[:, :BRF_BINS] - pb) / np.where(pb > 0, pb, 1.0)
    sig = 1.0 / np.sqrt(np.maximum(nbin, 1.0))
    if pred_photons is not None:
        pp = np.asarray(pred_photons, float)
        walked = pp > 0
        ratio = np.where(walked, photons / np.maximum(pp, 1.0), 0.0)      # noise of the prediction / of the judge
        sig = sig * np.sqrt(1.0 + ratio)[:, None]
        a = np.clip(mc[:, BRF_BINS], 1e-12, 1.0)
        sa = np.sqrt(a * (1 - a) * (1.0 / photons + np.where(walked, 1.0 / np.maximum(pp, 1.0), 0.0))) / a
        alb_err = np.where(walked, np.maximum(alb_err - 3.0 * sa, 0.0), alb_err)
    ok = nbin >= min_bin_photons
    arel = np.abs(rel)
    fails = ok & (arel > brf_bar + z_bin * sig)
    excess = np.where(ok, arel - z_bin * sig, -np.inf).max(axis=1)
    excess = np.where(np.isfinite(excess), excess, 0.0)
    cnt = ok.sum(axis=1)
    msq = np.where(ok, rel * rel, 0).sum(axis=1) / np.maximum(cnt, 1)
    mns = np.where(ok, sig * sig, 0).sum(axis=1) / np.maximum(cnt, 1)
    rms_corr = np.sqrt(np.maximum(0.0, msq - mns))
    med = np.array([np.median(arel[i, ok[i]]) if cnt[i] else 0.0 for i in range(pred.shape[0])])
    return dict(albedo=alb_err, brf_bins_fail=fails.sum(axis=1), brf_bin_excess=excess,
                brf_rms_excess=rms_corr, brf_median=med, bins_judged=cnt,
                rel=rel.astype(np.float32), sig2=(sig * sig).astype(np.float32), ok=ok)


def per_bin_rms_excess(err, cases=None, min_cases=20, z=4.0, bar=0.03):
    """Per BRF bin, over a set of cases: the mean squared relative error with
    the judging walks' noise removed, m_k = mean(rel^2) - mean(sigma^2), from
    the cases where the bin is judged.  Returns (rms, over): rms_k =
    sqrt(max(m_k, 0)) (NaN for bins judged in fewer than `min_cases` cases),
    and over_k = m_k - z SE_k > bar^2 with SE_k the standard error of m_k
    over those cases -- the bin's rms error exceeds `bar` with z-sigma
    confidence.  Averaging over cases resolves per-bin errors well below one
    walk's per-bin noise (1-3 % at 1e6 photons)."""
    rel, s2, ok = err['rel'], err['sig2'], err['ok']
    if cases is not None:
        rel, s2, ok = rel[cases], s2[cases], ok[cases]
    n = ok.sum(axis=0)
    d = np.where(ok, rel.astype(np.float64) ** 2 - s2.astype(np.float64), 0.0)
    nn = np.maximum(n, 1)
    m = d.sum(axis=0) / nn
    var = np.where(ok, (d - m[None, :]) ** 2, 0.0).sum(axis=0) / np.maximum(n - 1, 1)
    se = np.sqrt(var / nn)
    enough = n >= min_cases
    rms = np.where(enough, np.sqrt(np.maximum(m, 0.0)), np.nan)
    over = enough & (m - z * se > bar * bar)
    return rms, over


def case_fails(err, alb_bar=0.01):
    """The fallback bars: albedo error > alb_bar or any BRF bin failing."""
    return (err['albedo'] > alb_bar) | (err['brf_bins_fail'] > 0)


def region_of(inp, n_tau=8, n_eps=6, n_mu=3):
    """Input-space region: a grid over log10 tau x log10(1 - omega) x cos theta0."""
    lt = np.log10(inp[:, 4]) / 4.0
    le = (np.log10(1 - inp[:, 1]) + 5.0) / (5.0 + np.log10(5e-2))
    lm = (np.cos(np.radians(inp[:, 3])) - np.cos(np.radians(70.0))) / (1.0 - np.cos(np.radians(70.0)))
    a = np.clip((lt * n_tau).astype(int), 0, n_tau - 1)
    b = np.clip((le * n_eps).astype(int), 0, n_eps - 1)
    c = np.clip((lm * n_mu).astype(int), 0, n_mu - 1)
    return (a * n_eps + b) * n_mu + c


def _pct(x):
    x = np.asarray(x, float)
    return dict(median=float(np.median(x)), p95=float(np.quantile(x, 0.95)), max=float(x.max()))


def serve(sur: Surrogate, walker: GpuWalker, queries, audit_frac=0.01, audit_photons=1_000_000,
          seed=0, alb_bar=0.01, brf_bar=0.03, region_fail_frac=0.05, fallback_photons=1_000_000,
          run_fallback=True, validate=0):
    """Serve every query: the network on the GPU for all of them, a seeded
    audit subset re-run by the plain walk (its answers replace the
    network's), and the walk for every query of an input region where more
    than `region_fail_frac` of the audited cases miss a bar (albedo error >
    alb_bar, or a BRF bin beyond brf_bar plus the walk's noise allowance).
    With `validate` > 0, that many further queries (not audited) are walked
    at audit_photons after serving, and the served answers -- network or
    fallback -- are judged against them by the same bars.
    Returns (answers as a device tensor n x 651, report)."""
    import torch
    dev = torch.device('cuda', getattr(walker, 'device', 0))
    n = queries.shape[0]
    t0 = time.perf_counter()
    q_dev = torch.from_numpy(np.ascontiguousarray(queries, dtype=np.float32)).to(dev)
    answers = sur.predict_device(q_dev)
    t_pred = time.perf_counter() - t0
    pred_dev_s = sur.predict_ms() * 1e-3
    rng = np.random.Generator(np.random.PCG64(seed))
    n_audit = max(1, int(round(audit_frac * n)))
    perm = rng.permutation(n)
    aud = np.sort(perm[:n_audit])
    t1 = time.perf_counter()
    mc, mct, aud_dev_s = run_mc(walker, queries[aud], audit_photons, seed=seed + 7)
    aud_idx = torch.from_numpy(aud).to(dev)
    pred_aud = answers.index_select(0, aud_idx).cpu().numpy()
    err = case_errors(pred_aud, mc, audit_photons, brf_bar=brf_bar)
    fail = case_fails(err, alb_bar)
    # a region falls back when more than region_fail_frac of its audited cases
    # miss a bar, or when some BRF bin's rms error over its audited cases
    # (walk noise removed) exceeds the 3 % bar with 4-sigma confidence
    reg_q = region_of(queries)
    reg_a = reg_q[aud]
    bad_regions = []
    region_bin_rms = {}
    for r in np.unique(reg_a):
        m = reg_a == r
        rb, over = per_bin_rms_excess(err, m, bar=brf_bar)
        region_bin_rms[int(r)] = int(over.sum())
        if np.mean(fail[m]) > region_fail_frac or over.any():
            bad_regions.append(int(r))
    fb = np.isin(reg_q, bad_regions)
    fb[aud] = False
    answers.index_copy_(0, aud_idx, torch.from_numpy(mc.astype(np.float32)).to(dev))   # audited: the walk's answer
    fb_dev_s = 0.0
    if run_fallback and fb.any():
        fidx = np.nonzero(fb)[0]
        fmc, _, fb_dev_s = run_mc(walker, queries[fidx], fallback_photons, seed=seed + 99)
        answers.index_copy_(0, torch.from_numpy(fidx).to(dev), torch.from_numpy(fmc.astype(np.float32)).to(dev))
    torch.cuda.synchronize(dev)
    t_served = time.perf_counter() - t0
    dec = np.clip(np.log10(queries[aud, 4]).astype(int), 0, 3)
    diag = dict(fail_albedo=int(np.sum(err['albedo'] > alb_bar)),
                fail_brf_bins=int(np.sum(err['brf_bins_fail'] > 0)),
                fail_frac_by_tau_decade=[float(np.mean(fail[dec == d])) if np.any(dec == d) else None
                                         for d in range(4)],
                bins_judged_median=float(np.median(err['bins_judged'])))
    rep = dict(n_queries=int(n), n_audit=int(n_audit), audit_fail_frac=float(np.mean(fail)),
               diagnostics=diag,
               fallback_regions=bad_regions, fallback_frac=float(np.mean(fb)),
               fallback_cases=int(fb.sum()),
               predict_s_host=t_pred, predict_s_device=pred_dev_s, audit_s_device=aud_dev_s,
               fallback_s_device=fb_dev_s, served_s_host=t_served, audit_walk_s_host=time.perf_counter() - t1,
               albedo_err=_pct(err['albedo']), brf_bin_excess=_pct(err['brf_bin_excess']),
               brf_rms_excess=_pct(err['brf_rms_excess']),
               brf_median_abs=dict(median=float(np.median(err['brf_median']))),
               audit_bins_over_bar=int(per_bin_rms_excess(err, bar=brf_bar)[1].sum()),
               region_bins_over_bar={k: v for k, v in region_bin_rms.items() if v},
               bars=dict(albedo=alb_bar, brf_bin=brf_bar, z_bin=Z_BIN, min_bin_photons=MIN_BIN_PHOTONS,
                         region_fail_frac=region_fail_frac))
    rep['cases_per_s_surrogate'] = n / max(pred_dev_s, 1e-12)
    rep['cases_per_s_served'] = n / max(pred_dev_s + aud_dev_s + fb_dev_s, 1e-12)
    rep['cases_per_s_served_host'] = n / max(t_served, 1e-12)
    rep['plain_mc_cases_per_s_batched'] = n_audit / max(aud_dev_s, 1e-12)
    if validate:
        val = np.sort(perm[n_audit:n_audit + int(validate)])
        vmc, vmct, _ = run_mc(walker, queries[val], audit_photons, seed=seed + 555)
        got = answers.index_select(0, torch.from_numpy(val).to(dev)).cpu().numpy()
        verr = case_errors(got, vmc, audit_photons, brf_bar=brf_bar,
                           pred_photons=np.where(fb[val], fallback_photons, 0))
        vfail = case_fails(verr, alb_bar)
        vnet = ~fb[val]
        vdec = np.clip(np.log10(queries[val, 4]).astype(int), 0, 3)
        vbin, vover = per_bin_rms_excess(verr, vnet, bar=brf_bar)
        rep['validation'] = dict(cases=int(val.size), served_by_walk=int(fb[val].sum()),
                                 fail_frac=float(np.mean(vfail)),
                                 fail_frac_network=float(np.mean(vfail[vnet])) if vnet.any() else None,
                                 fail_frac_walk=float(np.mean(vfail[~vnet])) if (~vnet).any() else None,
                                 fail_frac_by_tau_decade=[float(np.mean(vfail[vdec == d])) if np.any(vdec == d)
                                                          else None for d in range(4)],
                                 network_bins_over_bar=int(vover.sum()),
                                 network_bin_rms=dict(median=float(np.nanmedian(vbin)),
                                                      p95=float(np.nanquantile(vbin, 0.95)),
                                                      max=float(np.nanmax(vbin))),
                                 fail_albedo_frac=float(np.mean(verr['albedo'] > alb_bar)),
                                 fail_brf_frac=float(np.mean(verr['brf_bins_fail'] > 0)),
                                 albedo_err=_pct(verr['albedo']), brf_bin_excess=_pct(verr['brf_bin_excess']),
                                 brf_rms_excess=_pct(verr['brf_rms_excess']))
    return answers, rep


def plain_mc_rate(walker, inp, photons=1_000_000, seed=5):
    """Cases/s of the plain walk at `photons` per case, all cases in one
    batched launch (device time)."""
    _, _, dev_s = run_mc(walker, inp, photons, seed)
    return inp.shape[0] / dev_s


def plain_mc_rate_per_case(walker, inp, photons=1_000_000, seed=5):
    """Cases/s of the plain walk run the way the drop-in run() runs it: one
    launch per case (device time of each walk kernel, summed)."""
    dev_s = 0.0
    for k, case in enumerate(mc_cases(inp, photons)):
        walker.set_cases([case])
        walker.walk(mode='native', seed=seed + k, records=False)
        dev_s += walker.stats()['kernel_ms'] * 1e-3
    return inp.shape[0] / dev_s


def train(walker, n_cases=200_000, photons=100_000, epochs=200, seed=0, thin_frac=0.0, log_every=0):
    """Config-4 training: walk-generated targets for n_cases sampled inputs
    (plus a `thin_frac` share of extra thin layers), then Adam on the GPU.
    Returns (surrogate, report)."""
    t0 = time.perf_counter()
    inp = sample_training_inputs(n_cases, seed, thin_frac=thin_frac)
    tg, tl, gen_s = run_mc(walker, inp, photons, seed=seed * 1000 + 1)
    var = target_variance(tl, tg, photons)
    del tl
    t_data = time.perf_counter() - t0
    sur = Surrogate()
    t1 = time.perf_counter()
    sur.fit(inp, tg, epochs=epochs, log_every=log_every, seed=seed, var=var)
    train_s = time.perf_counter() - t1
    sur.enable_serving()
    return sur, dict(train_cases=int(inp.shape[0]), train_photons=int(photons), data_walk_s_device=gen_s,
                     data_s_host=t_data, train_s_host=train_s, epochs=int(epochs),
                     adam_steps=int(epochs * (inp.shape[0] // 4096)), final_loss=float(sur.history[-1]))


def main():
    import argparse
    ap = argparse.ArgumentParser(description='train / serve the BRF surrogate on the GPU walk')
    ap.add_argument('--train-cases', type=int, default=200_000)
    ap.add_argument('--train-photons', type=int, default=100_000)
    ap.add_argument('--epochs', type=int, default=200)
    ap.add_argument('--thin-frac', type=float, default=0.0)
    ap.add_argument('--queries', type=int, default=1_000_000)
    ap.add_argument('--audit-frac', type=float, default=0.01)
    ap.add_argument('--audit-photons', type=int, default=1_000_000)
    ap.add_argument('--validate', type=int, default=2000)
    ap.add_argument('--rate-cases', type=int, default=200)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--save', type=str, default='')
    ap.add_argument('--report', type=str, default='')
    a = ap.parse_args()
    w = GpuWalker(0)
    w.set_launch(256, 4, 4096, 8)
    sur, trep = train(w, a.train_cases, a.train_photons, a.epochs, a.seed, a.thin_frac,
                      log_every=max(1, a.epochs // 10))
    print('trained: %s' % json.dumps(trep), flush=True)
    if a.save:
        sur.save(a.save)
    q = sample_inputs(a.queries, a.seed + 12345)
    _, rep = serve(sur, w, q, audit_frac=a.audit_frac, audit_photons=a.audit_photons, seed=a.seed + 3,
                   validate=a.validate)
    per_case = plain_mc_rate_per_case(w, sample_inputs(a.rate_cases, a.seed + 777), a.audit_photons)
    rep.update(trep)
    rep['plain_mc_cases_per_s_per_case'] = per_case
    rep['gain_surrogate'] = rep['cases_per_s_surrogate'] / rep['plain_mc_cases_per_s_batched']
    rep['gain_served_vs_batched_mc'] = rep['cases_per_s_served'] / rep['plain_mc_cases_per_s_batched']
    rep['gain_served_vs_per_case_mc'] = rep['cases_per_s_served'] / per_case
    print(json.dumps(rep, indent=1), flush=True)
    if a.report:
        with open(a.report, 'w') as fh:
            json.dump(rep, fh, indent=1)
    w.close()


if __name__ == '__main__':
    main()
