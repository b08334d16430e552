"""mcs_walk_noalloc: the native walk into caller-owned device memory with its
scratch in the library's static device array.  Checked against mcs_walk on
the same cases, seed and launch shape (tail mode off, so records are
launch-independent): records bit for bit (photon order), counts exactly,
FP64 tallies within 1e-12 relative (atomic summation order), on two calls in
a row, the second with fewer cases and photons."""
import numpy as np
import pytest

from monte_carlompi_amd._lib import McsError, NOALLOC_MAX_CASES
from monte_carlompi_amd.gpu import Band, Case, GpuWalker

pytestmark = pytest.mark.gpu


def cases_a():
    return [Case(tau_tot=40.0, theta0_deg=60.0, n_photons=300_000, lambertian_bottom=True, r_lambertian=0.5,
                 bands=[Band(wvl=1.3, g=0.89, ssa_ice=0.999, ext_cff=4907.3)]),
            Case(tau_tot=5.0, theta0_deg=20.0, n_photons=123_457,
                 bands=[Band(wvl=1.0, g=0.85, ssa_ice=0.99, ext_cff=3000.0, ssa_imp=0.3, p_ext_imp=0.01,
                             weight=2.0),
                        Band(wvl=1.1, g=0.86, ssa_ice=0.98, ext_cff=3100.0, weight=1.0)]),
            Case(tau_tot=2.0, theta0_deg=0.0, n_photons=77_777, lambertian_surface=True, r_lambertian=0.7,
                 bands=[Band(wvl=0.9, g=0.8, ssa_ice=0.995, ext_cff=2000.0)])]


def cases_b():
    return [Case(tau_tot=8.0, theta0_deg=45.0, n_photons=50_001,
                 bands=[Band(wvl=1.2, g=0.87, ssa_ice=0.997, ext_cff=4000.0)])]


def device_out(torch, nc, n):
    dev = torch.device('cuda:0')
    return dict(tallies=torch.empty((nc, 656), dtype=torch.float64, device=dev),
                counts=torch.empty((nc, 16), dtype=torch.int64, device=dev),
                records=dict(n_scat=torch.empty(n, dtype=torch.int32, device=dev),
                             cond_band=torch.empty(n, dtype=torch.int32, device=dev),
                             theta_n=torch.empty(n, dtype=torch.float32, device=dev),
                             phi_n=torch.empty(n, dtype=torch.float32, device=dev),
                             path_length=torch.empty(n, dtype=torch.float32, device=dev)))


def compare(w, out, n, nc, seed):
    """mcs_walk on the same cases and seed, then the no-alloc outputs
    (the first n photons / nc cases of the device buffers) against it."""
    import torch
    w.walk(mode='native', seed=seed, records=True)
    ref = w.records()
    t_ref, c_ref = w.tallies(), w.counts()
    w.walk_noalloc(out['tallies'], out['counts'], records=out['records'], seed=seed)
    with pytest.raises(McsError, match='STATE'):
        w.tallies()                                   # results are in the caller's buffers
    torch.cuda.synchronize()
    r = {k: v[:n].cpu().numpy() for k, v in out['records'].items()}
    t = out['tallies'][:nc].cpu().numpy()
    c = out['counts'][:nc].cpu().numpy().astype(np.uint64)
    cb = r['cond_band'].view(np.uint32)
    np.testing.assert_array_equal(r['n_scat'].astype(np.int64), ref['n_scat'])
    np.testing.assert_array_equal((cb & 0xFF).astype(np.int8), ref['condition'])
    np.testing.assert_array_equal((cb >> 8).astype(np.uint8), ref['band'])
    for f in ('theta_n', 'phi_n', 'path_length'):
        np.testing.assert_array_equal(r[f], ref[f].astype(np.float32), err_msg=f)
    np.testing.assert_array_equal(c, c_ref)
    np.testing.assert_allclose(t, t_ref, rtol=1e-12, atol=0)
    assert int(c[:, 0].sum()) == n


def test_noalloc_walk_equals_mcs_walk_twice_second_smaller():
    import torch
    w = GpuWalker(0)
    try:
        w.set_launch(512, 2, 2048, 8, 0)
        ca = cases_a()
        na = sum(c.n_photons for c in ca)
        out = device_out(torch, len(ca), na)
        w.set_cases(ca)
        compare(w, out, na, len(ca), seed=12345)
        cb = cases_b()
        w.set_cases(cb)
        compare(w, out, cb[0].n_photons, 1, seed=777)   # same (larger) buffers, smaller walk
        # without records: tallies and counts only
        w.walk(mode='native', seed=9, records=False)
        c_ref = w.counts()
        w.walk_noalloc(out['tallies'], out['counts'], seed=9)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out['counts'][:1].cpu().numpy().astype(np.uint64), c_ref)
    finally:
        w.close()


def test_noalloc_walk_rejects_what_it_cannot_do():
    import torch
    w = GpuWalker(0)
    try:
        dev = torch.device('cuda:0')
        t = torch.empty((1, 656), dtype=torch.float64, device=dev)
        c = torch.empty((1, 16), dtype=torch.int64, device=dev)
        with pytest.raises(McsError, match='STATE'):
            w.walk_noalloc(t, c)                       # no cases yet
        w.set_cases(cases_b())
        with pytest.raises(McsError, match='ARG'):
            w.walk_noalloc(None, c)                    # tallies required
        with pytest.raises(McsError, match='ARG'):
            w.walk_noalloc(t, c, records=dict(n_scat=c))   # all five record arrays
        many = [Case(tau_tot=1.0, theta0_deg=0.0, n_photons=1,
                     bands=[Band(wvl=1.0, g=0.8, ssa_ice=0.9, ext_cff=1000.0)])] * (NOALLOC_MAX_CASES + 1)
        w.set_cases(many)
        big_t = torch.empty((len(many), 656), dtype=torch.float64, device=dev)
        big_c = torch.empty((len(many), 16), dtype=torch.int64, device=dev)
        with pytest.raises(McsError, match='ARG'):
            w.walk_noalloc(big_t, big_c)               # beyond the scratch's chunk table
        w.set_cases(many[:NOALLOC_MAX_CASES])          # the largest allowed
        w.walk_noalloc(big_t, big_c, seed=3)
        torch.cuda.synchronize()
        assert int(big_c[:NOALLOC_MAX_CASES, 0].sum().item()) == NOALLOC_MAX_CASES
    finally:
        w.close()
