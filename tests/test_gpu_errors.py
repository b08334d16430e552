"""Error handling through the C ABI (every check raises before any launch, or
is reported after a normal kernel exit): invalid cases and bands, call order,
replay inputs, the per-photon step cap, and the context staying usable."""
import ctypes

import numpy as np
import pytest

from monte_carlompi_amd import _lib
from monte_carlompi_amd._lib import McsError
from monte_carlompi_amd.gpu import Band, Case, GpuWalker

pytestmark = pytest.mark.gpu


def good_case(n=1000, **kw):
    b = dict(wvl=1.0, g=0.85, ssa_ice=0.99, ext_cff=5000.0)
    b.update({k: kw.pop(k) for k in list(kw) if k in b or k in ('ssa_imp', 'p_ext_imp')})
    c = dict(tau_tot=5.0, theta0_deg=30.0, n_photons=n)
    c.update(kw)
    return Case(bands=[Band(**b)], **c)


@pytest.fixture(scope='module')
def walker():
    w = GpuWalker(0)
    yield w
    w.close()


@pytest.mark.parametrize('bad', [dict(g=1.0), dict(g=float('nan')), dict(ssa_ice=1.5), dict(ext_cff=-1.0),
                                 dict(tau_tot=0.0), dict(tau_tot=float('inf')), dict(p_ext_imp=-0.1),
                                 dict(r_lambertian=1.2), dict(n_photons=-1)])
def test_invalid_inputs_are_rejected(walker, bad):
    with pytest.raises(McsError, match='ARG'):
        walker.set_cases([good_case(**bad)])


def test_call_order_and_arguments(walker):
    w2 = GpuWalker(0)
    with pytest.raises(McsError, match='STATE'):
        w2.walk()                                  # no cases yet
    w2.set_cases([good_case(100)])
    with pytest.raises(McsError, match='STATE'):
        w2.records()                               # nothing walked with records
    w2.walk(records=False)
    with pytest.raises(McsError, match='STATE'):
        w2.records()
    with pytest.raises(McsError, match='ARG'):
        w2.walk(photon_band=np.full(100, 3, np.uint8))     # band 3 of a 1-band case
    with pytest.raises(McsError, match='ARG'):
        w2.walk(mode='replay')                     # replay needs deviates + bands
    with pytest.raises(McsError, match='ARG'):
        w2.set_launch(block=100)                   # not a multiple of 64
    with pytest.raises(McsError, match='ARG'):
        w2.set_launch(coop_live=33)                # tail mode serves at most 32 photons
    with pytest.raises(McsError, match='ARG'):
        w2.set_launch(chunk=(1 << 24) + 1)         # chunks hold at most 2^24 photons
    w2.close()


def test_replay_slice_shorter_than_triple(walker):
    walker.set_cases([good_case(2)])
    vals = np.full(5, 0.5)
    offs = np.array([0, 3, 5], np.int64)           # second slice has 2 deviates
    with pytest.raises(McsError, match='ARG'):
        walker.walk(mode='replay', replay=(vals, offs), photon_band=np.zeros(2, np.uint8))


def test_step_cap_is_reported_and_context_recovers(walker):
    # a non-absorbing thick layer: walks are long; cap them at 50 steps
    walker.set_cases([good_case(20_000, ssa_ice=1.0, tau_tot=1e4)])
    with pytest.raises(McsError, match='STEP_CAP'):
        walker.walk(records=True, max_steps=50)
    c = walker.counts()[0]
    assert c[6] > 0 and c[0] == 20_000            # capped photons counted, all accounted for
    r = walker.records()
    assert np.all(r['n_scat'][r['condition'] == 0] == 49)
    walker.set_cases([good_case(20_000)])
    walker.walk(records=True)                      # usable afterwards
    assert walker.counts()[0][0] == 20_000


def test_destroy_null_and_errors_text():
    L = _lib.lib()
    assert L.mcs_destroy(None) == -1
    ctx = ctypes.c_void_p()
    assert L.mcs_create(999, ctypes.byref(ctx)) == -1 and not ctx.value


@pytest.mark.parametrize('coop', [0, 32])
@pytest.mark.parametrize('cap', [1, 2, 7])
def test_step_cap_with_lambertian_reemission(walker, coop, cap):
    """A thin layer over a white Lambertian bottom: photons bounce often; no
    record exceeds max_steps steps and capped photons stop at exactly
    max_steps (n_scat = max_steps - 1), in the one-lane walk (coop_live 0) and
    with most steps cooperative (one chunk per wave)."""
    case = good_case(64 * walker.cus, tau_tot=0.3, r_lambertian=1.0, lambertian_bottom=True, ssa_ice=1.0)
    walker.set_cases([case])
    walker.set_launch(64, 1, 64, 1, coop)
    try:
        try:
            walker.walk(records=True, max_steps=cap)
        except McsError as e:
            assert 'STEP_CAP' in str(e)
        r = walker.records()
        c = walker.counts()[0]
        assert c[0] == case.n_photons
        assert np.all(r['n_scat'] <= cap - 1)
        capped = r['condition'] == 0
        assert int(capped.sum()) == int(c[6])
        assert np.all(r['n_scat'][capped] == cap - 1)
        assert int(c[8]) == int(np.sum(r['n_scat'] + 1))
    finally:
        walker.set_launch()


def test_cases_sharing_band_rows(walker):
    """Cases may share band rows only as a whole range (each row's cdf and
    alias entry belong to one band distribution)."""
    carr, barr = _lib.McsCase * 2, _lib.McsBand * 3
    bands = barr(*[_lib.McsBand(1.0, 0.85, 0.99, 0.0, 5000.0, 0.0, 0.001, c) for c in (0.5, 1.0, 1.0)])
    ok = carr(_lib.McsCase(5.0, 0.5, -np.sqrt(0.75), 0.0, 0, 0, 2, 0, 100),
              _lib.McsCase(5.0, 0.5, -np.sqrt(0.75), 0.0, 0, 0, 2, 0, 100))
    assert _lib.lib().mcs_set_cases(walker._ctx, ok, 2, bands, 3) == 0
    bad = carr(_lib.McsCase(5.0, 0.5, -np.sqrt(0.75), 0.0, 0, 0, 2, 0, 100),
               _lib.McsCase(5.0, 0.5, -np.sqrt(0.75), 0.0, 0, 1, 2, 0, 100))
    assert _lib.lib().mcs_set_cases(walker._ctx, bad, 2, bands, 3) == -1   # MCS_ERR_ARG


def test_out_of_memory_is_reported_and_context_recovers(monkeypatch):
    """Device allocations that fail (MCS_FAIL_ALLOC_BYTES makes every buffer
    of that size or more fail like hipMalloc would) come back as
    MCS_ERR_NOMEM from the call that needed them, nothing is launched, and
    the same context works again once memory is available: growing the case
    table fails (the context then holds no cases: MCS_ERR_STATE), growing the
    record buffers fails (the previous records are no longer readable), and
    a walk without records still runs at the failing size."""
    w = GpuWalker(0)
    try:
        w.set_cases([good_case(1000)])
        w.walk(records=True)
        assert w.counts()[0][0] == 1000
        # case table: 1000 cases need more than the one-case buffer holds
        monkeypatch.setenv('MCS_FAIL_ALLOC_BYTES', '4096')
        with pytest.raises(McsError, match='NOMEM'):
            w.set_cases([good_case(10) for _ in range(1000)])
        with pytest.raises(McsError, match='STATE'):
            w.walk()
        monkeypatch.delenv('MCS_FAIL_ALLOC_BYTES')
        w.set_cases([good_case(10) for _ in range(1000)])
        w.walk(records=True)
        assert int(w.counts()[:, 0].sum()) == 10_000
        # record buffers: 4 bytes x 2e7 photons = 80 MB each
        w.set_cases([good_case(20_000_000, tau_tot=0.5)])
        monkeypatch.setenv('MCS_FAIL_ALLOC_BYTES', str(64 << 20))
        with pytest.raises(McsError, match='NOMEM'):
            w.walk(records=True)
        with pytest.raises(McsError, match='STATE'):
            w.records()
        w.walk(records=False)                      # needs no new buffer
        assert w.counts()[0][0] == 20_000_000
        monkeypatch.delenv('MCS_FAIL_ALLOC_BYTES')
        w.walk(records=True)
        r = w.records()
        assert r['condition'].size == 20_000_000 and np.all(r['condition'] >= 1)
    finally:
        w.close()
