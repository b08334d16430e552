"""The record-file writer (mcs_write_records_text) against CPython: the
reference writes rows with '%d %r %r %r %d %r %r\\n' of numpy-1 float64
scalars, whose %r is Python's float repr (the shortest string that reads
back to the same double, mc3D:1626-1646).  CPython's own formatting of the
same values is the checker; no GPU."""
import ctypes

import numpy as np
import pytest

from monte_carlompi_amd import _lib
from monte_carlompi_amd.monte_carlo3D import MonteCarlo


def c_repr(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    buf = ctypes.create_string_buffer(26 * max(len(x), 1))
    k = _lib.lib().mcs_repr_doubles(x.ctypes.data, len(x), buf, len(buf))
    assert k >= 0
    return buf.raw[:k].decode().split('\n')[:-1]


def assert_same_repr(values):
    got = c_repr(values)
    exp = [repr(float(v)) for v in values]
    bad = [(float(v).hex(), g, e) for v, g, e in zip(values, got, exp) if g != e]
    assert not bad, bad[:10]


def test_repr_of_float32_walk_values():
    """What the native walk produces: FP32 angles and paths widened to
    float64 (16-17 significant digits)."""
    rng = np.random.default_rng(1)
    for scale in (np.pi, 2 * np.pi, 10.0, 1e-3):
        assert_same_repr((rng.random(50_000) * scale).astype(np.float32).astype(np.float64))


def test_repr_of_arbitrary_doubles_and_magnitudes():
    rng = np.random.default_rng(2)
    assert_same_repr(rng.random(50_000))
    assert_same_repr(10.0 ** rng.uniform(-320, 308, 50_000) * rng.choice([-1.0, 1.0], 50_000))
    bits = rng.integers(0, 2 ** 63, 50_000, dtype=np.uint64).view(np.float64)
    assert_same_repr(bits[np.isfinite(bits)])


def test_repr_special_values_and_format_switches():
    """Zeros, infinities, nan, subnormals, the fixed/exponent switch at
    1e-4 / 1e16, integers ('.0'), and every power of two (where the
    rounding interval is one-sided narrower)."""
    vals = [0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
            1e-4, 9.999999999999999e-05, 1e-5, 0.00012345, 1e15, 999999999999999.9, 1e16,
            9999999999999998.0, 1.2345678901234567e16, 123456789012345678.0, 0.1, 0.2, 0.3, 1 / 3,
            0.7692307692307692, 100.0, 1e22, 1e23, 5e-5, 4.35, 0.5e-4]
    assert_same_repr(np.array(vals))
    assert_same_repr(np.ldexp(1.0, np.arange(-1074, 1024)))
    ints = np.arange(-3000, 3000, dtype=np.float64)
    assert_same_repr(ints)
    assert_same_repr(ints / 8)
    assert_same_repr(ints / 10)
    # values next to powers of two and of ten
    p2 = np.ldexp(1.0, np.arange(-60, 60))
    assert_same_repr(np.nextafter(p2, 0))
    assert_same_repr(np.nextafter(p2, np.inf))
    p10 = 10.0 ** np.arange(-20, 22)
    assert_same_repr(np.nextafter(p10, 0))
    assert_same_repr(np.nextafter(p10, np.inf))


def python_rows(rec):
    """The reference's row text for the same columns (CPython formatting)."""
    cond, nsc = rec['condition'].tolist(), rec['n_scat'].tolist()
    cols = [rec[k].tolist() for k in ('wvn', 'theta_n', 'phi_n', 'path_length', 'snow_depth')]
    if 'Q_n' not in rec:
        return ''.join('%d %r %r %r %d %r %r\n' % r for r in
                       zip(cond, cols[0], cols[1], cols[2], nsc, cols[3], cols[4]))
    st = [rec[k].tolist() for k in ('Q_n', 'U_n', 'V_n')]
    return ''.join('%d %r %r %r %d %r %r %r %r %r\n' % r for r in
                   zip(cond, cols[0], cols[1], cols[2], nsc, cols[3], cols[4], st[0], st[1], st[2]))


def random_records(n, rng, stokes=False, bands=3):
    band = rng.integers(0, bands, n)
    wv = 1.0 / rng.uniform(0.9, 1.6, bands)
    dp = rng.uniform(0.05, 1.0, bands)
    rec = dict(condition=rng.integers(0, 6, n).astype(np.int8),
               wvn=wv[band], theta_n=(rng.random(n) * np.pi).astype(np.float32).astype(np.float64),
               phi_n=(rng.random(n) * 2 * np.pi).astype(np.float32).astype(np.float64),
               n_scat=rng.integers(0, 20000, n).astype(np.int64),
               path_length=(rng.exponential(0.3, n)).astype(np.float32).astype(np.float64),
               snow_depth=dp[band])
    if stokes:
        for k in ('Q_n', 'U_n', 'V_n'):
            rec[k] = rng.uniform(-1, 1, n).astype(np.float32).astype(np.float64)
    return rec


@pytest.mark.parametrize('stokes', [False, True])
def test_record_file_is_the_reference_text(tmp_path, stokes):
    """Byte-equal to the reference's loop over the same records: the header
    (with its stray comma) and every row, over several host threads."""
    rng = np.random.default_rng(3)
    rec = random_records(70_001, rng, stokes)
    p = tmp_path / 'r.txt'
    MonteCarlo.write_records(str(p), rec)
    header = MonteCarlo.HEADER_STOKES if stokes else MonteCarlo.HEADER
    assert p.read_text() == header + python_rows(rec)


def test_record_file_edge_cases(tmp_path):
    rng = np.random.default_rng(4)
    p = tmp_path / 'e.txt'
    MonteCarlo.write_records(str(p), random_records(0, rng))
    assert p.read_text() == MonteCarlo.HEADER
    MonteCarlo.write_records(str(p), random_records(1, rng))
    assert p.read_text().count('\n') == 2
    rec = random_records(5, rng)
    rec['theta_n'][:] = [0.0, np.pi / 2, 1e-7, 3.0, 1e-20]
    MonteCarlo.write_records(str(p), rec)
    assert p.read_text() == MonteCarlo.HEADER + python_rows(rec)
    bad = random_records(4, rng)
    bad['phi_n'] = bad['phi_n'][:3]
    with pytest.raises(ValueError):
        MonteCarlo.write_records(str(p), bad)
    with pytest.raises(OSError):
        MonteCarlo.write_records(str(tmp_path / 'no_such_dir' / 'x.txt'), random_records(3, rng))
