"""Philox4x32-10: host reference on published known-answer vectors (Random123
kat_vectors, Salmon et al. SC'11), device generator equal to the host one."""
import numpy as np
import pytest

from tests.philox_ref import philox4x32_10

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize('ctr,key,want', KAT)
def test_host_philox_known_answers(ctr, key, want):
    got = philox4x32_10(np.array(ctr, dtype=np.uint32), key)[0]
    assert tuple(int(x) for x in got) == want


@pytest.mark.gpu
def test_device_philox_equals_host():
    from monte_carlompi_amd.gpu import GpuWalker
    rng = np.random.default_rng(5)
    ctr = rng.integers(0, 2**32, size=(4096, 4), dtype=np.uint64).astype(np.uint32)
    ctr[:3] = [k[0] for k in KAT]
    key = (0xa4093822, 0x299f31d0)
    with GpuWalker(0) as w:
        got = w.philox(ctr, key)
    np.testing.assert_array_equal(got, philox4x32_10(ctr, key))
