"""Host Philox4x32-10 (Salmon et al., SC'11) in numpy, for checking the
device generator.  Pinned by the known-answer vectors in test_philox.py."""
import numpy as np

M = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    c = np.asarray(ctr, dtype=np.uint64).reshape(-1, 4).copy()
    k0 = np.uint64(key[0]); k1 = np.uint64(key[1])
    for r in range(10):
        if r:
            k0 = (k0 + np.uint64(0x9E3779B9)) & M
            k1 = (k1 + np.uint64(0xBB67AE85)) & M
        p0 = np.uint64(0xD2511F53) * c[:, 0]
        p1 = np.uint64(0xCD9E8D57) * c[:, 2]
        c = np.stack([(p1 >> np.uint64(32)) ^ c[:, 1] ^ k0, p1 & M,
                      (p0 >> np.uint64(32)) ^ c[:, 3] ^ k1, p0 & M], axis=1)
    return c.astype(np.uint32)
