"""The mathematics of the cooperative tail mode (walk.hip, coop_round), restated
in numpy and checked on the CPU:

* the per-step rotation quaternion L = Rz(phi) Ry(theta) built from half angles
  turns e_z into the scattering direction (sin t cos p, sin t sin p, cos t);
* frame_of(d) is a unit quaternion whose rotation takes e_z to d, on both of its
  branches (d_z >= 0 and d_z < 0, down to d = -e_z);
* the DPP schedule of group_scan_q (row_shr 1/2/4/8 with the lane-in-group
  masks, row_bcast:15 on rows 1 and 3, row_bcast:31 on rows 2 and 3) yields,
  for every group size G = 2..64, the ordered prefix products q_0 (x) ... (x) q_j;
* a walk whose azimuths are measured in the parallel-transported frame has the
  law of the reference's lab-frame rotation (mc3D:1255-1281): the distributions
  of the direction after k scatterings agree (Kolmogorov-Smirnov).
The GPU tests (test_gpu_coop.py) check the kernel itself against the oracle."""
import numpy as np
import pytest
from scipy import stats


def qmul(A, B):
    """Hamilton product of quaternions stored as (..., 4) = (w, x, y, z)."""
    w1, x1, y1, z1 = np.moveaxis(A, -1, 0)
    w2, x2, y2, z2 = np.moveaxis(B, -1, 0)
    return np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                     w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                     w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], -1)


def frame_dir(q):
    w, x, y, z = np.moveaxis(q, -1, 0)
    return np.stack([2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)], -1)


def local_quat(cost, rev):
    """L = Rz(phi) Ry(theta), phi = 2 pi rev, from the half angles."""
    hc, hs = np.sqrt(0.5 * (1 + cost)), np.sqrt(np.maximum(0.5 * (1 - cost), 0))
    ha, hb = np.cos(np.pi * rev), np.sin(np.pi * rev)
    return np.stack([ha * hc, -hb * hs, ha * hs, hb * hc], -1)


def frame_of(d):
    dx, dy, dz = np.moveaxis(d, -1, 0)
    up = dz >= 0
    f = np.where(up[..., None], np.stack([1 + dz, -dy, dx, np.zeros_like(dz)], -1),
                 np.stack([-dy, 1 - dz, np.zeros_like(dz), dx], -1))
    return f / np.linalg.norm(f, axis=-1, keepdims=True)


def dpp(old, v, ctrl, rows):
    """gfx9 DPP lane moves over a wave of 64 (v: (64, 4)); disabled rows and
    lanes without a source keep `old`."""
    out = old.copy()
    for lane in range(64):
        row = lane >> 4
        if not (rows >> row) & 1:
            continue
        if 0x111 <= ctrl <= 0x11F:            # row_shr:n
            src = lane - (ctrl - 0x110)
            if src < 0 or (src >> 4) != row:
                continue
        elif ctrl == 0x142:                    # row_bcast:15
            if row == 0:
                continue
            src = 16 * row - 1
        elif ctrl == 0x143:                    # row_bcast:31
            if row < 2:
                continue
            src = 31
        out[lane] = v[src]
    return out


def group_scan_q(q, lg):
    """The schedule of walk.hip's group_scan_q on a (64, 4) wave: groups under
    16 lanes select per level; groups of whole rows take every level as is
    (lanes without a DPP source read the identity)."""
    lanes = np.arange(64)
    jj = lanes & ((1 << lg) - 1)
    ident = np.tile([1.0, 0.0, 0.0, 0.0], (64, 1))
    if lg < 4:
        levels = [(1, 0x111, 0xF, jj >= 1), (2, 0x112, 0xF, jj >= 2), (3, 0x114, 0xF, jj >= 4)]
    else:
        every = np.ones(64, bool)
        levels = [(1, 0x111, 0xF, every), (1, 0x112, 0xF, every), (1, 0x114, 0xF, every),
                  (1, 0x118, 0xF, every), (5, 0x142, 0xA, every), (6, 0x143, 0xC, every)]
    for need, ctrl, rows, take in levels:
        if lg >= need:
            y = dpp(ident, q, ctrl, rows)
            q = np.where(take[:, None], qmul(y, q), q)
    return q


def test_local_rotation_gives_the_scattering_direction():
    rng = np.random.default_rng(1)
    cost, rev = rng.uniform(-1, 1, 1000), rng.uniform(0, 1, 1000)
    st, ph = np.sqrt(1 - cost ** 2), 2 * np.pi * rev
    np.testing.assert_allclose(frame_dir(local_quat(cost, rev)),
                               np.stack([st * np.cos(ph), st * np.sin(ph), cost], -1), atol=1e-12)


def test_frame_of_takes_ez_to_d_on_both_branches():
    rng = np.random.default_rng(2)
    d = rng.normal(size=(2000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = np.concatenate([d, [[0, 0, 1.0], [0, 0, -1.0], [1.0, 0, 0], [0, 1e-9, -1.0]]])
    f = frame_of(d)
    np.testing.assert_allclose(np.linalg.norm(f, axis=1), 1, atol=1e-12)
    np.testing.assert_allclose(frame_dir(f), d / np.linalg.norm(d, axis=1, keepdims=True), atol=1e-9)
    # the next direction makes the scattering angle with d
    cost = rng.uniform(-1, 1, d.shape[0])
    nd = frame_dir(qmul(f, local_quat(cost, rng.uniform(0, 1, d.shape[0]))))
    np.testing.assert_allclose(np.sum(nd * d, axis=1) / np.linalg.norm(d, axis=1), cost, atol=1e-9)


def test_group_scan_is_the_ordered_prefix_product_for_every_group_size():
    rng = np.random.default_rng(3)
    for lg in range(1, 7):
        q = rng.normal(size=(64, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        got = group_scan_q(q.copy(), lg)
        G = 1 << lg
        for lane in range(64):
            acc = q[lane & ~(G - 1)]
            for m in range((lane & ~(G - 1)) + 1, lane + 1):
                acc = qmul(acc, q[m])              # lower lanes on the left
            np.testing.assert_allclose(got[lane], acc, atol=1e-12, err_msg='G=%d lane %d' % (G, lane))
    # a reversed product (upper lanes on the left) is a different rotation: the
    # order matters and the check above sees it
    a, b = q[0], q[1]
    assert not np.allclose(qmul(a, b), qmul(b, a))


def hg_cos(u, g):
    """Henyey-Greenstein inverse CDF (mc3D:790-800)."""
    t = (1 - g * g) / (1 - g + 2 * g * u)
    return (1 + g * g - t * t) / (2 * g)


def lab_rotation(d, cost, ph):
    """The reference's rotation about the lab z axis (mc3D:1255-1281)."""
    ux, uy, uz = d.T
    st = np.sqrt(1 - cost ** 2)
    s2 = np.sqrt(np.maximum(1 - uz ** 2, 1e-300))
    nx = st * (ux * uz * np.cos(ph) - uy * np.sin(ph)) / s2 + ux * cost
    ny = st * (uy * uz * np.cos(ph) + ux * np.sin(ph)) / s2 + uy * cost
    nz = -st * np.cos(ph) * s2 + uz * cost
    return np.stack([nx, ny, nz], 1)


def test_transported_frame_walk_has_the_lab_frame_law():
    """Directions after k = 1..6 HG scatterings from theta0 = 60 deg: the lab-frame
    rotation and the quaternion frame product (independent streams) give the same
    distributions of u_z and of the azimuth atan2(u_y, u_x)."""
    n, g, K = 200_000, 0.85, 6
    th0 = np.radians(60.0)
    d0 = np.tile([np.sin(th0), 0.0, -np.cos(th0)], (n, 1))
    ra, rb = np.random.default_rng(10), np.random.default_rng(20)
    da = d0.copy()
    fb = frame_of(d0)
    for k in range(1, K + 1):
        da = lab_rotation(da, hg_cos(ra.uniform(size=n), g), 2 * np.pi * ra.uniform(size=n))
        fb = qmul(fb, local_quat(hg_cos(rb.uniform(size=n), g), rb.uniform(size=n)))
        db = frame_dir(fb)
        assert stats.ks_2samp(da[:, 2], db[:, 2]).pvalue > 1e-3, k
        assert stats.ks_2samp(np.arctan2(da[:, 1], da[:, 0]), np.arctan2(db[:, 1], db[:, 0])).pvalue > 1e-3, k
    # the check is sensitive: a frame that ignores the incoming direction's roll
    # (scattering always about the lab z axis) fails it at once
    bad = frame_dir(qmul(frame_of(d0), local_quat(hg_cos(rb.uniform(size=n), g), rb.uniform(size=n))))
    wrong = lab_rotation(d0, hg_cos(ra.uniform(size=n), g), 2 * np.pi * ra.uniform(size=n))
    assert stats.ks_2samp(bad[:, 2], wrong[:, 2]).pvalue > 1e-3   # one step: same law
    swapped = frame_dir(qmul(local_quat(hg_cos(rb.uniform(size=n), g), rb.uniform(size=n)), frame_of(d0)))
    assert stats.ks_2samp(swapped[:, 2], wrong[:, 2]).pvalue < 1e-6  # L on the left: wrong law


def test_two_steps_per_lane_give_every_frame_and_depth():
    """Lane j of a group walks steps 2j+1 and 2j+2: the scan runs over
    Q_j = L_a(j) (x) L_b(j) (the carried frame f on lane 0's left), the frame
    after step 2j+2 is the inclusive product and the frame after step 2j+1 is
    that product (x) conj(L_b(j)); depths follow from one prefix sum of
    dtau_a uz_a + dtau_b uz_b.  Checked against the sequential walk for every
    group size."""
    rng = np.random.default_rng(4)
    conj = lambda q: q * np.array([1.0, -1.0, -1.0, -1.0])
    for lg in range(0, 7):
        G = 1 << lg
        d0 = rng.normal(size=3)
        d0 /= np.linalg.norm(d0)
        f = frame_of(d0[None])[0]
        cost, rev = rng.uniform(-1, 1, (2, 64)), rng.uniform(0, 1, (2, 64))
        La, Lb = local_quat(cost[0], rev[0]), local_quat(cost[1], rev[1])
        dta, dtb = rng.exponential(size=64), rng.exponential(size=64)
        lanes = np.arange(64)
        Q = qmul(La, Lb)
        Q[(lanes & (G - 1)) == 0] = qmul(np.broadcast_to(f, (64, 4)), Q)[(lanes & (G - 1)) == 0]
        Fb = group_scan_q(Q.copy(), lg) if lg > 0 else Q
        Fa = qmul(Fb, conj(Lb))
        uza, uzb = frame_dir(Fa)[:, 2], frame_dir(Fb)[:, 2]
        z0 = -1.5
        for g0 in range(0, 64, G):
            F = f.copy()
            z = z0
            for j in range(g0, g0 + G):
                F = qmul(F, La[j])
                np.testing.assert_allclose(Fa[j], F, atol=1e-12)
                z += dta[j] * frame_dir(F)[2]
                za = z
                F = qmul(F, Lb[j])
                np.testing.assert_allclose(Fb[j], F, atol=1e-12)
                z += dtb[j] * frame_dir(F)[2]
                # the kernel's depths: prefix sum of both steps, first step by difference
                zb_k = z0 + np.sum(dta[g0:j + 1] * uza[g0:j + 1] + dtb[g0:j + 1] * uzb[g0:j + 1])
                np.testing.assert_allclose(zb_k, z, atol=1e-10)
                np.testing.assert_allclose(zb_k - dtb[j] * uzb[j], za, atol=1e-10)


@pytest.mark.parametrize('K', [2, 3, 4, 5])
def test_k_steps_per_lane_give_every_frame_and_depth(K):
    """The kernel's round with K steps per lane (coop_round<K>): lane j of a
    group walks steps K j + 1 .. K j + K; the scan runs over the product of
    its K local rotations (the carried frame on lane 0's left), the frame
    after its last step is the inclusive product, the earlier frames follow
    backwards, F_{m-1} = F_m (x) conj(L_m), and the depths backwards from one
    prefix sum of the K steps' dtau uz.  Checked against the sequential walk
    for every group size."""
    rng = np.random.default_rng(40 + K)
    conj = lambda q: q * np.array([1.0, -1.0, -1.0, -1.0])
    for lg in range(0, 7):
        G = 1 << lg
        d0 = rng.normal(size=3)
        d0 /= np.linalg.norm(d0)
        f = frame_of(d0[None])[0]
        L = [local_quat(rng.uniform(-1, 1, 64), rng.uniform(0, 1, 64)) for _ in range(K)]
        dt = [rng.exponential(size=64) for _ in range(K)]
        lanes = np.arange(64)
        Q = L[0]
        for m in range(1, K):
            Q = qmul(Q, L[m])
        first = (lanes & (G - 1)) == 0
        Q[first] = qmul(np.broadcast_to(f, (64, 4)), Q)[first]
        F = group_scan_q(Q.copy(), lg) if lg > 0 else Q
        frames = [None] * K
        frames[K - 1] = F
        for m in range(K - 1, 0, -1):
            frames[m - 1] = qmul(frames[m], conj(L[m]))
        uz = [frame_dir(fr)[:, 2] for fr in frames]
        sz = sum(dt[m] * uz[m] for m in range(K))
        z0 = -2.5
        for g0 in range(0, 64, G):
            Fs = f.copy()
            z = z0
            for j in range(g0, g0 + G):
                seq_z = []
                for m in range(K):
                    Fs = qmul(Fs, L[m][j])
                    np.testing.assert_allclose(frames[m][j], Fs, atol=1e-11, err_msg='K=%d G=%d' % (K, G))
                    z += dt[m][j] * frame_dir(Fs)[2]
                    seq_z.append(z)
                # kernel: depth after the last step from the prefix sum, earlier ones backwards
                zk = z0 + np.sum(sz[g0:j + 1])
                for m in range(K - 1, -1, -1):
                    np.testing.assert_allclose(zk, seq_z[m], atol=1e-9)
                    zk -= dt[m][j] * uz[m][j]
