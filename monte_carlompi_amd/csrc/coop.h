// coop.h -- the wave-cooperative tail mode of the HG walk and the hand-over of its photons between the waves of a block
// (part of walk.hip's translation unit; included from there in this order)
#pragma once

namespace mcs {

// ------------------------------------------------- wave-cooperative tail mode
// Once a block's queue is empty, a wave's last photons would walk one step
// per wave instruction stream (~1,100 cycles per step for a lone wave), and
// the longest walks of a launch (~1e4 steps) set its duration.  Here the
// G = 64 / L lanes of a group walk 2G consecutive steps of one of the L live
// photons at once:
//  * lane j of the group draws the Philox blocks of steps i + 1 + 2j and
//    i + 2 + 2j -- the counters the one-lane walk uses, so every step keeps
//    its uniforms;
//  * a step is "scatter by (theta_k, phi_k) about the current direction, then
//    move dtau_k": with the direction carried as the third column of a frame
//    quaternion F, F_k = F_{k-1} (x) L_k (L_k = Rz(phi_k) Ry(theta_k)).  Each
//    lane multiplies its two steps' L, one inclusive product scan over the
//    group (DPP row shifts + row broadcasts, lower lanes on the left) gives
//    the frame after each lane's second step, and the frame after its first
//    is that one times the conjugate of its second L; depth and optical path
//    are one prefix sum of both steps' dtau_k * uz_k and dtau_k;
//  * the first step (in walk order) that exits or absorbs (or reaches
//    max_steps) ends the segment; the photon's lane takes that step's state,
//    otherwise the last lane's, and continues.
// Azimuths are measured in the transported frame instead of the lab-axis
// frame of mc3D:1255-1281.  phi_k is uniform and independent of the history,
// so the walk has the same law; events use the uniforms of their own step,
// so no step is re-drawn.

struct Quat {
    float w, x, y, z;
};

__device__ __forceinline__ Quat qmul(const Quat &A, const Quat &B)
{
    return Quat{fmaf(A.w, B.w, -fmaf(A.x, B.x, fmaf(A.y, B.y, A.z * B.z))),
                fmaf(A.w, B.x, fmaf(A.x, B.w, fmaf(A.y, B.z, -A.z * B.y))),
                fmaf(A.w, B.y, fmaf(-A.x, B.z, fmaf(A.y, B.w, A.z * B.x))),
                fmaf(A.w, B.z, fmaf(A.x, B.y, fmaf(-A.y, B.x, A.z * B.w)))};
}

// DPP lane move: CTRL 0x111..0x118 row_shr:1..8, 0x142 row_bcast:15,
// 0x143 row_bcast:31; rows outside ROWS and lanes without a source get `old`.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dppf(float old, float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, ROWS,
                                                      0xF, false));
}

// One scan level.  Lanes without a source (outside the row, or rows not in
// ROWS) read the identity, so they keep q; `take` is needed only where the
// source lane is in the row but in another group (groups under 16 lanes).
template <int CTRL, int ROWS, bool SEL>
__device__ __forceinline__ void qlevel(Quat &q, bool take)
{
    const Quat y{dppf<CTRL, ROWS>(1.f, q.w), dppf<CTRL, ROWS>(0.f, q.x), dppf<CTRL, ROWS>(0.f, q.y),
                 dppf<CTRL, ROWS>(0.f, q.z)};
    const Quat m = qmul(y, q);
    if constexpr (SEL) {
        q.w = take ? m.w : q.w;
        q.x = take ? m.x : q.x;
        q.y = take ? m.y : q.y;
        q.z = take ? m.z : q.z;
    } else {
        q = m;
    }
}

template <int CTRL, int ROWS, bool SEL>
__device__ __forceinline__ void slevel(float &s0, float &s1, bool take)
{
    const float t0 = dppf<CTRL, ROWS>(0.f, s0), t1 = dppf<CTRL, ROWS>(0.f, s1);
    if constexpr (SEL) {
        s0 = take ? s0 + t0 : s0;
        s1 = take ? s1 + t1 : s1;
    } else {
        s0 += t0;
        s1 += t1;
    }
}

// Inclusive scans within groups of G = 2^lg lanes (jj = lane within group):
// q <- q_0 (x) q_1 (x) ... (x) q_jj.  Levels 1..8 stay inside 16-lane rows;
// row_bcast:15 then joins row pairs (rows 1, 3) and row_bcast:31 the halves.
__device__ __forceinline__ void group_scan_q(Quat &q, int lg, int jj)
{
    if (lg < 4) {
        if (lg >= 1) qlevel<0x111, 0xF, true>(q, jj >= 1);
        if (lg >= 2) qlevel<0x112, 0xF, true>(q, jj >= 2);
        if (lg >= 3) qlevel<0x114, 0xF, true>(q, jj >= 4);
    } else {
        // groups of whole rows: row scans, then row pairs and halves joined
        qlevel<0x111, 0xF, false>(q, true);
        qlevel<0x112, 0xF, false>(q, true);
        qlevel<0x114, 0xF, false>(q, true);
        qlevel<0x118, 0xF, false>(q, true);
        if (lg >= 5) qlevel<0x142, 0xA, false>(q, true);
        if (lg >= 6) qlevel<0x143, 0xC, false>(q, true);
    }
}

// the same for two prefix sums
__device__ __forceinline__ void group_scan_s(float &s0, float &s1, int lg, int jj)
{
    if (lg < 4) {
        if (lg >= 1) slevel<0x111, 0xF, true>(s0, s1, jj >= 1);
        if (lg >= 2) slevel<0x112, 0xF, true>(s0, s1, jj >= 2);
        if (lg >= 3) slevel<0x114, 0xF, true>(s0, s1, jj >= 4);
    } else {
        slevel<0x111, 0xF, false>(s0, s1, true);
        slevel<0x112, 0xF, false>(s0, s1, true);
        slevel<0x114, 0xF, false>(s0, s1, true);
        slevel<0x118, 0xF, false>(s0, s1, true);
        if (lg >= 5) slevel<0x142, 0xA, false>(s0, s1, true);
        if (lg >= 6) slevel<0x143, 0xC, false>(s0, s1, true);
    }
}

__device__ __forceinline__ float bperm(int src_lane, float v)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src_lane << 2, __float_as_int(v)));
}
__device__ __forceinline__ uint32_t bperm(int src_lane, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src_lane << 2, (int)v);
}

// Per-wave state of the cooperative mode.  On entry every live photon of the
// wave writes its constants once into its own entry of the wave's LDS table
// (index = its rank at entry), so the walker's constants and its pipelined
// Philox block need no registers during the phase.  Groups are sticky: while
// no photon of the wave exits, is absorbed or joins, each group keeps its
// photon's state in its own lanes and carries it from its last lane to the
// next segment; the photon's lane is updated (and groups rebuilt) only after a
// round with an event.
struct CoopEntry {
    uint32_t ck0, h0, l0, t_abs;     // Philox words, absorption threshold
    float a, b, cc, d;               // HG constants
    float e, tau_tot, z, tau;        // HG constant, layer depth; state at grouping
    float dx, dy, dz;
    uint32_t i;
};

struct CoopGroups {
    unsigned long long em;   // photon lanes of the current grouping (0: rebuild)
    int lg;                  // G = 2^lg lanes per photon
    int ent;                 // table entry of this lane's group's photon
    PhiloxPre pre;           // that photon's Philox words
    float z0, tau0;          // its state before the next segment: depth, path,
    Quat f;                  // direction frame (third column = direction),
    uint32_t i0;             // steps taken
};

// direction = third column of the rotation of unit quaternion q
__device__ __forceinline__ void frame_dir(const Quat &q, float &x, float &y, float &z)
{
    x = 2.f * fmaf(q.x, q.z, q.w * q.y);
    y = 2.f * fmaf(q.y, q.z, -q.w * q.x);
    z = fmaf(-2.f, fmaf(q.x, q.x, q.y * q.y), 1.f);
}

// a frame whose third column is the unit vector d: the rotation taking e_z to d
__device__ __forceinline__ Quat frame_of(float dx, float dy, float dz)
{
    const bool up = dz >= 0.f;
    const Quat f = up ? Quat{1.f + dz, -dy, dx, 0.f} : Quat{-dy, 1.f - dz, 0.f, dx};
    const float n = __builtin_amdgcn_rsqf(fmaf(f.w, f.w, fmaf(f.x, f.x, fmaf(f.y, f.y, f.z * f.z))));
    return Quat{f.w * n, f.x * n, f.y * n, f.z * n};
}

__device__ __forceinline__ float rdlane63(float v)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Groups carrying a photon's state (no event in their last rounds) hand it
// back to the photon's lane.
__device__ __forceinline__ void coop_handback(NativeWalker &w, const CoopGroups &cg, unsigned long long *coop_steps)
{
    const int lane = lane_id();
    const bool was = (cg.em >> lane) & 1ull;
    const int src = was ? (lane_rank(cg.em) << cg.lg) : lane;
    const float z = bperm(src, cg.z0), t = bperm(src, cg.tau0);
    float x, y, u;
    frame_dir(cg.f, x, y, u);
    x = bperm(src, x); y = bperm(src, y); u = bperm(src, u);
    const uint32_t i = bperm(src, cg.i0);
    if (was) {
        atomicAdd(coop_steps, (unsigned long long)(i - w.i));
        w.i = i; w.z = z; w.tau_acc = t; w.mux = x; w.muy = y; w.muz = u;
    }
}

// A step's free path and local rotation L = Rz(phi) Ry(theta) (from the half
// angles) from its Philox block, with the HG constants of the photon (the
// same draws as NativeWalker::step).
__device__ __forceinline__ void coop_local(const u32x4 &cur, float ca, float cb, float ccc, float cd, float ce,
                                           float &dtau, Quat &L)
{
    dtau = -LN2_F * __builtin_amdgcn_logf(u01f(cur.z));
    const float xf = (float)cur.x;
    const float t = ccc * __builtin_amdgcn_rcpf(fmaf(ce, xf, cd));
    // isotropic bands carry e = 0
    const float cost = (ce == 0.f) ? fmaf(-0x1p-31f, xf, 1.f) : fminf(fmaxf(ca * fmaf(-t, t, cb), -1.f), 1.f);
    const float hc = __builtin_amdgcn_sqrtf(fmaf(0.5f, cost, 0.5f));    // cos(theta/2); |cost| <= 1
    const float hs = __builtin_amdgcn_sqrtf(fmaf(-0.5f, cost, 0.5f));   // sin(theta/2)
    const float hrev = 0.5f * u01f(cur.y);                                          // phi/2 in revolutions
    const float ha = __builtin_amdgcn_cosf(hrev), hb = __builtin_amdgcn_sinf(hrev);
    L = Quat{ha * hc, -hb * hs, ha * hs, hb * hc};
}

// Steps per lane and round of the cooperative mode, by the build's waves per
// SIMD: the frame and depth scans run once per round, so more steps per lane
// spread their cost (and give a lane more independent work while the round's
// dependent chain runs), as far as the register budget holds them.  Measured
// on MI355X for the 1e6-photon single case at 4 waves/SIMD (128 VGPRs):
// 5 steps 0.704 ms, 4 steps 0.712, 6 steps 0.725 (spills), 8 steps slower
// still; the 96-VGPR build (5 waves) with 4 steps 0.746.
template <int WPE>
constexpr int coop_k() { return WPE <= 4 ? 5 : (WPE == 5 ? 4 : 2); }

__device__ __forceinline__ Quat qconj(const Quat &q) { return Quat{q.w, -q.x, -q.y, -q.z}; }
__device__ __forceinline__ float frame_uz(const Quat &q) { return fmaf(-2.f, fmaf(q.x, q.x, q.y * q.y), 1.f); }

// One cooperative round of a wave (all 64 lanes call; no divergence around
// the lane permutes).  `tab`: the wave's entries; `map`: group -> entry;
// `me`: this lane's photon's entry.  Lane jj of a group walks the K steps
// k0 .. k0 + K - 1, k0 = i0 + 1 + K jj (the Philox blocks the one-lane walk
// draws for them): it multiplies their local rotations, one inclusive scan
// over the group gives the frame after its last step, and the frames after
// its earlier steps follow by multiplying back with the conjugates; depth and
// optical path are one prefix sum of the K steps' dtau uz and dtau, earlier
// steps by difference.
template <int K>
__device__ __forceinline__ void coop_round(NativeWalker &w, int &state, const KParams &P, CoopEntry *tab,
                                           int *map, int me, CoopGroups &cg, unsigned long long *coop_steps)
{
    const int lane = lane_id();
    const bool elig = state == LIVE;
    const unsigned long long em = __ballot(elig);
    if (em != cg.em) {
        // a photon joined: hand the carried state back before regrouping
        if (cg.em) coop_handback(w, cg, coop_steps);
        cg.em = em;
        if (em == 0ull) return;
        const int L = __popcll(em);
        cg.lg = L <= 1 ? 6 : 6 - (32 - __clz(L - 1));
        // group g walks the live photon of rank g: it posts its entry and state
        if (elig) {
            CoopEntry &E = tab[me];
            E.z = w.z; E.tau = w.tau_acc; E.dx = w.mux; E.dy = w.muy; E.dz = w.muz; E.i = w.i;
            __atomic_store_n(&map[lane_rank(em)], me, __ATOMIC_RELAXED);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        cg.ent = __atomic_load_n(&map[min(lane >> cg.lg, L - 1)], __ATOMIC_RELAXED);
        const CoopEntry &E = tab[cg.ent];
        cg.pre = PhiloxPre{E.ck0, E.h0, E.l0};
        cg.z0 = E.z; cg.tau0 = E.tau; cg.i0 = E.i;
        cg.f = frame_of(E.dx, E.dy, E.dz);
    }
    const int lg = cg.lg;
    const int G = 1 << lg;
    const int jj = lane & (G - 1);
    const bool member = (lane >> lg) < __popcll(em);
    const CoopEntry &E = tab[cg.ent];
    const float ca = E.a, cb = E.b, ccc = E.cc, cd = E.d, ce = E.e, ctau = E.tau_tot;
    const uint32_t ct_abs = E.t_abs;
    const uint32_t k0 = cg.i0 + 1u + (uint32_t)K * (uint32_t)jj;
    float dt[K];
    uint32_t wv[K];
    Quat Lr[K];
    Quat q;
#pragma unroll
    for (int m = 0; m < K; ++m) {
        const u32x4 u = philox_step(cg.pre, k0 + (uint32_t)m, P.key0, P.key1);
        coop_local(u, ca, cb, ccc, cd, ce, dt[m], Lr[m]);
        wv[m] = u.w;
        q = (m == 0) ? Lr[0] : qmul(q, Lr[m]);
    }
    if (jj == 0) q = qmul(cg.f, q);   // the group's carried frame on the left
    // frames first: the depth changes need the directions
    group_scan_q(q, lg, jj);
    float uz[K];
    float sz = 0.f, sp = 0.f;
    {
        Quat F = q;                    // frame after the lane's last step
#pragma unroll
        for (int m = K - 1; m >= 0; --m) {
            uz[m] = frame_uz(F);
            sz = fmaf(dt[m], uz[m], sz);
            sp += dt[m];
            if (m > 0) F = qmul(F, qconj(Lr[m]));
        }
    }
    group_scan_s(sz, sp, lg, jj);
    // depths and paths after each step, from the last backwards; the first
    // step with an event (exit, absorption, step cap) is kept
    int A = -1;
    bool top = false, bot = false;
    float zk = 0.f, pk = 0.f, dtau = 0.f, uzk = 0.f;
    uint32_t wk = 0u;
    {
        float z = cg.z0 + sz, pth = cg.tau0 + sp;
#pragma unroll
        for (int m = K - 1; m >= 0; --m) {
            const bool t_ = z > 0.f && uz[m] > 0.f, b_ = z < -ctau && uz[m] < 0.f;
            const bool e_ = member && (t_ || b_ || wv[m] < ct_abs || k0 + (uint32_t)m >= P.max_steps);
            if (e_) {
                A = m; top = t_; bot = b_; zk = z; pk = pth; dtau = dt[m]; uzk = uz[m]; wk = wv[m];
            }
            z -= dt[m] * uz[m];
            pth -= dt[m];
        }
    }
    const unsigned long long ev = __ballot(A >= 0);
    if (ev == 0ull) {
        // no event: every group walked K G steps; carry its last lane's state
        // (the frame renormalised against drift)
        const float zb = cg.z0 + sz, pb = cg.tau0 + sp;
        Quat f;
        if (lg == 6) {
            cg.z0 = rdlane63(zb); cg.tau0 = rdlane63(pb);
            f = Quat{rdlane63(q.w), rdlane63(q.x), rdlane63(q.y), rdlane63(q.z)};
        } else {
            const int src = lane | (G - 1);
            cg.z0 = bperm(src, zb); cg.tau0 = bperm(src, pb);
            f = Quat{bperm(src, q.w), bperm(src, q.x), bperm(src, q.y), bperm(src, q.z)};
        }
        const float n = __builtin_amdgcn_rsqf(fmaf(f.w, f.w, fmaf(f.x, f.x, fmaf(f.y, f.y, f.z * f.z))));
        cg.f = Quat{f.w * n, f.x * n, f.y * n, f.z * n};
        cg.i0 += (uint32_t)(K * G);
        return;
    }
    // this lane's outcome, should one of its steps end the segment: its first
    // step with an event (else its last); the frame after that step
    if (A < 0) {
        A = K - 1;
        zk = cg.z0 + sz; pk = cg.tau0 + sp; dtau = dt[K - 1]; uzk = uz[K - 1]; wk = wv[K - 1];
    }
    Quat qe = q;
#pragma unroll
    for (int m = K - 1; m >= 1; --m)
        if (m > A) qe = qmul(qe, qconj(Lr[m]));
    const uint32_t k = k0 + (uint32_t)A;
    const bool ab = wk < ct_abs, cap = k >= P.max_steps;
    const float ox = 2.f * fmaf(qe.x, qe.z, qe.w * qe.y);
    const float oy = 2.f * fmaf(qe.y, qe.z, -qe.w * qe.x);
    const int type = top ? 1 : (bot ? 2 : (ab ? 3 : (cap ? 4 : 0)));
    const float corr = top ? zk : zk + ctau;
    const float pout = (top || bot) ? pk - fminf(corr * __builtin_amdgcn_rcpf(uzk), dtau) : pk;
    const float zout = bot ? -ctau : zk;
    // each photon's lane reads the first event lane of its group (or the last)
    int src = lane;
    if (elig) {
        const int base = lane_rank(em) << lg;
        const unsigned long long gm = (ev >> base) & (lg == 6 ? ~0ull : ((1ull << G) - 1ull));
        src = base + (gm ? (__ffsll((long long)gm) - 1) : G - 1);
    }
    const int r_type = (int)bperm(src, (uint32_t)type);
    const uint32_t r_k = bperm(src, k);
    const uint32_t r_w = bperm(src, wk);
    const float r_z = bperm(src, zout), r_p = bperm(src, pout);
    const float r_x = bperm(src, ox), r_y = bperm(src, oy), r_uz = bperm(src, uzk);
    cg.em = 0ull;   // photons' lanes hold the state again: regroup next round
    if (elig) {
        atomicAdd(coop_steps, (unsigned long long)(r_k - w.i));
        w.i = r_k;
        w.z = r_z;
        w.tau_acc = r_p;
        w.mux = r_x; w.muy = r_y; w.muz = r_uz;
        if (r_type == 1) {
            w.cond = 1;
            state = DONE;
        } else if (r_type == 2) {
            if (P.cases[w.case_id(P)].lamb && w.reflects(P)) {
                // re-emitted from the bottom: the next step, taken here
                w.r = philox_step(w.pre, w.i + 1u, P.key0, P.key1);
                if (w.reemit(P)) {
                    state = DONE;
                } else if (w.i >= P.max_steps) {
                    w.cond = 0;
                    state = DONE;
                }
            } else {
                w.cond = 2;
                state = DONE;
            }
        } else if (r_type == 3) {
            const uint32_t cs = w.case_id(P);
            w.cond = (r_w < P.bands[P.cases[cs].band_begin + w.band].t_abs5) ? 5 : 4;
            state = DONE;
        } else if (r_type == 4) {
            w.cond = 0;
            state = DONE;
        }
    }
}

// Hand-over within the block.  A wave that has finished its photons sets its
// bit in `idle_mask` and polls its own inbox; a wave in the cooperative phase
// with L >= 2 photons claims a waiting wave (clears its bit), hands its
// carried group state back to the photon lanes, writes the upper half of its
// photons' walk state into their table entries and posts entry mask and wave
// to the claimed wave's inbox; those lanes are emptied.  Each wave has its own
// inbox, so hand-overs to different waves proceed at once.  The photons keep
// their chunk slot (found again from case and photon id), so tallies, records
// and chunk completion are unchanged; each photon still walks its own Philox
// stream.  Spread over the block's SIMDs, the last long walks get more lanes.
__device__ __forceinline__ void coop_donate(BlockState &bs, NativeWalker &w, int &state, CoopEntry *tab, int me,
                                            CoopGroups &cg, unsigned long long *coop_steps, unsigned long long *moved,
                                            int wave)
{
    const int lane = lane_id();
    const unsigned long long live = __ballot(state == LIVE);
    const int L = __popcll(live);
    if (L < 2) return;
    unsigned int im = lds_load(&bs.idle_mask);
    if (im == 0u) return;
    int v = -1;
    if (lane == 0) {
        while (im) {
            const int c = __ffs(im) - 1;
            const unsigned int old = atomicAnd(&bs.idle_mask, ~(1u << c));
            if ((old >> c) & 1u) {
                v = c;
                break;
            }
            im = old & ~(1u << c);
        }
    }
    v = bcast(v, 0);
    if (v < 0) return;
    if (cg.em) coop_handback(w, cg, coop_steps);
    cg.em = 0ull;   // regroup next round
    const int nd = L >> 1;
    if (lane == 0) {
        bs.ib_mask[v] = 0u;
        bs.ib_src[v] = wave;
        atomicAdd(&bs.out_posts[wave], 1);
    }
    if (state == LIVE && lane_rank(live) >= L - nd) {
        CoopEntry &E = tab[me];
        E.z = w.z; E.tau = w.tau_acc; E.dx = w.mux; E.dy = w.muy; E.dz = w.muz; E.i = w.i;
        atomicOr(&bs.ib_mask[v], 1u << me);
        atomicAdd(moved, 1ull);
        state = EMPTY;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __atomic_store_n(&bs.ib_state[v], 1, __ATOMIC_RELAXED);
}

// The cooperative phase of a wave: every photon the wave still holds is
// finished (or, with HAND, handed over) when it returns.  `tab`: coop_live
// entries; `map`: coop_live group slots.
template <bool HAND, int K>
__device__ __forceinline__ void coop_phase(BlockState &bs, NativeWalker &w, int &state, const KParams &P,
                                           CoopEntry *tab, int *map, unsigned long long *coop_steps,
                                           unsigned long long *moved, int wave)
{
    const unsigned long long live0 = __ballot(state == LIVE);
    const int me = lane_rank(live0);
    if (state == LIVE) {
        CoopEntry &E = tab[me];
        E.ck0 = w.pre.ck0; E.h0 = w.pre.h0; E.l0 = w.pre.l0; E.t_abs = w.t_abs_all;
        E.a = w.a; E.b = w.b; E.cc = w.cc; E.d = w.d;
        E.e = w.e; E.tau_tot = w.tau_tot;
    }
    CoopGroups cg;
    cg.em = 0ull;
    cg.lg = 0;
    cg.ent = 0;
    for (uint32_t round = 1; __ballot(state == LIVE); ++round) {
        coop_round<K>(w, state, P, tab, map, me, cg, coop_steps);
        if constexpr (HAND) {
            if ((round & 1u) == 0u) coop_donate(bs, w, state, tab, me, cg, coop_steps, moved, wave);
        }
    }
}

// Wait for photons posted to this wave's inbox (all lanes EMPTY on entry).
// True with the posted photons LIVE on the first lanes; false once no sibling
// holds photons (then none can post) and the wave has left the idle set before
// a donor claimed it.
__device__ bool coop_receive(BlockState &bs, const KParams &P, NativeWalker &w, int &state, int &my_slot,
                             uint64_t &my_gid, const CoopEntry *coop_tab, int wave)
{
    const int lane = lane_id();
    if (lane == 0) {
        atomicSub(&bs.holders, 1);
        atomicOr(&bs.idle_mask, 1u << wave);
    }
    __builtin_amdgcn_s_setprio(0);
    for (int spin = 0; spin < (1 << 24); ++spin) {
        if (lds_load(&bs.ib_state[wave]) == 1) {
            if (lane == 0) atomicAdd(&bs.holders, 1);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const int dw = lds_load(&bs.ib_src[wave]);
            const unsigned int mask = lds_load(&bs.ib_mask[wave]);
            if (lane < __popc(mask)) {
                unsigned int m = mask;
                for (int k = 0; k < lane; ++k) m &= m - 1u;
                const CoopEntry &E = coop_tab[dw * P.coop_live + (__ffs(m) - 1)];
                const PhiloxPre pre{E.ck0, E.h0, E.l0};
                const uint32_t cs = pre.ck0 ^ P.key0, ph = pre.l0 * 0x991A7CDBu;
                int slot = -1;
                for (int s = 0; s < SLOTS; ++s) {
                    const SlotHdr &H = bs.hdr[s];
                    if (lds_load(&H.state) == 1 && (uint32_t)H.case_id == cs &&
                        ph - H.p0 < (uint32_t)lds_load(&H.seqcount))
                        slot = s;
                }
                if (slot >= 0) {
                    my_slot = slot;
                    my_gid = bs.hdr[slot].gid0 + (ph - bs.hdr[slot].p0);
                    w.resume(P, pre, my_gid, E.z, E.tau, E.dx, E.dy, E.dz, E.i);
                    state = LIVE;
                } else {
                    atomicAdd(P.ctl + 1, 1ull);   // cannot happen: the chunk waits for the photon
                }
            }
            // the donor's entries are read before it may rewrite its table
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (lane == 0) {
                bs.ib_state[wave] = 0;
                atomicSub(&bs.out_posts[dw], 1);
            }
            return true;
        }
        if (lds_load(&bs.holders) == 0) {
            // nobody can post any more -- unless a donor claimed this wave just
            // before: then its post is on the way
            int left = 0;
            if (lane == 0) left = (atomicAnd(&bs.idle_mask, ~(1u << wave)) >> wave) & 1u;
            if (bcast(left, 0)) return false;
            continue;
        }
        __builtin_amdgcn_s_sleep(4);
    }
    if (lane == 0) atomicAdd(P.ctl + 1, 1ull);   // cannot happen: siblings finish
    return false;
}

// Before a wave rewrites its coop table for photons it received: wait until
// the waves it posted photons to have read them (they are polling their
// inboxes, so this is short).
__device__ __forceinline__ void await_own_posts(BlockState &bs, const KParams &P, int wave)
{
    for (int spin = 0; spin < (1 << 24); ++spin) {
        if (lds_load(&bs.out_posts[wave]) == 0) return;
        __builtin_amdgcn_s_sleep(1);
    }
    if (lane_id() == 0) atomicAdd(P.ctl + 1, 1ull);   // cannot happen
}


}  // namespace mcs
