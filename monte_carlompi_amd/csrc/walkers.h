// walkers.h -- the photon walkers: NativeWalkerT (FP32, Philox; polarised for aspherical grains) and ReplayWalker (FP64, the reference's deviates)
// (part of walk.hip's translation unit; included from there in this order)
#pragma once

namespace mcs {

// ----------------------------------------------------------- native walker
// FP32 walk with one Philox4x32-10 call per step (4 uniforms: HG / Lambertian
// cosine, azimuth, free path, absorption).  The common path is straight-line
// code -- exits, absorption and near-vertical rotations sit behind rarely
// taken branches (the last one wave-uniform) -- and the next step's Philox call
// is issued in the same basic block as the current step's arithmetic: the two
// chains are independent, so their latencies overlap.
// np.interp(v, x, y) on an ascending grid of n points (what the reference's
// interp1d(cos_theta[::-1], P[::-1]) evaluates, mc3D:842-857): y at the ends
// and at grid points, else slope (v - x_j) + y_j with x_j <= v < x_j+1.
template <class T>
__device__ __forceinline__ int grid_index(const T *x, int n, T v)
{
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (x[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

template <class T>
__device__ __forceinline__ T interp_at(const T *x, const T *y, int n, int j, T v)
{
#pragma clang fp contract(off)
    if (v <= x[0]) return y[0];
    if (v >= x[n - 1]) return y[n - 1];
    if (x[j] == v) return y[j];
    const T slope = (y[j + 1] - y[j]) / (x[j + 1] - x[j]);
    T r = slope * (v - x[j]) + y[j];
    if (r != r) {
        r = slope * (v - x[j + 1]) + y[j + 1];
        if (r != r && y[j] == y[j + 1]) r = y[j];
    }
    return r;
}

// The state-independent part of one step of the polarised walk (exact
// sampler): free path and absorption word from the step's main Philox block,
// cos theta, sin theta and the phase matrix there, the first 2 phi candidate
// (with the sines and cosines of phi and 2 phi) and its acceptance uniform,
// and the second candidate's two uniforms.  One lane can build it for any
// step of any photon.  80 bytes, 16-aligned: the tail mode reads a packet
// from LDS as five 16-byte loads.  (The candidate's 2 phi and the table
// segment, which only the Lambertian azimuth needs, come out of make_packet
// and phi_retry beside it.)
struct alignas(16) PolarPacket {
    float dtau;
    uint32_t w;
    float ct, acc;
    float p11, p12, p22, p33, p43, p44;
    float sint, sinp, cosp, c2r, s2r;
    float omc2;                          // max(1 - ct^2, 0): the rotation back uses it
    float s2s, m2s;                      // the rotation signs of the candidate (phi_signs)
    float r2b, accb;                     // the second candidate (phi_second)
};
static_assert(sizeof(PolarPacket) == 80, "PolarPacket: five 16-byte LDS loads");

// The state-independent part of a step of the HG walk: free path, absorption
// word, cos and sin theta, the azimuth's sine and cosine (hg_packet, from the
// step's Philox block with the expressions of NativeWalkerT::step, so that a
// step formed from a packet built in another lane rounds alike).  32 bytes,
// two 16-byte LDS loads.
struct alignas(16) HgPacket {
    float dtau;
    uint32_t w;
    float cost, sint, sinp, cosp;
    float pad0, pad1;
};
static_assert(sizeof(HgPacket) == 32, "HgPacket: two 16-byte LDS loads");

// sin and cos of phi and of 2 phi for 2 phi = r2 revolutions (one
// transcendental pair; no implicit contraction, so every caller rounds alike)
__device__ __forceinline__ void phi_trig(float r2, float &sinp, float &cosp, float &c2r, float &s2r)
{
#pragma clang fp contract(off)
    sinp = __builtin_amdgcn_sinf(0.5f * r2);
    cosp = __builtin_amdgcn_cosf(0.5f * r2);
    c2r = fmaf(cosp, cosp, -sinp * sinp);
    s2r = 2.f * sinp * cosp;
}

// The signs the reference's folded azimuth phi_sca = arccos(cos phi) puts on
// the Stokes rotations (mc3D:1291-1337): s2s = sin 2 phi_sca (-sin 2 phi for
// phi > pi) into the scattering plane, and m2s = -2 sign for the return
// rotation (sign = +1 only at phi_sca = pi).  2 phi = r2 revolutions.
__device__ __forceinline__ void phi_signs(float r2, float s2r, float &s2s, float &m2s)
{
    s2s = (r2 > 1.f ? -1.f : 1.f) * s2r;
    m2s = (r2 == 1.f) ? -2.f : 2.f;
}

template <bool POLAR>
struct NativeWalkerT {
    using Real = float;
    static constexpr bool COOP = !POLAR;      // has the wave-cooperative tail mode below
    static constexpr bool STOKES = POLAR;     // carries a Stokes vector (aspherical particles)
    // step() itself ends a walk at the step cap (condition 0); the hot loop
    // tests the cap after each step for walkers without this
    static constexpr bool CAP_IN_STEP = !POLAR;
    // the polarised walk is built for 3 waves per SIMD (168 VGPRs), which
    // caps its blocks at 512 threads
    static constexpr int MAX_BLOCK = POLAR ? 512 : 1024;
    // k_walk is built for 5 (96 VGPRs) and 6 (80 VGPRs) waves per SIMD
    float z, mux, muy, muz, tau_acc;
    uint32_t i;
    int band;
    u32x4 r;                     // uniforms of the next step
    // per-photon parameters (the photon and case ids live in `pre`; the
    // bottom reflectance and 1 / ext_cff are read from the tables when needed)
    float a, b, cc, d, e, tau_tot;
    uint32_t t_abs5, t_abs_all;  // absorbed as condition 5 iff bits < t_abs5, else 4 iff < t_abs_all
    bool iso, lamb;
    int cond;
    PhiloxPre pre;
    // aspherical particles (POLAR): Stokes vector normalised to I = 1 and the
    // band's phase-matrix table (-1: Henyey-Greenstein)
    float sI, sQ, sU, sV;
    int pm;
    bool pm_exact;               // the band's exact sampler applies (see DevBand)
    float pm_mass;               // integral of S11 over the grid
    uint32_t n_rej;              // scatterings drawn by the two-area rejection

    // Set up photon `photon` of case `case_id` and take its first step (or
    // its whole walk in the Lambertian surface mode); true when it is done.
    __device__ bool start(const KParams &P, int case_id, uint32_t photon, uint64_t gid)
    {
        const DevCase &C = P.cases[case_id];
        pre = philox_pre(photon, (uint32_t)case_id, 0u, P.key0, P.key1);
        r = philox_step(pre, 1u, P.key0, P.key1);
        const int bnd = select_band(P, C, gid, r.x);
        band = bnd;
        load_params(P, C, bnd);
        z = 0.f; mux = C.mux0_f; muy = 0.f; muz = C.muz0_f; tau_acc = 0.f;
        i = 0;
        cond = 0;
        if constexpr (POLAR) {
            const DevBand &PB = P.bands[C.band_begin + bnd];
            pm = C.polar ? PB.pm : -1;
            // per-photon constants stay in registers: a step reads only the tables
            pm_exact = PB.pm_exact != 0 && C.s0_exact != 0;
            pm_mass = PB.fM;
            n_rej = 0u;
            sI = C.s0f[0]; sQ = C.s0f[1]; sU = C.s0f[2]; sV = C.s0f[3];
        }
        if (C.surf) return surface_walk(P, C);
        // first step: along the incident direction (mc3D:1344-1347)
        ++i;
        const u32x4 cur = r;
        r = philox_step(pre, 2u, P.key0, P.key1);
        const float dtau = free_path(cur);
        z = dtau * muz;
        tau_acc = dtau;
        bool done = false;
        if (z < -tau_tot) {                          // unscattered through the layer
            tau_acc -= fminf((z + tau_tot) / muz, dtau);
            z = -tau_tot;
            if (lamb && reflects(P)) {
                done = reemit(P);
            } else {
                cond = 3;
                done = true;
            }
        } else if (cur.w < t_abs_all) {
            cond = (cur.w < t_abs5) ? 5 : 4;
            done = true;
        }
        if (!done && i >= P.max_steps) {             // capped, as after any step
            cond = 0;
            done = true;
        }
        return done;
    }

    // the photon's wavelength bin: given per photon, or drawn from the case's
    // band distribution with the first step's polar word x (the first step
    // goes along the incident direction and draws no angle): the high bits of
    // x n_bands pick an alias-table column, the low bits accept it or take its
    // alias -- one table read instead of a search
    __device__ static int select_band(const KParams &P, const DevCase &C, uint64_t gid, uint32_t x)
    {
        int bnd = 0;
        if (P.photon_band) {
            bnd = P.photon_band[gid];
        } else if (C.n_bands > 1) {
            const uint64_t m = (uint64_t)x * (uint32_t)C.n_bands;
            const int j = (int)(m >> 32);
            const DevBand &B = P.bands[C.band_begin + j];
            bnd = (uint32_t)m < B.alias_t ? j : B.alias_k;
        }
        return bnd;
    }

    // Continue a photon handed over by another wave of the block: its Philox
    // words and walk state; band and constants are derived again.
    __device__ void resume(const KParams &P, const PhiloxPre &p, uint64_t gid, float z_, float tau_,
                           float dx, float dy, float dz, uint32_t i_)
    {
        pre = p;
        const uint32_t cs = case_id(P);
        const DevCase &C = P.cases[cs];
        band = select_band(P, C, gid, philox_step(pre, 1u, P.key0, P.key1).x);
        load_params(P, C, band);
        z = z_; tau_acc = tau_; mux = dx; muy = dy; muz = dz; i = i_;
        cond = 0;
        r = philox_step(pre, i + 1u, P.key0, P.key1);
    }

    // per-photon constants of (case, band)
    __device__ __forceinline__ void load_params(const KParams &P, const DevCase &C, int bnd)
    {
        const DevBand &B = P.bands[C.band_begin + bnd];
        a = B.hg_a; b = B.hg_b; cc = B.hg_c; d = B.hg_d; e = B.hg_e;
        iso = B.iso != 0;
        if (C.surf) {
            // Lambertian surface: absorbed with probability 1 - R, split by species
            const double q = 1.0 - C.r_lamb;
            t_abs5 = prob_bits_dev(B.p_ext_imp * q);
            t_abs_all = prob_bits_dev(q);
        } else {
            t_abs5 = B.t_abs5;
            t_abs_all = B.t_abs_all;
        }
        tau_tot = C.tau_tot_f;
        lamb = C.lamb != 0;
    }

    __device__ __forceinline__ static float free_path(const u32x4 &q)
    {
        return -LN2_F * __builtin_amdgcn_logf(u01f(q.z));
    }

    // Cosine-weighted upward direction, cos = sqrt(u): the pdf 2 sin cos of
    // the rejection loop in mc3D:1242-1250 (Lambertian re-emission).
    __device__ __forceinline__ void lambertian_direction(const u32x4 &q)
    {
        const float cost = __builtin_amdgcn_sqrtf(u01f(q.x));
        const float sint = __builtin_amdgcn_sqrtf(fmaxf(fmaf(-cost, cost, 1.f), 0.f));
        const float rev = u01f(q.y);
        mux = sint * __builtin_amdgcn_cosf(rev);
        muy = sint * __builtin_amdgcn_sinf(rev);
        muz = cost;
    }

    // After a Lambertian bottom reflection at step i (z = -tau_tot): the next
    // step re-emits cosine-weighted upward and moves (mc3D:1420-1459); true
    // when the photon is done.
    __device__ bool reemit(const KParams &P)
    {
        if (i >= P.max_steps) {
            cond = 0;
            return true;
        }
        ++i;
        const u32x4 cur = r;
        r = philox_step(pre, i + 1u, P.key0, P.key1);
        lambertian_direction(cur);
        if constexpr (POLAR) {
            // the step after a bottom reflection still draws its angles from
            // the phase matrix (the azimuth is used), then resets the Stokes
            // vector (mc3D:1238-1285)
            if (pm >= 0) polar_lambertian(P, false);
            sI = 1.f; sQ = 0.f; sU = 0.f; sV = 0.f;
        }
        const float dtau = free_path(cur);
        z = fmaf(dtau, muz, z);
        tau_acc += dtau;
        if (z > 0.f) {
            tau_acc -= fminf(z / muz, dtau);
            cond = 1;
            return true;
        }
        if (cur.w < t_abs_all) {
            cond = (cur.w < t_abs5) ? 5 : 4;
            return true;
        }
        return false;
    }

    // Lambertian surface (mc3D:1228-1238, 1385-1387): first event at z = 0,
    // absorbed with probability 1 - R, else re-emitted cosine-weighted upward
    // until the photon leaves.
    __device__ bool surface_walk(const KParams &P, const DevCase &C)
    {
        (void)C;
        ++i;
        u32x4 cur = r;
        if (cur.w < t_abs_all) {
            cond = (cur.w < t_abs5) ? 5 : 4;
            return true;
        }
        for (;;) {
            if (i >= P.max_steps) {
                cond = 0;
                return true;
            }
            ++i;
            cur = philox_step(pre, i, P.key0, P.key1);
            lambertian_direction(cur);
            if constexpr (POLAR) {
                // Lambertian surface of aspherical grains: azimuth from the
                // phase matrix, Stokes vector scattered at the Lambertian
                // angle (mc3D:1238-1250, 1287-1342)
                if (pm >= 0) polar_lambertian(P, true);
            }
            const float dtau = free_path(cur);
            z = fmaf(dtau, muz, z);
            tau_acc += dtau;
            if (z > 0.f) {
                tau_acc -= fminf(z / muz, dtau);
                cond = 1;
                return true;
            }
            if (cur.w < t_abs_all) {
                cond = (cur.w < t_abs5) ? 5 : 4;
                return true;
            }
        }
    }

    // one step (i >= 1: after the first); true when the photon has left or
    // been absorbed.  PIN_I: how the next Philox block is held back until the
    // step's uniforms are read (below); the two forms differ only in register
    // allocation (the 128-VGPR build issues one move less per step with the
    // step count pinned; the 80-VGPR build then spills at the hot-loop entry)
    template <bool PIN_I = false>
    __device__ __forceinline__ bool step(const KParams &P)
    {
        if constexpr (POLAR) {
            if (pm >= 0 && pm_exact) {
                const u32x4 cur = r;
                r = philox_step(pre, i + 2u, P.key0, P.key1);
                const u32x4 pb = philox4x32_10(u32x4{photon_id(), case_id(P), i + 1u, 2u}, P.key0, P.key1);
                PolarPacket k;
                float r2;
                int seg;
                make_packet(cur, pb, pm_table(P), P.pm_n, pm_mass, k, r2, seg);
                return polar_step(P, k);
            }
        }
        ++i;
        const u32x4 cur = r;
        const float dtau = free_path(cur);
        const float xf = (float)cur.x;
        // azimuth: v_sin/v_cos take revolutions, 2 pi u -> u
        float rev = u01f(cur.y);
        // The next block's counter (or the step count) passes an empty asm
        // that reads the three uniforms above, so the block is built after
        // they are: the words of `cur` are then dead and the new block can
        // take their registers (without this every step ended with register
        // copies).
        if constexpr (PIN_I) {
            asm volatile("" : "+v"(i) : "v"(dtau), "v"(xf), "v"(rev));
            r = philox_step(pre, i + 1u, P.key0, P.key1);
        } else {
            uint32_t ctr = i + 1u;
            asm volatile("" : "+v"(ctr) : "v"(dtau), "v"(xf), "v"(rev));
            r = philox_step(pre, ctr, P.key0, P.key1);
        }
        // polar cosine: HG inverse CDF (mc3D:790-800) with the uniform's
        // scale and offset folded into d and e; isotropic for g = 0
        const float t = cc * __builtin_amdgcn_rcpf(fmaf(e, xf, d));
        float cost = iso ? fmaf(-0x1p-31f, xf, 1.f) : fminf(fmaxf(a * fmaf(-t, t, b), -1.f), 1.f);
        float rev2 = 0.f;
        int seg = -1;
        if constexpr (POLAR) {
            if (pm >= 0) {   // aspherical: phase-matrix sampler
                polar_angles(P, cost, rev2, seg);
                rev = 0.5f * rev2;
            }
        }
        const float sint = __builtin_amdgcn_sqrtf(fmaf(-cost, cost, 1.f));   // |cost| <= 1
        const float sinp = __builtin_amdgcn_sinf(rev);
        const float cosp = __builtin_amdgcn_cosf(rev);
        // rotation (mc3D:1255-1281); near-vertical incoming directions take the
        // axis form, behind a wave-uniform branch so that the common path
        // issues none of its instructions
        const float s2 = fmaf(-muz, muz, 1.f);
        const float st = sint * __builtin_amdgcn_rsqf(s2);
        const float mz0 = muz, mx0 = mux, my0 = muy;
        const float gx = fmaf(st, fmaf(mux * muz, cosp, -muy * sinp), mux * cost);
        const float gy = fmaf(st, fmaf(muy * muz, cosp, mux * sinp), muy * cost);
        muz = fmaf(-s2 * st, cosp, muz * cost);
        mux = gx; muy = gy;
        if (__builtin_expect(__ballot(s2 < 1e-7f) != 0ull, 0)) {
            if (s2 < 1e-7f) {
                const float sg = mz0 > 0.f ? 1.f : -1.f;
                mux = sint * cosp;
                muy = sg * sint * sinp;
                muz = sg * cost;
            }
        }
        if constexpr (POLAR) {
            if (pm >= 0) stokes_update(P, cost, rev2, mx0, my0, mz0, seg);
        }
        z = fmaf(dtau, muz, z);
        tau_acc += dtau;
        if (!(z > 0.f || z < -tau_tot || cur.w < t_abs_all || (CAP_IN_STEP && i >= P.max_steps))) return false;
        return hg_event(P, dtau, cur.w);
    }

    // The rare end of a step (its free path dtau and absorption word w):
    // exits, absorption and the step cap; true when the walk has ended,
    // false after a Lambertian re-emission below the cap (r must hold the
    // next step's block).  The depth passes an empty asm so that the tests
    // are issued here, not hoisted into every step.
    __device__ __forceinline__ bool hg_event(const KParams &P, float dtau, uint32_t w)
    {
        float ze = z;
        asm volatile("" : "+v"(ze));
        if (ze > 0.f) {                              // mc3D:1390-1397
            tau_acc -= fminf(z / muz, dtau);
            cond = 1;
            return true;
        }
        if (ze < -tau_tot) {                         // mc3D:1399-1459
            tau_acc -= fminf((z + tau_tot) / muz, dtau);
            z = -tau_tot;
            if (lamb && reflects(P)) {
                if (reemit(P)) return true;
                if (CAP_IN_STEP && i >= P.max_steps) {   // re-emitted at the cap
                    cond = 0;
                    return true;
                }
                return false;
            }
            cond = 2;
            return true;
        }
        // absorbed: one uniform decides species and absorption jointly, with
        // P(cond 5) = P_imp (1 - omega_imp), P(cond 4) = (1 - P_imp)(1 - omega_ice)
        // (mc3D:1374-1383, 1461-1466); otherwise the step cap (condition 0)
        cond = (w < t_abs5) ? 5 : (CAP_IN_STEP && w >= t_abs_all) ? 0 : 4;
        return true;
    }

    // Packet of a step of the HG walk from its Philox block and the band's
    // constants (the expressions of step(); tail mode of the reproducible walk)
    __device__ __forceinline__ static void hg_packet(const u32x4 &cur, float a, float b, float cc, float d, float e,
                                                     bool iso, HgPacket &o)
    {
        o.dtau = free_path(cur);
        o.w = cur.w;
        const float xf = (float)cur.x;
        const float rev = u01f(cur.y);
        const float t = cc * __builtin_amdgcn_rcpf(fmaf(e, xf, d));
        o.cost = iso ? fmaf(-0x1p-31f, xf, 1.f) : fminf(fmaxf(a * fmaf(-t, t, b), -1.f), 1.f);
        o.sint = __builtin_amdgcn_sqrtf(fmaf(-o.cost, o.cost, 1.f));
        o.sinp = __builtin_amdgcn_sinf(rev);
        o.cosp = __builtin_amdgcn_cosf(rev);
        o.pad0 = o.pad1 = 0.f;
    }

    // The move of a step from its packet (the expressions of step()); true
    // when an event follows (exit through either face, or absorbed)
    __device__ __forceinline__ bool hg_move(const HgPacket &k)
    {
        ++i;
        const float cost = k.cost, sint = k.sint, sinp = k.sinp, cosp = k.cosp;
        const float s2 = fmaf(-muz, muz, 1.f);
        const float st = sint * __builtin_amdgcn_rsqf(s2);
        const float mz0 = muz;
        const float gx = fmaf(st, fmaf(mux * muz, cosp, -muy * sinp), mux * cost);
        const float gy = fmaf(st, fmaf(muy * muz, cosp, mux * sinp), muy * cost);
        muz = fmaf(-s2 * st, cosp, muz * cost);
        mux = gx; muy = gy;
        if (__builtin_expect(__ballot(s2 < 1e-7f) != 0ull, 0)) {
            if (s2 < 1e-7f) {
                const float sg = mz0 > 0.f ? 1.f : -1.f;
                mux = sint * cosp;
                muy = sg * sint * sinp;
                muz = sg * cost;
            }
        }
        z = fmaf(k.dtau, muz, z);
        tau_acc += k.dtau;
        return z > 0.f || z < -tau_tot || k.w < t_abs_all;
    }

    // ids from the Philox words: ck0 = case ^ k0, l0 = lo(M0 * photon) with
    // M0 odd, so photon = l0 * M0^-1 (mod 2^32), M0^-1 = 0x991A7CDB
    __device__ __forceinline__ uint32_t case_id(const KParams &P) const { return pre.ck0 ^ P.key0; }
    __device__ __forceinline__ uint32_t photon_id() const { return pre.l0 * 0x991A7CDBu; }

    // Lambertian bottom: reflected with probability R, from Philox block
    // (photon, case, step, 1)
    __device__ __forceinline__ bool reflects(const KParams &P) const
    {
        const uint32_t cs = case_id(P);
        const u32x4 q = philox4x32_10(u32x4{photon_id(), cs, i, 1u}, P.key0, P.key1);
        return (unsigned long long)q.y < P.cases[cs].t_refl;
    }

    // ---- aspherical particles (POLAR only)
    __device__ __forceinline__ const float *pm_table(const KParams &P) const
    {
        extern __shared__ float pm_lds_tab[];
        const size_t off = (size_t)pm * (size_t)P.pm_stride;
        return (P.pm_lds_floats > 0 ? pm_lds_tab : P.pmf) + off;
    }

    // Scattering angles of mc3D:905-1004: rejection sampling of
    // I S11(cos) + S12(cos) (Q cos 2phi + U sin 2phi) over cos theta in two
    // areas split at the cutoff (each bounded by |I| max|P11| + max|P12|
    // sqrt(Q^2 + U^2), the reference's A + B C) and 2 phi uniform in
    // [0, 4 pi).  Attempt k of step i draws Philox block (photon, case, i,
    // 2 + k): area, cos theta, 2 phi, acceptance.  Returns cos theta and
    // 2 phi in revolutions.
    __device__ void polar_angles(const KParams &P, float &cost, float &rev2, int &seg)
    {
        const uint32_t cs = case_id(P);
        const float *T = pm_table(P);
        const int n = P.pm_n;
        seg = -1;
        if (pm_exact) {
            // The same law without the two-area rejection loop (whose
            // acceptance is ~10 % on a forward-peaked P11): cos theta from
            // the marginal S11 by inverting its piecewise-quadratic integral
            // (make_packet), then 2 phi from 1 + (S12 / S11) (Q cos 2phi +
            // U sin 2phi) by rejection against 1 + |S12 / S11| sqrt(Q^2 +
            // U^2) (phi_ok, phi_retry: acceptance >= 1/2, no table reads).  Philox
            // block (photon, case, i, 2 + k), k = 0, 1, ...
            const u32x4 q = philox4x32_10(u32x4{photon_id(), cs, i, 2u}, P.key0, P.key1);
            PolarPacket k;
            float r2;
            make_packet(q, q, T, n, pm_mass, k, r2, seg);   // its free path and absorption word go unused
            // (no second candidate here: with one block for both, its words
            // are the ones cos theta and the first 2 phi came from)
            if (!phi_ok(k)) r2 = phi_retry(P, k);
            cost = k.ct;
            rev2 = r2;
            return;
        }
        ++n_rej;
        const DevCase &C = P.cases[cs];
        const DevBand &B = P.bands[C.band_begin + band];
        const float cq = __builtin_amdgcn_sqrtf(fmaf(sQ, sQ, sU * sU));
        const float ai = fabsf(sI);
        const float mv1 = fmaf(ai, B.fA1, B.fB1 * cq), mv2 = fmaf(ai, B.fA2, B.fB2 * cq);
        const float w1 = C.fcmax - C.fccut, w2 = C.fccut - C.fcmin;
        const float a1 = w1 * mv1, a2 = w2 * mv2;
        const float pct = a1 / (a1 + a2);
        const uint32_t ph = photon_id();
        for (uint32_t k = 0;; ++k) {
            const u32x4 q = philox4x32_10(u32x4{ph, cs, i, 2u + k}, P.key0, P.key1);
            const bool in1 = u01f(q.x) < pct;
            const float u = u01f(q.y);
            const float ct = in1 ? fmaf(u, w1, C.fccut) : fmaf(-u, w2, C.fccut);
            const float r2 = 2.f * u01f(q.z);
            const float r3 = 1.000001f * u01f(q.w) * (in1 ? mv1 : mv2);
            const int j = grid_index(T, n, ct);
            const float s11 = interp_at(T, T + n, n, j, ct);
            const float s12 = interp_at(T, T + 2 * n, n, j, ct);
            const float val = fmaf(sI, s11, s12 * fmaf(sQ, __builtin_amdgcn_cosf(r2), sU * __builtin_amdgcn_sinf(r2)));
            if (r3 <= val) {
                cost = ct;
                rev2 = r2;
                return;
            }
            if (k + 1 >= POLAR_MAX_TRIES) {
                atomicAdd(P.ctl + 2, 1ull);
                cost = ct;
                rev2 = r2;
                return;
            }
        }
    }

    // Stokes vector after scattering by (theta, phi) from direction z-cosine
    // mz0 to muz_n (mc3D:1287-1342): rotate into the scattering plane by
    // phi_sca = arccos(cos phi) (phi folded into [0, pi]: the double angle's
    // sine changes sign for phi > pi), apply the phase matrix at cos theta,
    // rotate back by -i_2 with cos i_2 = (muz_n cos - mz0) / (sign den)
    // (sign -1 unless phi_sca = pi, as in the reference; cos i_2 = 0 where
    // den = 0), normalise to I = 1.  The doubled angles of the return
    // rotation come from cos i_2 directly: cos 2i = 2c^2 - 1, sin 2i = 2c s.
    __device__ void stokes_update(const KParams &P, float cost, float rev2, float mx0, float my0, float mz0, int seg)
    {
        const float *T = pm_table(P);
        const int n = P.pm_n;
        const int j = seg >= 0 ? seg : grid_index(T, n, cost);
        // linear in the segment (slopes precomputed), clamped to the grid
        const float dx = fminf(fmaxf(cost, T[0]), T[n - 1]) - T[j];
        const float p11 = fmaf(T[8 * n + j], dx, T[n + j]);
        const float p12 = fmaf(T[9 * n + j], dx, T[2 * n + j]);
        const float p22 = fmaf(T[10 * n + j], dx, T[3 * n + j]);
        const float p33 = fmaf(T[11 * n + j], dx, T[4 * n + j]);
        const float p43 = fmaf(T[12 * n + j], dx, T[5 * n + j]);
        const float p44 = fmaf(T[13 * n + j], dx, T[6 * n + j]);
        float s2s, m2s;
        phi_signs(rev2, __builtin_amdgcn_sinf(rev2), s2s, m2s);
        stokes_apply(cost, fmaxf(fmaf(-cost, cost, 1.f), 0.f), __builtin_amdgcn_cosf(rev2), s2s, m2s, mx0, my0,
                     mz0, p11, p12, p22, p33, p43, p44);
    }

    // (mx0, my0, mz0): the direction before the scattering; the photon's
    // direction members hold the one after it.  c2 = cos 2 phi, s2s and m2s
    // from phi_signs; omc2 = max(1 - cost^2, 0)
    __device__ __forceinline__ void stokes_apply(float cost, float omc2, float c2, float s2s, float m2s, float mx0,
                                                 float my0, float mz0, float p11, float p12, float p22, float p33,
                                                 float p43, float p44)
    {
#pragma clang fp contract(off)
        const float Qs = fmaf(sQ, c2, sU * s2s);
        const float Us = fmaf(-sQ, s2s, sU * c2);
        const float Ic = fmaf(sI, p11, Qs * p12);
        const float Qc = fmaf(sI, p12, Qs * p22);
        const float Uc = fmaf(Us, p33, -sV * p43);
        const float Vc = fmaf(Us, p43, sV * p44);
        const float muz_n = muz;
        const float den2 = omc2 * fmaxf(fmaf(-muz_n, muz_n, 1.f), 0.f);
        const float rd = __builtin_amdgcn_rsqf(den2);   // 1 / den
        // i_2 in [0, pi]: cos i_2 as the reference forms it, up to its sign,
        // which only the return rotation's sine uses (m2s; the clamp is odd);
        // sin i_2 >= 0 from its exact identity (den sin i_2)^2 = (e_z . (u_0 x
        // u_n))^2 -- the Gram determinant 1 + 2 mu_0 mu_n cos - mu_0^2 - mu_n^2
        // - cos^2 -- instead of sqrt(1 - cos^2 i_2), which near |cos i_2| = 1
        // (scattering plane close to the meridian plane) would turn an FP32
        // rounding of cos i_2 into a rotation of ~1e-3 rad.  den = 0: cos i_2 =
        // 0.  (straight-line: both values are formed and den = 0 selects 0 and 1)
        const float ci0 = fminf(fmaxf(fmaf(muz_n, cost, -mz0) * rd, -1.f), 1.f);
        const float si0 = fminf(fabsf(fmaf(mx0, muy, -my0 * mux)) * rd, 1.f);
        const float ci = (den2 != 0.f) ? ci0 : 0.f;
        const float si = (den2 != 0.f) ? si0 : 1.f;
        const float cr = fmaf(-2.f * si, si, 1.f);     // cos(-2 i_2) = 1 - 2 sin^2 i_2
        const float sr = m2s * ci * si;                // sin(-2 i_2) = -2 cos i_2 sin i_2
        const float inv = __builtin_amdgcn_rcpf(Ic);
        const float Qm = fmaf(Qc, cr, Uc * sr);
        const float Um = fmaf(-Qc, sr, Uc * cr);
        sI = 1.f;
        sQ = Qm * inv;
        sU = Um * inv;
        sV = Vc * inv;
    }

    // Packet of one step (see PolarPacket) from its main Philox block and its
    // block (photon, case, step, 2), for a band with the exact sampler: cos
    // theta by inverting the integral of the linear S11 (guided search), the
    // six elements there.
    __device__ static void make_packet(const u32x4 &mainb, const u32x4 &pb, const float *T, int n, float mass,
                                       PolarPacket &o, float &r2, int &seg)
    {
        // no implicit contraction in the packet or the step that uses it: a
        // packet built in another lane and read from LDS must give the same
        // roundings as one built in place (explicit fmaf only)
#pragma clang fp contract(off)
        o.dtau = free_path(mainb);
        o.w = mainb.w;
        const float *cum = T + 7 * n;
        const float u = u01f(pb.x);
        const float t = u * mass;
        const int *guide = reinterpret_cast<const int *>(T + 14 * n);
        const int gk = min((int)(u * (float)PM_GUIDE), PM_GUIDE - 1);
        int lo = guide[gk], hi = guide[gk + 1] + 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (cum[mid] <= t) lo = mid; else hi = mid;
        }
        const int j = lo;
        const float x0 = T[j], h = T[j + 1] - T[j];
        const float y0 = T[n + j];
        const float slope = T[8 * n + j];
        const float dm = fmaxf(t - cum[j], 0.f);
        const float s = 2.f * dm *
                        __builtin_amdgcn_rcpf(y0 + __builtin_amdgcn_sqrtf(fmaxf(fmaf(2.f * slope, dm, y0 * y0), 0.f)));
        const float ct = fminf(x0 + fminf(fmaxf(s, 0.f), h), T[n - 1]);
        const float dx = ct - x0;
        o.ct = ct;
        seg = j;
        o.p11 = fmaf(slope, dx, y0);
        o.p12 = fmaf(T[9 * n + j], dx, T[2 * n + j]);
        o.p22 = fmaf(T[10 * n + j], dx, T[3 * n + j]);
        o.p33 = fmaf(T[11 * n + j], dx, T[4 * n + j]);
        o.p43 = fmaf(T[12 * n + j], dx, T[5 * n + j]);
        o.p44 = fmaf(T[13 * n + j], dx, T[6 * n + j]);
        r2 = 2.f * u01f(pb.y);
        o.acc = u01f(pb.z);
        o.sint = __builtin_amdgcn_sqrtf(fmaf(-ct, ct, 1.f));
        o.omc2 = fmaxf(fmaf(-ct, ct, 1.f), 0.f);
        phi_trig(r2, o.sinp, o.cosp, o.c2r, o.s2r);
        phi_signs(r2, o.s2r, o.s2s, o.m2s);
        // the main block's first two words (the exact sampler draws its angles
        // from the other block) give the second candidate
        o.r2b = 2.f * u01f(mainb.x);
        o.accb = u01f(mainb.y);
    }

    // 2 phi given cos theta, by rejection: target I S11 + S12 (Q cos 2phi + U
    // sin 2phi), bound |I| S11 + |S12| sqrt(Q^2 + U^2) (I = 1 after the first
    // scattering; the initial vector keeps its own I); acceptance >= 1/2.
    // The packet carries the first candidate (r2 = 2 phi in revolutions, its
    // acceptance uniform, the sines and cosines of phi and 2 phi and the
    // rotation signs) and the second's two uniforms.  phi_ok tests the
    // packet's candidate against the current Stokes vector; phi_second makes
    // the second candidate the packet's; phi_retry then draws until one is
    // accepted (blocks (photon, case, i, 2 + k), k >= 1), replacing the
    // packet's candidate in place and returning its r2.
    __device__ __forceinline__ bool phi_ok(const PolarPacket &k) const
    {
#pragma clang fp contract(off)
        const float s11 = k.p11, s12 = k.p12;
        const float bound = fmaf(fabsf(sI), s11, fabsf(s12) * __builtin_amdgcn_sqrtf(fmaf(sQ, sQ, sU * sU)));
        return k.acc * bound <= fmaf(sI, s11, s12 * fmaf(sQ, k.c2r, sU * k.s2r));
    }
    __device__ __forceinline__ static void phi_second(PolarPacket &k)
    {
        k.acc = k.accb;
        phi_trig(k.r2b, k.sinp, k.cosp, k.c2r, k.s2r);
        phi_signs(k.r2b, k.s2r, k.s2s, k.m2s);
    }
    __device__ float phi_retry(const KParams &P, PolarPacket &k) const
    {
#pragma clang fp contract(off)
        const float s11 = k.p11, s12 = k.p12;
        const float bound = fmaf(fabsf(sI), s11, fabsf(s12) * __builtin_amdgcn_sqrtf(fmaf(sQ, sQ, sU * sU)));
        for (uint32_t kk = 1;; ++kk) {
            const u32x4 q = philox4x32_10(u32x4{photon_id(), case_id(P), i, 2u + kk}, P.key0, P.key1);
            const float r2 = 2.f * u01f(q.x);
            const float acc = u01f(q.y);
            phi_trig(r2, k.sinp, k.cosp, k.c2r, k.s2r);
            phi_signs(r2, k.s2r, k.s2s, k.m2s);
            if (acc * bound <= fmaf(sI, s11, s12 * fmaf(sQ, k.c2r, sU * k.s2r))) return r2;
            if (kk + 1 >= POLAR_MAX_TRIES) {
                atomicAdd(P.ctl + 2, 1ull);
                return r2;
            }
        }
    }

    // One step of an exact-sampler photon given its packet: 2 phi, rotation,
    // Stokes update and move (polar_move_seq), then the events (polar_event).
    // The tail mode's polar_move takes the same candidate by the same tests
    // and forms the step with the same arithmetic, so a photon's record does
    // not depend on which of them walked it.  Returns true when the walk has
    // ended.
    __device__ __forceinline__ bool polar_step(const KParams &P, PolarPacket &k)
    {
        return polar_move_seq(P, k) && polar_event(P, k.dtau, k.w);
    }

    // The one-lane walk's move: the candidate is chosen first (the first of
    // the packet's two, then retries, that is accepted), then the step is
    // formed.  In a wave of 64 photons some lane rejects its first candidate
    // at nearly every step, so the second is ready without a Philox block.
    __device__ __forceinline__ bool polar_move_seq(const KParams &P, PolarPacket &k)
    {
#pragma clang fp contract(off)
        ++i;
        bool ok = phi_ok(k);
        if (!ok) {
            phi_second(k);
            ok = phi_ok(k);
        }
        if (__builtin_expect(__ballot(!ok) != 0ull, 0)) {
            if (!ok) phi_retry(P, k);
        }
        polar_scatter(k);
        const float dtau = k.dtau;
        z = fmaf(dtau, muz, z);
        tau_acc += dtau;
        return z > 0.f || z < -tau_tot || k.w < t_abs_all;
    }

    // The scattering of a step with the packet's candidate: rotation of the
    // direction (mc3D:1255-1281) and Stokes update, from the state before it.
    __device__ __forceinline__ void polar_scatter(const PolarPacket &k)
    {
#pragma clang fp contract(off)
        const float sinp = k.sinp, cosp = k.cosp;
        const float cost = k.ct;
        const float sint = k.sint;
        const float s2 = fmaf(-muz, muz, 1.f);
        const float st = sint * __builtin_amdgcn_rsqf(s2);
        const float mz0 = muz, mx0 = mux, my0 = muy;
        const float gx = fmaf(st, fmaf(mux * muz, cosp, -muy * sinp), mux * cost);
        const float gy = fmaf(st, fmaf(muy * muz, cosp, mux * sinp), muy * cost);
        muz = fmaf(-s2 * st, cosp, muz * cost);
        mux = gx; muy = gy;
        // near-vertical incoming directions: the axis form, behind a
        // wave-uniform branch (the common path issues none of it)
        if (__builtin_expect(__ballot(s2 < 1e-7f) != 0ull, 0)) {
            if (s2 < 1e-7f) {
                const float sg = mz0 > 0.f ? 1.f : -1.f;
                mux = sint * cosp;
                muy = sg * sint * sinp;
                muz = sg * cost;
            }
        }
        stokes_apply(cost, k.omc2, k.c2r, k.s2s, k.m2s, mx0, my0, mz0, k.p11, k.p12, k.p22, k.p33, k.p43, k.p44);
    }

    // The tail mode's move; true when an event follows (the photon left the
    // slab through either face, or its absorption word says absorbed).  The
    // step is formed with the packet's first candidate while that candidate's
    // acceptance test (against the state before the step) runs beside it, so
    // the direction and Stokes chains do not wait for the test; a rejected
    // candidate (rare with a few photons per wave: behind a wave-uniform
    // branch) restores the state, takes the second candidate or draws until
    // one is accepted (replacing the packet's candidate) and forms the step
    // again.  Results equal polar_move_seq's.
    __device__ __forceinline__ bool polar_move(const KParams &P, PolarPacket &k)
    {
#pragma clang fp contract(off)
        ++i;
        const bool ok = phi_ok(k);
        const float mx0 = mux, my0 = muy, mz0 = muz;
        const float i0 = sI, q0 = sQ, u0 = sU, v0 = sV;
        polar_scatter(k);
        if (__builtin_expect(__ballot(!ok) != 0ull, 0)) {
            if (!ok) {
                mux = mx0; muy = my0; muz = mz0;
                sI = i0; sQ = q0; sU = u0; sV = v0;
                phi_second(k);
                if (!phi_ok(k)) phi_retry(P, k);
                polar_scatter(k);
            }
        }
        const float dtau = k.dtau;
        z = fmaf(dtau, muz, z);
        tau_acc += dtau;
        return z > 0.f || z < -tau_tot || k.w < t_abs_all;
    }

    // The event after polar_move returned true (dtau, w: that step's free
    // path and absorption word).  Returns true when the walk has ended; false
    // after a Lambertian re-emission, which walks a step of its own.
    __device__ bool polar_event(const KParams &P, float dtau, uint32_t w)
    {
        if (z > 0.f) {                               // mc3D:1390-1397
            tau_acc -= fminf(z / muz, dtau);
            cond = 1;
            return true;
        }
        if (z < -tau_tot) {                          // mc3D:1399-1459
            tau_acc -= fminf((z + tau_tot) / muz, dtau);
            z = -tau_tot;
            if (lamb && reflects(P)) {
                r = philox_step(pre, i + 1u, P.key0, P.key1);   // the re-emission step's block
                return reemit(P);
            }
            cond = 2;
            return true;
        }
        cond = (w < t_abs5) ? 5 : 4;
        return true;
    }

    // Lambertian re-emission of an aspherical-grain photon (cos theta already
    // set by lambertian_direction): the azimuth is the phase-matrix
    // sampler's (the reference keeps phi_rand there); with `update` the
    // Stokes vector scatters at the Lambertian angle from mz0 = 1.
    __device__ void polar_lambertian(const KParams &P, bool update)
    {
        float ct, rev2;
        int seg;
        polar_angles(P, ct, rev2, seg);
        const float cost = muz;
        const float sint = __builtin_amdgcn_sqrtf(fmaxf(fmaf(-cost, cost, 1.f), 0.f));
        const float rv = 0.5f * rev2;
        mux = sint * __builtin_amdgcn_cosf(rv);
        muy = sint * __builtin_amdgcn_sinf(rv);
        if (update) stokes_update(P, cost, rev2, 0.f, 0.f, 1.f, -1);
    }

    __device__ void result(const KParams &P, int case_id, float &theta, float &phi, float &path, bool &axis) const
    {
        theta = acosf(fminf(fmaxf(muz, -1.f), 1.f));
        path = tau_acc * P.bands[P.cases[case_id].band_begin + band].inv_ext;
        axis = false;
        if (i == 1) {
            phi = 0.f;
        } else {
            float ph = atan2f(muy, mux);
            if (ph < 0.f) ph += TWO_PI_F;
            phi = ph;
            axis = (mux == 0.f) || (muy == 0.f);
        }
    }
};

using NativeWalker = NativeWalkerT<false>;
using PolarWalker = NativeWalkerT<true>;

// ----------------------------------------------------------- replay walker
// FP64, the reference's operation order; contraction off so that every
// product and sum rounds as numpy's float64 scalars do.
struct ReplayWalker {
    using Real = double;
    static constexpr bool COOP = false;
    static constexpr bool STOKES = true;
    static constexpr bool CAP_IN_STEP = false;
    static constexpr int MAX_BLOCK = 256;
    static constexpr int WAVES_PER_EU = 2;
    double z, mux, muy, muz, muz2, path;
    double u_ssa, u_ext, dtau0;
    double g, ssa_ice, ssa_imp, ext_cff, p_ext_imp, tau_tot, R;
    int64_t pos, end;
    uint32_t i;
    int band;
    bool bounce, lamb, err, surf;
    int cond;
    // aspherical particles: Stokes vector and the band's phase-matrix table
    double sI, sQ, sU, sV;
    int pm;
    int case_id;

    __device__ double draw(const KParams &P)
    {
        if (pos >= end) {
            err = true;
            return 0.5;
        }
        return P.rdev[pos++];
    }

    // set up the photon; its steps all run in step() (returns false)
    __device__ bool start(const KParams &P, int case_id, uint32_t photon, uint64_t gid)
    {
        (void)photon;
        const DevCase &C = P.cases[case_id];
        band = P.photon_band ? P.photon_band[gid] : 0;
        const DevBand &B = P.bands[C.band_begin + band];
        g = B.g; ssa_ice = B.ssa_ice; ssa_imp = B.ssa_imp; ext_cff = B.ext_cff;
        p_ext_imp = B.p_ext_imp;
        tau_tot = C.tau_tot; R = C.r_lamb; lamb = C.lamb != 0;
        surf = C.surf != 0;
        pos = P.roff[gid];
        end = P.roff[gid + 1];
        err = false;
        this->case_id = case_id;
        pm = C.polar ? B.pm : -1;
        sI = C.s0[0]; sQ = C.s0[1]; sU = C.s0[2]; sV = C.s0[3];
        dtau0 = -log(draw(P));
        u_ssa = draw(P);
        u_ext = draw(P);
        z = 0.0; mux = C.mux0; muy = 0.0; muz = C.muz0;
        muz2 = muz * muz;
        path = 0.0;
        i = 0;
        bounce = false;
        cond = 0;
        return false;
    }

    __device__ bool step(const KParams &P)
    {
#pragma clang fp contract(off)
        ++i;
        double dtau, u_hg = 0.0, u_phi = 0.0, p_cos = 0.0, two_phi = 0.0;
        if (i >= 2) {
            if (pm < 0 && pos + 5 <= end) {
                // the step's five deviates are consecutive: five independent
                // loads in flight at once instead of five dependent draws
                const double *d = P.rdev + pos;
                const double d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3], d4 = d[4];
                pos += 5;
                u_hg = d0;
                u_phi = d1;
                dtau = -log(d2);
                u_ssa = d3;
                u_ext = d4;
            } else {
                if (pm >= 0) {
                    polar_draws(P, p_cos, two_phi);
                } else {
                    u_hg = draw(P);
                    u_phi = draw(P);
                }
                dtau = -log(draw(P));
                u_ssa = draw(P);
                u_ext = draw(P);
            }
        } else {
            dtau = surf ? 0.0 : dtau0;   // mc3D:1228-1229
        }
        double mxn = mux, myn = muy, mzn = muz;
        if (i > 1) {
            double cost, sint;
            if (bounce || surf) {
                mux = 0.0; muy = 0.0; muz = 1.0;
                for (;;) {
                    const double th = 0.0 + (PI_D / 2) * draw(P);
                    const double r1 = draw(P);
                    if (err) { cost = 1.0; sint = 0.0; break; }
                    if (r1 < 2.0 * sin(th) * cos(th)) {
                        cost = cos(th);
                        sint = sqrt(1.0 - cost * cost);
                        break;
                    }
                }
            } else if (pm >= 0) {
                cost = p_cos;
                sint = sqrt(1.0 - cost * cost);
            } else if (g == 0.0) {
                cost = 1.0 - 2.0 * u_hg;
                sint = sqrt(1.0 - cost * cost);
            } else {
                const double t = (1.0 - g * g) / ((1.0 - g) + (2.0 * g) * u_hg);
                cost = (1.0 / (2.0 * g)) * ((1.0 + g * g) - t * t);
                sint = sqrt(1.0 - cost * cost);
            }
            const double ph = (pm >= 0) ? two_phi / 2.0 : u_phi * (2.0 * PI_D);
            const double cosp = cos(ph), sinp = sin(ph);
            if (muz == 1.0) {
                mxn = sint * cosp; myn = sint * sinp; mzn = cost;
            } else if (muz == -1.0) {
                mxn = sint * cosp; myn = -sint * sinp; mzn = -cost;
            } else {
                const double sq = sqrt(1.0 - muz2);
                mxn = (sint * (mux * muz * cosp - muy * sinp)) / sq + mux * cost;
                myn = (sint * (muy * muz * cosp + mux * sinp)) / sq + muy * cost;
                mzn = -sq * sint * cosp + muz * cost;
            }
            if (bounce) {                        // mc3D:1283-1285
                sI = 1.0; sQ = 0.0; sU = 0.0; sV = 0.0;
            } else if (pm >= 0) {                // mc3D:1287-1342
                stokes_update(P, cost, cosp, mzn, muz);
            }
            bounce = false;
        }
        const double z_prev = z;
        const double z_i = z_prev + dtau * mzn;
        z = z_i;
        if (i > 1) {
            mux = mxn; muy = myn; muz = mzn;
            muz2 = muz * muz;
        }
        path += dtau / ext_cff;
        const bool ice = u_ext > p_ext_imp;
        const double ssa_event = surf ? R : (ice ? ssa_ice : ssa_imp);   // mc3D:1374-1387
        if (z_i > 0.0) {
            path += -((z_i * dtau) / ((z_i - z_prev) * ext_cff));
            cond = 1;
            return true;
        }
        if (z_i < -tau_tot) {
            path += -(((z_i + tau_tot) * dtau) / ((z_i - z_prev) * ext_cff));
            const double dcorr = -(((z_i + tau_tot) / (z_prev - z_i)) * dtau);
            z = z_i - (mzn * dcorr);
            const int exit_cond = (i == 1) ? 3 : 2;
            if (lamb) {
                if (draw(P) <= R) {
                    bounce = true;
                    return err;
                }
            }
            cond = exit_cond;
            return true;
        }
        if (u_ssa >= ssa_event) {
            cond = ice ? 4 : 5;
            return true;
        }
        return err;
    }

    // populate_pdfs for an aspherical particle (mc3D:898-1004): area, cos
    // theta and phi, then the rejection loop (r3; after a rejection: area,
    // cos theta, 2 phi, r3) against I S11 + S12 (Q cos 2phi + U sin 2phi).
    // The reference's operation order and float64 functions.
    __device__ void polar_draws(const KParams &P, double &ct, double &two_phi)
    {
#pragma clang fp contract(off)
        const DevCase &C = P.cases[case_id];
        const DevBand &B = P.bands[C.band_begin + band];
        const int n = P.pm_n;
        const double *x = P.pm + (size_t)pm * 7 * (size_t)n;
        const double I = sI, Q = sQ, U = sU;
        const double two_beta = (Q == 0.0) ? atan(U / 1e-15) : atan(U / Q);
        double A = fabs(I) * B.pA1;
        double Bm = B.pB1;
        const double Cc = fabs(Q * cos(two_beta) + U * sin(two_beta));
        const double mv1 = A + Bm * Cc;
        A = fabs(I) * B.pA2;
        Bm = B.pB2;
        const double mv2 = A + Bm * Cc;
        const double a1 = (C.cmax - C.ccut) * mv1;
        const double a2 = (C.ccut - C.cmin) * mv2;
        const double pct = a1 / (a1 + a2);
        double ar = draw(P);
        ct = draw(P);
        two_phi = 2.0 * (draw(P) * (2.0 * PI_D));
        int area;
        if (ar < pct) {
            area = 1;
            ct = ct * (C.cmax - C.ccut) + C.ccut;
        } else {
            area = 2;
            ct = -ct * (C.ccut - C.cmin) + C.ccut;
        }
        double r3 = 1.0, val = 0.0;
        for (int k = 0; r3 > val; ++k) {
            if (k > 0) {
                ar = draw(P);
                if (ar < pct) {
                    area = 1;
                    ct = draw(P) * (C.cmax - C.ccut) + C.ccut;
                } else {
                    area = 2;
                    ct = -draw(P) * (C.ccut - C.cmin) + C.ccut;
                }
                two_phi = draw(P) * (4.0 * PI_D);
            }
            r3 = 1.000001 * draw(P);
            r3 = (area == 1) ? r3 * mv1 : r3 * mv2;
            const int j = grid_index(x, n, ct);
            const double S11 = interp_at(x, x + n, n, j, ct);
            const double S12 = interp_at(x, x + 2 * n, n, j, ct);
            val = I * S11 + S12 * (Q * cos(two_phi) + U * sin(two_phi));
            if (err) break;
        }
    }

    // mc3D:1287-1342 and rotate_stokes_vector (mc3D:1090-1109), the
    // reference's order and float64 functions.
    __device__ void stokes_update(const KParams &P, double costheta, double cosphi, double muz_n, double muz_0)
    {
#pragma clang fp contract(off)
        const int n = P.pm_n;
        const double *x = P.pm + (size_t)pm * 7 * (size_t)n;
        const double cos2theta = costheta * costheta;
        const double phi_sca = acos(cosphi);
        double da = 2.0 * phi_sca;
        double c = cos(da), s = sin(da);
        const double Isp = sI, Qsp = sQ * c + sU * s, Usp = -sQ * s + sU * c, Vsp = sV;
        const int j = grid_index(x, n, costheta);
        const double P11 = interp_at(x, x + n, n, j, costheta);
        const double P12 = interp_at(x, x + 2 * n, n, j, costheta);
        const double P22 = interp_at(x, x + 3 * n, n, j, costheta);
        const double P33 = interp_at(x, x + 4 * n, n, j, costheta);
        const double P43 = interp_at(x, x + 5 * n, n, j, costheta);
        const double P44 = interp_at(x, x + 6 * n, n, j, costheta);
        const double Ic = Isp * P11 + Qsp * P12;
        const double Qc = Isp * P12 + Qsp * P22;
        const double Uc = Usp * P33 - Vsp * P43;
        const double Vc = Usp * P43 + Vsp * P44;
        const double den = sqrt((1.0 - cos2theta) * (1.0 - muz_n * muz_n));
        double ci;
        if (den == 0.0) {
            ci = 0.0;
        } else {
            const double num = muz_n * costheta - muz_0;
            const double sg = (phi_sca >= PI_D) ? 1.0 : -1.0;
            ci = num / (sg * den);
            if (ci < -1.0) ci = -1.0;
            else if (ci > 1.0) ci = 1.0;
        }
        const double i2 = acos(ci);
        da = 2.0 * (-i2);
        c = cos(da);
        s = sin(da);
        const double Qm = Qc * c + Uc * s, Um = -Qc * s + Uc * c;
        sI = Ic / Ic;
        sQ = Qm / Ic;
        sU = Um / Ic;
        sV = Vc / Ic;
    }

    __device__ void result(const KParams &, int, double &theta, double &phi, double &pth, bool &axis) const
    {
#pragma clang fp contract(off)
        theta = acos(muz);
        pth = path;
        axis = false;
        if (i == 1) phi = 0.0;
        else if (mux > 0 && muy > 0) phi = atan(muy / mux);
        else if (mux < 0 && muy > 0) phi = atan(muy / mux) + PI_D;
        else if (mux < 0 && muy < 0) phi = atan(muy / mux) + PI_D;
        else if (mux > 0 && muy < 0) phi = atan(muy / mux) + 2.0 * PI_D;
        else { phi = __builtin_nan(""); axis = true; }
    }
};


// Lanes are EMPTY (no photon), LIVE or DONE (finished, record pending).
enum { EMPTY = 0, LIVE = 1, DONE = 2 };


}  // namespace mcs
