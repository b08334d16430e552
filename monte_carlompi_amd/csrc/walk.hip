// walk.hip -- photon random walk in a snow layer for MI355X (gfx950).
//
// One photon per lane in a persistent grid.  Work is a queue of chunks
// (case, photon range); the waves of a block draw photons from the block's
// current chunk through an LDS counter and refill idle lanes wave-wide
// (ballot + mbcnt rank, one LDS atomic per refill).  Each chunk owns one of
// SLOTS LDS tally slots (FP64 BRF histogram 18 x 36 + sums + counts); the wave
// that finishes a chunk's last photon flushes the slot to the case's global
// tallies.  Finished photons are compacted per wave (ballot/mbcnt + one LDS
// atomic) into the chunk's contiguous id range, so record stores of lanes that
// finish together are contiguous; a second kernel puts each chunk's records in
// photon order through LDS.
//
// Two walkers share that skeleton:
//  * NativeWalker: FP32, Philox4x32-10 keyed by the seed, counter
//    (photon, case, step, stream); once the queue is empty, sparse waves
//    finish their photons in the wave-cooperative mode below and hand them
//    to waves of the block that ran out of work;
//  * ReplayWalker: FP64 with the reference's exact step semantics
//    (monte_carloMPI/monte_carlo3D.py:1212-1490), consuming the reference's
//    uniform deviates from CSR slices.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "mcsnow.h"
#include "philox.h"

namespace mcs {

constexpr int SLOTS = 4;
constexpr int MAX_WAVES = 16;     // waves per block (1024 threads)
static_assert((SLOTS & (SLOTS - 1)) == 0, "SLOTS is a power of two");
constexpr int TF = MCS_TALLY_F64;
constexpr int TU = MCS_TALLY_U64;
constexpr double PI_D = 3.141592653589793;
constexpr float TWO_PI_F = 6.283185307179586f;
constexpr float LN2_F = 0.6931471805599453f;

// tally layout (doubles), see mcs_read_tallies
constexpr int T_QDOWN = MCS_BRF_BINS;        // 648
constexpr int T_QCOND = MCS_BRF_BINS + 1;    // 649..653 by condition 1..5
constexpr int T_QCAP = MCS_BRF_BINS + 6;     // 654
constexpr int T_QPATH = MCS_BRF_BINS + 7;    // 655
// counts (uint64)
constexpr int U_N = 0, U_CAP = 6, U_AXIS = 7, U_STEPS = 8, U_COOP = 9, U_MOVED = 10;
// rejection attempts per scattering angle before the walk reports an error
// (ctl[2]); valid tables accept within a few
constexpr uint32_t POLAR_MAX_TRIES = 1u << 20;
constexpr int PM_GUIDE = 256;   // guide cells of the FP32 phase-matrix tables

struct DevCase {
    double tau_tot, mux0, muz0, r_lamb;
    unsigned long long t_refl;  // reflect iff 32 random bits < t_refl (R * 2^32)
    float tau_tot_f, mux0_f, muz0_f;
    int32_t lamb, band_begin, n_bands, surf;
    int64_t n_photons, gid_base;
    // aspherical particles: initial Stokes vector and the rejection sampler's
    // cos(theta) range and split (mc3D:858-883, 1132)
    int32_t polar;
    double s0[4], cmax, ccut, cmin;
    float s0f[4], fcmax, fccut, fcmin;
};

struct DevBand {
    double wvn, g, ssa_ice, ssa_imp, ext_cff, p_ext_imp, cdf;
    // HG inverse CDF from the 32 Philox bits x: t = hg_c / (hg_e x + hg_d),
    // cos = hg_a (hg_b - t^2), with hg_e = 2g 2^-32 and hg_d = 1 - g + 2g 2^-33
    // (u = (x + 1/2) 2^-32 folded in); isotropic bands: hg_e = 0, hg_d = 1
    float hg_a, hg_b, hg_c, hg_d, hg_e;
    float inv_ext;                       // 1/ext_cff [m]
    uint32_t t_abs5;                     // bits < P_imp (1 - ssa_imp) 2^32: absorbed by impurity
    uint32_t t_abs_all;                  // bits < [P_imp (1 - ssa_imp) + (1 - P_imp)(1 - ssa_ice)] 2^32: absorbed
    int32_t iso;                         // g == 0
    // alias table of the case's band distribution (Walker): column j of the
    // case keeps band j iff the fraction bits < alias_t, else takes alias_k
    uint32_t alias_t;
    int32_t alias_k;
    // phase-matrix table of the band (-1: Henyey-Greenstein) and the maxima
    // of |P11| and |P12| on either side of the cutoff (mc3D:926-941)
    int32_t pm;
    double pA1, pB1, pA2, pB2;
    float fA1, fB1, fA2, fB2;
    // native mode: |P12| <= P11 on the whole grid, so I S11 + S12 (Q cos 2phi
    // + U sin 2phi) >= 0 for |(Q, U)| <= I and its cos(theta) marginal is S11:
    // the walk samples it exactly by inversion (fM = integral of S11)
    int32_t pm_exact;
    float fM;
};

// probability -> 32-bit threshold (bits < t has probability p)
__host__ __device__ inline uint32_t prob_bits_dev(double p)
{
    const double t = p * 4294967296.0;
    if (!(t > 0.0)) return 0u;
    if (t >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)(t + 0.5);
}

struct KParams {
    const DevCase *cases;
    const DevBand *bands;
    const int32_t *case_chunk0;  // n_cases + 1: first chunk of each case
    int32_t n_cases;
    int32_t chunk;
    uint32_t n_chunks;
    uint32_t refill_min;
    uint32_t key0, key1;
    uint32_t max_steps;
    int32_t coop_live;           // tail mode at <= this many live photons per wave (0: off)
    const uint8_t *photon_band;  // nullable
    const double *rdev;          // replay
    const int64_t *roff;
    double *tally_f;
    unsigned long long *tally_u;
    unsigned int *chunk_counter;
    unsigned long long *ctl;     // [0] replay slice errors, [1] failed hand-overs
    uint32_t *rec_gid, *rec_nscat, *rec_cb;
    void *rec_theta, *rec_phi, *rec_path;
    // phase-matrix tables.  FP64 (replay): 7 rows of pm_n -- ascending cos
    // theta, P11, P12, P22, P33, P43, P44.  FP32 (native), pm_stride floats
    // per table: those 7 rows, the integral of the linear P11 from the first
    // grid point, the six elements' slopes per grid segment, and a guide of
    // PM_GUIDE + 1 segment indices (int bits) where that integral crosses
    // k / PM_GUIDE of its total.  Final Stokes (Q, U, V) per photon in photon
    // order (3 x n_total, nullable).
    const double *pm;
    const float *pmf;
    int32_t pm_n;
    int32_t pm_lds_floats;       // > 0: the FP32 tables are copied into dynamic LDS at block start
    int32_t pm_stride;
    int32_t pad1;
    void *rec_stokes;
    uint64_t n_total;
};

// ---------------------------------------------------------------- LDS state
struct SlotHdr {
    unsigned long long seqcount;  // (block-local chunk seq << 32) | count
    unsigned long long gid0;      // global id of the chunk's first photon
    unsigned int p0;              // first photon (within its case)
    int case_id;
    int remaining;                // photons of the chunk not finished
    int rec_next;                 // records written so far
    int state;                    // 0 free, 1 live
    int pad;
};

struct SlotTally {
    double f[TF];
    unsigned long long u[TU];
};

struct BlockState {
    unsigned long long cur;  // (seq << 32) | next photon offset in chunk
    int lock;
    int exhausted;
    // hand-over of cooperative-tail photons between the block's waves
    int holders;                  // waves not waiting for photons
    unsigned int idle_mask;       // waves waiting for photons (bit per wave)
    int ib_state[MAX_WAVES];      // inbox of each wave: 1 = photons posted
    int ib_src[MAX_WAVES];        //   posting wave (its coop table holds them)
    unsigned int ib_mask[MAX_WAVES];  //   their table entries
    int out_posts[MAX_WAVES];     // posts of each wave not yet taken
    SlotHdr hdr[SLOTS];
    SlotTally tal[SLOTS];
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int lane_rank(unsigned long long mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

__device__ __forceinline__ int bcast(int v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ unsigned long long bcast64(unsigned long long v, int src)
{
    return __shfl(v, src, 64);
}

template <class T>
__device__ __forceinline__ T lds_load(const T *p)
{
    return __atomic_load_n(p, __ATOMIC_RELAXED);
}

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
// cos theta and the phase matrix there, the first 2 phi candidate and its
// acceptance uniform.  One lane can build it for any step of any photon.
struct PolarPacket {
    float dtau;
    uint32_t w;
    float ct, r2, acc;
    float p11, p12, p22, p33, p43, p44;
    int seg;
};

template <bool POLAR>
struct NativeWalkerT {
    using Real = float;
    static constexpr bool COOP = !POLAR;      // has the wave-cooperative tail mode below
    static constexpr bool STOKES = POLAR;     // carries a Stokes vector (aspherical particles)
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
            pm_exact = PB.pm_exact != 0;   // per-photon constants stay in registers:
            pm_mass = PB.fM;               // a step reads only the tables
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
    // been absorbed
    __device__ __forceinline__ bool step(const KParams &P)
    {
        if constexpr (POLAR) {
            if (pm >= 0 && pm_exact) {
                const u32x4 cur = r;
                r = philox_step(pre, i + 2u, P.key0, P.key1);
                const u32x4 pb = philox4x32_10(u32x4{photon_id(), case_id(P), i + 1u, 2u}, P.key0, P.key1);
                PolarPacket k;
                make_packet(cur, pb, pm_table(P), P.pm_n, pm_mass, k);
                return polar_step(P, k);
            }
        }
        ++i;
        const u32x4 cur = r;
        r = philox_step(pre, i + 1u, P.key0, P.key1);
        const float dtau = free_path(cur);
        // polar cosine: HG inverse CDF (mc3D:790-800) with the uniform's
        // scale and offset folded into d and e; isotropic for g = 0
        const float xf = (float)cur.x;
        const float t = cc * __builtin_amdgcn_rcpf(fmaf(e, xf, d));
        float cost = iso ? fmaf(-0x1p-31f, xf, 1.f) : fminf(fmaxf(a * fmaf(-t, t, b), -1.f), 1.f);
        // azimuth: v_sin/v_cos take revolutions, 2 pi u -> u
        float rev = u01f(cur.y);
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
        if (!(z > 0.f || z < -tau_tot || cur.w < t_abs_all)) return false;
        // ---- rare: exits and absorption.  The depth passes an empty asm so
        // that the tests below are issued here, not hoisted into every step.
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
            if (lamb && reflects(P)) return reemit(P);
            cond = 2;
            return true;
        }
        // absorbed: one uniform decides species and absorption jointly, with
        // P(cond 5) = P_imp (1 - omega_imp), P(cond 4) = (1 - P_imp)(1 - omega_ice)
        // (mc3D:1374-1383, 1461-1466)
        cond = (cur.w < t_abs5) ? 5 : 4;
        return true;
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
    __device__ void polar_angles(const KParams &P, float &cost, float &rev2, int &seg) const
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
            // U^2) (polar_phi: acceptance >= 1/2, no table reads).  Philox
            // block (photon, case, i, 2 + k), k = 0, 1, ...
            const u32x4 q = philox4x32_10(u32x4{photon_id(), cs, i, 2u}, P.key0, P.key1);
            PolarPacket k;
            make_packet(q, q, T, n, pm_mass, k);   // its free path and absorption word go unused
            float r2 = k.r2, sinp, cosp, c2r, s2r;
            polar_phi(P, k.p11, k.p12, k.acc, r2, sinp, cosp, c2r, s2r);
            cost = k.ct;
            rev2 = r2;
            seg = k.seg;
            return;
        }
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
        stokes_apply(cost, rev2, __builtin_amdgcn_cosf(rev2), __builtin_amdgcn_sinf(rev2), mx0, my0, mz0, p11, p12,
                     p22, p33, p43, p44);
    }

    // (mx0, my0, mz0): the direction before the scattering; the photon's
    // direction members hold the one after it
    // c2r, s2r: cos and sin of 2 phi (rev2 revolutions)
    __device__ __forceinline__ void stokes_apply(float cost, float rev2, float c2r, float s2r, float mx0, float my0,
                                                 float mz0, float p11, float p12, float p22, float p33, float p43,
                                                 float p44)
    {
#pragma clang fp contract(off)
        const float c2 = c2r;
        const float s2 = (rev2 > 1.f ? -1.f : 1.f) * s2r;
        const float Qs = fmaf(sQ, c2, sU * s2);
        const float Us = fmaf(-sQ, s2, sU * c2);
        const float Ic = fmaf(sI, p11, Qs * p12);
        const float Qc = fmaf(sI, p12, Qs * p22);
        const float Uc = fmaf(Us, p33, -sV * p43);
        const float Vc = fmaf(Us, p43, sV * p44);
        const float muz_n = muz;
        const float den = __builtin_amdgcn_sqrtf(fmaxf(fmaf(-cost, cost, 1.f), 0.f) *
                                                 fmaxf(fmaf(-muz_n, muz_n, 1.f), 0.f));
        // i_2 in [0, pi]: cos i_2 as the reference forms it; sin i_2 >= 0 from
        // its exact identity (den sin i_2)^2 = (e_z . (u_0 x u_n))^2 -- the
        // Gram determinant 1 + 2 mu_0 mu_n cos - mu_0^2 - mu_n^2 - cos^2 --
        // instead of sqrt(1 - cos^2 i_2), which near |cos i_2| = 1 (scattering
        // plane close to the meridian plane) would turn an FP32 rounding of
        // cos i_2 into a rotation of ~1e-3 rad.  den = 0: cos i_2 = 0.
        float ci = 0.f, si = 1.f;
        if (den != 0.f) {
            const float rd = __builtin_amdgcn_rcpf(den);
            const float sg = (rev2 == 1.f) ? 1.f : -1.f;
            ci = fminf(fmaxf(sg * fmaf(muz_n, cost, -mz0) * rd, -1.f), 1.f);
            si = fminf(fabsf(fmaf(mx0, muy, -my0 * mux)) * rd, 1.f);
        }
        const float cr = fmaf(-2.f * si, si, 1.f);     // cos(-2 i_2) = 1 - 2 sin^2 i_2
        const float sr = -2.f * ci * si;               // sin(-2 i_2)
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
                                       PolarPacket &o)
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
        o.seg = j;
        o.p11 = fmaf(slope, dx, y0);
        o.p12 = fmaf(T[9 * n + j], dx, T[2 * n + j]);
        o.p22 = fmaf(T[10 * n + j], dx, T[3 * n + j]);
        o.p33 = fmaf(T[11 * n + j], dx, T[4 * n + j]);
        o.p43 = fmaf(T[12 * n + j], dx, T[5 * n + j]);
        o.p44 = fmaf(T[13 * n + j], dx, T[6 * n + j]);
        o.r2 = 2.f * u01f(pb.y);
        o.acc = u01f(pb.z);
    }

    // 2 phi given cos theta: the first candidate r2 (revolutions) and its
    // acceptance uniform come with the packet, retries draw blocks (photon,
    // case, i, 2 + k); target I S11 + S12 (Q cos 2phi + U sin 2phi), bound
    // S11 + |S12| sqrt(Q^2 + U^2).  Returns the accepted r2 with the sine and
    // cosine of the azimuth (r2 / 2) and of 2 phi (one transcendental pair).
    __device__ __forceinline__ void polar_phi(const KParams &P, float s11, float s12, float acc, float &r2, float &sinp,
                                              float &cosp, float &c2r, float &s2r) const
    {
#pragma clang fp contract(off)
        const float bound = s11 + fabsf(s12) * __builtin_amdgcn_sqrtf(fmaf(sQ, sQ, sU * sU));
        for (uint32_t kk = 1;; ++kk) {
            sinp = __builtin_amdgcn_sinf(0.5f * r2);
            cosp = __builtin_amdgcn_cosf(0.5f * r2);
            c2r = fmaf(cosp, cosp, -sinp * sinp);
            s2r = 2.f * sinp * cosp;
            const float val = fmaf(sI, s11, s12 * fmaf(sQ, c2r, sU * s2r));
            if (acc * bound <= val) break;
            if (kk >= POLAR_MAX_TRIES) {
                atomicAdd(P.ctl + 2, 1ull);
                break;
            }
            const u32x4 q = philox4x32_10(u32x4{photon_id(), case_id(P), i, 2u + kk}, P.key0, P.key1);
            r2 = 2.f * u01f(q.x);
            acc = u01f(q.y);
        }
    }

    // One step of an exact-sampler photon given its packet: 2 phi by
    // rejection against the current Stokes vector (retries draw blocks
    // (photon, case, i, 2 + k)), rotation, Stokes update, move, events.  The
    // one-lane walk and the tail mode below both step through here, so a
    // photon's record does not depend on which of them walked it.
    __device__ bool polar_step(const KParams &P, const PolarPacket &k)
    {
#pragma clang fp contract(off)
        ++i;
        float r2 = k.r2, sinp, cosp, c2r, s2r;
        polar_phi(P, k.p11, k.p12, k.acc, r2, sinp, cosp, c2r, s2r);
        const float cost = k.ct;
        const float sint = __builtin_amdgcn_sqrtf(fmaf(-cost, cost, 1.f));
        const float s2 = fmaf(-muz, muz, 1.f);
        const float st = sint * __builtin_amdgcn_rsqf(s2);
        const float mz0 = muz, mx0 = mux, my0 = muy;
        const float gx = fmaf(st, fmaf(mux * muz, cosp, -muy * sinp), mux * cost);
        const float gy = fmaf(st, fmaf(muy * muz, cosp, mux * sinp), muy * cost);
        muz = fmaf(-s2 * st, cosp, muz * cost);
        mux = gx; muy = gy;
        if (s2 < 1e-7f) {
            const float sg = mz0 > 0.f ? 1.f : -1.f;
            mux = sint * cosp;
            muy = sg * sint * sinp;
            muz = sg * cost;
        }
        stokes_apply(cost, r2, c2r, s2r, mx0, my0, mz0, k.p11, k.p12, k.p22, k.p33, k.p43, k.p44);
        const float dtau = k.dtau;
        z = fmaf(dtau, muz, z);
        tau_acc += dtau;
        if (!(z > 0.f || z < -tau_tot || k.w < t_abs_all)) return false;
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
        cond = (k.w < t_abs5) ? 5 : 4;
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
            if (pm >= 0) {
                polar_draws(P, p_cos, two_phi);
            } else {
                u_hg = draw(P);
                u_phi = draw(P);
            }
            dtau = -log(draw(P));
            u_ssa = draw(P);
            u_ext = draw(P);
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

// One cooperative round of a wave (all 64 lanes call; no divergence around
// the lane permutes).  `tab`: the wave's entries; `map`: group -> entry;
// `me`: this lane's photon's entry.
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
    // lane jj walks steps ka = i0 + 1 + 2 jj and kb = ka + 1 (the same Philox
    // blocks the one-lane walk draws for them)
    const uint32_t ka = cg.i0 + 1u + 2u * (uint32_t)jj, kb = ka + 1u;
    const u32x4 ua = philox_step(cg.pre, ka, P.key0, P.key1);
    const u32x4 ub = philox_step(cg.pre, kb, P.key0, P.key1);
    float dta, dtb;
    Quat La, Lb;
    coop_local(ua, ca, cb, ccc, cd, ce, dta, La);
    coop_local(ub, ca, cb, ccc, cd, ce, dtb, Lb);
    Quat q = qmul(La, Lb);
    if (jj == 0) q = qmul(cg.f, q);   // the group's carried frame on the left
    // frames first: the depth changes need the directions
    group_scan_q(q, lg, jj);
    // frame after step ka: F_b (x) conj(L_b)
    const Quat qa = qmul(q, Quat{Lb.w, -Lb.x, -Lb.y, -Lb.z});
    const float uza = fmaf(-2.f, fmaf(qa.x, qa.x, qa.y * qa.y), 1.f);
    const float uzb = fmaf(-2.f, fmaf(q.x, q.x, q.y * q.y), 1.f);
    const float dzb = dtb * uzb;
    float sz = fmaf(dta, uza, dzb), sp = dta + dtb;
    group_scan_s(sz, sp, lg, jj);
    const float zb = cg.z0 + sz, pb = cg.tau0 + sp;
    const float za = zb - dzb, pa = pb - dtb;
    const bool top_a = za > 0.f && uza > 0.f, bot_a = za < -ctau && uza < 0.f;
    const bool top_b = zb > 0.f && uzb > 0.f, bot_b = zb < -ctau && uzb < 0.f;
    const bool eva = member && (top_a || bot_a || ua.w < ct_abs || ka >= P.max_steps);
    const bool evb = member && (top_b || bot_b || ub.w < ct_abs || kb >= P.max_steps);
    const unsigned long long ev = __ballot(eva || evb);
    if (ev == 0ull) {
        // no event: every group walked 2G steps; carry its last lane's state
        // (the frame renormalised against drift)
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
        cg.i0 += 2u * (uint32_t)G;
        return;
    }
    // this lane's outcome, should one of its steps end the segment: its first
    // step if that one has an event, else its second
    const bool A = eva;
    const bool top = A ? top_a : top_b, bot = A ? bot_a : bot_b;
    const uint32_t wv = A ? ua.w : ub.w, k = A ? ka : kb;
    const bool ab = wv < ct_abs, cap = k >= P.max_steps;
    const float zk = A ? za : zb, pk = A ? pa : pb, dtau = A ? dta : dtb, uz = A ? uza : uzb;
    const Quat qe{A ? qa.w : q.w, A ? qa.x : q.x, A ? qa.y : q.y, A ? qa.z : q.z};
    const float ox = 2.f * fmaf(qe.x, qe.z, qe.w * qe.y);
    const float oy = 2.f * fmaf(qe.y, qe.z, -qe.w * qe.x);
    const int type = top ? 1 : (bot ? 2 : (ab ? 3 : (cap ? 4 : 0)));
    const float corr = top ? zk : zk + ctau;
    const float pout = (top || bot) ? pk - fminf(corr * __builtin_amdgcn_rcpf(uz), dtau) : pk;
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
    const uint32_t r_w = bperm(src, wv);
    const float r_z = bperm(src, zout), r_p = bperm(src, pout);
    const float r_x = bperm(src, ox), r_y = bperm(src, oy), r_uz = bperm(src, uz);
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
template <bool HAND>
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
        coop_round(w, state, P, tab, map, me, cg, coop_steps);
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

// ------------------------------------------------------------ chunk queue
// The block's current chunk: BlockState::cur = (seq << 32) | (slot << 28) |
// next photon offset (chunks are at most 2^24 photons, so hand-outs past the
// end stay below 2^28).
constexpr int CUR_SLOT_SHIFT = 28;
constexpr uint32_t CUR_OFF_MASK = (1u << CUR_SLOT_SHIFT) - 1u;

// Take the block's chunk number `seq` into free slot `slot` (whole wave calls).
// Returns false when the global queue is empty.
__device__ bool acquire_chunk(BlockState &bs, uint32_t seq, int slot, const KParams &P)
{
    const int lane = lane_id();
    unsigned int j = 0;
    if (lane == 0) j = atomicAdd(P.chunk_counter, 1u);
    j = (unsigned int)bcast((int)j, 0);
    if (j >= P.n_chunks) {
        if (lane == 0) __atomic_store_n(&bs.exhausted, 1, __ATOMIC_RELAXED);
        return false;
    }
    // case of chunk j: last c with case_chunk0[c] <= j
    int lo = 0, hi = P.n_cases;  // invariant: chunk0[lo] <= j < chunk0[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((unsigned int)P.case_chunk0[mid] <= j) lo = mid; else hi = mid;
    }
    const DevCase &C = P.cases[lo];
    const unsigned long long p0 = (unsigned long long)(j - (unsigned int)P.case_chunk0[lo]) * (unsigned)P.chunk;
    const long long left = C.n_photons - (long long)p0;
    const int count = (int)(left < P.chunk ? left : P.chunk);
    SlotTally &T = bs.tal[slot];
    for (int w = lane; w < TF; w += 64) T.f[w] = 0.0;
    for (int w = lane; w < TU; w += 64) T.u[w] = 0ull;
    if (lane == 0) {
        SlotHdr &H = bs.hdr[slot];
        H.gid0 = (unsigned long long)C.gid_base + p0;
        H.p0 = (unsigned int)p0;
        H.case_id = lo;
        H.remaining = count;
        H.rec_next = 0;
        H.state = 1;
        __atomic_store_n(&H.seqcount, ((unsigned long long)seq << 32) | (unsigned)count,
                         __ATOMIC_RELAXED);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    return true;
}

// Flush slot s to its case's global tallies and free it (whole wave calls).
__device__ void flush_slot(BlockState &bs, int s, const KParams &P)
{
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const int lane = lane_id();
    const int cid = bs.hdr[s].case_id;
    const SlotTally &T = bs.tal[s];
    double *gf = P.tally_f + (size_t)cid * TF;
    unsigned long long *gu = P.tally_u + (size_t)cid * TU;
    for (int w = lane; w < TF; w += 64) {
        const double v = T.f[w];
        if (v != 0.0) unsafeAtomicAdd(gf + w, v);
    }
    for (int w = lane; w < TU; w += 64) {
        const unsigned long long v = T.u[w];
        if (v) atomicAdd(gu + w, v);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __atomic_store_n(&bs.hdr[s].state, 0, __ATOMIC_RELAXED);
}

template <class W>
__device__ __forceinline__ void write_record(const KParams &P, size_t pos, uint64_t gid,
                                             const W &w, typename W::Real th,
                                             typename W::Real ph, typename W::Real pa)
{
    using R = typename W::Real;
    P.rec_gid[pos] = (uint32_t)gid;
    P.rec_nscat[pos] = w.i - 1;
    P.rec_cb[pos] = (uint32_t)w.cond | ((uint32_t)w.band << 8);
    ((R *)P.rec_theta)[pos] = th;
    ((R *)P.rec_phi)[pos] = ph;
    ((R *)P.rec_path)[pos] = pa;
}

// Record + tally the DONE lanes of the wave; flush chunks that complete.
template <class W>
__device__ __forceinline__ void retire(BlockState &bs, const KParams &P, W &w, int &state,
                                       int my_slot, uint64_t my_gid)
{
    const int lane = lane_id();
    unsigned long long fin = __ballot(state == DONE);
    while (fin) {
        const int l0 = __ffsll((long long)fin) - 1;
        const int s = bcast(my_slot, l0);
        const unsigned long long same = __ballot(state == DONE && my_slot == s) & fin;
        fin &= ~same;
        const int cnt = __popcll(same);
        const bool mine = (same >> lane) & 1ull;
        SlotHdr &H = bs.hdr[s];
        SlotTally &T = bs.tal[s];
        int base = 0;
        if (lane == 0) base = atomicAdd(&H.rec_next, cnt);
        base = bcast(base, 0);
        if (mine) {
            typename W::Real th, ph, pa;
            bool axis;
            w.result(P, H.case_id, th, ph, pa, axis);
            if (P.rec_gid) write_record(P, (size_t)H.gid0 + base + lane_rank(same), my_gid, w, th, ph, pa);
            if constexpr (W::STOKES) {
                if (P.rec_stokes) {   // photon order: Q, U, V planes of n_total
                    typename W::Real *st = (typename W::Real *)P.rec_stokes;
                    st[my_gid] = w.sQ;
                    st[P.n_total + my_gid] = w.sU;
                    st[2 * P.n_total + my_gid] = w.sV;
                }
            }
            const DevBand &B = P.bands[P.cases[H.case_id].band_begin + w.band];
            const double wvn = B.wvn;
            unsafeAtomicAdd(&T.f[T_QDOWN], wvn);
            atomicAdd(&T.u[U_N], 1ull);
            atomicAdd(&T.u[U_STEPS], (unsigned long long)w.i);
            if (w.cond >= 1) {
                unsafeAtomicAdd(&T.f[T_QCOND + w.cond - 1], wvn);
                atomicAdd(&T.u[w.cond], 1ull);
            } else {
                unsafeAtomicAdd(&T.f[T_QCAP], wvn);
                atomicAdd(&T.u[U_CAP], 1ull);
            }
            if (axis) atomicAdd(&T.u[U_AXIS], 1ull);
            if (w.cond == 1) {
                const double thd = (double)th, phd = (double)ph;
                int jt = (int)(thd * (MCS_THETA_BINS / (PI_D / 2)));
                int kp = (int)(phd * (MCS_PHI_BINS / (2 * PI_D)));
                jt = jt < 0 ? 0 : (jt > MCS_THETA_BINS - 1 ? MCS_THETA_BINS - 1 : jt);
                kp = kp < 0 ? 0 : (kp > MCS_PHI_BINS - 1 ? MCS_PHI_BINS - 1 : kp);
                if (phd != phd) kp = 0;
                unsafeAtomicAdd(&T.f[jt * MCS_PHI_BINS + kp], wvn);
                unsafeAtomicAdd(&T.f[T_QPATH], wvn * (double)pa);
            }
            if constexpr (sizeof(typename W::Real) == 8) {
                if (w.err || w.pos != w.end) atomicAdd(P.ctl, 1ull);
            }
            state = EMPTY;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        int old = 0;
        if (lane == 0) old = atomicSub(&H.remaining, cnt);
        old = bcast(old, 0);
        if (old == cnt) flush_slot(bs, s, P);
    }
}

// Give photons of the block's current chunk to EMPTY lanes, advancing the
// block to the next chunk of the global queue when the current one is used up.
template <class W>
__device__ __forceinline__ void refill(BlockState &bs, const KParams &P, W &w, int &state,
                                      int &my_slot, uint64_t &my_gid)
{
    const int lane = lane_id();
    for (int attempt = 0; attempt < 2; ++attempt) {
        const unsigned long long idle = __ballot(state == EMPTY);
        const int n = __popcll(idle);
        if (n == 0) break;
        const unsigned long long cur = lds_load(&bs.cur);
        uint32_t seq = (uint32_t)(cur >> 32);
        int cslot = (int)((uint32_t)cur >> CUR_SLOT_SHIFT) & (SLOTS - 1);
        uint32_t off = (uint32_t)cur & CUR_OFF_MASK;
        unsigned long long meta = lds_load(&bs.hdr[cslot].seqcount);
        uint32_t count = ((uint32_t)(meta >> 32) == seq) ? (uint32_t)meta : 0u;
        int got = 0;
        if (off < count) {
            unsigned long long old = 0;
            if (lane == 0) old = atomicAdd(&bs.cur, (unsigned long long)n);
            old = bcast64(old, 0);
            seq = (uint32_t)(old >> 32);
            cslot = (int)((uint32_t)old >> CUR_SLOT_SHIFT) & (SLOTS - 1);
            off = (uint32_t)old & CUR_OFF_MASK;
            meta = lds_load(&bs.hdr[cslot].seqcount);
            count = ((uint32_t)(meta >> 32) == seq) ? (uint32_t)meta : 0u;
            got = off < count ? (int)min((uint32_t)n, count - off) : 0;
            const int rk = lane_rank(idle);
            if (state == EMPTY && rk < got) {
                const int s = cslot;
                const SlotHdr &H = bs.hdr[s];
                const uint32_t photon = H.p0 + off + rk;
                my_gid = H.gid0 + off + rk;
                my_slot = s;
                state = w.start(P, H.case_id, photon, my_gid) ? DONE : LIVE;
            }
        }
        if (got == n) break;
        // current chunk used up: one wave advances the block to the next
        int lk = 1;
        if (lane == 0) lk = atomicCAS(&bs.lock, 0, 1);
        lk = bcast(lk, 0);
        if (lk != 0) break;
        bool advanced = false;
        const unsigned long long now = lds_load(&bs.cur);
        const uint32_t seq_now = (uint32_t)(now >> 32);
        if (seq_now == seq && !lds_load(&bs.exhausted)) {
            // any free slot: a chunk whose last photons still walk holds only its own
            int ns = -1;
            for (int d = 1; d <= SLOTS && ns < 0; ++d) {
                const int s = (cslot + d) & (SLOTS - 1);
                if (lds_load(&bs.hdr[s].state) == 0) ns = s;
            }
            if (ns >= 0 && acquire_chunk(bs, seq_now + 1, ns, P)) {
                if (lane == 0)
                    __atomic_store_n(&bs.cur,
                                     ((unsigned long long)(seq_now + 1) << 32) | ((unsigned long long)ns << CUR_SLOT_SHIFT),
                                     __ATOMIC_RELAXED);
                advanced = true;
            } else if (ns < 0) {
                // every slot still holds a walking chunk: no acquisition can
                // tell this block that the queue ran dry, so read the counter
                // (otherwise small chunks would keep the block out of the
                // cooperative tail until its last chunk ends)
                unsigned int taken = 0u;
                if (lane == 0) taken = atomicOr(P.chunk_counter, 0u);
                taken = (unsigned int)bcast((int)taken, 0);
                if (taken >= P.n_chunks && lane == 0) __atomic_store_n(&bs.exhausted, 1, __ATOMIC_RELAXED);
            }
        } else if (seq_now != seq) {
            advanced = true;  // someone else advanced: retry the handout
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __atomic_store_n(&bs.lock, 0, __ATOMIC_RELAXED);
        if (!advanced) break;
    }
}

// Tail mode of the polarised walk.  Once the queue is empty and a wave holds
// at most 16 photons (all with the exact sampler), the wave works in rounds:
// the G = 64 / 2^ceil(log2 L) lanes of group g build the packets of the next
// G steps of the group's photon (Philox blocks, cos theta, phase matrix,
// first phi candidate: everything that does not depend on the photon's
// state), then each photon's own lane walks those G steps through
// polar_step.  The serial chain of a step keeps only the phi acceptance,
// rotation, Stokes update and events.  Records equal the one-lane walk's.
// LDS after the tables: 16 owner lanes and 64 packets per wave.
__device__ void polar_tail(PolarWalker &w, int &state, const KParams &P, int wave)
{
    extern __shared__ float pm_lds_tab[];
    int *const owners = reinterpret_cast<int *>(pm_lds_tab + P.pm_lds_floats);
    PolarPacket *const pk = reinterpret_cast<PolarPacket *>(owners + (blockDim.x >> 6) * 16) + wave * 64;
    int *const own = owners + wave * 16;
    const int lane = lane_id();
    for (;;) {
        const unsigned long long live = __ballot(state == LIVE);
        if (live == 0ull) break;
        const int L = __popcll(live);
        const int lg = (L <= 1) ? 0 : 32 - __clz(L - 1);
        const int G = 64 >> lg;
        if (state == LIVE) own[lane_rank(live)] = lane;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const int g = lane >> (6 - lg);
        const int o = own[g < L ? g : 0];
        const uint32_t ck0 = bperm(o, w.pre.ck0), h0 = bperm(o, w.pre.h0), l0 = bperm(o, w.pre.l0);
        const uint32_t io = bperm(o, w.i);
        const int pmo = (int)bperm(o, (uint32_t)w.pm);
        const float mass = bperm(o, w.pm_mass);
        if (g < L) {
            const PhiloxPre pre{ck0, h0, l0};
            const uint32_t st = io + 1u + (uint32_t)(lane & (G - 1));
            const u32x4 mainb = philox_step(pre, st, P.key0, P.key1);
            const u32x4 pb = philox4x32_10(u32x4{l0 * 0x991A7CDBu, ck0 ^ P.key0, st, 2u}, P.key0, P.key1);
            const float *T = (P.pm_lds_floats > 0 ? pm_lds_tab : P.pmf) + (size_t)pmo * (size_t)P.pm_stride;
            PolarPacket q;
            PolarWalker::make_packet(mainb, pb, T, P.pm_n, mass, q);
            pk[lane] = q;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (state == LIVE) {
            const int gme = lane_rank(live);
            for (int k = 0; k < G; ++k) {
                const uint32_t i0 = w.i;
                const PolarPacket q = pk[gme * G + k];
                const bool d = w.polar_step(P, q);
                if (d || w.i >= P.max_steps) {
                    if (!d) w.cond = 0;
                    state = DONE;
                    break;
                }
                if (w.i != i0 + 1u) break;   // a bottom re-emission took a step of its own
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

// Wave issue priority from the oldest live photon (i = its step count, 0 for
// other lanes): long walks set the tail of a launch, so their waves win issue
// slots over young ones.
__device__ __forceinline__ void set_priority(uint32_t i)
{
    if (__ballot(i > 6000u)) __builtin_amdgcn_s_setprio(3);
    else if (__ballot(i > 3000u)) __builtin_amdgcn_s_setprio(2);
    else if (__ballot(i > 1500u)) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// WPE: waves per SIMD the register allocation targets (the launch picks the
// instance that fits its shape).  The hand-over of cooperative-tail photons
// between waves is left out of the 6-wave build: at 80 VGPRs its code would
// spill in the step loop.
template <class W, int WPE>
__global__ void __launch_bounds__(W::MAX_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE)))
k_walk(const KParams P)
{
    constexpr bool HAND = W::COOP && WPE <= 5;
    __shared__ BlockState bs;
    // native mode: coop_live table entries per wave, then coop_live group slots per wave
    extern __shared__ CoopEntry coop_tab[];
    const int lane = lane_id();
    const int wave = threadIdx.x >> 6;
    // ---- init: zero LDS, wave 0 takes the first chunk
    {
        unsigned long long *w64 = reinterpret_cast<unsigned long long *>(&bs);
        for (int k = threadIdx.x; k < (int)(sizeof(BlockState) / 8); k += blockDim.x) w64[k] = 0ull;
    }
    if constexpr (W::STOKES && W::COOP == false && sizeof(typename W::Real) == 4) {
        // polarised native walk: phase-matrix tables into dynamic LDS
        if (P.pm_lds_floats > 0) {
            extern __shared__ float pm_lds_tab[];
            for (int k = threadIdx.x; k < P.pm_lds_floats; k += blockDim.x) pm_lds_tab[k] = P.pmf[k];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) bs.holders = blockDim.x >> 6;
    if (wave == 0) {
        if (!acquire_chunk(bs, 0u, 0, P)) {
            if (lane == 0) bs.hdr[0].seqcount = 0ull;  // seq 0, count 0
        }
    }
    __syncthreads();

    int *const coop_map = reinterpret_cast<int *>(coop_tab + (blockDim.x >> 6) * P.coop_live);
    W w;
    int state = EMPTY;
    int my_slot = 0;
    uint64_t my_gid = 0;
    int spins = 0;
    bool to_coop = false;

    for (;;) {
        // ---------------- cold: retire finished lanes, refill empty ones
        if (__ballot(state == DONE)) retire(bs, P, w, state, my_slot, my_gid);
        const bool exhausted = lds_load(&bs.exhausted) != 0;
        if (!exhausted && __ballot(state == EMPTY)) refill(bs, P, w, state, my_slot, my_gid);
        const unsigned long long live = __ballot(state == LIVE);
        if (live == 0ull) {
            // photons can finish at hand-out (first step, surface mode): retire first
            if (__ballot(state == DONE)) continue;
            if (lds_load(&bs.exhausted)) break;
            __builtin_amdgcn_s_sleep(2);
            if (++spins > (1 << 28)) break;  // cannot happen: chunks always finish
            continue;
        }
        if constexpr (W::COOP) {
            // queue empty and few photons left: walk them cooperatively
            if (lds_load(&bs.exhausted) && __popcll(live) <= P.coop_live) {
                if constexpr (HAND) {
                    to_coop = true;   // below, with the hand-over
                } else {
                    __builtin_amdgcn_s_setprio(3);
                    coop_phase<false>(bs, w, state, P, coop_tab + wave * P.coop_live, coop_map + wave * P.coop_live,
                                      &bs.tal[my_slot].u[U_COOP], &bs.tal[my_slot].u[U_MOVED], wave);
                    retire(bs, P, w, state, my_slot, my_gid);
                }
                break;
            }
        }
        constexpr bool PTAIL = W::STOKES && !W::COOP && sizeof(typename W::Real) == 4;
        const int ptail = PTAIL ? min(P.coop_live, 16) : 0;
        if constexpr (PTAIL) {
            // polarised walk: queue empty, few photons, all with the exact sampler
            if (ptail > 0 && lds_load(&bs.exhausted) && __popcll(live) <= ptail &&
                __ballot(state == LIVE && !(w.pm >= 0 && w.pm_exact)) == 0ull) {
                __builtin_amdgcn_s_setprio(3);
                polar_tail(w, state, P, wave);
                retire(bs, P, w, state, my_slot, my_gid);
                break;
            }
        }
        // leave the hot loop once `thresh` lanes are not live (with the queue
        // empty: once the cooperative tail can take over)
        const int thresh = lds_load(&bs.exhausted) ? (W::COOP ? 64 - P.coop_live : 64 - ptail)
                                                   : min(64, (64 - __popcll(live)) + (int)P.refill_min);
        // ---------------- hot: step every live photon
        for (uint32_t iter = 1;; ++iter) {
            if (state == LIVE) {
                const bool d = w.step(P);
                if (d || w.i >= P.max_steps) {
                    if (!d) w.cond = 0;
                    state = DONE;
                }
            }
            const unsigned long long nl = __ballot(state == LIVE);
            if (nl == 0ull || 64 - __popcll(nl) >= thresh) break;
            if ((iter & 255u) == 0u) set_priority(state == LIVE ? w.i : 0u);
        }
    }
    if constexpr (HAND) {
        // the cooperative tail: the wave's last photons, then photons taken
        // over from siblings (at most coop_live / 2 at a time) until the block
        // has none left
        if (P.coop_live > 0) {
            for (;;) {
                if (to_coop) {
                    __builtin_amdgcn_s_setprio(3);
                    await_own_posts(bs, P, wave);
                    coop_phase<true>(bs, w, state, P, coop_tab + wave * P.coop_live, coop_map + wave * P.coop_live,
                                     &bs.tal[my_slot].u[U_COOP], &bs.tal[my_slot].u[U_MOVED], wave);
                    // every photon the wave held is finished or handed over
                    retire(bs, P, w, state, my_slot, my_gid);
                }
                if (!coop_receive(bs, P, w, state, my_slot, my_gid, coop_tab, wave)) break;
                to_coop = true;
            }
        }
    }
}

// records: compacted order -> photon order
template <class R>
__global__ void k_scatter_records(const uint32_t *gid, const uint32_t *nscat, const uint32_t *cb,
                                  const R *th, const R *ph, const R *pa, uint64_t n,
                                  uint32_t *o_nscat, uint32_t *o_cb, R *o_th, R *o_ph, R *o_pa)
{
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t g = gid[j];
        if (g >= n) continue;
        o_nscat[g] = nscat[j];
        o_cb[g] = cb[j];
        o_th[g] = th[j];
        o_ph[g] = ph[j];
        o_pa[g] = pa[j];
    }
}

// records: compacted order -> photon order, one chunk per block pass.  A
// chunk's records sit at its own id range [gid0, gid0 + count) in any order
// (retirement order), so the block gathers them into LDS by id and writes
// them out contiguously: coalesced reads and writes.  LDS: chunk x (8 + 3
// sizeof(R)) bytes.
template <class R>
__global__ void k_order_records(const DevCase *cases, const int32_t *chunk0, int n_cases, int chunk,
                                uint32_t n_chunks, const uint32_t *gid, const uint32_t *nscat, const uint32_t *cb,
                                const R *th, const R *ph, const R *pa, uint32_t *o_nscat, uint32_t *o_cb, R *o_th,
                                R *o_ph, R *o_pa)
{
    extern __shared__ uint32_t ord_lds[];
    uint32_t *s_ns = ord_lds, *s_cb = ord_lds + chunk;
    R *s_th = reinterpret_cast<R *>(ord_lds + 2 * chunk), *s_ph = s_th + chunk, *s_pa = s_ph + chunk;
    for (uint32_t j = blockIdx.x; j < n_chunks; j += gridDim.x) {
        int lo = 0, hi = n_cases;   // chunk0[lo] <= j < chunk0[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if ((uint32_t)chunk0[mid] <= j) lo = mid; else hi = mid;
        }
        const DevCase &C = cases[lo];
        const int64_t p0 = (int64_t)(j - (uint32_t)chunk0[lo]) * chunk;
        const int count = (int)min<int64_t>(chunk, C.n_photons - p0);
        const uint64_t g0 = (uint64_t)(C.gid_base + p0);
        for (int t = threadIdx.x; t < count; t += blockDim.x) {
            const uint64_t pos = g0 + t;
            const uint32_t k = gid[pos] - (uint32_t)g0;
            if (k < (uint32_t)count) {
                s_ns[k] = nscat[pos];
                s_cb[k] = cb[pos];
                s_th[k] = th[pos];
                s_ph[k] = ph[pos];
                s_pa[k] = pa[pos];
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < count; t += blockDim.x) {
            const uint64_t pos = g0 + t;
            o_nscat[pos] = s_ns[t];
            o_cb[pos] = s_cb[t];
            o_th[pos] = s_th[t];
            o_ph[pos] = s_ph[t];
            o_pa[pos] = s_pa[t];
        }
        __syncthreads();
    }
}

// Zero the tallies, counts and control words of a walk (one launch; a kernel,
// not a memset, so that the walk can be captured in a graph).
__global__ void k_clear(unsigned long long *a, uint64_t na, unsigned long long *b, uint64_t nb,
                        unsigned long long *c, uint64_t nc)
{
    const uint64_t n = na + nb + nc;
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n;
         j += (uint64_t)gridDim.x * blockDim.x) {
        if (j < na) a[j] = 0ull;
        else if (j < na + nb) b[j - na] = 0ull;
        else c[j - na - nb] = 0ull;
    }
}

// The walk's checks for the host: photons, photon-steps, capped, phi on an
// axis (summed over cases), control words 0 and 1, and a written-flag, into
// pinned host memory (one block of 256 threads).
__global__ void k_summary(const unsigned long long *tu, int nc, const unsigned long long *ctl,
                          unsigned long long *out)
{
    __shared__ unsigned long long red[4][256];
    const int t = threadIdx.x;
    unsigned long long a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int k = t; k < nc; k += 256) {
        const unsigned long long *u = tu + (size_t)k * TU;
        a0 += u[U_N]; a1 += u[U_STEPS]; a2 += u[U_CAP]; a3 += u[U_AXIS];
    }
    red[0][t] = a0; red[1][t] = a1; red[2][t] = a2; red[3][t] = a3;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
            for (int q = 0; q < 4; ++q) red[q][t] += red[q][t + h];
        __syncthreads();
    }
    if (t == 0) {
        out[0] = red[0][0]; out[1] = red[1][0]; out[2] = red[2][0]; out[3] = red[3][0];
        out[4] = ctl[0]; out[5] = ctl[1]; out[6] = ctl[2];
        out[7] = 1ull;
    }
}

__global__ void k_philox(const uint32_t *ctr, uint32_t k0, uint32_t k1, uint32_t *out, uint64_t n)
{
    const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    const u32x4 r = philox4x32_10(u32x4{ctr[4 * j], ctr[4 * j + 1], ctr[4 * j + 2], ctr[4 * j + 3]}, k0, k1);
    out[4 * j] = r.x; out[4 * j + 1] = r.y; out[4 * j + 2] = r.z; out[4 * j + 3] = r.w;
}

}  // namespace mcs

// =================================================================== host
using namespace mcs;

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct mcs_ctx {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    char err[512] = {0};
    mcs_launch launch{256, 4, 2048, 8, 16};
    // host copies
    std::vector<mcs_case> cases;
    std::vector<mcs_band> bands;
    std::vector<DevCase> dcases;
    std::vector<DevBand> dbands;
    int64_t total_photons = 0;
    bool cases_set = false;
    // device buffers
    DevBuf d_cases, d_bands, d_chunk0, d_tally_f, d_tally_u, d_ctl, d_pband, d_rdev, d_roff;
    DevBuf d_rec[6], d_out[5];
    // aspherical particles: phase-matrix tables (FP64, FP32), final Stokes
    bool polar = false;
    int32_t pm_n = 0;
    size_t pm_floats = 0;        // FP32 table floats (n_tables x pm_stride)
    DevBuf d_pm, d_pmf, d_stokes;
    bool last_stokes = false;
    int last_mode = MCS_NATIVE;
    bool last_records = false;
    int64_t last_n = 0;
    bool pending = false;
    mcs_stats stats{};
    // graph cache
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t gkey = 0;
    // chunk table on the device (uploaded when it changes)
    std::vector<int32_t> chunk0;
    bool chunk0_valid = false;
    // pinned host words written by k_summary (h_sum) and its device address
    unsigned long long *h_sum = nullptr, *d_sum = nullptr;
};

static int set_err(mcs_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof c->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(ctx, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return set_err(ctx, e_ == hipErrorOutOfMemory ? MCS_ERR_NOMEM : MCS_ERR_HIP, \
                           "%s: %s", #expr, hipGetErrorString(e_));                    \
    } while (0)

static int ensure(mcs_ctx *c, DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return MCS_OK;
    if (b.p) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess)
        return set_err(c, MCS_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    b.cap = bytes;
    c->gkey = 0;  // buffer addresses changed: recapture any graph
    return MCS_OK;
}

// Walker/Vose alias table over a case's bands.  Band k's probability is what
// "first k with u < cdf_k" gives (cdf clamped to [0, 1], mass beyond the
// running maximum; the last band takes the rest), so any cdf column keeps its
// meaning.
static void build_alias(DevBand *db, const mcs_band *b, int n)
{
    std::vector<double> q(n);
    double m = 0.0;
    for (int k = 0; k < n - 1; ++k) {
        const double ck = std::min(std::max(b[k].cdf, 0.0), 1.0);
        q[k] = std::max(ck - m, 0.0) * n;
        m = std::max(m, ck);
    }
    q[n - 1] = (1.0 - m) * n;
    std::vector<int> small, large;
    for (int k = 0; k < n; ++k) (q[k] < 1.0 ? small : large).push_back(k);
    for (int k = 0; k < n; ++k) { db[k].alias_k = k; db[k].alias_t = 0xFFFFFFFFu; }
    while (!small.empty() && !large.empty()) {
        const int s = small.back(), l = large.back();
        small.pop_back();
        db[s].alias_t = (uint32_t)std::min(q[s] * 4294967296.0, 4294967295.0);
        db[s].alias_k = l;
        q[l] -= 1.0 - q[s];
        if (q[l] < 1.0) { large.pop_back(); small.push_back(l); }
    }
}

static uint32_t prob_bits(double p)
{
    double t = p * 4294967296.0;
    if (!(t > 0.0)) return 0u;
    if (t >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)llround(t);
}

extern "C" {

int mcs_create(int device, mcs_ctx **out)
{
    if (!out) return MCS_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MCS_ERR_HIP;
    if (device < 0 || device >= n) return MCS_ERR_ARG;
    mcs_ctx *c = new (std::nothrow) mcs_ctx();
    if (!c) return MCS_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return MCS_ERR_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return MCS_ERR_HIP; }
    c->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipHostMalloc((void **)&c->h_sum, 8 * sizeof(unsigned long long), hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void **)&c->d_sum, c->h_sum, 0) != hipSuccess) {
        (void)mcs_destroy(c);
        return MCS_ERR_HIP;
    }
    *out = c;
    return MCS_OK;
}

int mcs_destroy(mcs_ctx *c)
{
    if (!c) return MCS_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    DevBuf *all[] = {&c->d_cases, &c->d_bands, &c->d_chunk0, &c->d_tally_f, &c->d_tally_u,
                     &c->d_ctl, &c->d_pband, &c->d_rdev, &c->d_roff, &c->d_pm, &c->d_pmf, &c->d_stokes};
    for (DevBuf *b : all) if (b->p) (void)hipFree(b->p);
    for (auto &b : c->d_rec) if (b.p) (void)hipFree(b.p);
    for (auto &b : c->d_out) if (b.p) (void)hipFree(b.p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->h_sum) (void)hipHostFree(c->h_sum);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MCS_OK;
}

const char *mcs_last_error(const mcs_ctx *c) { return c ? c->err : "null context"; }

int mcs_device_cus(const mcs_ctx *c, int32_t *out)
{
    if (!c || !out) return MCS_ERR_ARG;
    *out = c->cus;
    return MCS_OK;
}

int mcs_set_cases(mcs_ctx *c, const mcs_case *cases, int32_t n_cases, const mcs_band *bands,
                  int32_t n_bands)
{
    if (!c) return MCS_ERR_ARG;
    if (!cases || !bands || n_cases <= 0 || n_bands <= 0)
        return set_err(c, MCS_ERR_ARG, "cases/bands must be non-empty");
    std::vector<DevCase> dc(n_cases);
    std::vector<DevBand> db(n_bands);
    for (int b = 0; b < n_bands; ++b) {
        const mcs_band &B = bands[b];
        const double v[] = {B.wvn, B.g, B.ssa_ice, B.ssa_imp, B.ext_cff, B.p_ext_imp, B.snow_depth, B.cdf};
        for (double x : v)
            if (!std::isfinite(x)) return set_err(c, MCS_ERR_ARG, "band %d: non-finite value", b);
        if (!(B.g > -1.0 && B.g < 1.0)) return set_err(c, MCS_ERR_ARG, "band %d: |g| must be < 1", b);
        if (B.ssa_ice < 0 || B.ssa_ice > 1 || B.ssa_imp < 0 || B.ssa_imp > 1)
            return set_err(c, MCS_ERR_ARG, "band %d: ssa outside [0, 1]", b);
        if (!(B.ext_cff > 0)) return set_err(c, MCS_ERR_ARG, "band %d: ext_cff must be > 0", b);
        if (B.p_ext_imp < 0 || B.p_ext_imp > 1) return set_err(c, MCS_ERR_ARG, "band %d: p_ext_imp outside [0, 1]", b);
        if (B.wvn <= 0) return set_err(c, MCS_ERR_ARG, "band %d: wvn must be > 0", b);
        DevBand &D = db[b];
        D.wvn = B.wvn; D.g = B.g; D.ssa_ice = B.ssa_ice; D.ssa_imp = B.ssa_imp;
        D.ext_cff = B.ext_cff; D.p_ext_imp = B.p_ext_imp; D.cdf = B.cdf;
        D.iso = (B.g == 0.0);
        const double g = D.iso ? 0.5 : B.g;
        D.hg_a = (float)(1.0 / (2.0 * g)); D.hg_b = (float)(1.0 + g * g); D.hg_c = (float)(1.0 - g * g);
        D.hg_d = D.iso ? 1.f : (float)(1.0 - g + 2.0 * g * 0x1p-33);
        D.hg_e = D.iso ? 0.f : (float)(2.0 * g * 0x1p-32);
        D.inv_ext = (float)(1.0 / B.ext_cff);
        const double q5 = B.p_ext_imp * (1.0 - B.ssa_imp);
        D.t_abs5 = prob_bits(q5);
        D.t_abs_all = prob_bits(q5 + (1.0 - B.p_ext_imp) * (1.0 - B.ssa_ice));
        D.pm = -1;   // Henyey-Greenstein until mcs_set_polar
    }
    int64_t total = 0;
    std::vector<int> row_case(n_bands, -1);   // first case using each band row
    for (int k = 0; k < n_cases; ++k) {
        const mcs_case &C = cases[k];
        if (!(C.tau_tot > 0) || !std::isfinite(C.tau_tot)) return set_err(c, MCS_ERR_ARG, "case %d: tau_tot must be finite and > 0", k);
        if (!std::isfinite(C.mux0) || !std::isfinite(C.muz0) || !(C.muz0 < 0) ||
            std::fabs(C.mux0 * C.mux0 + C.muz0 * C.muz0 - 1.0) > 1e-9)
            return set_err(c, MCS_ERR_ARG, "case %d: (mux0, muz0) must be a downward unit vector", k);
        if (!(C.r_lambertian >= 0 && C.r_lambertian <= 1)) return set_err(c, MCS_ERR_ARG, "case %d: r_lambertian outside [0, 1]", k);
        if (C.n_bands < 1 || C.n_bands > MCS_MAX_BANDS || C.band_begin < 0 || C.band_begin + C.n_bands > n_bands)
            return set_err(c, MCS_ERR_ARG, "case %d: band range [%d, %d) invalid", k, C.band_begin, C.band_begin + C.n_bands);
        if (C.n_photons < 0 || C.n_photons > 0xFFFFFFFFll) return set_err(c, MCS_ERR_ARG, "case %d: n_photons outside [0, 2^32)", k);
        // a band row's cdf (and alias entry) belongs to one band range
        for (int j = C.band_begin; j < C.band_begin + C.n_bands; ++j) {
            const int o = row_case[j];
            if (o < 0) { row_case[j] = k; continue; }
            if (cases[o].band_begin != C.band_begin || cases[o].n_bands != C.n_bands)
                return set_err(c, MCS_ERR_ARG, "cases %d and %d share band %d with different band ranges", o, k, j);
        }
        DevCase &D = dc[k];
        D.tau_tot = C.tau_tot; D.mux0 = C.mux0; D.muz0 = C.muz0; D.r_lamb = C.r_lambertian;
        D.t_refl = (unsigned long long)llround(C.r_lambertian * 4294967296.0);
        D.tau_tot_f = (float)C.tau_tot; D.mux0_f = (float)C.mux0; D.muz0_f = (float)C.muz0;
        D.lamb = C.lambertian_bottom ? 1 : 0; D.band_begin = C.band_begin; D.n_bands = C.n_bands;
        D.surf = C.lambertian_surface ? 1 : 0;
        D.polar = 0;
        D.s0[0] = 1.0; D.s0f[0] = 1.f;   // (1, 0, 0, 0): unpolarised until mcs_set_polar
        D.n_photons = C.n_photons; D.gid_base = total;
        total += C.n_photons;
        build_alias(&db[C.band_begin], &bands[C.band_begin], C.n_bands);
    }
    (void)hipSetDevice(c->device);
    if (c->pending) (void)hipStreamSynchronize(c->stream), c->pending = false;
    int r;
    if ((r = ensure(c, c->d_cases, sizeof(DevCase) * n_cases))) return r;
    if ((r = ensure(c, c->d_bands, sizeof(DevBand) * n_bands))) return r;
    if ((r = ensure(c, c->d_tally_f, sizeof(double) * TF * n_cases))) return r;
    if ((r = ensure(c, c->d_tally_u, sizeof(unsigned long long) * TU * n_cases))) return r;
    if ((r = ensure(c, c->d_chunk0, sizeof(int32_t) * (n_cases + 1)))) return r;
    c->chunk0_valid = false;
    if ((r = ensure(c, c->d_ctl, sizeof(unsigned long long) * 16))) return r;
    HIPCHK(c, hipMemcpyAsync(c->d_cases.p, dc.data(), sizeof(DevCase) * n_cases, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_bands.p, db.data(), sizeof(DevBand) * n_bands, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->cases.assign(cases, cases + n_cases);
    c->bands.assign(bands, bands + n_bands);
    c->dcases.swap(dc);
    c->dbands.swap(db);
    c->total_photons = total;
    c->cases_set = true;
    c->polar = false;
    c->gkey = 0;
    return MCS_OK;
}

int mcs_set_polar(mcs_ctx *c, const mcs_polar_case *pc, int32_t n_cases, const double *tables,
                  int32_t n_tables, int32_t n_angles)
{
    if (!c) return MCS_ERR_ARG;
    if (!c->cases_set) return set_err(c, MCS_ERR_STATE, "mcs_set_cases must come first");
    const int nc = (int)c->cases.size();
    std::vector<DevCase> dc = c->dcases;
    std::vector<DevBand> db = c->dbands;
    for (auto &D : dc) {
        D.polar = 0;
        D.s0[0] = 1.0; D.s0[1] = D.s0[2] = D.s0[3] = 0.0;
        D.s0f[0] = 1.f; D.s0f[1] = D.s0f[2] = D.s0f[3] = 0.f;
    }
    for (auto &D : db) D.pm = -1;
    bool any = false;
    const int n = n_angles;
    if (pc) {
        if (n_cases != nc) return set_err(c, MCS_ERR_ARG, "n_cases = %d, %d cases are set", n_cases, nc);
        if (n_tables < 0 || (n_tables > 0 && (!tables || n < 2)))
            return set_err(c, MCS_ERR_ARG, "tables: need n_tables >= 0 and n_angles >= 2");
        if ((int64_t)n_tables * 7 * n > ((int64_t)1 << 31))
            return set_err(c, MCS_ERR_ARG, "phase-matrix tables too large");
        for (int t = 0; t < n_tables; ++t) {
            const double *x = tables + (size_t)t * 7 * n;
            for (int k = 0; k < 7 * n; ++k)
                if (!std::isfinite(x[k])) return set_err(c, MCS_ERR_ARG, "table %d: non-finite value", t);
            for (int k = 0; k < n; ++k) {
                if (x[k] < -1.0 || x[k] > 1.0) return set_err(c, MCS_ERR_ARG, "table %d: cos theta outside [-1, 1]", t);
                if (k && !(x[k] > x[k - 1])) return set_err(c, MCS_ERR_ARG, "table %d: cos theta not ascending", t);
            }
        }
        for (int k = 0; k < nc; ++k) {
            const mcs_polar_case &Q = pc[k];
            if (Q.table < 0) continue;
            DevCase &D = dc[k];
            if (Q.table + D.n_bands > n_tables)
                return set_err(c, MCS_ERR_ARG, "case %d: tables [%d, %d) beyond %d", k, Q.table, Q.table + D.n_bands, n_tables);
            if (Q.cutoff < 0 || Q.cutoff >= n) return set_err(c, MCS_ERR_ARG, "case %d: cutoff outside the grid", k);
            for (int q = 0; q < 4; ++q)
                if (!std::isfinite(Q.stokes0[q])) return set_err(c, MCS_ERR_ARG, "case %d: non-finite Stokes vector", k);
            if (!(Q.stokes0[0] != 0.0)) return set_err(c, MCS_ERR_ARG, "case %d: Stokes I must be non-zero", k);
            const double *x0 = tables + (size_t)Q.table * 7 * n;
            const int jc = n - 1 - Q.cutoff;   // the cutoff in ascending order
            D.polar = 1;
            for (int q = 0; q < 4; ++q) { D.s0[q] = Q.stokes0[q]; D.s0f[q] = (float)Q.stokes0[q]; }
            D.cmax = x0[n - 1]; D.cmin = x0[0]; D.ccut = x0[jc];
            D.fcmax = (float)D.cmax; D.fcmin = (float)D.cmin; D.fccut = (float)D.ccut;
            for (int b = 0; b < D.n_bands; ++b) {
                const int t = Q.table + b;
                const double *x = tables + (size_t)t * 7 * n;
                for (int k2 = 0; k2 < n; ++k2)
                    if (x[k2] != x0[k2]) return set_err(c, MCS_ERR_ARG, "case %d: its tables have different grids", k);
                DevBand &B = db[D.band_begin + b];
                if (B.pm >= 0 && B.pm != t)
                    return set_err(c, MCS_ERR_ARG, "band row %d: two phase-matrix tables", D.band_begin + b);
                B.pm = t;
                // |P11|, |P12| maxima over theta <= theta_cut (area 1) and >= (area 2)
                double a1 = 0, b1 = 0, a2 = 0, b2 = 0;
                for (int k2 = 0; k2 < n; ++k2) {
                    const double p11 = std::fabs(x[n + k2]), p12 = std::fabs(x[2 * n + k2]);
                    if (k2 >= jc) { a1 = std::max(a1, p11); b1 = std::max(b1, p12); }
                    if (k2 <= jc) { a2 = std::max(a2, p11); b2 = std::max(b2, p12); }
                }
                if (!(a1 > 0) || !(a2 > 0)) return set_err(c, MCS_ERR_ARG, "table %d: P11 must be positive", t);
                B.pA1 = a1; B.pB1 = b1; B.pA2 = a2; B.pB2 = b2;
                B.fA1 = (float)a1; B.fB1 = (float)b1; B.fA2 = (float)a2; B.fB2 = (float)b2;
                bool exact = true;
                double mass = 0.0;
                for (int k2 = 0; k2 < n; ++k2) {
                    if (!(x[n + k2] >= 0.0) || std::fabs(x[2 * n + k2]) > x[n + k2]) exact = false;
                    if (k2 + 1 < n) mass += 0.5 * (x[n + k2] + x[n + k2 + 1]) * (x[k2 + 1] - x[k2]);
                }
                B.pm_exact = (exact && mass > 0.0) ? 1 : 0;
                B.fM = (float)mass;
            }
            any = true;
        }
    }
    (void)hipSetDevice(c->device);
    if (c->pending) (void)hipStreamSynchronize(c->stream), c->pending = false;
    int r;
    if (any) {
        const size_t nv = (size_t)n_tables * 7 * n;
        const size_t stride = (size_t)14 * n + PM_GUIDE + 1;
        std::vector<float> tf((size_t)n_tables * stride, 0.f);
        std::vector<double> cum(n);
        for (int t = 0; t < n_tables; ++t) {
            const double *x = tables + (size_t)t * 7 * n;
            float *f = tf.data() + (size_t)t * stride;
            for (int k = 0; k < 7 * n; ++k) f[k] = (float)x[k];
            cum[0] = 0.0;   // integral of the linear P11 from x[0]
            for (int k = 0; k + 1 < n; ++k) cum[k + 1] = cum[k] + 0.5 * (x[n + k] + x[n + k + 1]) * (x[k + 1] - x[k]);
            for (int k = 0; k < n; ++k) f[7 * n + k] = (float)cum[k];
            for (int e = 0; e < 6; ++e)
                for (int k = 0; k + 1 < n; ++k)
                    f[(8 + e) * n + k] = (float)((x[(1 + e) * n + k + 1] - x[(1 + e) * n + k]) / (x[k + 1] - x[k]));
            int *guide = reinterpret_cast<int *>(f + 14 * n);
            int j = 0;
            for (int g = 0; g <= PM_GUIDE; ++g) {
                const double target = cum[n - 1] * g / PM_GUIDE;
                while (j < n - 2 && cum[j + 1] <= target) ++j;
                guide[g] = j;
            }
            guide[PM_GUIDE] = n - 2;
        }
        if ((r = ensure(c, c->d_pm, sizeof(double) * nv))) return r;
        if ((r = ensure(c, c->d_pmf, sizeof(float) * tf.size()))) return r;
        HIPCHK(c, hipMemcpyAsync(c->d_pm.p, tables, sizeof(double) * nv, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_pmf.p, tf.data(), sizeof(float) * tf.size(), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(c->d_cases.p, dc.data(), sizeof(DevCase) * nc, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_bands.p, db.data(), sizeof(DevBand) * db.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->dcases.swap(dc);
    c->dbands.swap(db);
    c->polar = any;
    c->pm_n = any ? n : 0;
    c->pm_floats = any ? (size_t)n_tables * ((size_t)14 * n + PM_GUIDE + 1) : 0;
    c->gkey = 0;
    return MCS_OK;
}

int mcs_set_launch(mcs_ctx *c, const mcs_launch *l)
{
    if (!c || !l) return MCS_ERR_ARG;
    if (l->block < 64 || l->block > 1024 || l->block % 64 || l->blocks_per_cu < 1 || l->blocks_per_cu > 64 ||
        l->chunk < 64 || l->chunk > (1 << 24) || l->refill_min < 1 || l->refill_min > 64 || l->coop_live < 0 || l->coop_live > 32)
        return set_err(c, MCS_ERR_ARG, "launch shape out of range");
    c->launch = *l;
    c->gkey = 0;
    return MCS_OK;
}

int mcs_get_launch(const mcs_ctx *c, mcs_launch *out)
{
    if (!c || !out) return MCS_ERR_ARG;
    *out = c->launch;
    return MCS_OK;
}

static int finish_walk(mcs_ctx *c);

static uint64_t mix64(uint64_t h, uint64_t v)
{
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    return h;
}

// enqueue fill + walk + scatter on the context stream
static int enqueue_walk(mcs_ctx *c, int mode, const KParams &P, uint32_t n_chunks, int64_t n, bool records,
                        bool events)
{
    const int nc = (int)c->cases.size();
    const uint64_t nclear = (uint64_t)(TF + TU) * nc + 16;
    const unsigned blocks_clear = (unsigned)std::min<uint64_t>((nclear + 255) / 256, 1024);
    hipLaunchKernelGGL(k_clear, dim3(blocks_clear), dim3(256), 0, c->stream,
                       (unsigned long long *)c->d_tally_f.p, (uint64_t)TF * nc,
                       (unsigned long long *)c->d_tally_u.p, (uint64_t)TU * nc,
                       (unsigned long long *)c->d_ctl.p, (uint64_t)16);
    int block = (mode == MCS_REPLAY) ? std::min(c->launch.block, ReplayWalker::MAX_BLOCK)
                                     : c->launch.block;
    int grid = c->launch.blocks_per_cu * c->cus;
    if (mode == MCS_NATIVE && c->polar && block > PolarWalker::MAX_BLOCK) {
        grid *= block / PolarWalker::MAX_BLOCK;   // same threads per CU in smaller blocks
        block = PolarWalker::MAX_BLOCK;
    }
    if ((uint32_t)grid > n_chunks) grid = (int)std::max<uint32_t>(n_chunks, 1u);
    // events bracket the walk kernel alone: kernel_ms is its device time
    if (events && hipEventRecord(c->ev0, c->stream) != hipSuccess)
        return set_err(c, MCS_ERR_HIP, "hipEventRecord failed");
    if (mode == MCS_NATIVE && c->polar) {
        // aspherical particles: one photon per lane, then the tail mode of
        // polar_tail; tables up to 16 KB ride in LDS
        // + the tail mode's owner lanes and packets (16 ints and 64 packets a wave)
        const size_t lds = (size_t)P.pm_lds_floats * sizeof(float) +
                           (size_t)(block / 64) * (16 * sizeof(int) + 64 * sizeof(PolarPacket));
        hipLaunchKernelGGL((k_walk<PolarWalker, 3>), dim3(grid), dim3(block), lds, c->stream, P);
    } else if (mode == MCS_NATIVE) {
        const size_t tab = (size_t)(block / 64) * c->launch.coop_live * (sizeof(CoopEntry) + sizeof(int));
        // shapes asking for more than 5 waves per SIMD get the 80-VGPR instance
        // (6 waves per SIMD; its spills sit outside the step loop)
        const int waves_per_simd = (c->launch.blocks_per_cu * (block / 64) + 3) / 4;
        if (waves_per_simd > 5)
            hipLaunchKernelGGL((k_walk<NativeWalker, 6>), dim3(grid), dim3(block), tab, c->stream, P);
        else
            hipLaunchKernelGGL((k_walk<NativeWalker, 5>), dim3(grid), dim3(block), tab, c->stream, P);
    } else {
        hipLaunchKernelGGL((k_walk<ReplayWalker, ReplayWalker::WAVES_PER_EU>), dim3(grid), dim3(block), 0,
                           c->stream, P);
    }
    if (events && hipEventRecord(c->ev1, c->stream) != hipSuccess)
        return set_err(c, MCS_ERR_HIP, "hipEventRecord failed");
    const size_t rsz = (mode == MCS_NATIVE) ? sizeof(float) : sizeof(double);
    const size_t ord_lds = (size_t)c->launch.chunk * (8 + 3 * rsz);
    if (records && n > 0 && ord_lds <= 65536) {
        const unsigned g2 = std::min<uint32_t>(n_chunks, 4096u);
        if (mode == MCS_NATIVE)
            hipLaunchKernelGGL(k_order_records<float>, dim3(g2), dim3(256), ord_lds, c->stream,
                               P.cases, P.case_chunk0, nc, c->launch.chunk, n_chunks,
                               (uint32_t *)c->d_rec[0].p, (uint32_t *)c->d_rec[1].p, (uint32_t *)c->d_rec[2].p,
                               (float *)c->d_rec[3].p, (float *)c->d_rec[4].p, (float *)c->d_rec[5].p,
                               (uint32_t *)c->d_out[0].p, (uint32_t *)c->d_out[1].p, (float *)c->d_out[2].p,
                               (float *)c->d_out[3].p, (float *)c->d_out[4].p);
        else
            hipLaunchKernelGGL(k_order_records<double>, dim3(g2), dim3(256), ord_lds, c->stream,
                               P.cases, P.case_chunk0, nc, c->launch.chunk, n_chunks,
                               (uint32_t *)c->d_rec[0].p, (uint32_t *)c->d_rec[1].p, (uint32_t *)c->d_rec[2].p,
                               (double *)c->d_rec[3].p, (double *)c->d_rec[4].p, (double *)c->d_rec[5].p,
                               (uint32_t *)c->d_out[0].p, (uint32_t *)c->d_out[1].p, (double *)c->d_out[2].p,
                               (double *)c->d_out[3].p, (double *)c->d_out[4].p);
    } else if (records && n > 0) {
        // chunks too large for LDS: scatter record by record
        const unsigned g2 = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
        if (mode == MCS_NATIVE)
            hipLaunchKernelGGL(k_scatter_records<float>, dim3(g2), dim3(256), 0, c->stream,
                               (uint32_t *)c->d_rec[0].p, (uint32_t *)c->d_rec[1].p, (uint32_t *)c->d_rec[2].p,
                               (float *)c->d_rec[3].p, (float *)c->d_rec[4].p, (float *)c->d_rec[5].p, (uint64_t)n,
                               (uint32_t *)c->d_out[0].p, (uint32_t *)c->d_out[1].p, (float *)c->d_out[2].p,
                               (float *)c->d_out[3].p, (float *)c->d_out[4].p);
        else
            hipLaunchKernelGGL(k_scatter_records<double>, dim3(g2), dim3(256), 0, c->stream,
                               (uint32_t *)c->d_rec[0].p, (uint32_t *)c->d_rec[1].p, (uint32_t *)c->d_rec[2].p,
                               (double *)c->d_rec[3].p, (double *)c->d_rec[4].p, (double *)c->d_rec[5].p, (uint64_t)n,
                               (uint32_t *)c->d_out[0].p, (uint32_t *)c->d_out[1].p, (double *)c->d_out[2].p,
                               (double *)c->d_out[3].p, (double *)c->d_out[4].p);
    }
    hipLaunchKernelGGL(k_summary, dim3(1), dim3(256), 0, c->stream, (const unsigned long long *)c->d_tally_u.p, nc,
                       (const unsigned long long *)c->d_ctl.p, c->d_sum);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(c, MCS_ERR_HIP, "kernel launch: %s", hipGetErrorString(e));
    return MCS_OK;
}

int mcs_walk(mcs_ctx *c, const mcs_walk_args *a)
{
    if (!c || !a) return MCS_ERR_ARG;
    if (!c->cases_set) return set_err(c, MCS_ERR_STATE, "mcs_set_cases must come first");
    if (a->mode != MCS_NATIVE && a->mode != MCS_REPLAY) return set_err(c, MCS_ERR_ARG, "unknown mode %d", a->mode);
    if (a->max_steps < 0 || a->max_steps > 0xFFFFFFFFll) return set_err(c, MCS_ERR_ARG, "max_steps outside [0, 2^32)");
    (void)hipSetDevice(c->device);
    if (c->pending) {
        int r = finish_walk(c);
        if (r) return r;
    }
    const int64_t n = c->total_photons;
    const bool records = a->want_records != 0;
    if (records && n > 0xFFFFFFFFll) return set_err(c, MCS_ERR_ARG, "records need fewer than 2^32 photons");
    if (a->mode == MCS_REPLAY) {
        if (!a->replay_dev || !a->replay_off || !a->photon_band)
            return set_err(c, MCS_ERR_ARG, "replay needs replay_dev, replay_off and photon_band");
    }
    const int nc = (int)c->cases.size();
    // photon bands: validate against each case's band count
    if (a->photon_band) {
        for (int k = 0; k < nc; ++k) {
            const DevCase &D = c->dcases[k];
            for (int64_t j = 0; j < D.n_photons; ++j)
                if (a->photon_band[D.gid_base + j] >= D.n_bands)
                    return set_err(c, MCS_ERR_ARG, "photon %lld: band %d >= n_bands %d of case %d",
                                   (long long)(D.gid_base + j), a->photon_band[D.gid_base + j], D.n_bands, k);
        }
    }
    int64_t rlen = 0;
    if (a->mode == MCS_REPLAY) {
        if (a->replay_off[0] != 0) return set_err(c, MCS_ERR_ARG, "replay_off[0] must be 0");
        for (int64_t j = 0; j < n; ++j)
            if (a->replay_off[j + 1] < a->replay_off[j] + 3)
                return set_err(c, MCS_ERR_ARG, "replay slice %lld shorter than the initial triple", (long long)j);
        rlen = a->replay_off[n];
    }
    // chunking
    const int chunk = c->launch.chunk;
    std::vector<int32_t> chunk0(nc + 1);
    int64_t nch = 0;
    for (int k = 0; k < nc; ++k) {
        chunk0[k] = (int32_t)nch;
        nch += (c->dcases[k].n_photons + chunk - 1) / chunk;
        if (nch > 0x7FFFFFFF) return set_err(c, MCS_ERR_ARG, "too many chunks; raise the chunk size");
    }
    chunk0[nc] = (int32_t)nch;
    int r;
    if (!c->chunk0_valid || c->chunk0 != chunk0) {
        c->chunk0 = chunk0;
        c->chunk0_valid = false;
        HIPCHK(c, hipMemcpyAsync(c->d_chunk0.p, c->chunk0.data(), sizeof(int32_t) * (nc + 1), hipMemcpyHostToDevice,
                                 c->stream));
        c->chunk0_valid = true;
    }
    if (a->photon_band) {
        if ((r = ensure(c, c->d_pband, (size_t)n))) return r;
        HIPCHK(c, hipMemcpyAsync(c->d_pband.p, a->photon_band, (size_t)n, hipMemcpyHostToDevice, c->stream));
    }
    if (a->mode == MCS_REPLAY) {
        if ((r = ensure(c, c->d_rdev, sizeof(double) * (size_t)rlen))) return r;
        if ((r = ensure(c, c->d_roff, sizeof(int64_t) * (size_t)(n + 1)))) return r;
        HIPCHK(c, hipMemcpyAsync(c->d_rdev.p, a->replay_dev, sizeof(double) * (size_t)rlen, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_roff.p, a->replay_off, sizeof(int64_t) * (size_t)(n + 1), hipMemcpyHostToDevice, c->stream));
    }
    const size_t rsz = (a->mode == MCS_NATIVE) ? 4 : 8;
    if (records) {
        for (int k = 0; k < 3; ++k) if ((r = ensure(c, c->d_rec[k], 4 * (size_t)n))) return r;
        for (int k = 3; k < 6; ++k) if ((r = ensure(c, c->d_rec[k], rsz * (size_t)n))) return r;
        for (int k = 0; k < 2; ++k) if ((r = ensure(c, c->d_out[k], 4 * (size_t)n))) return r;
        for (int k = 2; k < 5; ++k) if ((r = ensure(c, c->d_out[k], rsz * (size_t)n))) return r;
    }
    KParams P;
    memset(&P, 0, sizeof P);
    P.cases = (const DevCase *)c->d_cases.p;
    P.bands = (const DevBand *)c->d_bands.p;
    P.case_chunk0 = (const int32_t *)c->d_chunk0.p;
    P.n_cases = nc;
    P.chunk = chunk;
    P.n_chunks = (uint32_t)nch;
    P.refill_min = (uint32_t)c->launch.refill_min;
    P.coop_live = c->launch.coop_live;
    P.key0 = (uint32_t)a->seed;
    P.key1 = (uint32_t)(a->seed >> 32);
    P.max_steps = a->max_steps ? (uint32_t)a->max_steps : (1u << 24);
    P.photon_band = a->photon_band ? (const uint8_t *)c->d_pband.p : nullptr;
    P.rdev = (const double *)c->d_rdev.p;
    P.roff = (const int64_t *)c->d_roff.p;
    P.tally_f = (double *)c->d_tally_f.p;
    P.tally_u = (unsigned long long *)c->d_tally_u.p;
    P.ctl = (unsigned long long *)c->d_ctl.p;
    P.chunk_counter = (unsigned int *)((unsigned long long *)c->d_ctl.p + 8);
    P.pm = (const double *)c->d_pm.p;
    P.pmf = (const float *)c->d_pmf.p;
    P.pm_n = c->pm_n;
    P.pm_stride = c->pm_n * 14 + PM_GUIDE + 1;
    P.pm_lds_floats = (c->polar && c->pm_floats * sizeof(float) <= 16 * 1024) ? (int32_t)c->pm_floats : 0;
    P.n_total = (uint64_t)n;
    if (records && c->polar) {
        if ((r = ensure(c, c->d_stokes, 3 * rsz * (size_t)std::max<int64_t>(n, 1)))) return r;
        P.rec_stokes = c->d_stokes.p;
    }
    if (records) {
        P.rec_gid = (uint32_t *)c->d_rec[0].p;
        P.rec_nscat = (uint32_t *)c->d_rec[1].p;
        P.rec_cb = (uint32_t *)c->d_rec[2].p;
        P.rec_theta = c->d_rec[3].p;
        P.rec_phi = c->d_rec[4].p;
        P.rec_path = c->d_rec[5].p;
    }
    c->h_sum[7] = 0ull;   // set again by this walk's k_summary
    if (a->use_graph) {
        uint64_t key = 0x1234;
        const uint64_t *w = (const uint64_t *)&P;
        for (size_t k = 0; k < sizeof(P) / 8; ++k) key = mix64(key, w[k]);
        key = mix64(key, (uint64_t)a->mode);
        key = mix64(key, (uint64_t)records);
        key = mix64(key, (uint64_t)n);
        key = mix64(key, (uint64_t)c->launch.block * 131 + c->launch.blocks_per_cu);
        if (!c->gexec || c->gkey != key) {
            if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
            if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int er = enqueue_walk(c, a->mode, P, (uint32_t)nch, n, records, false);
            hipGraph_t g = nullptr;
            hipError_t ee = hipStreamEndCapture(c->stream, &g);
            if (er) { if (g) (void)hipGraphDestroy(g); return er; }
            if (ee != hipSuccess) return set_err(c, MCS_ERR_HIP, "graph capture: %s", hipGetErrorString(ee));
            c->graph = g;
            HIPCHK(c, hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
            c->gkey = key;
        }
        // a graph launch is timed as a whole (fills + walk + record scatter)
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    } else {
        if ((r = enqueue_walk(c, a->mode, P, (uint32_t)nch, n, records, true))) return r;
    }
    c->last_mode = a->mode;
    c->last_records = records;
    c->last_stokes = records && c->polar;
    c->last_n = n;
    c->pending = true;
    if (a->async) return MCS_OK;
    return finish_walk(c);
}

static int finish_walk(mcs_ctx *c)
{
    c->pending = false;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    // k_summary wrote the sums into pinned host memory before the stream drained
    const volatile unsigned long long *h = c->h_sum;
    if (h[7] != 1ull) return set_err(c, MCS_ERR_HIP, "internal: walk summary not written");
    unsigned long long ctl[3] = {h[4], h[5], h[6]};
    mcs_stats s{};
    s.photons = h[0];
    s.steps = h[1];
    s.capped = h[2];
    s.phi_unbound = h[3];
    s.kernel_ms = ms;
    c->stats = s;
    if (ctl[1])
        return set_err(c, MCS_ERR_HIP, "internal: %llu failed photon hand-overs between waves",
                       (unsigned long long)ctl[1]);
    if (ctl[2])
        return set_err(c, MCS_ERR_ARG, "%llu scattering angles not accepted in %u phase-matrix rejection tries "
                       "(P11 + P12 terms not positive?)", (unsigned long long)ctl[2], POLAR_MAX_TRIES);
    if (s.photons != (uint64_t)c->last_n)
        return set_err(c, MCS_ERR_HIP, "walk finished %llu of %lld photons",
                       (unsigned long long)s.photons, (long long)c->last_n);
    if (c->last_mode == MCS_REPLAY && ctl[0])
        return set_err(c, MCS_ERR_REPLAY, "%llu photons did not consume exactly their replay slice",
                       (unsigned long long)ctl[0]);
    if (s.capped) return set_err(c, MCS_ERR_STEP_CAP, "%llu photons reached max_steps", (unsigned long long)s.capped);
    if (c->last_mode == MCS_REPLAY && s.phi_unbound)
        return set_err(c, MCS_ERR_PHI_UNBOUND,
                       "%llu photons end on an axis: the reference leaves phi_n unbound there",
                       (unsigned long long)s.phi_unbound);
    return MCS_OK;
}

int mcs_sync(mcs_ctx *c)
{
    if (!c) return MCS_ERR_ARG;
    if (!c->pending) return MCS_OK;
    (void)hipSetDevice(c->device);
    return finish_walk(c);
}

int mcs_read_records(mcs_ctx *c, const mcs_records *o, int64_t n)
{
    if (!c || !o) return MCS_ERR_ARG;
    if (c->pending) return set_err(c, MCS_ERR_STATE, "walk still pending: call mcs_sync");
    if (!c->last_records) return set_err(c, MCS_ERR_STATE, "last walk kept no records");
    if (n != c->last_n) return set_err(c, MCS_ERR_ARG, "n = %lld, last walk had %lld photons", (long long)n, (long long)c->last_n);
    (void)hipSetDevice(c->device);
    std::vector<uint32_t> nscat(n), cb(n);
    HIPCHK(c, hipMemcpy(nscat.data(), c->d_out[0].p, 4 * (size_t)n, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(cb.data(), c->d_out[1].p, 4 * (size_t)n, hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < n; ++j) {
        if (o->n_scat) o->n_scat[j] = nscat[j];
        if (o->condition) o->condition[j] = (int8_t)(cb[j] & 0xff);
        if (o->band) o->band[j] = (uint8_t)(cb[j] >> 8);
    }
    double *dst[3] = {o->theta_n, o->phi_n, o->path_length};
    if (c->last_mode == MCS_REPLAY) {
        for (int k = 0; k < 3; ++k)
            if (dst[k]) HIPCHK(c, hipMemcpy(dst[k], c->d_out[2 + k].p, 8 * (size_t)n, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(n);
        for (int k = 0; k < 3; ++k) {
            if (!dst[k]) continue;
            HIPCHK(c, hipMemcpy(tmp.data(), c->d_out[2 + k].p, 4 * (size_t)n, hipMemcpyDeviceToHost));
            for (int64_t j = 0; j < n; ++j) dst[k][j] = tmp[j];
        }
    }
    return MCS_OK;
}

int mcs_read_stokes(mcs_ctx *c, double *q, double *u, double *v, int64_t n)
{
    if (!c) return MCS_ERR_ARG;
    if (c->pending) return set_err(c, MCS_ERR_STATE, "walk still pending: call mcs_sync");
    if (!c->last_stokes) return set_err(c, MCS_ERR_STATE, "last walk kept no Stokes vectors (records off or no polar cases)");
    if (n != c->last_n) return set_err(c, MCS_ERR_ARG, "n = %lld, last walk had %lld photons", (long long)n, (long long)c->last_n);
    (void)hipSetDevice(c->device);
    double *dst[3] = {q, u, v};
    if (c->last_mode == MCS_REPLAY) {
        for (int k = 0; k < 3; ++k)
            if (dst[k] && n) HIPCHK(c, hipMemcpy(dst[k], (double *)c->d_stokes.p + (size_t)k * n, 8 * (size_t)n, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(n);
        for (int k = 0; k < 3; ++k) {
            if (!dst[k] || !n) continue;
            HIPCHK(c, hipMemcpy(tmp.data(), (float *)c->d_stokes.p + (size_t)k * n, 4 * (size_t)n, hipMemcpyDeviceToHost));
            for (int64_t j = 0; j < n; ++j) dst[k][j] = tmp[j];
        }
    }
    return MCS_OK;
}

int mcs_read_tallies(mcs_ctx *c, double *out, int32_t n_cases)
{
    if (!c || !out) return MCS_ERR_ARG;
    if (c->pending) return set_err(c, MCS_ERR_STATE, "walk still pending: call mcs_sync");
    if (n_cases != (int32_t)c->cases.size()) return set_err(c, MCS_ERR_ARG, "n_cases mismatch");
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipMemcpy(out, c->d_tally_f.p, sizeof(double) * TF * n_cases, hipMemcpyDeviceToHost));
    return MCS_OK;
}

int mcs_read_counts(mcs_ctx *c, uint64_t *out, int32_t n_cases)
{
    if (!c || !out) return MCS_ERR_ARG;
    if (c->pending) return set_err(c, MCS_ERR_STATE, "walk still pending: call mcs_sync");
    if (n_cases != (int32_t)c->cases.size()) return set_err(c, MCS_ERR_ARG, "n_cases mismatch");
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipMemcpy(out, c->d_tally_u.p, sizeof(uint64_t) * TU * n_cases, hipMemcpyDeviceToHost));
    return MCS_OK;
}

int mcs_last_stats(const mcs_ctx *c, mcs_stats *out)
{
    if (!c || !out) return MCS_ERR_ARG;
    *out = c->stats;
    return MCS_OK;
}

// Put back the phase-matrix state of the cases (mcs_set_cases clears it):
// the polar fields of the case and band rows, re-uploaded; the device tables
// themselves are untouched by mcs_set_cases.
static int restore_polar(mcs_ctx *c, const std::vector<DevCase> &dc, const std::vector<DevBand> &db, bool polar,
                         int32_t pm_n, size_t pm_floats)
{
    if (!polar) return MCS_OK;
    for (size_t k = 0; k < c->dcases.size(); ++k) {
        DevCase &D = c->dcases[k];
        const DevCase &S = dc[k];
        D.polar = S.polar;
        for (int q = 0; q < 4; ++q) { D.s0[q] = S.s0[q]; D.s0f[q] = S.s0f[q]; }
        D.cmax = S.cmax; D.ccut = S.ccut; D.cmin = S.cmin;
        D.fcmax = S.fcmax; D.fccut = S.fccut; D.fcmin = S.fcmin;
    }
    for (size_t k = 0; k < c->dbands.size(); ++k) {
        DevBand &B = c->dbands[k];
        const DevBand &S = db[k];
        B.pm = S.pm; B.pm_exact = S.pm_exact; B.fM = S.fM;
        B.pA1 = S.pA1; B.pB1 = S.pB1; B.pA2 = S.pA2; B.pB2 = S.pB2;
        B.fA1 = S.fA1; B.fB1 = S.fB1; B.fA2 = S.fA2; B.fB2 = S.fB2;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_cases.p, c->dcases.data(), sizeof(DevCase) * c->dcases.size(), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_bands.p, c->dbands.data(), sizeof(DevBand) * c->dbands.size(), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->polar = true;
    c->pm_n = pm_n;
    c->pm_floats = pm_floats;
    return MCS_OK;
}

int mcs_autotune(mcs_ctx *c, const mcs_launch *cand, int32_t n, int64_t photons_cap, int32_t repeats,
                 uint64_t seed, double *times_ms, mcs_launch *best)
{
    if (!c || !cand || n <= 0 || repeats <= 0 || !best) return MCS_ERR_ARG;
    if (!c->cases_set) return set_err(c, MCS_ERR_STATE, "mcs_set_cases must come first");
    // a capped copy of the current cases
    std::vector<mcs_case> saved = c->cases;
    std::vector<mcs_band> bands = c->bands;
    std::vector<mcs_case> capped = saved;
    if (photons_cap > 0)
        for (auto &k : capped) k.n_photons = std::min<int64_t>(k.n_photons, photons_cap);
    const mcs_launch keep = c->launch;
    // the phase-matrix state survives the capped copy (aspherical cases are
    // timed with their own kernel)
    const std::vector<DevCase> pdc = c->dcases;
    const std::vector<DevBand> pdb = c->dbands;
    const bool polar = c->polar;
    const int32_t pm_n = c->pm_n;
    const size_t pm_floats = c->pm_floats;
    int r = mcs_set_cases(c, capped.data(), (int32_t)capped.size(), bands.data(), (int32_t)bands.size());
    if (r) return r;
    if ((r = restore_polar(c, pdc, pdb, polar, pm_n, pm_floats))) return r;
    double best_t = 1e300;
    int best_i = -1;
    for (int k = 0; k < n; ++k) {
        if ((r = mcs_set_launch(c, &cand[k]))) break;
        mcs_walk_args a;
        memset(&a, 0, sizeof a);
        a.mode = MCS_NATIVE;
        a.seed = seed;
        if ((r = mcs_walk(c, &a))) break;  // warm-up
        // the same seeds for every candidate, a different walk for every repeat
        std::vector<double> t;
        for (int q = 0; q < repeats; ++q) {
            a.seed = seed + 1 + (uint64_t)q;
            if ((r = mcs_walk(c, &a))) break;
            t.push_back(c->stats.kernel_ms);
        }
        if (r) break;
        std::sort(t.begin(), t.end());
        const double med = t[t.size() / 2];
        if (times_ms) times_ms[k] = med;
        if (med < best_t) { best_t = med; best_i = k; }
    }
    int r2 = mcs_set_cases(c, saved.data(), (int32_t)saved.size(), bands.data(), (int32_t)bands.size());
    if (!r2) r2 = restore_polar(c, pdc, pdb, polar, pm_n, pm_floats);
    if (r) { c->launch = keep; return r; }
    if (r2) return r2;
    *best = cand[best_i];
    c->launch = cand[best_i];
    return MCS_OK;
}

// Diagnostic: Philox4x32-10 of n counters (4 x u32 each) with key (k0, k1),
// computed on the device (checks the device generator against a host one).
int mcs_philox(mcs_ctx *c, const uint32_t *ctr, uint32_t k0, uint32_t k1, uint32_t *out, int64_t n)
{
    if (!c || !ctr || !out || n <= 0) return MCS_ERR_ARG;
    (void)hipSetDevice(c->device);
    void *dc = nullptr, *dout = nullptr;
    HIPCHK(c, hipMalloc(&dc, 16 * (size_t)n));
    hipError_t e = hipMalloc(&dout, 16 * (size_t)n);
    if (e != hipSuccess) { (void)hipFree(dc); return set_err(c, MCS_ERR_NOMEM, "hipMalloc"); }
    e = hipMemcpy(dc, ctr, 16 * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_philox, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                           (const uint32_t *)dc, k0, k1, (uint32_t *)dout, (uint64_t)n);
        e = hipStreamSynchronize(c->stream);
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, 16 * (size_t)n, hipMemcpyDeviceToHost);
    (void)hipFree(dc);
    (void)hipFree(dout);
    if (e != hipSuccess) return set_err(c, MCS_ERR_HIP, "mcs_philox: %s", hipGetErrorString(e));
    return MCS_OK;
}

}  // extern "C"
