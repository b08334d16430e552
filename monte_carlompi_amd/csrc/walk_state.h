// walk_state.h -- constants, case / band tables, kernel parameters, LDS block state and wave helpers of the photon walk
// (part of walk.hip's translation unit; included from there in this order)
#pragma once

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
constexpr int U_N = 0, U_CAP = 6, U_AXIS = 7, U_STEPS = 8, U_COOP = 9, U_MOVED = 10, U_REJ = 11;
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
    // the initial Stokes vector is physical (I > 0, Q^2 + U^2 + V^2 <= I^2):
    // required by the native walk's exact sampler (see DevBand::pm_exact)
    int32_t s0_exact;
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
    // + U sin 2phi) >= 0 for |(Q, U)| <= I and its cos(theta) marginal is I S11:
    // the walk samples it exactly by inversion (fM = integral of S11).  The
    // matrix must also keep every physical Stokes vector physical (degree of
    // polarisation <= 1 after scattering), so that |(Q, U)| <= I holds at
    // every later scattering; photons of a case whose initial vector is not
    // physical use the two-area rejection (DevCase::s0_exact)
    int32_t pm_exact;
    float fM;
    double snow_depth;                   // record column of the band's photons [m]
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
    int32_t rec_direct;          // 1: records go straight to photon order (pos = gid), no compaction
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

template <int S>
struct BlockStateT {
    static constexpr int NSLOT = S;
    static_assert((S & (S - 1)) == 0, "slots: a power of two");
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
    SlotHdr hdr[S];
    SlotTally tal[S];
};
// the HG walk's block state (the polarised walk in blocks of up to 256
// threads runs with two slots: k_walk's SL)
using BlockState = BlockStateT<SLOTS>;

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


}  // namespace mcs
