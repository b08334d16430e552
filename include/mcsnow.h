/*
 * mcsnow.h -- C ABI of the MI355X (gfx950) photon random-walk library
 * (libmcsnow.so, built from monte_carlompi_amd/csrc/).
 *
 * Drop-in boundary.  The reference has no FFI; its seam is the Python region
 *     par_wvls = Parallel(wvls)                                  mc3D:1521
 *     for i, wvl in enumerate(par_wvls.working_set):             mc3D:1613-1616
 *         self.photon = i; answer.append(self.monte_carlo3D(wvl))
 *     all_answers = par_wvls.answer_and_reduce(answer, ...)      mc3D:1618-1619
 * with the mpi4py scatter/gather of monte_carloMPI/parallelize.py:14-38
 * (mc3D = monte_carloMPI/monte_carlo3D.py).  These entry points replace that
 * region: the host uploads the per-photon parameter tables (the scatter),
 * one call walks every photon on the GPU (the per-rank photon loop), and the
 * records come back in photon order (the gather + flatten_list, mc3D:1845).
 *
 * Conventions: plain C types and pointers only; host buffers are owned by
 * the caller; device buffers by the context.  Every function returns an
 * int status (MCS_OK or a negative MCS_ERR_*); mcs_last_error() gives the
 * text.  No function calls exit() or abort().  A context binds one HIP
 * device and one stream; calls on one context must not overlap.
 */
#ifndef MCSNOW_H
#define MCSNOW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCS_OK 0
#define MCS_ERR_ARG (-1)        /* invalid argument (NULL, size, NaN, range) */
#define MCS_ERR_HIP (-2)        /* HIP runtime error                          */
#define MCS_ERR_NOMEM (-3)      /* device allocation failed                   */
#define MCS_ERR_STEP_CAP (-4)   /* a photon reached max_steps                 */
#define MCS_ERR_REPLAY (-5)     /* replay slice too short or not consumed     */
#define MCS_ERR_STATE (-6)      /* call out of order (e.g. no cases set)      */
#define MCS_ERR_PHI_UNBOUND (-7) /* replay: final direction on an axis; the
                                    reference raises UnboundLocalError there */

#define MCS_NATIVE 0            /* FP32 walk, Philox4x32-10 per photon        */
#define MCS_REPLAY 1            /* FP64 walk consuming the reference's
                                   uniform-deviate stream (CSR slices)         */

#define MCS_THETA_BINS 18       /* BRF zenith bins over [0, pi/2), 5 deg      */
#define MCS_PHI_BINS 36         /* BRF azimuth bins over [0, 2 pi), 10 deg    */
#define MCS_BRF_BINS (MCS_THETA_BINS * MCS_PHI_BINS)
#define MCS_TALLY_F64 656       /* doubles per case, see mcs_read_tallies     */
#define MCS_TALLY_U64 16        /* uint64 per case, see mcs_read_counts       */

typedef struct mcs_ctx mcs_ctx;

/* One case: a snowpack + illumination configuration (the scalar part of
 * MonteCarlo.run's state, mc3D:1504-1512, 1612).  Its photons take their
 * single-scattering properties from n_bands wavelength bins (rows of the band
 * table starting at band_begin): the per-photon arrays g, ssa_ice, ssa_imp,
 * ext_cff_mss, P_ext_imp of mc3D:1577-1588, which depend on the photon only
 * through its wavelength (rounded to 0.01 um, mc3D:1519-1520). */
typedef struct {
    double tau_tot;            /* layer optical depth (> 0)                   */
    double mux0;               /* initial direction x: sin(theta0)  mc3D:1119 */
    double muz0;               /* initial direction z: -cos(theta0) mc3D:1121 */
    double r_lambertian;       /* bottom reflectance in [0, 1]                */
    int32_t lambertian_bottom; /* 1: Lambertian bottom (mc3D:1420-1459)       */
    int32_t band_begin;        /* first band row of this case                 */
    int32_t n_bands;           /* number of band rows (1..MCS_MAX_BANDS)      */
    int32_t lambertian_surface; /* 1: the layer is a Lambertian reflector of
                                   reflectance r_lambertian: first event at
                                   the surface, re-emission cosine-weighted
                                   (mc3D:1228-1238, 1385-1387)                */
    int64_t n_photons;         /* photon packets of this case (>= 0)          */
} mcs_case;

#define MCS_MAX_BANDS 256

/* One wavelength bin of a case. */
typedef struct {
    double wvn;                /* 1/wvl [1/um]: record field, tally weight    */
    double g;                  /* HG asymmetry parameter, |g| < 1             */
    double ssa_ice;            /* single-scattering albedo of ice             */
    double ssa_imp;            /* single-scattering albedo of the impurity    */
    double ext_cff;            /* ext_cff_mss * rho_snw [1/m] (> 0)           */
    double p_ext_imp;          /* P(extinction by impurity), mc3D:1580        */
    double snow_depth;         /* tau_tot / ext_cff [m] (record field)        */
    double cdf;                /* cumulative photon fraction up to and
                                  including this bin (last = 1); used when
                                  the GPU samples photon wavelengths          */
} mcs_band;

typedef struct {
    int32_t mode;              /* MCS_NATIVE or MCS_REPLAY                    */
    int32_t want_records;      /* 1: keep per-photon records for
                                  mcs_read_records                            */
    uint64_t seed;             /* Philox key (native mode)                    */
    int64_t max_steps;         /* per-photon step cap, 0 = 2^24               */
    /* optional: band index (within its case) of every photon, all cases
     * concatenated in case order (sum of n_photons entries).  NULL: native
     * mode samples bands from the cdf column; replay requires it.           */
    const uint8_t *photon_band;
    /* replay: CSR uniform deviates, slice k = photon k's initial triple
     * (tau0, ssa0, ext0 uniforms, mc3D:1035-1038) then every uniform its
     * walk draws in the reference's order (mc3D:915-1023, 1245-1246, 1422,
     * 1453).  replay_off has sum(n_photons) + 1 entries.                    */
    const double *replay_dev;
    const int64_t *replay_off;
    int32_t async;             /* 1: return after enqueueing; mcs_sync()
                                  waits and reports errors                    */
    int32_t use_graph;         /* 1: launch through a captured hipGraph       */
} mcs_walk_args;

/* Launch shape of the persistent walk kernel (autotuned).  Each photon's
 * Philox stream is keyed by (seed, case, photon): with coop_live = 0 records
 * are bit-identical for every shape.  With coop_live > 0, the photons that are
 * still walking when the queue runs dry finish in the cooperative mode, whose
 * azimuths are measured in a transported frame and whose depths are prefix
 * sums: the same law (statistics agree), but which photons take that path
 * depends on scheduling, so records are not bit-reproducible. */
typedef struct {
    int32_t block;             /* threads per block: 64..1024, multiple of 64 */
    int32_t blocks_per_cu;     /* grid = blocks_per_cu * CUs                  */
    int32_t chunk;             /* photons per work-queue chunk, 64..2^24;
                                  a walk raises it to `block` when smaller.  A
                                  block keeps 4 chunks in flight and a chunk
                                  holds its tally slot until its longest walk
                                  ends, so chunks much smaller than the block
                                  would leave its lanes idle while the queue
                                  still has chunks (a block whose slots all
                                  hold walking chunks reads the queue counter
                                  to enter the cooperative tail once it is
                                  dry).  mcs_get_launch returns the value set */
    int32_t refill_min;        /* idle lanes that trigger a wave refill 1..64 */
    int32_t coop_live;         /* native mode, once the work queue is empty: a
                                  wave with at most this many live photons
                                  (0..32) walks them cooperatively, 64 / L
                                  lanes per photon stepping it in parallel;
                                  0 = off (records bit-reproducible; with at
                                  most 4 waves per SIMD a dry wave with <= 16
                                  photons builds their next steps' Philox and
                                  angles in its idle lanes, records
                                  unchanged).  Default 16.  Polar cases
                                  (mcs_set_polar): a wave with at most
                                  min(coop_live, 16) photons builds their next
                                  steps' state-independent part in its idle
                                  lanes (records unchanged by this)          */
} mcs_launch;

/* Host output arrays (each sum(n_photons) long, photon order; any may be
 * NULL).  The CSV record of mc3D:1489-1490 / 1626-1636 is
 * (condition, wvn[band], theta_n, phi_n, n_scat, path_length,
 *  snow_depth[band]).  The device writes the requested columns in these
 * types; an array in pinned host memory (hipHostMalloc, torch pin_memory) is
 * filled by one DMA copy, any other through the library's pinned staging. */
typedef struct {
    int8_t *condition;         /* 1 top, 2 bottom diffuse, 3 bottom direct,
                                  4 absorbed by ice, 5 by impurity, 0 capped  */
    uint8_t *band;             /* band index within the photon's case         */
    int64_t *n_scat;
    double *theta_n;
    double *phi_n;
    double *path_length;
    double *wvn;               /* mcs_band.wvn of the photon's band           */
    double *snow_depth;        /* mcs_band.snow_depth of the photon's band    */
} mcs_records;

typedef struct {
    uint64_t photons;          /* photons walked by the last call             */
    uint64_t steps;            /* photon-steps: sum over photons of n_scat+1  */
    uint64_t capped;           /* photons stopped at max_steps                */
    uint64_t phi_unbound;      /* photons whose final direction lies on an
                                  axis (phi_n from atan2; replay: error)      */
    double kernel_ms;          /* device time of the last walk kernel (HIP
                                  events on the context stream around the
                                  walk kernel; around the whole graph when
                                  use_graph is set)                           */
} mcs_stats;

/* replaces: process start + MonteCarlo() construction on every rank (drv:104) */
int mcs_create(int device, mcs_ctx **out);
int mcs_destroy(mcs_ctx *ctx);
const char *mcs_last_error(const mcs_ctx *ctx);
/* number of compute units of the context's device */
int mcs_device_cus(const mcs_ctx *ctx, int32_t *out);

/* replaces: Parallel(wvls) scatter of the per-photon parameter arrays
 * (parallelize.py:28-38, mc3D:1521-1612) */
int mcs_set_cases(mcs_ctx *ctx, const mcs_case *cases, int32_t n_cases,
                  const mcs_band *bands, int32_t n_bands);

int mcs_set_launch(mcs_ctx *ctx, const mcs_launch *launch);
int mcs_get_launch(const mcs_ctx *ctx, mcs_launch *out);

/* replaces: the photon loop of run() (mc3D:1613-1616) calling
 * monte_carlo3D (mc3D:1111-1490) on every rank */
int mcs_walk(mcs_ctx *ctx, const mcs_walk_args *args);
int mcs_sync(mcs_ctx *ctx);

/* replaces: comm.gather + flatten_list (parallelize.py:17-26, mc3D:1618) */
int mcs_read_records(mcs_ctx *ctx, const mcs_records *out, int64_t n_photons);

/* Host record columns for the record file (photon order, n each; q, u, v
 * all set for the aspherical record or all NULL). */
typedef struct {
    const int8_t *condition;
    const double *wvn;
    const double *theta_n;
    const double *phi_n;
    const int64_t *n_scat;
    const double *path_length;
    const double *snow_depth;
    const double *q, *u, *v;
} mcs_text_records;

/* replaces: the record-file loop of run() (mc3D:1621-1648): `header`, then
 * per photon '%d %r %r %r %d %r %r\n' (condition, wvn, theta_n, phi_n,
 * n_scat, path_length, snow_depth; aspherical: then ' %r %r %r' of Q_n, U_n,
 * V_n before the newline), each float as Python's shortest round-trip repr
 * (numpy float64's %r before numpy 2).  Formats 8192-row blocks on up to 16
 * host threads, each block written at its offset in the file (pwrite);
 * MCS_ERR_ARG for bad arguments or a file that cannot be written (the file
 * may then hold part of the text). */
int mcs_write_records_text(const char *path, const char *header, const mcs_text_records *records,
                           int64_t n_photons);

/* Python repr of n doubles, each followed by '\n', into out (cap bytes; 25
 * per value always suffice).  Returns the bytes written, -1 if cap is short.
 * No counterpart in the reference: the float formatting of the record file,
 * exposed for checking against CPython. */
int64_t mcs_repr_doubles(const double *x, int64_t n, char *out, int64_t cap);

/* Caller-owned DEVICE arrays that mcs_walk_noalloc writes (photon order). */
typedef struct {
    double *tallies;           /* n_cases x MCS_TALLY_F64 (mcs_read_tallies)  */
    uint64_t *counts;          /* n_cases x MCS_TALLY_U64 (mcs_read_counts)   */
    /* records (want_records = 1; all five required), sum(n_photons) each:   */
    uint32_t *n_scat;
    uint32_t *cond_band;       /* condition | band << 8                       */
    float *theta_n, *phi_n, *path_length;
    float *stokes;             /* aspherical cases: Q, U, V planes of
                                  sum(n_photons) each (NULL: not written)     */
} mcs_dev_out;

#define MCS_NOALLOC_MAX_CASES 65536

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc -- the native walk of mcs_walk with every buffer
 * supplied: inputs are the tables of mcs_set_cases / mcs_set_polar, outputs
 * the caller's device arrays in `out`; the walk's scratch (control words,
 * work-queue counter, chunk table: 256.1 KiB) is a static device array of
 * the library, one per device, kept until the process ends.  Allocates
 * nothing (no hipMalloc, hipFree or allocating library call), copies nothing
 * between host and device.  args: mode MCS_NATIVE, photon_band, replay_*
 * NULL, use_graph 0 (MCS_ERR_ARG otherwise); at most MCS_NOALLOC_MAX_CASES
 * cases.  Records go straight to photon order (no compaction step).  Calls
 * on one device share its scratch, so two calls that overlap on one device
 * (two contexts, two host threads) race; sequential calls are safe (each
 * clears what it uses).  Afterwards mcs_read_tallies / mcs_read_counts /
 * mcs_read_records report MCS_ERR_STATE (the results are in `out`). */
int mcs_walk_noalloc(mcs_ctx *ctx, const mcs_walk_args *args, const mcs_dev_out *out);

/* Aspherical ice grains (Yang et al. 2013 tables, shape != 'sphere' and not
 * HG): the scattering phase matrix replaces Henyey-Greenstein and every
 * photon carries a Stokes vector.  One entry per case set by mcs_set_cases. */
typedef struct {
    int32_t table;             /* phase-matrix table of the case's band 0
                                  (band j uses table + j); -1: HG (spheres,
                                  or HG forced for aspherical grains)         */
    int32_t cutoff;            /* cutoff_idx of interpolate_phase_matrix
                                  (mc3D:858-883): index into the angle grid
                                  in file order (theta ascending)             */
    double stokes0[4];         /* initial Stokes vector (I, Q, U, V), I != 0
                                  (run(stokes_params=...), mc3D:1132)         */
} mcs_polar_case;

/* replaces: interpolate_phase_matrix (mc3D:819-883) and the per-wavelength
 * P11..P44 of get_aspherical_SSPs (mc3D:267-496).  `tables` holds n_tables
 * tables of 7 x n_angles doubles: row 0 the ascending cos(theta) grid
 * (cos(theta_deg * pi / 180) of the files, reversed), rows 1..6 P11, P12,
 * P22, P33, P43, P44 on that grid (P12..P44 already multiplied by P11), i.e.
 * the arrays of the reference's interp1d(cos_theta[::-1], P[::-1]).  The
 * tables of one case share one grid.  Call after mcs_set_cases; pc = NULL
 * clears the polar state (mcs_set_cases also does).  Walks then sample the
 * angles by the reference's rejection method (mc3D:905-1004; native mode:
 * Philox blocks (photon, case, step, 2 + attempt)) and update the Stokes
 * vector at every scattering (mc3D:1283-1342).  The native walk draws the
 * same law by inversion of the S11 marginal plus an azimuth rejection when
 * the band's table has P11 >= |P12| >= 0 and keeps physical vectors physical
 * at every grid point and the case's initial vector is physical (I > 0,
 * Q^2 + U^2 + V^2 <= I^2); otherwise it runs the two-area rejection
 * (counted in mcs_read_counts slot 11). */
int mcs_set_polar(mcs_ctx *ctx, const mcs_polar_case *pc, int32_t n_cases,
                  const double *tables, int32_t n_tables, int32_t n_angles);

/* replaces: the Q_n, U_n, V_n columns of the aspherical record
 * (mc3D:1637-1646): final Stokes vector (normalised to I = 1) of every
 * photon of the last walk, photon order (needs want_records and polar cases;
 * HG cases of the same walk read (0, 0, 0)). */
int mcs_read_stokes(mcs_ctx *ctx, double *q, double *u, double *v, int64_t n_photons);

/* Per-case float64 tallies (n_cases x MCS_TALLY_F64), the sums that
 * post_processing.py computes from the CSV (pp:46-100, 125-128, 269-275):
 *   [0, 648)   sum of wvn over reflected photons (condition 1) binned by
 *              theta_n (18 x 5 deg, row) and phi_n (36 x 10 deg, column)
 *   648        Q_down = sum of wvn over all photons
 *   649..653   sum of wvn by condition 1..5
 *   654        sum of wvn over capped photons (condition 0)
 *   655        sum of wvn * path_length over reflected photons [m] */
int mcs_read_tallies(mcs_ctx *ctx, double *out, int32_t n_cases);

/* Per-case counts (n_cases x MCS_TALLY_U64):
 *   0 photons, 1..5 photons by condition, 6 capped, 7 phi on an axis,
 *   8 photon-steps, 9 photon-steps walked in the cooperative tail mode
 *   (mcs_launch.coop_live), 10 photons handed from one wave's cooperative
 *   tail to an idle wave of the same block, 11 scattering angles of the
 *   native polarised walk drawn by the reference's two-area rejection
 *   (tables or initial Stokes vectors the exact sampler does not cover, see
 *   mcs_set_polar), 12..15 reserved (0) */
int mcs_read_counts(mcs_ctx *ctx, uint64_t *out, int32_t n_cases);

int mcs_last_stats(const mcs_ctx *ctx, mcs_stats *out);

/* Time each launch shape of `candidates` on the cases currently set (native
 * mode, records off, `photons_cap` photons per case at most, `repeats`
 * timed runs each) and keep the fastest.  times_ms (optional, n entries)
 * receives the median time of each candidate.  No reference counterpart:
 * the reference has a fixed rank count chosen by the user (README:54). */
int mcs_autotune(mcs_ctx *ctx, const mcs_launch *candidates, int32_t n,
                 int64_t photons_cap, int32_t repeats, uint64_t seed,
                 double *times_ms, mcs_launch *best);

/* Diagnostic: Philox4x32-10 (Salmon et al., SC'11) of n counters (4 x u32
 * each, row-major) under key (k0, k1), computed on the device; lets tests
 * check the device generator against known-answer vectors.  No reference
 * counterpart (the reference draws from numpy's MT19937). */
int mcs_philox(mcs_ctx *ctx, const uint32_t *ctr, uint32_t k0, uint32_t k1,
               uint32_t *out, int64_t n);

/* ---------------------------------------------------------------------
 * Learned surrogate: an MLP n_in -> n_hidden_layers x hidden (SiLU) -> n_out
 * trained and served with hand-written BF16 MFMA GEMM kernels (FP32
 * accumulate) and an FP32 Adam.  No reference counterpart: the reference
 * always runs the Monte Carlo; these entry points serve cases the plain
 * walk would otherwise have to simulate, and the walk stays the arbiter of accuracy.
 *
 * Parameter layout (flat float32, unpadded): for layer l = 0..n_hidden_layers
 * the weight W_l [rows_l x cols_l] row-major then its bias b_l [rows_l];
 * rows = hidden (hidden layers) or n_out (last), cols = n_in (first layer)
 * or hidden.  Gradients use the same layout.
 * ------------------------------------------------------------------- */
typedef struct mcs_mlp mcs_mlp;

typedef struct {
    int32_t n_in;              /* inputs, 1..32                               */
    int32_t hidden;            /* hidden width, multiple of 64 (<= 1024)      */
    int32_t n_hidden_layers;   /* >= 1                                        */
    int32_t n_out;             /* outputs, >= 1                               */
    int32_t max_batch;         /* rows per device pass, multiple of 64        */
    int32_t reserved;
} mcs_mlp_dims;

typedef struct {
    float lr, beta1, beta2, eps, weight_decay;
    int32_t step;              /* Adam step count before this update (>= 0)   */
} mcs_adam;

int mcs_mlp_create(int device, const mcs_mlp_dims *dims, mcs_mlp **out);
int mcs_mlp_destroy(mcs_mlp *mlp);
const char *mcs_mlp_last_error(const mcs_mlp *mlp);
/* number of float32 parameters of the layout above */
int mcs_mlp_param_count(const mcs_mlp *mlp, int64_t *out);
int mcs_mlp_set_params(mcs_mlp *mlp, const float *params, int64_t n);
int mcs_mlp_get_params(mcs_mlp *mlp, float *params, int64_t n);

/* y [n x n_out] = MLP(x [n x n_in]); host buffers, any n >= 0 (processed in
 * max_batch slices).  forward_ms of mcs_mlp_last_ms = device time of the
 * GEMM / epilogue kernels (HIP events), excluding host copies. */
int mcs_mlp_forward(mcs_mlp *mlp, const float *x, int64_t n, float *y);
int mcs_mlp_last_ms(const mcs_mlp *mlp, double *forward_ms);

/* Serving on device memory.  A feature map turns raw inputs into the
 * network's inputs: MCS_FEATURES_IDENTITY (inputs are the n_in features) or
 * MCS_FEATURES_SNOW_BRF (rows of 5 raw case inputs r_eff [um], omega, g,
 * theta0 [deg], tau -> 9 features: log r_eff, log10(1 - omega), g,
 * cos theta0, log10 tau, exp(-tau / mu0), sqrt((1 - omega) / (1 - omega g)),
 * log10(tau (1 - omega g)), log10(tau / mu0), each standardised as
 * (f - feat_mu) / feat_sd; needs n_in = 9).  The outputs pass through
 * out_k = exp(out_scale_k y_k + out_shift_k) - out_offset_k (targets learned
 * in log space), fused into the last GEMM.  No reference counterpart: the
 * network answers cases the walk would otherwise simulate. */
#define MCS_FEATURES_IDENTITY 0
#define MCS_FEATURES_SNOW_BRF 1
int mcs_mlp_set_serving(mcs_mlp *mlp, int32_t feature_map, const float *feat_mu, const float *feat_sd,
                        const float *out_scale, const float *out_shift, const float *out_offset);
/* y [n x n_out] on the device from raw inputs x [n x 5 or n_in] on the
 * device (caller-owned device memory; returns when the results are written).
 * mcs_mlp_last_ms then gives the device time of the whole call. */
int mcs_mlp_serve(mcs_mlp *mlp, const float *d_x, int64_t n, float *d_y);

/* Loss = sum over rows and outputs of w_o (y - t)^2 / n (w: n_out weights).
 * Gradients of that loss for one batch (n <= max_batch, multiple of 64),
 * flat layout above; no update.  */
int mcs_mlp_grads(mcs_mlp *mlp, const float *x, const float *t, const float *w, int32_t n,
                  float *grads, float *loss);

/* Keep a training set on the device: x [n x n_in], t [n x n_out]. */
int mcs_mlp_set_dataset(mcs_mlp *mlp, const float *x, const float *t, int64_t n);
/* n_steps (1..1000) Adam steps; step s trains on rows idx[s*batch .. (s+1)*batch) of
 * the resident dataset (batch multiple of 64, <= max_batch); w as above;
 * losses (optional, n_steps) receives each step's loss. */
int mcs_mlp_train_steps(mcs_mlp *mlp, const int32_t *idx, int32_t n_steps, int32_t batch,
                        const float *w, const mcs_adam *adam, float *losses);

#ifdef __cplusplus
}
#endif
#endif /* MCSNOW_H */
