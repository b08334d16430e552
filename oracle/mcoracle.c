/*
 * mcoracle.c -- plain-C float64 oracle of the photon random walk.
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, smoke() and bench.py's
 * cpu_baseline leg as the checker; never by the product library.
 *
 * Restates monte_carloMPI/monte_carlo3D.py (mc3D) for the sphere /
 * Henyey-Greenstein path, in the same operation order as oracle/walk.py:
 *   initial_pdfs            mc3D:1027-1044
 *   populate_pdfs (HG)      mc3D:911-921, 1010-1023
 *   Henyey_Greenstein2      mc3D:790-800
 *   monte_carlo3D           mc3D:1111-1490
 * The uniform stream is numpy's legacy RandomState MT19937 (state layout of
 * RandomState.get_state(): 624 words + position), so that a walk started from
 * a numpy state consumes exactly the deviates the reference would draw.
 * Pins: draw-for-draw equality with numpy (tests/test_oracle_c.py) and
 * record-for-record equality with oracle/walk.py, which the golden records
 * written by the reference pin in turn.
 *
 * The MT19937 generator is the published Matsumoto-Nishimura (1998)
 * algorithm; 53-bit doubles are (a>>5, b>>6) -> (a*2^26 + b) / 2^53, the
 * conversion numpy's legacy random_sample uses.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MT_N 624
#define MT_M 397

typedef struct {
    uint32_t key[MT_N];
    int32_t pos;
} mco_mt;

typedef struct {
    double tau_tot;            /* optical depth of the layer */
    double rho_snw;            /* snow density [kg/m3] */
    double mux0, muz0;         /* sin(theta0), -cos(theta0) (mc3D:1119-1121) */
    int32_t lambertian_bottom; /* 1: Lambertian bottom boundary (mc3D:1420-1428) */
    int32_t lambertian_surface; /* 1: Lambertian reflector (mc3D:1228-1238, 1385-1387) */
    double r_lambertian;       /* its reflectance */
    int64_t max_steps;         /* 0: unlimited (the reference has no cap) */
} mco_case;

static void mt_regen(mco_mt *s)
{
    const uint32_t upper = 0x80000000u, lower = 0x7fffffffu, matrix = 0x9908b0dfu;
    int i;
    uint32_t y;
    for (i = 0; i < MT_N - MT_M; i++) {
        y = (s->key[i] & upper) | (s->key[i + 1] & lower);
        s->key[i] = s->key[i + MT_M] ^ (y >> 1) ^ ((0u - (y & 1u)) & matrix);
    }
    for (; i < MT_N - 1; i++) {
        y = (s->key[i] & upper) | (s->key[i + 1] & lower);
        s->key[i] = s->key[i + MT_M - MT_N] ^ (y >> 1) ^ ((0u - (y & 1u)) & matrix);
    }
    y = (s->key[MT_N - 1] & upper) | (s->key[0] & lower);
    s->key[MT_N - 1] = s->key[MT_M - 1] ^ (y >> 1) ^ ((0u - (y & 1u)) & matrix);
    s->pos = 0;
}

static inline uint32_t mt_next32(mco_mt *s)
{
    if (s->pos >= MT_N)
        mt_regen(s);
    uint32_t y = s->key[s->pos++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

static inline double mt_next_double(mco_mt *s)
{
    uint32_t a = mt_next32(s) >> 5, b = mt_next32(s) >> 6;
    return (a * 67108864.0 + b) / 9007199254740992.0;
}

/* ---- deviate source: MT stream, optionally recorded, or a replay slice ---- */
typedef struct {
    mco_mt *mt;
    const double *rep;   /* replay slice (NULL: use mt) */
    int64_t rep_len, rep_used;
    double *rec;         /* recording buffer (NULL: none) */
    int64_t rec_cap, rec_len;
    int rec_overflow;
    int err;
} src_t;

static inline double draw(src_t *s)
{
    double v;
    if (s->rep) {
        if (s->rep_used >= s->rep_len) {
            s->err = -3;
            return 0.5;
        }
        v = s->rep[s->rep_used++];
    } else {
        v = mt_next_double(s->mt);
    }
    if (s->rec) {
        if (s->rec_len < s->rec_cap)
            s->rec[s->rec_len] = v;
        else
            s->rec_overflow = 1;
        s->rec_len++;
    }
    return v;
}

static inline double hg2(double g, double u)
{
    if (g == 0.0)
        return 1.0 - 2.0 * u;
    double t = (1.0 - g * g) / ((1.0 - g) + (2.0 * g) * u);
    return (1.0 / (2.0 * g)) * ((1.0 + g * g) - t * t);
}

int mco_mt_doubles(mco_mt *mt, double *out, int64_t n)
{
    for (int64_t i = 0; i < n; i++)
        out[i] = mt_next_double(mt);
    return 0;
}

/*
 * Walk n photons.  Stream: replay slices (replay/replay_off, CSR with n+1
 * offsets; slice k = initial triple then walk draws) if given, else the MT
 * state (initial triples for all n photons first, as initial_pdfs draws
 * them, then each walk in turn).  rec/rec_off (optional) receive the CSR
 * replay stream of what was consumed.
 * Returns 0; -1 bad args; -2 rec capacity exceeded (rec_off[n] = needed);
 * -3 replay slice exhausted or not fully consumed; -4 some phi_n unbound
 * (set to NaN; the reference raises UnboundLocalError there); -5 step cap.
 */
int mco_walk(const mco_case *c, int64_t n,
             const double *g, const double *ssa_ice, const double *ssa_imp,
             const double *ext_cff_mss, const double *p_ext_imp,
             mco_mt *mt, const double *replay, const int64_t *replay_off,
             double *rec, int64_t rec_cap, int64_t *rec_off,
             int8_t *cond, int64_t *nscat, double *theta, double *phi, double *path,
             int64_t *steps_total)
{
    if (!c || n < 0 || !g || !ssa_ice || !ssa_imp || !ext_cff_mss || !p_ext_imp ||
        !cond || !nscat || !theta || !phi || !path)
        return -1;
    if (!replay && !mt)
        return -1;
    if (rec && !rec_off)
        return -1;
    const double two_pie = 2.0 * M_PI;
    const double half_pi = M_PI / 2.0;
    int ret = 0;
    int64_t steps = 0;
    double *first = NULL;
    src_t s;
    memset(&s, 0, sizeof s);
    s.mt = mt;
    s.rec = rec;
    s.rec_cap = rec_cap;
    if (!replay) {
        first = (double *)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
        if (!first)
            return -1;
        for (int64_t k = 0; k < 3 * n; k++)
            first[k] = mt_next_double(mt);
    }
    for (int64_t k = 0; k < n; k++) {
        double u_tau0, u_ssa0, u_ext0;
        if (replay) {
            s.rep = replay + replay_off[k];
            s.rep_len = replay_off[k + 1] - replay_off[k];
            s.rep_used = 0;
            if (s.rec) rec_off[k] = s.rec_len;
            u_tau0 = draw(&s);
            u_ssa0 = draw(&s);
            u_ext0 = draw(&s);
        } else {
            u_tau0 = first[3 * k];
            u_ssa0 = first[3 * k + 1];
            u_ext0 = first[3 * k + 2];
            if (s.rec) {
                rec_off[k] = s.rec_len;
                for (int j = 0; j < 3; j++) {
                    if (s.rec_len < s.rec_cap) s.rec[s.rec_len] = first[3 * k + j];
                    else s.rec_overflow = 1;
                    s.rec_len++;
                }
            }
        }
        const double gk = g[k];
        const double ext_cff = ext_cff_mss[k] * c->rho_snw;
        double mux0 = c->mux0, muy0 = 0.0, muz0 = c->muz0;
        double muz2 = muz0 * muz0;
        double z = 0.0, path_length = 0.0;
        double dtau = c->lambertian_surface ? 0.0 : -log(u_tau0), u_ssa = u_ssa0, u_ext = u_ext0;
        double u_hg = 0.0, u_phi = 0.0;
        int bottom_reflection = 0, condition = 0;
        int64_t i = 0;
        while (condition == 0) {
            i++;
            if (c->max_steps > 0 && i > c->max_steps) {
                ret = -5;
                break;
            }
            if (i >= 2) {
                u_hg = draw(&s);
                u_phi = draw(&s);
                dtau = -log(draw(&s));
                u_ssa = draw(&s);
                u_ext = draw(&s);
            }
            double mux_n, muy_n, muz_n;
            if (i == 1) {
                mux_n = mux0; muy_n = muy0; muz_n = muz0;
            } else {
                double costheta, sintheta;
                if (bottom_reflection || c->lambertian_surface) {
                    mux0 = 0.0; muy0 = 0.0; muz0 = 1.0;
                    for (;;) {
                        double th = 0.0 + half_pi * draw(&s);
                        double r1 = draw(&s);
                        if (s.err) break;
                        if (r1 < 2.0 * sin(th) * cos(th)) {
                            costheta = cos(th);
                            sintheta = sqrt(1.0 - costheta * costheta);
                            break;
                        }
                    }
                    if (s.err) { costheta = 1.0; sintheta = 0.0; }
                } else {
                    costheta = hg2(gk, u_hg);
                    sintheta = sqrt(1.0 - costheta * costheta);
                }
                double ph = u_phi * two_pie;
                double cosphi = cos(ph), sinphi = sin(ph);
                if (muz0 == 1.0) {
                    mux_n = sintheta * cosphi;
                    muy_n = sintheta * sinphi;
                    muz_n = costheta;
                } else if (muz0 == -1.0) {
                    mux_n = sintheta * cosphi;
                    muy_n = -sintheta * sinphi;
                    muz_n = -costheta;
                } else {
                    double sq = sqrt(1.0 - muz2);
                    mux_n = (sintheta * (mux0 * muz0 * cosphi - muy0 * sinphi)) / sq + mux0 * costheta;
                    muy_n = (sintheta * (muy0 * muz0 * cosphi + mux0 * sinphi)) / sq + muy0 * costheta;
                    muz_n = -sq * sintheta * cosphi + muz0 * costheta;
                }
                bottom_reflection = 0;
            }
            double z_prev = z;
            double z_i = z_prev + dtau * muz_n;
            z = z_i;
            if (i > 1) {
                mux0 = mux_n; muy0 = muy_n; muz0 = muz_n;
                muz2 = muz0 * muz0;
            }
            path_length += dtau / ext_cff;
            int ice = u_ext > p_ext_imp[k];
            double ssa_event = c->lambertian_surface ? c->r_lambertian : (ice ? ssa_ice[k] : ssa_imp[k]);
            if (z_i > 0.0) {
                condition = 1;
                path_length += -((z_i * dtau) / ((z_i - z_prev) * ext_cff));
            } else if (z_i < -c->tau_tot) {
                path_length += -(((z_i + c->tau_tot) * dtau) / ((z_i - z_prev) * ext_cff));
                double dtau_corr = -(((z_i + c->tau_tot) / (z_prev - z_i)) * dtau);
                z = z_i - (muz_n * dtau_corr);
                int exit_cond = (i == 1) ? 3 : 2;
                if (c->lambertian_bottom) {
                    if (draw(&s) <= c->r_lambertian)
                        bottom_reflection = 1;
                    else
                        condition = exit_cond;
                } else {
                    condition = exit_cond;
                }
            } else if (u_ssa >= ssa_event) {
                condition = ice ? 4 : 5;
            }
            if (s.err) break;
        }
        steps += i;
        if (s.err) {
            ret = s.err;
            break;
        }
        if (replay && s.rep_used != s.rep_len) {
            ret = -3;
            break;
        }
        cond[k] = (int8_t)condition;
        nscat[k] = i - 1;
        theta[k] = acos(muz0);
        path[k] = path_length;
        double p;
        if (i == 1) p = 0.0;
        else if (mux0 > 0 && muy0 > 0) p = atan(muy0 / mux0);
        else if (mux0 < 0 && muy0 > 0) p = atan(muy0 / mux0) + M_PI;
        else if (mux0 < 0 && muy0 < 0) p = atan(muy0 / mux0) + M_PI;
        else if (mux0 > 0 && muy0 < 0) p = atan(muy0 / mux0) + two_pie;
        else { p = NAN; if (ret == 0) ret = -4; }
        phi[k] = p;
        if (ret == -5) break;
    }
    if (s.rec) rec_off[n] = s.rec_len;
    if (s.rec_overflow && ret == 0) ret = -2;
    if (steps_total) *steps_total = steps;
    free(first);
    return ret;
}
