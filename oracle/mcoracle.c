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
 * and, for aspherical grains (`pol` given), the phase-matrix rejection
 * sampler and Stokes update of oracle/polar.py:
 *   populate_pdfs (aspherical) mc3D:898-1004
 *   rotate_stokes_vector       mc3D:1090-1109
 *   Stokes update              mc3D:1283-1342
 * The uniform stream is numpy's legacy RandomState MT19937 (state layout of
 * RandomState.get_state(): 624 words + position), so that a walk started from
 * a numpy state consumes exactly the deviates the reference would draw.
 * Pins: draw-for-draw equality with numpy (tests/test_oracle_c.py) and
 * record-for-record equality with oracle/walk.py, which the golden records
 * written by the reference pin in turn; for aspherical grains the golden
 * phase-matrix records (conditions, scattering counts and the uniform
 * stream exactly, angles and Q, U, V to libm-vs-numpy rounding), and a
 * planted P12 = 0 that moves the Stokes vectors.  np.interp is restated
 * after numpy's compiled_base.c rules (interp_np).
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

/* Aspherical grains: one phase-matrix table per photon wavelength (the
 * arrays of interp1d(cos_theta[::-1], P[::-1]), mc3D:842-857) and the
 * rejection sampler's constants (mc3D:858-883, 926-941). */
typedef struct {
    int32_t n;                 /* grid points */
    int32_t pad;
    const double *x;           /* ascending cos theta [n] */
    const double *P;           /* P11, P12, P22, P33, P43, P44 on x [6 n] */
    double A1, B1, A2, B2;     /* max |P11|, |P12| over theta <= / >= cutoff */
} mco_pm;

typedef struct {
    double cmax, ccut, cmin;   /* cos theta range and the cutoff */
    double stokes0[4];         /* initial Stokes vector */
    const mco_pm *const *pm;   /* per photon: its wavelength's table */
    double *q, *u, *v;         /* out: final Q, U, V per photon */
} mco_polar;

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

/* np.interp(v, x, y) (numpy's compiled_base.c rules): the end values
 * outside the grid and at its last point, y_j at grid points, else
 * slope (v - x_j) + y_j with x_j <= v < x_j+1 */
static double interp_np(double v, const double *x, const double *y, int n)
{
    if (v > x[n - 1]) return y[n - 1];
    if (v < x[0]) return y[0];
    int lo = 0, hi = n - 1;
    if (v == x[n - 1]) return y[n - 1];
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (x[mid] <= v) lo = mid; else hi = mid;
    }
    if (x[lo] == v) return y[lo];
    const double slope = (y[lo + 1] - y[lo]) / (x[lo + 1] - x[lo]);
    double r = slope * (v - x[lo]) + y[lo];
    if (r != r) {
        r = slope * (v - x[lo + 1]) + y[lo + 1];
        if (r != r && y[lo] == y[lo + 1]) r = y[lo];
    }
    return r;
}

/* populate_pdfs for an aspherical grain (mc3D:898-1004): area, cos theta,
 * phi, then r3 and, after each rejection, area, cos theta, 2 phi, r3. */
static void polar_draws(src_t *s, const mco_polar *pl, const mco_pm *t, const double st[4],
                        double *ct_out, double *phi_out)
{
    const double two_pie = 2.0 * M_PI, four_pie = 4.0 * M_PI;
    const double I = st[0], Q = st[1], U = st[2];
    const double two_beta = (Q == 0.0) ? atan(U / 1e-15) : atan(U / Q);
    double A = fabs(I) * t->A1;
    double B = t->B1;
    const double C = fabs(Q * cos(two_beta) + U * sin(two_beta));
    const double mv1 = A + B * C;
    A = fabs(I) * t->A2;
    B = t->B2;
    const double mv2 = A + B * C;
    const double a1 = (pl->cmax - pl->ccut) * mv1;
    const double a2 = (pl->ccut - pl->cmin) * mv2;
    const double pct = a1 / (a1 + a2);
    double ar = draw(s);
    double ct = draw(s);
    double two_phi = 2.0 * (draw(s) * two_pie);
    int area;
    if (ar < pct) { area = 1; ct = ct * (pl->cmax - pl->ccut) + pl->ccut; }
    else { area = 2; ct = -ct * (pl->ccut - pl->cmin) + pl->ccut; }
    double r3 = 1.0, val = 0.0;
    for (int k = 0; r3 > val; ++k) {
        if (k > 0) {
            ar = draw(s);
            if (ar < pct) { area = 1; ct = draw(s) * (pl->cmax - pl->ccut) + pl->ccut; }
            else { area = 2; ct = -draw(s) * (pl->ccut - pl->cmin) + pl->ccut; }
            two_phi = draw(s) * four_pie;
        }
        r3 = 1.000001 * draw(s);
        r3 = (area == 1) ? r3 * mv1 : r3 * mv2;
        const double S11 = interp_np(ct, t->x, t->P, t->n);
        const double S12 = interp_np(ct, t->x, t->P + t->n, t->n);
        val = I * S11 + S12 * (Q * cos(two_phi) + U * sin(two_phi));
        if (s->err) break;
    }
    *ct_out = ct;
    *phi_out = two_phi / 2.0;
}

/* mc3D:1287-1342 with rotate_stokes_vector (mc3D:1090-1109). */
static void stokes_update(const mco_pm *t, double st[4], double costheta, double cosphi, double muz_n,
                          double muz_0)
{
    const int n = t->n;
    const double cos2theta = costheta * costheta;
    const double phi_sca = acos(cosphi);
    double da = 2.0 * phi_sca;
    double c = cos(da), sn = sin(da);
    const double Isp = st[0], Qsp = st[1] * c + st[2] * sn, Usp = -st[1] * sn + st[2] * c, Vsp = st[3];
    const double P11 = interp_np(costheta, t->x, t->P, n), P12 = interp_np(costheta, t->x, t->P + n, n);
    const double P22 = interp_np(costheta, t->x, t->P + 2 * n, n), P33 = interp_np(costheta, t->x, t->P + 3 * n, n);
    const double P43 = interp_np(costheta, t->x, t->P + 4 * n, n), P44 = interp_np(costheta, t->x, t->P + 5 * n, n);
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
        const double sg = (phi_sca >= M_PI) ? 1.0 : -1.0;
        ci = num / (sg * den);
        if (ci < -1.0) ci = -1.0;
        else if (ci > 1.0) ci = 1.0;
    }
    const double i2 = acos(ci);
    da = 2.0 * (-i2);
    c = cos(da);
    sn = sin(da);
    const double Qm = Qc * c + Uc * sn, Um = -Qc * sn + Uc * c;
    st[0] = Ic / Ic;
    st[1] = Qm / Ic;
    st[2] = Um / Ic;
    st[3] = Vc / Ic;
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
             int64_t *steps_total, const mco_polar *pol)
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
        double u_hg = 0.0, u_phi = 0.0, p_cos = 0.0, p_phi = 0.0;
        const mco_pm *tab = pol ? pol->pm[k] : NULL;
        double st[4] = {1.0, 0.0, 0.0, 0.0};
        if (pol) memcpy(st, pol->stokes0, sizeof st);
        int bottom_reflection = 0, condition = 0;
        int64_t i = 0;
        while (condition == 0) {
            i++;
            if (c->max_steps > 0 && i > c->max_steps) {
                ret = -5;
                break;
            }
            if (i >= 2) {
                if (tab) {
                    polar_draws(&s, pol, tab, st, &p_cos, &p_phi);
                } else {
                    u_hg = draw(&s);
                    u_phi = draw(&s);
                }
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
                } else if (tab) {
                    costheta = p_cos;
                    sintheta = sqrt(1.0 - costheta * costheta);
                } else {
                    costheta = hg2(gk, u_hg);
                    sintheta = sqrt(1.0 - costheta * costheta);
                }
                double ph = tab ? p_phi : u_phi * two_pie;
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
                if (bottom_reflection) {             /* mc3D:1283-1285 */
                    st[0] = 1.0; st[1] = 0.0; st[2] = 0.0; st[3] = 0.0;
                } else if (tab) {                    /* mc3D:1287-1342 */
                    stokes_update(tab, st, costheta, cosphi, muz_n, muz0);
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
        if (pol) {
            pol->q[k] = st[1];
            pol->u[k] = st[2];
            pol->v[k] = st[3];
        }
        if (ret == -5) break;
    }
    if (s.rec) rec_off[n] = s.rec_len;
    if (s.rec_overflow && ret == 0) ret = -2;
    if (steps_total) *steps_total = steps;
    free(first);
    return ret;
}
