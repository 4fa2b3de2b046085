/* CPU oracle for CFDwavemaker's second-order irregular-wave kinematics path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): an FP64
 * restatement of the reference algorithm, written from scratch in plain C.
 * Each function cites the reference file:line it restates.  Expression order
 * follows the reference so that, built without -ffast-math (oracle/Makefile),
 * results agree with the compiled reference to rounding level.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - golden vectors written from the compiled reference (tests/golden/, npz files)
 *     for every or_eval kind, the pair cache, the s-grid fill/update/gradients
 *     and the spline query;
 *   - closed forms: Airy first order, Stokes second-order self-interaction,
 *     unittests/irregular_tests.cpp:53,69,73 known answers.
 */
#include "cfdwm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>

#define OR_PI 3.1415926535897        /* truncated PI of src/Irregular.h:13-15 */
#define OR_SMALL 1.e-12              /* src/Irregular.h:7-9 */
#define OR_HUGE 1.e12                /* src/Irregular.h:10-12 */

struct or_sea {
    int n, order, normalize;
    double *omega, *A, *k, *phase, *theta, *rt, *ts, *du;
    double depth, g, rho, swl, mtheta, tofmax, x0, y0, bw, cutoff, ampl;
    /* component cache: Irregular::rebuild_component_cache (Irregular.cpp:805-838) */
    double *cth, *sth, *R, *Rs, *ag, *cosh2kd, *ph0;
    double kmax;
    /* pair cache: Irregular::rebuild_pair_cache (Irregular.cpp:840-967) */
    long np;
    int *pi, *pm;
    unsigned char *pflag;
    double *gam, *kp, *km, *Dp, *Dm, *ap, *am;
    double *Hp, *Hm;              /* Smit et al. H coefficients (Irregular.cpp:919-960) */
    int pairs_built;
    int vt;                       /* vertical_theory: 0 Sharma & Dean, 1 Smit (Irregular.h:83) */
};

static double* dup_or(const double* p, int n, double dflt) {
    double* r = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) r[i] = p ? p[i] : dflt;
    return r;
}

static double dmin(double a, double b) { return (b < a) ? b : a; }   /* std::min */
static double dmax(double a, double b) { return (a < b) ? b : a; }   /* std::max */

/* Irregular::freq_ramp, Irregular.cpp:200-204 */
double or_freq_ramp(double t, double ramptime, double tstart, double duration) {
    double down = dmax(dmin(1. - ((t - tstart - duration + ramptime) / (ramptime + OR_SMALL)), 1.), 0.);
    double up = dmax(dmin(((t - tstart) / (ramptime + OR_SMALL)), 1.), 0.);
    return up * down;
}

static double fr(const or_sea* s, int i, double t) {
    return or_freq_ramp(t, s->rt[i], s->ts[i], s->du[i]);
}

/* Ramp::ramp1d / Ramp::ramp, Utils.cpp:26-68 */
static double ramp1d(double x, double xs, double xe, int inv) {
    if (inv) return dmax(dmin(1. - ((x - xs) / (xe - xs)), 1.), 0.);
    return dmax(dmin(((x - xs) / (xe - xs)), 1.), 0.);
}
double or_ramp(int enabled, int flags, const double* se, double t, double x, double y) {
    if (!enabled) return 1.0;
    double r[6] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
    if (flags & 1) r[0] = ramp1d(t, se[0], se[1], 0);
    if (flags & 2) r[1] = ramp1d(t, se[2], se[3], 1);
    if (flags & 4) r[2] = ramp1d(x, se[4], se[5], 0);
    if (flags & 8) r[3] = ramp1d(x, se[6], se[7], 1);
    if (flags & 16) r[4] = ramp1d(y, se[8], se[9], 0);
    if (flags & 32) r[5] = ramp1d(y, se[10], se[11], 1);
    double m = r[0];                       /* std::min_element: first smallest */
    for (int q = 1; q < 6; ++q) if (r[q] < m) m = r[q];
    return m;
}

/* Irregular::spectral_moments, Irregular.cpp:1336-1369 (note: loop stops at nfreq-1) */
void or_spectral_moments(const or_sea* s, double* m) {
    m[0] = m[1] = m[2] = 0.;
    for (int i = 0; i < s->n - 1; ++i) {
        m[0] += s->A[i] * s->A[i] / 2.;
        m[1] += s->omega[i] * s->A[i] * s->A[i] / 2.;
        m[2] += s->omega[i] * s->omega[i] * s->A[i] * s->A[i] / 2.;
    }
}

/* Irregular::interpolate, Irregular.cpp:1311-1334 */
static double interp_lin(const double* xd, const double* yd, int size, double x) {
    int i = 0;
    if (x >= xd[size - 2]) i = size - 2;
    else while (x > xd[i + 1]) i++;
    double xL = xd[i], yL = yd[i], xR = xd[i + 1], yR = yd[i + 1];
    double dydx = (yR - yL) / (xR - xL);
    return yL + dydx * (x - xL);
}

/* Irregular::mean_wave_period / phase_speed / mean_wave_length, Irregular.cpp:1372-1436 */
double or_mean_period(const or_sea* s, int opt) {
    double m[3]; or_spectral_moments(s, m);
    double t1 = 2. * OR_PI * (m[0] / m[1]);
    double t2 = 2. * OR_PI * sqrt(m[0] / m[2]);
    return opt == 1 ? t1 : (opt == 2 ? t2 : 0.);
}
double or_phase_speed(const or_sea* s, int opt) {
    double m[3]; or_spectral_moments(s, m);
    double t1 = 2. * OR_PI * (m[0] / m[1]);
    double t2 = 2. * OR_PI * sqrt(m[0] / m[2]);
    double w1 = 2 * OR_PI / t1, w2 = 2 * OR_PI / t2;
    double k1 = interp_lin(s->omega, s->k, s->n, w1);
    double k2 = interp_lin(s->omega, s->k, s->n, w2);
    return opt == 1 ? w1 / k1 : (opt == 2 ? w2 / k2 : 0.);
}
double or_mean_length(const or_sea* s, int opt) {
    double m[3]; or_spectral_moments(s, m);
    double t1 = 2. * OR_PI * (m[0] / m[1]);
    double t2 = 2. * OR_PI * sqrt(m[0] / m[2]);
    double w1 = 2 * OR_PI / t1, w2 = 2 * OR_PI / t2;
    double l1 = (2. * OR_PI) / interp_lin(s->omega, s->k, s->n, w1);
    double l2 = (2. * OR_PI) / interp_lin(s->omega, s->k, s->n, w2);
    return opt == 1 ? l1 : (opt == 2 ? l2 : 0.);
}

static void build_component_cache(or_sea* s) {
    const int N = s->n;
    const double mt = s->mtheta * OR_PI / 180.0;
    s->cth = dup_or(NULL, N, 0.); s->sth = dup_or(NULL, N, 0.); s->R = dup_or(NULL, N, 0.);
    s->Rs = dup_or(NULL, N, 0.); s->ag = dup_or(NULL, N, 0.); s->cosh2kd = dup_or(NULL, N, 0.);
    s->ph0 = dup_or(NULL, N, 0.);
    s->kmax = 0.0;
    for (int i = 0; i < N; ++i) {
        const double th = s->theta[i] + mt;
        s->cth[i] = cos(th);
        s->sth[i] = sin(th);
        const double R = s->k[i] * tanh(s->k[i] * s->depth);
        s->R[i] = R;
        s->Rs[i] = sqrt(R);
        s->ag[i] = s->A[i] * s->g / s->omega[i];
        s->cosh2kd[i] = cosh(2.0 * s->k[i] * s->depth);
        s->ph0[i] = s->omega[i] * s->tofmax + s->phase[i];
        if (s->k[i] > s->kmax) s->kmax = s->k[i];
    }
}

/* Sharma-Dean sum/difference transfer coefficients, in the order of
 * Irregular.cpp:903-960, with the Smit H coefficients of the same entries
 * (Irregular.cpp:919-927, :950-958: Eq. 10 inverted with C = alpha / 2). */
static void build_pair_cache(or_sea* s) {
    const int N = s->n;
    long cap = (long)N * (N - 1) / 2 + 1;
    s->pi = (int*)malloc(sizeof(int) * cap); s->pm = (int*)malloc(sizeof(int) * cap);
    s->pflag = (unsigned char*)malloc(cap);
    s->gam = dup_or(NULL, cap, 0.); s->kp = dup_or(NULL, cap, 0.); s->km = dup_or(NULL, cap, 0.);
    s->Dp = dup_or(NULL, cap, 0.); s->Dm = dup_or(NULL, cap, 0.);
    s->ap = dup_or(NULL, cap, 0.); s->am = dup_or(NULL, cap, 0.);
    s->Hp = dup_or(NULL, cap, 0.); s->Hm = dup_or(NULL, cap, 0.);
    long e = 0;
    for (int i = 0; i < N; ++i) {
        if (s->omega[i] > s->cutoff) continue;
        const double ki = s->k[i], oi = s->omega[i], Rni = s->R[i], Rsni = s->Rs[i];
        for (int m = i + 1; m < N; ++m) {
            if (s->omega[m] > s->cutoff) continue;
            if (fabs(oi - s->omega[m]) > s->bw) continue;
            const double km = s->k[m], Rnm = s->R[m], Rsm = s->Rs[m];
            const double gamma = s->cth[i] * s->cth[m] + s->sth[i] * s->sth[m];
            const double ki2 = ki * ki, km2 = km * km, kikm = ki * km;
            const double kplus = sqrt(ki2 + km2 + 2.0 * kikm * gamma);
            const double kminus = sqrt(ki2 + km2 - 2.0 * kikm * gamma);
            const double sp = Rsni + Rsm, sm = Rsni - Rsm;
            const double R2p = sp * sp, R2m = sm * sm;
            const double om = s->omega[m];
            int hp = 0, hm = 0;
            double Dp = 0.0, Dm = 0.0, ap = 0.0, am = 0.0, Hp = 0.0, Hm = 0.0;
            if ((ki + km) <= s->kmax) {
                const double den = R2p - kplus * tanh(kplus * s->depth);
                Dp = (2.0 * R2p * (kikm * gamma - Rni * Rnm)) / den +
                     (sp * (Rsni * (km2 - Rnm * Rnm) + Rsm * (ki2 - Rni * Rni))) / den;
                ap = (Dp - (kikm * gamma - Rni * Rnm)) / sqrt(Rni * Rnm) + (Rni + Rnm);
                hp = 1;
                const double om12 = oi + om, k1dk2 = kikm * gamma;
                Hp = -s->g / om12 * (ap / 2.0 - (oi * oi + om * om + oi * om) / (2.0 * s->g)
                                     + s->g * k1dk2 / (2.0 * oi * om));
            }
            if (ki != km) {
                const double den = R2m - kminus * tanh(kminus * s->depth);
                Dm = (2.0 * R2m * (kikm * gamma + Rni * Rnm)) / den +
                     (sm * (-Rsni * (km2 - Rnm * Rnm) + Rsm * (ki2 - Rni * Rni))) / den;
                am = (Dm - (kikm * gamma + Rni * Rnm)) / sqrt(Rni * Rnm) + (Rni + Rnm);
                hm = 1;
                const double om12 = oi - om, k1dk2 = kikm * gamma;
                if (fabs(om12) > OR_SMALL)
                    Hm = -s->g / om12 * (am / 2.0 - (oi * oi + om * om - oi * om) / (2.0 * s->g)
                                         + s->g * k1dk2 / (2.0 * oi * om));
            }
            if (hp || hm) {
                s->pi[e] = i; s->pm[e] = m; s->pflag[e] = (unsigned char)(hp | (hm << 1));
                s->gam[e] = gamma; s->kp[e] = kplus; s->km[e] = kminus;
                s->Dp[e] = Dp; s->Dm[e] = Dm; s->ap[e] = ap; s->am[e] = am;
                s->Hp[e] = Hp; s->Hm[e] = Hm;
                ++e;
            }
        }
    }
    s->np = e;
    s->pairs_built = 1;
}

or_sea* or_sea_create(int n, const double* omega, const double* A, const double* k,
                      const double* phase, const double* theta,
                      const double* ramptime, const double* starttime, const double* duration,
                      double depth, double g, double rho, double swl, double mtheta,
                      double tofmax, double x0, double y0, double bandwidth, double cutoff,
                      int order, int normalize, double ampl, int bw_auto)
{
    or_sea* s = (or_sea*)calloc(1, sizeof(or_sea));
    s->n = n;
    s->omega = dup_or(omega, n, 0.); s->A = dup_or(A, n, 0.); s->k = dup_or(k, n, 0.);
    s->phase = dup_or(phase, n, 0.); s->theta = dup_or(theta, n, 0.);
    s->rt = dup_or(ramptime, n, 0.); s->ts = dup_or(starttime, n, -OR_SMALL);
    s->du = dup_or(duration, n, OR_HUGE);
    s->depth = depth; s->g = g; s->rho = rho; s->swl = swl; s->mtheta = mtheta;
    s->tofmax = tofmax; s->x0 = x0; s->y0 = y0; s->bw = bandwidth; s->cutoff = cutoff;
    s->order = order; s->normalize = normalize; s->ampl = ampl;
    /* bandwidth_estimator (Irregular.cpp:1439-1453): 0.7 * m1 / m0, on raw amplitudes */
    if (bw_auto) {
        double m[3]; or_spectral_moments(s, m);
        s->bw = 0.7 * (m[1] / m[0]);
    }
    /* normalize_data (Irregular.cpp:1282-1296) */
    double sumA = 0.;
    for (int i = 0; i < n; ++i) sumA += s->A[i];
    for (int i = 0; i < n; ++i) {
        if (normalize) { s->A[i] = ampl * s->A[i] / sumA; s->phase[i] = 0.; }
        else s->A[i] = ampl * s->A[i];
    }
    build_component_cache(s);
    if (order == 2) build_pair_cache(s);
    return s;
}

void or_sea_destroy(or_sea* s) {
    if (!s) return;
    double* arrs[] = {s->omega, s->A, s->k, s->phase, s->theta, s->rt, s->ts, s->du, s->cth, s->sth,
                      s->R, s->Rs, s->ag, s->cosh2kd, s->ph0, s->gam, s->kp, s->km, s->Dp, s->Dm,
                      s->ap, s->am, s->Hp, s->Hm};
    for (size_t q = 0; q < sizeof(arrs) / sizeof(arrs[0]); ++q) free(arrs[q]);
    free(s->pi); free(s->pm); free(s->pflag);
    free(s);
}

void or_sea_components(const or_sea* s, double* A, double* phase, double* bandwidth) {
    for (int i = 0; i < s->n; ++i) { A[i] = s->A[i]; phase[i] = s->phase[i]; }
    *bandwidth = s->bw;
}
double or_sea_kmax(const or_sea* s) { return s->kmax; }
void or_sea_set_vertical_theory(or_sea* s, int vt) {
    s->vt = vt;
    /* the Smit members build the second-order caches when absent
     * (Irregular.cpp:1536, :1634) */
    if (vt == 1 && !s->pairs_built) build_pair_cache(s);
}
void or_pair_export_h(const or_sea* s, double* h2) {
    for (long e = 0; e < s->np; ++e) { h2[2 * e] = s->Hp[e]; h2[2 * e + 1] = s->Hm[e]; }
}
long or_pair_count(const or_sea* s) { return s->np; }
void or_pair_export(const or_sea* s, int* im, unsigned char* flags, double* v) {
    for (long e = 0; e < s->np; ++e) {
        im[2 * e] = s->pi[e]; im[2 * e + 1] = s->pm[e]; flags[e] = s->pflag[e];
        v[7 * e + 0] = s->gam[e]; v[7 * e + 1] = s->kp[e]; v[7 * e + 2] = s->km[e];
        v[7 * e + 3] = s->Dp[e]; v[7 * e + 4] = s->Dm[e]; v[7 * e + 5] = s->ap[e]; v[7 * e + 6] = s->am[e];
    }
}

/* ---------------------------------------------------------------- first order */
/* Linear phase argument as written in eta1/u1/... (Irregular.cpp:187-189):
 * k*(cos(th)*(x-x0) + sin(th)*(y-y0)) - omega*t + (omega*tofmax + phase) */
static double arg1(const or_sea* s, int i, double t, double x, double y) {
    const double th = s->theta[i] + (s->mtheta * OR_PI / 180.);
    const double phi = s->omega[i] * s->tofmax + s->phase[i];
    return s->k[i] * (cos(th) * (x - s->x0) + sin(th) * (y - s->y0)) - s->omega[i] * t + phi;
}
static double thp(const or_sea* s, int i) { return s->theta[i] + (s->mtheta * OR_PI / 180.); }
/* profileX / profileZ (Irregular.cpp:98-116), sloping_bottom == false */
static double pX(const or_sea* s, int i, double z) { return cosh(s->k[i] * (z + s->depth)) / cosh(s->k[i] * s->depth); }
static double pZ(const or_sea* s, int i, double z) { return sinh(s->k[i] * (z + s->depth)) / cosh(s->k[i] * s->depth); }

static double eta1(const or_sea* s, double t, double x, double y) {            /* :178-198 */
    double w = 0.0;
    for (int i = 0; i < s->n; ++i) w += s->A[i] * cos(arg1(s, i, t, x, y)) * fr(s, i, t);
    return w;
}
static double u1(const or_sea* s, double t, double x, double y, double z) {    /* :394-410 */
    double u = 0.0;
    for (int i = 0; i < s->n; ++i)
        u += s->A[i] * s->g * s->k[i] / s->omega[i] * pX(s, i, z) * cos(thp(s, i)) * cos(arg1(s, i, t, x, y)) * fr(s, i, t);
    return u;
}
static double v1(const or_sea* s, double t, double x, double y, double z) {    /* :414-429 */
    double v = 0.0;
    for (int i = 0; i < s->n; ++i)
        v += s->A[i] * s->g * s->k[i] / s->omega[i] * pX(s, i, z) * sin(thp(s, i)) * cos(arg1(s, i, t, x, y)) * fr(s, i, t);
    return v;
}
static double w1(const or_sea* s, double t, double x, double y, double z) {    /* :433-447 */
    double w = 0.0;
    for (int i = 0; i < s->n; ++i)
        w += s->A[i] * s->g * s->k[i] / s->omega[i] * pZ(s, i, z) * sin(arg1(s, i, t, x, y)) * fr(s, i, t);
    return w;
}
static double dp1(const or_sea* s, double t, double x, double y, double z) {   /* :1107-1122 */
    double p = 0.0;
    for (int i = 0; i < s->n; ++i)
        p += s->A[i] * s->rho * s->g * (cosh(s->k[i] * (z + s->depth)) / cosh(s->k[i] * s->depth)) * cos(arg1(s, i, t, x, y)) * fr(s, i, t);
    return p;
}
static double phi1_dxdz(const or_sea* s, double t, double x, double y) {      /* :1125-1138 */
    double u = 0.0;
    for (int i = 0; i < s->n; ++i)
        u += cos(thp(s, i)) * s->A[i] * s->omega[i] * s->k[i] * cos(arg1(s, i, t, x, y)) * fr(s, i, t);
    return u;
}
static double dp1_dz(const or_sea* s, double t, double x, double y) {         /* :1141-1154 */
    double p = 0.0;
    for (int i = 0; i < s->n; ++i)
        p += s->A[i] * s->rho * s->g * s->k[i] * cos(arg1(s, i, t, x, y)) * fr(s, i, t);
    return p;
}
static double phi1_dydz(const or_sea* s, double t, double x, double y) {      /* :1158-1171 */
    double v = 0.0;
    for (int i = 0; i < s->n; ++i)
        v += sin(thp(s, i)) * s->A[i] * s->omega[i] * s->k[i] * cos(arg1(s, i, t, x, y)) * fr(s, i, t);
    return v;
}
static double phi1_dzdz(const or_sea* s, double t, double x, double y) {      /* :1175-1192 */
    double w = 0.0;
    for (int i = 0; i < s->n; ++i)
        w += s->A[i] * s->omega[i] * s->k[i] * (cosh(s->k[i] * s->depth) / sinh(s->k[i] * s->depth)) * sin(arg1(s, i, t, x, y)) * fr(s, i, t);
    return w;
}

/* ------------------------------------------------ second order, pair-cached */
/* Per-component phase with the cached factors, Irregular.cpp:237-246 / :994-1004 */
static void cached_phases(const or_sea* s, double t, double x, double y,
                          double* c, double* sn, double* r) {
    const double dxp = x - s->x0, dyp = y - s->y0;
    for (int i = 0; i < s->n; ++i) {
        const double ph = s->k[i] * (s->cth[i] * dxp + s->sth[i] * dyp) - s->omega[i] * t + s->ph0[i];
        c[i] = cos(ph); sn[i] = sin(ph); r[i] = fr(s, i, t);
    }
}

static double eta2(const or_sea* s, double t, double x, double y) {           /* :223-293 */
    const int N = s->n;
    double* c = (double*)malloc(sizeof(double) * 3 * (N > 0 ? N : 1));
    double* sn = c + N; double* r = c + 2 * N;
    cached_phases(s, t, x, y, c, sn, r);
    double tot = 0.0;
    for (int i = 0; i < N; ++i) {
        if (s->omega[i] > s->cutoff) continue;
        const double ki = s->k[i], Rn = s->R[i], Rsn = s->Rs[i];
        const double Dnn = (6.0 * Rsn * Rsn * (ki * ki - Rn * Rn)) / (2.0 * Rsn * Rsn - ki * tanh(2.0 * ki * s->depth));
        const double ann = (Dnn - (ki * ki - Rn * Rn)) / Rn + 2.0 * Rn;
        if (2.0 * ki <= s->kmax) {
            const double c2 = c[i] * c[i] - sn[i] * sn[i];
            tot += 0.25 * s->A[i] * s->A[i] * ann * c2 * r[i];
        }
    }
    for (long e = 0; e < s->np; ++e) {
        const int i = s->pi[e], m = s->pm[e];
        const double amp = 0.5 * s->A[i] * s->A[m];
        const double ri = r[i];
        if (s->pflag[e] & 1) tot += amp * s->ap[e] * (c[i] * c[m] - sn[i] * sn[m]) * ri;
        if (s->pflag[e] & 2) tot += amp * s->am[e] * (c[i] * c[m] + sn[i] * sn[m]) * ri;
    }
    free(c);
    return tot;
}

/* Irregular::uvw2, Irregular.cpp:975-1102 (pair-cached, filtered; p without rho) */
static void uvw2(const or_sea* s, double t, double x, double y, double z, double* o4) {
    const int N = s->n;
    const double zpd = z + s->depth, gg = s->g * s->g;
    double* c = (double*)malloc(sizeof(double) * 5 * (N > 0 ? N : 1));
    double* sn = c + N; double* r = c + 2 * N; double* px = c + 3 * N; double* pz = c + 4 * N;
    cached_phases(s, t, x, y, c, sn, r);
    for (int i = 0; i < N; ++i) { px[i] = pX(s, i, z); pz[i] = pZ(s, i, z); }
    double us = 0.0, vs = 0.0, ws = 0.0, ps = 0.0;
    for (int i = 0; i < N; ++i) {
        if (s->omega[i] > s->cutoff) continue;
        const double ki = s->k[i], oi = s->omega[i], cti = s->cth[i], sti = s->sth[i];
        const double Rn = s->R[i], Rsn = s->Rs[i], agi = s->ag[i];
        const double Dnn = (6.0 * Rsn * Rsn * (ki * ki - Rn * Rn)) / (2.0 * Rsn * Rsn - ki * tanh(2.0 * ki * s->depth));
        const double AA = 0.25 * agi * agi * (Dnn / (2.0 * oi));
        const double c2 = c[i] * c[i] - sn[i] * sn[i];
        const double s2 = 2.0 * sn[i] * c[i];
        const double chd = cosh(2.0 * ki * zpd) / s->cosh2kd[i];
        const double shd = sinh(2.0 * ki * zpd) / s->cosh2kd[i];
        double u = 0.0, v = 0.0, w = 0.0, p = 0.0;
        u += AA * (2.0 * ki * cti) * c2 * chd;
        v += AA * (2.0 * ki * sti) * c2 * chd;
        w += AA * (2.0 * ki) * s2 * shd;
        p += AA * (2.0 * oi) * c2 * chd;
        p -= (gg * (2.0 * ki * cti + 2.0 * ki * sti) * 2.0 * px[i]) / (2.0 * oi);
        p += (gg * (2.0 * ki) * 2.0 * pz[i]) / (2.0 * oi);
        us += u * r[i]; vs += v * r[i]; ws += w * r[i]; ps += p * r[i];
    }
    for (long e = 0; e < s->np; ++e) {
        const int i = s->pi[e], m = s->pm[e];
        const double ki = s->k[i], km = s->k[m], oi = s->omega[i], om = s->omega[m];
        const double cti = s->cth[i], sti = s->sth[i], ctm = s->cth[m], stm = s->sth[m];
        const double ri = r[i];
        const double AA = 0.5 * s->ag[i] * s->ag[m];
        if (s->pflag[e] & 1) {
            const double BB = s->Dp[e] / (oi + om);
            const double cp = c[i] * c[m] - sn[i] * sn[m];
            const double sp = sn[i] * c[m] + c[i] * sn[m];
            const double den = cosh(s->kp[e] * s->depth);
            const double ch = cosh(s->kp[e] * zpd) / den;
            const double sh = sinh(s->kp[e] * zpd) / den;
            us += AA * BB * (ki * cti + km * ctm) * cp * ch * ri;
            vs += AA * BB * (ki * sti + km * stm) * cp * ch * ri;
            ws += AA * BB * s->kp[e] * sp * sh * ri;
            ps += AA * BB * (oi + om) * cp * ch * ri;
            ps -= (gg * (ki * cti * km * ctm + ki * sti * km * stm) * px[i] * px[m]) / (oi * om) * ri;
            ps += (gg * (ki * km) * pz[i] * pz[m]) / (oi * om) * ri;
        }
        if (s->pflag[e] & 2) {
            const double CC = s->Dm[e] / (oi - om);
            const double cm = c[i] * c[m] + sn[i] * sn[m];
            const double sm = sn[i] * c[m] - c[i] * sn[m];
            const double den = cosh(s->km[e] * s->depth);
            const double ch = cosh(s->km[e] * zpd) / den;
            const double sh = sinh(s->km[e] * zpd) / den;
            us += AA * CC * (ki * cti - km * ctm) * cm * ch * ri;
            vs += AA * CC * (ki * sti - km * stm) * cm * ch * ri;
            ws += AA * CC * s->km[e] * sm * sh * ri;
            ps += AA * CC * (oi - om) * cm * ch * ri;
        }
    }
    free(c);
    o4[0] = us; o4[1] = vs; o4[2] = ws; o4[3] = ps;
}

/* ------------------------------------------- second order, unfiltered direct */
/* Phase of u2/v2/w2/dp2 (Irregular.cpp:456): ... - omega*t + omega*tofmax + phase */
static double arg2(const or_sea* s, int i, double t, double x, double y) {
    const double th = thp(s, i);
    return s->k[i] * (cos(th) * (x - s->x0) + sin(th) * (y - s->y0)) - s->omega[i] * t + s->omega[i] * s->tofmax + s->phase[i];
}

/* D+ / D- recomputed per pair as in u2 (Irregular.cpp:474-486), with pow(.,2.) */
static void direct_D(const or_sea* s, int i, int m, double Rn, double Rsn,
                     double* kpl, double* kmi, double* Dp, double* Dm) {
    const double thi = thp(s, i), thm = thp(s, m);
    const double ki = s->k[i], km = s->k[m];
    const double gamma = cos(thi - thm);
    const double kp = sqrt(ki * ki + km * km + (2. * ki * km * gamma));
    const double kn = sqrt(ki * ki + km * km - (2. * ki * km * gamma));
    const double Rm = km * tanh(km * s->depth);
    const double Rsm = sqrt(Rm);
    const double R2p = pow(Rsn + Rsm, 2.);
    const double R2m = pow(Rsn - Rsm, 2.);
    *Dp = (2. * R2p * (ki * km * gamma - Rn * Rm)) / (R2p - kp * tanh(kp * s->depth)) +
          ((Rsn + Rsm) * (Rsn * (km * km - Rm * Rm) + Rsm * (ki * ki - Rn * Rn))) / (R2p - kp * tanh(kp * s->depth));
    *Dm = (2. * R2m * (ki * km * gamma + Rn * Rm)) / (R2m - kn * tanh(kn * s->depth)) +
          ((Rsn - Rsm) * (-Rsn * (km * km - Rm * Rm) + Rsm * (ki * ki - Rn * Rn))) / (R2m - kn * tanh(kn * s->depth));
    *kpl = kp; *kmi = kn;
}

/* comp: 0 = u2 (:450-506), 1 = v2 (:544-601), 2 = w2 (:638-695), 3 = dp2 (:739-803) */
static double direct2(const or_sea* s, int comp, double t, double x, double y, double z) {
    const int N = s->n;
    const double g = s->g, d = s->depth;
    double tot = 0.0;
    for (int i = 0; i < N; ++i) {
        double acc = 0.0;
        const double thi = thp(s, i);
        const double phi_i = arg2(s, i, t, x, y);
        const double ki = s->k[i], oi = s->omega[i], Ai = s->A[i];
        const double Rn = ki * tanh(ki * d);
        const double Rsn = sqrt(Rn);
        const double Dnn = (6. * Rsn * Rsn * (ki * ki - Rn * Rn)) / (2. * Rsn * Rsn - ki * tanh(2 * ki * d));
        const double AAd = 0.25 * (Ai * g / oi) * (Ai * g / oi) * (Dnn / (2 * oi));
        switch (comp) {
        case 0: acc += AAd * (2 * ki * cos(thi)) * cos(2 * phi_i) * cosh(2 * ki * (z + d)) / cosh(2 * ki * d); break;
        case 1: acc += AAd * (2 * ki * sin(thi)) * cos(2 * phi_i) * cosh(2 * ki * (z + d)) / cosh(2 * ki * d); break;
        case 2: acc += AAd * 2 * ki * sin(2 * phi_i) * sinh(2 * ki * (z + d)) / cosh(2 * ki * d); break;
        default:
            acc += AAd * (2 * oi) * cos(2 * phi_i) * cosh(2 * ki * (z + d)) / cosh(2 * ki * d);
            acc -= (g * g * (2 * ki * cos(thi) + 2 * ki * sin(thi)) * 2 * pX(s, i, z)) / (2 * oi);
            acc += (g * g * (2 * ki) * 2 * pZ(s, i, z)) / (2 * oi);
            break;
        }
        for (int m = i + 1; m < N; ++m) {
            const double thm = thp(s, m);
            const double phi_m = arg2(s, m, t, x, y);
            const double km = s->k[m], om = s->omega[m];
            double kp, kn, Dp, Dm;
            direct_D(s, i, m, Rn, Rsn, &kp, &kn, &Dp, &Dm);
            const double AA = 0.5 * (Ai * g / oi) * (s->A[m] * g / om);
            const double BB = (Dp / (oi + om));
            switch (comp) {
            case 0: {
                const double ch = cosh(kp * (z + d)) / cosh(kp * d);
                acc += AA * BB * (ki * cos(thi) + km * cos(thm)) * cos(phi_i + phi_m) * ch;
                if (ki != km) {
                    const double CC = (Dm / (oi - om));
                    const double chm = cosh(kn * (z + d)) / cosh(kn * d);
                    acc += AA * CC * (ki * cos(thi) - km * cos(thm)) * cos(phi_i - phi_m) * chm;
                }
                break;
            }
            case 1: {
                const double ch = cosh(kp * (z + d)) / cosh(kp * d);
                acc += AA * BB * (ki * sin(thi) + km * sin(thm)) * cos(phi_i + phi_m) * ch;
                if (ki != km) {
                    const double CC = (Dm / (oi - om));
                    const double chm = cosh(kn * (z + d)) / cosh(kn * d);
                    acc += AA * CC * (ki * sin(thi) - km * sin(thm)) * cos(phi_i - phi_m) * chm;
                }
                break;
            }
            case 2: {
                const double sh = sinh(kp * (z + d)) / cosh(kp * d);
                acc += AA * BB * kp * sin(phi_i + phi_m) * sh;
                if (ki != km) {
                    const double CC = (Dm / (oi - om));
                    const double shm = sinh(kn * (z + d)) / cosh(kn * d);
                    acc += AA * CC * kn * sin(phi_i - phi_m) * shm;
                }
                break;
            }
            default: {
                const double ch = cosh(kp * (z + d)) / cosh(kp * d);
                acc += AA * BB * (oi + om) * cos(phi_i + phi_m) * ch;
                acc -= (g * g * (ki * cos(thi) * km * cos(thm) + ki * sin(thi) * km * sin(thm)) * pX(s, i, z) * pX(s, m, z)) / (oi * om);
                acc += (g * g * (ki * km) * pZ(s, i, z) * pZ(s, m, z)) / (oi * om);
                if (ki != km) {
                    const double CC = (Dm / (oi - om));
                    const double chm = cosh(kn * (z + d)) / cosh(kn * d);
                    acc += AA * CC * (oi - om) * cos(phi_i - phi_m) * chm;
                }
                break;
            }
            }
        }
        tot += acc * fr(s, i, t);
    }
    return comp == 3 ? s->rho * tot : tot;
}

/* ------------------------------------------------------------ dispatch (a12) */
/* ------------------------------------- Smit et al. (2017) s-coordinate theory */
/* Stretched profiles, Irregular.cpp:122-141: the vertical argument k (z + h)
 * becomes k h (h + z) / (h + eta) (Wheeler stretching). */
static double sm_pX(const or_sea* s, double kk, double z, double eta) {
    const double sf = s->depth * (s->depth + z) / (s->depth + eta);
    return cosh(kk * sf) / cosh(kk * s->depth);
}
static double sm_pZ(const or_sea* s, double kk, double z, double eta) {
    const double sf = s->depth * (s->depth + z) / (s->depth + eta);
    return sinh(kk * sf) / cosh(kk * s->depth);
}
/* Diagonal H from the Sharma-Dean alpha_nn through Eq. 10, as each of
 * u_smit/v_smit/w_smit/dp_smit writes it (e.g. Irregular.cpp:1672-1678). */
static double sm_Hdiag(const or_sea* s, int i) {
    const double ki = s->k[i], oi = s->omega[i], om2 = 2.0 * oi;
    const double Rni = s->R[i], Rsni = s->Rs[i];
    const double D_nn = (6.0 * Rsni * Rsni * (ki * ki - Rni * Rni)) / (2.0 * Rsni * Rsni - ki * tanh(2.0 * ki * s->depth));
    const double alpha_nn = (D_nn - (ki * ki - Rni * Rni)) / Rni + 2.0 * Rni;
    return -s->g / om2 * (alpha_nn / 2.0 - (3.0 * oi * oi) / (2.0 * s->g) + s->g * ki * ki / (2.0 * oi * oi));
}
static double sm_phase(const or_sea* s, int i, double t, double x, double y) {
    return s->k[i] * (s->cth[i] * (x - s->x0) + s->sth[i] * (y - s->y0)) - s->omega[i] * t + s->ph0[i];
}

/* u_smit / v_smit / w_smit / dp_smit, Irregular.cpp:1623-2027.  comp: 0 u,
 * 1 v, 2 w, 3 dp.  The stretch height is the Sharma-Dean eta1 + eta2
 * (swl excluded); second-order terms for order >= 1. */
static double smit(const or_sea* s, int comp, double t, double x, double y, double zz) {
    const int N = s->n;
    double eta_val = eta1(s, t, x, y);
    if (s->order >= 1) eta_val += eta2(s, t, x, y);
    if (zz > eta_val + s->swl) return 0.0;
    zz = dmax(-s->depth, zz - s->swl);
    const double g = s->g, rho = s->rho, depth = s->depth;
    double v1 = 0.0;
    for (int i = 0; i < N; ++i) {
        const double phi = sm_phase(s, i, t, x, y);
        const double f = fr(s, i, t);
        if (comp == 0) v1 += s->A[i] * g * s->k[i] / s->omega[i] * sm_pX(s, s->k[i], zz, eta_val) * s->cth[i] * cos(phi) * f;
        else if (comp == 1) v1 += s->A[i] * g * s->k[i] / s->omega[i] * sm_pX(s, s->k[i], zz, eta_val) * s->sth[i] * cos(phi) * f;
        else if (comp == 2) v1 += s->A[i] * g * s->k[i] / s->omega[i] * sm_pZ(s, s->k[i], zz, eta_val) * sin(phi) * f;
        else v1 += s->A[i] * rho * g * sm_pX(s, s->k[i], zz, eta_val) * cos(phi) * f;
    }
    if (s->order < 1) return v1;
    double v2 = 0.0, dphi2ds = 0.0;
    const double sp1 = (depth + zz) / (depth + eta_val), g1s = g * sp1;
    for (int i = 0; i < N; ++i) {                                       /* diagonal */
        const double phi_i = sm_phase(s, i, t, x, y);
        const double f = fr(s, i, t);
        const double ki = s->k[i], oi = s->omega[i];
        const double AA = s->A[i] * s->A[i] * f * f;
        const double k2 = 2.0 * ki, om2 = 2.0 * oi;
        const double H_diag = sm_Hdiag(s, i);
        if (comp == 0) {                                                /* :1657-1689 */
            const double kx2 = 2.0 * ki * s->cth[i];
            const double Ch2 = sm_pX(s, k2, zz, eta_val);
            v2 += 0.5 * AA * H_diag * kx2 * (-cos(2.0 * phi_i)) * Ch2;
            const double Sh_i = sm_pZ(s, ki, zz, eta_val);
            v2 += 0.5 * AA * kx2 * g1s * (ki * Sh_i / oi) * cos(2.0 * phi_i);
            const double Ch_i = sm_pX(s, ki, zz, eta_val);
            const double Sh2 = sm_pZ(s, k2, zz, eta_val);
            dphi2ds += -0.5 * AA * H_diag * (k2 * depth * Sh2) * sin(2.0 * phi_i);
            dphi2ds += 0.5 * AA * g * (ki * Sh_i / oi + sp1 * ki * (ki * depth * Ch_i) / oi) * sin(2.0 * phi_i);
        } else if (comp == 1) {                                         /* :1797-1799 */
            const double ky2 = 2.0 * ki * s->sth[i];
            const double Ch2 = sm_pX(s, k2, zz, eta_val);
            v2 += 0.5 * AA * H_diag * ky2 * (-cos(2.0 * phi_i)) * Ch2;
        } else if (comp == 2) {                                         /* :1888-1890 */
            const double Sh2 = sm_pZ(s, k2, zz, eta_val);
            v2 += 0.5 * AA * H_diag * k2 * (-sin(2.0 * phi_i)) * Sh2;
        } else {                                                        /* :1961-1967 */
            const double Ch2 = sm_pX(s, k2, zz, eta_val);
            v2 += 0.5 * AA * rho * om2 * H_diag * (-cos(2.0 * phi_i)) * Ch2;
            const double Ch_i = sm_pX(s, ki, zz, eta_val);
            const double Sh_i = sm_pZ(s, ki, zz, eta_val);
            v2 -= 0.5 * AA * rho * g * g * ki * ki * (Ch_i * Ch_i - Sh_i * Sh_i) / (oi * oi) * cos(2.0 * phi_i);
        }
    }
    for (long e = 0; e < s->np; ++e) {                                  /* pairs */
        const int i = s->pi[e], m = s->pm[e];
        const double Ai = s->A[i], Am = s->A[m], ki = s->k[i], km = s->k[m];
        const double oi = s->omega[i], om = s->omega[m];
        const double phi_i = sm_phase(s, i, t, x, y), phi_m = sm_phase(s, m, t, x, y);
        const double AA = Ai * Am * fr(s, i, t) * fr(s, m, t);
        const double Sh_i = sm_pZ(s, ki, zz, eta_val), Sh_m = sm_pZ(s, km, zz, eta_val);
        const double Ch_i = sm_pX(s, ki, zz, eta_val), Ch_m = sm_pX(s, km, zz, eta_val);
        for (int br = 0; br < 2; ++br) {
            if (!(s->pflag[e] & (1 << br))) continue;
            const double sg = br == 0 ? 1.0 : -1.0;                     /* k2 -> -k2, omega2 -> -omega2 */
            const double ph = br == 0 ? phi_i + phi_m : phi_i - phi_m;
            const double H = br == 0 ? s->Hp[e] : s->Hm[e];
            const double knm = br == 0 ? s->kp[e] : s->km[e];
            if (comp == 0) {                                            /* :1718-1746 */
                const double kx12 = br == 0 ? ki * s->cth[i] + km * s->cth[m] : ki * s->cth[i] - km * s->cth[m];
                const double Ch12 = sm_pX(s, knm, zz, eta_val);
                v2 += AA * H * kx12 * (-cos(ph)) * Ch12;
                const double a1 = br == 0 ? ki * Sh_i / oi + km * Sh_m / om : ki * Sh_i / oi - km * Sh_m / om;
                v2 += AA * kx12 * g1s * a1 * cos(ph);
                const double Sh12 = sm_pZ(s, knm, zz, eta_val);
                dphi2ds += -AA * H * (knm * depth * Sh12) * sin(ph);
                const double a2 = br == 0 ? ki * (ki * depth * Ch_i) / oi + km * (km * depth * Ch_m) / om
                                          : ki * (ki * depth * Ch_i) / oi - km * (km * depth * Ch_m) / om;
                dphi2ds += AA * g * (a1 + sp1 * a2) * sin(ph);
            } else if (comp == 1) {                                     /* :1832-1844 */
                const double ky12 = br == 0 ? ki * s->sth[i] + km * s->sth[m] : ki * s->sth[i] - km * s->sth[m];
                const double Ch12 = sm_pX(s, knm, zz, eta_val);
                v2 += AA * H * ky12 * (-cos(ph)) * Ch12;
            } else if (comp == 2) {                                     /* :1911-1921 */
                const double Sh12 = sm_pZ(s, knm, zz, eta_val);
                v2 += AA * H * knm * (-sin(ph)) * Sh12;
            } else {                                                    /* :1986-2021 */
                const double om12 = oi + sg * om;
                const double Ch12 = sm_pX(s, knm, zz, eta_val);
                v2 += AA * rho * om12 * H * (-cos(ph)) * Ch12;
                const double k1dk2 = ki * km * s->gam[e];
                const double br2 = br == 0 ? Ch_i * Ch_m - Sh_i * Sh_m : Ch_i * Ch_m + Sh_i * Sh_m;
                v2 -= AA * rho * g * g * k1dk2 * br2 / (oi * om) * cos(ph);
            }
        }
    }
    if (comp == 0) {                                                    /* chain term, :1752-1769 */
        double eta_x = 0.0, phi_s = 0.0;
        for (int n = 0; n < N; ++n) {
            const double phi_n = sm_phase(s, n, t, x, y);
            const double f = fr(s, n, t);
            const double sphi = sin(phi_n);
            eta_x += -s->A[n] * s->k[n] * s->cth[n] * sphi * f;
            phi_s += (g * s->A[n] / s->omega[n]) * s->k[n] * depth * sm_pZ(s, s->k[n], zz, eta_val) * sphi * f;
        }
        phi_s += dphi2ds;
        const double oneplus_s_over_D = (depth + zz) / ((depth + eta_val) * (depth + eta_val));
        v2 += -oneplus_s_over_D * eta_x * phi_s;
    }
    return v1 + v2;
}

static double eval1(const or_sea* s, int kind, double t, double x, double y, double z) {
    double o4[4];
    switch (kind) {
    case OR_ETA1: return eta1(s, t, x, y);
    case OR_ETA2: return eta2(s, t, x, y);
    case OR_ETA:                                                      /* :9-22 */
        if (s->order == 2) return (s->swl + eta1(s, t, x, y) + eta2(s, t, x, y));
        return (s->swl + eta1(s, t, x, y));
    case OR_U1: return u1(s, t, x, y, z);
    case OR_V1: return v1(s, t, x, y, z);
    case OR_W1: return w1(s, t, x, y, z);
    case OR_DP1: return dp1(s, t, x, y, z);
    case OR_PHI1_DXDZ: return phi1_dxdz(s, t, x, y);
    case OR_PHI1_DYDZ: return phi1_dydz(s, t, x, y);
    case OR_PHI1_DZDZ: return phi1_dzdz(s, t, x, y);
    case OR_DP1_DZ: return dp1_dz(s, t, x, y);
    case OR_U2: return direct2(s, 0, t, x, y, z);
    case OR_V2: return direct2(s, 1, t, x, y, z);
    case OR_W2: return direct2(s, 2, t, x, y, z);
    case OR_DP2: return direct2(s, 3, t, x, y, z);
    case OR_U_SMIT: return smit(s, 0, t, x, y, z);
    case OR_V_SMIT: return smit(s, 1, t, x, y, z);
    case OR_W_SMIT: return smit(s, 2, t, x, y, z);
    case OR_DP_SMIT: return smit(s, 3, t, x, y, z);
    case OR_U: case OR_V: case OR_W: case OR_DP: {                   /* :24-157 */
        if (s->vt == 1) return smit(s, kind - OR_U, t, x, y, z);    /* :25, :53, :77, :143 */
        if (s->order == 2) {
            const double z0 = dmin(0., z - s->swl), za = dmax(0., z - s->swl);
            switch (kind) {
            case OR_U: return (u1(s, t, x, y, z0) + direct2(s, 0, t, x, y, z0) + phi1_dxdz(s, t, x, y) * za);
            case OR_V: return (v1(s, t, x, y, z0) + direct2(s, 1, t, x, y, z0) + phi1_dydz(s, t, x, y) * za);
            case OR_W: return (w1(s, t, x, y, z0) + direct2(s, 2, t, x, y, z0) + phi1_dzdz(s, t, x, y) * za);
            default: return (dp1(s, t, x, y, z0) + direct2(s, 3, t, x, y, z0) + dp1_dz(s, t, x, y) * za);
            }
        }
        if (s->order == 1) {
            const double z1 = dmin(0., z - s->swl);
            switch (kind) {
            case OR_U: return (u1(s, t, x, y, z1) + direct2(s, 0, t, x, y, z1));
            case OR_V: return (v1(s, t, x, y, z1) + direct2(s, 1, t, x, y, z1));
            case OR_W: return (w1(s, t, x, y, z1) + direct2(s, 2, t, x, y, z1));
            default: return (dp1(s, t, x, y, z1) + direct2(s, 3, t, x, y, z1));
            }
        }
        switch (kind) {
        case OR_U: return u1(s, t, x, y, z - s->swl);
        case OR_V: return v1(s, t, x, y, z - s->swl);
        case OR_W: return w1(s, t, x, y, z - s->swl);
        default: return dp1(s, t, x, y, z - s->swl);
        }
    }
    case OR_UVW2_U: uvw2(s, t, x, y, z, o4); return o4[0];
    case OR_UVW2_V: uvw2(s, t, x, y, z, o4); return o4[1];
    case OR_UVW2_W: uvw2(s, t, x, y, z, o4); return o4[2];
    case OR_UVW2_P: uvw2(s, t, x, y, z, o4); return o4[3];
    default: return 0.0;
    }
}

void or_eval(const or_sea* s, int kind, long n, const double* t, const double* x,
             const double* y, const double* z, double* out) {
    #pragma omp parallel for schedule(dynamic, 1)
    for (long p = 0; p < n; ++p) out[p] = eval1(s, kind, t[p], x[p], y[p], z ? z[p] : 0.0);
}

/* ===================================================================== s-grid */
/* lsGridSpline transforms, lsgrid_spline.cpp:89-122 */
static double z2s(double z, double e, double d, double swl) { return (z - (e - swl)) / (d + (e - swl)); }
static double s2z(double s, double e, double d, double swl) { return (e - swl) + s * (d + (e - swl)); }
double or_cart2sigmaS(double zpt, double e, double d, double swl, double a, double b) {
    const double z = dmin(zpt, e);
    double sigma = dmax(z2s(z, e, d, swl), -1.);
    /* With swl > 0, points in (e - swl, e] give sigma > 0, and the reference
     * takes pow() of a negative base: a NaN cell index and an out-of-bounds
     * read.  That band is mapped to the surface (sigma = 0) here. */
    if (sigma > 0.) sigma = 0.;
    return 1 - atan(pow(-sigma * pow(tan(a), b), 1. / b)) / a;
}
double or_sigmaS2cart(double s, double e, double d, double swl, double a, double b) {
    const double sigma = -pow(tan(-(s - 1) * a), b) / pow(tan(a), b);
    return s2z(sigma, e, d, swl);
}

struct or_grid {
    const or_sea* sea;
    int nx, ny, nl, tstep;
    double dom[4], domi[4], t0, dt, a, b, depth, swl, dx, dy, ds;
    int* ign;
    double *ETA, *ETAdx, *ETAdy, *ETAdt;
    double *F[3], *Fdt[3], *Fdx[3], *Fdy[3], *Fds[3];   /* UX/UY/UZ and derivatives */
};

/* The column body of initialize_kinematics / update (lsgrid_spline.cpp:697-722):
 * eta = eta1+eta2+swl, three Taylor gradients, then per level uvw2 + u1/v1/w1. */
static void column_body(const or_sea* s, double t, double x, double y, int nl, double ds,
                        double a, double b, double depth, double swl,
                        double* eta_out, double* ux, double* uy, double* uz) {
    const double eta = eta1(s, t, x, y) + eta2(s, t, x, y) + swl;
    *eta_out = eta;
    const double gx = phi1_dxdz(s, t, x, y), gy = phi1_dydz(s, t, x, y), gz = phi1_dzdz(s, t, x, y);
    for (int m = 0; m < nl; ++m) {
        const double z = or_sigmaS2cart(ds * m, eta, depth, swl, a, b);
        const double zmax = dmax(0., z - swl);
        const double zmin = dmin(0., z);          /* no -swl here (lsgrid_spline.cpp:711) */
        double o4[4];
        uvw2(s, t, x, y, zmin, o4);
        ux[m] = u1(s, t, x, y, zmin) + o4[0] + gx * zmax;
        uy[m] = v1(s, t, x, y, zmin) + o4[1] + gy * zmax;
        uz[m] = w1(s, t, x, y, zmin) + o4[2] + gz * zmax;
    }
}

static void fill_slot(or_grid* g, int slot, double t, int use_ignore) {
    const long c = (long)g->nx * g->ny;
    #pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int i = 0; i < g->nx; ++i)
        for (int j = 0; j < g->ny; ++j) {
            if (use_ignore && g->ign[i * g->ny + j]) continue;
            const double x = g->dom[0] + g->dx * i, y = g->dom[2] + g->dy * j;
            const long col = (long)i * g->ny + j;
            column_body(g->sea, t, x, y, g->nl, g->ds, g->a, g->b, g->depth, g->swl,
                        &g->ETA[slot * c + col], &g->F[0][(slot * c + col) * g->nl],
                        &g->F[1][(slot * c + col) * g->nl], &g->F[2][(slot * c + col) * g->nl]);
        }
}

/* update_gradient_dt / update_gradient_eta_dt, lsgrid_spline.cpp:124-161 */
static void grad_dt(const or_grid* g, const double* d, double* gd, long per) {
    if (g->tstep > 0) {
        const int tid = (g->tstep + 2) % 4, t1 = (g->tstep + 3) % 4, t0 = (g->tstep + 1) % 4;
        for (long i = 0; i < per; ++i) gd[tid * per + i] = (d[t1 * per + i] - d[t0 * per + i]) / (2 * g->dt);
    } else {
        for (long i = 0; i < per; ++i) {
            gd[1 * per + i] = (d[2 * per + i] - d[1 * per + i]) / g->dt;
            gd[2 * per + i] = (d[3 * per + i] - d[1 * per + i]) / (2 * g->dt);
        }
    }
}

/* update_gradient_eta_dxdy (:163-229) and update_gradient_dxdydz (:231-366):
 * central differences inside, one-sided at the boundaries, zero for a single
 * point in that direction.  Output slots 0/1 hold ring slots (tstep+1)%4 and (tstep+2)%4. */
static void grad_xyz(const or_grid* g, const double* d, double* gx, double* gy, double* gs, int L) {
    const int nx = g->nx, ny = g->ny;
    const long c = (long)nx * ny, per = c * L;
    const int tids[2] = {(g->tstep + 1) % 4, (g->tstep + 2) % 4};
    for (int o = 0; o < 2; ++o) {
        const double* D = d + tids[o] * per;
        double* GX = gx + o * per; double* GY = gy + o * per; double* GS = gs ? gs + o * per : NULL;
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int m = 0; m < L; ++m) {
                    const long q = ((long)i * ny + j) * L + m;
                    double vx, vy;
                    if (nx == 1) vx = 0.;
                    else if (i == 0) vx = (D[q + (long)ny * L] - D[q]) / g->dx;
                    else if (i == nx - 1) vx = (D[q] - D[q - (long)ny * L]) / g->dx;
                    else vx = (D[q + (long)ny * L] - D[q - (long)ny * L]) / (2 * g->dx);
                    if (ny == 1) vy = 0.;
                    else if (j == 0) vy = (D[q + L] - D[q]) / g->dy;
                    else if (j == ny - 1) vy = (D[q] - D[q - L]) / g->dy;
                    else vy = (D[q + L] - D[q - L]) / (2 * g->dy);
                    GX[q] = vx; GY[q] = vy;
                    if (GS) {
                        double vs;
                        if (L == 1) vs = 0.;
                        else if (m == 0) vs = (D[q + 1] - D[q]) / g->ds;
                        else if (m == L - 1) vs = (D[q] - D[q - 1]) / g->ds;
                        else vs = (D[q + 1] - D[q - 1]) / (2 * g->ds);
                        GS[q] = vs;
                    }
                }
    }
}

static void grid_gradients(or_grid* g) {
    const long c = (long)g->nx * g->ny, v = c * g->nl;
    grad_xyz(g, g->ETA, g->ETAdx, g->ETAdy, NULL, 1);
    grad_dt(g, g->ETA, g->ETAdt, c);
    for (int f = 0; f < 3; ++f) {
        grad_xyz(g, g->F[f], g->Fdx[f], g->Fdy[f], g->Fds[f], g->nl);
        grad_dt(g, g->F[f], g->Fdt[f], v);
    }
}

or_grid* or_grid_create(const or_sea* s, const double* domain, int nx, int ny, int nl,
                        double t0, double dt, double tan_a, double tan_b,
                        double water_depth, double swl, const double* domain_ignore,
                        int ignore_at_init) {
    or_grid* g = (or_grid*)calloc(1, sizeof(or_grid));
    g->sea = s; g->nx = nx; g->ny = ny; g->nl = nl; g->t0 = t0; g->dt = dt;
    g->a = tan_a; g->b = tan_b; g->depth = water_depth; g->swl = swl; g->tstep = 0;
    for (int q = 0; q < 4; ++q) { g->dom[q] = domain[q]; g->domi[q] = domain_ignore[q]; }
    const long c = (long)nx * ny, v = c * nl;
    g->ETA = (double*)calloc(4 * c, sizeof(double)); g->ETAdt = (double*)calloc(4 * c, sizeof(double));
    g->ETAdx = (double*)calloc(2 * c, sizeof(double)); g->ETAdy = (double*)calloc(2 * c, sizeof(double));
    for (int f = 0; f < 3; ++f) {
        g->F[f] = (double*)calloc(4 * v, sizeof(double)); g->Fdt[f] = (double*)calloc(4 * v, sizeof(double));
        g->Fdx[f] = (double*)calloc(2 * v, sizeof(double)); g->Fdy[f] = (double*)calloc(2 * v, sizeof(double));
        g->Fds[f] = (double*)calloc(2 * v, sizeof(double));
    }
    g->ign = (int*)calloc(c, sizeof(int));
    /* set_ignore (:1406-1425) */
    g->dx = (g->dom[1] - g->dom[0]) / dmax(1., (double)(nx - 1));
    g->dy = (g->dom[3] - g->dom[2]) / dmax(1., (double)(ny - 1));
    for (int i = 0; i < nx; ++i) {
        const double xp = g->dom[0] + g->dx * i;
        for (int j = 0; j < ny; ++j) {
            const double yp = g->dom[2] + g->dy * j;
            g->ign[i * ny + j] = (xp >= g->domi[0] && xp <= g->domi[1] && yp >= g->domi[2] && yp <= g->domi[3]) ? 1 : 0;
        }
    }
    /* initialize_kinematics (:666-753) / _with_ignore (:755-844) */
    g->ds = 1. / dmax(1., (double)(nl - 1));
    const double tt0a[4] = {t0 - dt, t0, t0 + dt, t0 + 2 * dt};
    const double tt0b[4] = {t0, t0, t0 + dt, t0 + 2 * dt};
    for (int tt = 0; tt < 4; ++tt) fill_slot(g, tt, ignore_at_init ? tt0b[tt] : tt0a[tt], ignore_at_init);
    grid_gradients(g);
    return g;
}

void or_grid_destroy(or_grid* g) {
    if (!g) return;
    free(g->ETA); free(g->ETAdt); free(g->ETAdx); free(g->ETAdy); free(g->ign);
    for (int f = 0; f < 3; ++f) { free(g->F[f]); free(g->Fdt[f]); free(g->Fdx[f]); free(g->Fdy[f]); free(g->Fds[f]); }
    free(g);
}

/* wave_force_update -> CheckTime / update (CFDwavemaker.cpp:2253-2261,
 * lsgrid_spline.cpp:610-617, :847-931) */
int or_grid_force_update(or_grid* g, double t) {
    int n = 0;
    while (t > g->t0 + g->dt) {
        g->tstep++;
        g->t0 += g->dt;
        const double t2 = g->t0 + 2 * g->dt;
        const int slot = (g->tstep + 3) % 4;
        fill_slot(g, slot, t2, 1);
        grid_gradients(g);
        ++n;
    }
    return n;
}

void or_grid_state(const or_grid* g, int* tstep, double* t0, double* dx, double* dy, double* ds) {
    *tstep = g->tstep; *t0 = g->t0; *dx = g->dx; *dy = g->dy; *ds = g->ds;
}

long or_grid_export(const or_grid* g, int which, double* out) {
    const long c = (long)g->nx * g->ny, v = c * g->nl;
    const double* src = NULL; long n = 0;
    if (which == 0) { src = g->ETA; n = 4 * c; }
    else if (which == 1) { src = g->ETAdx; n = 2 * c; }
    else if (which == 2) { src = g->ETAdy; n = 2 * c; }
    else if (which == 3) { src = g->ETAdt; n = 4 * c; }
    else if (which >= 4 && which <= 6) { src = g->F[which - 4]; n = 4 * v; }
    else if (which >= 7 && which <= 9) { src = g->Fdt[which - 7]; n = 4 * v; }
    else if (which >= 10 && which <= 18) {
        const int f = (which - 10) / 3, q = (which - 10) % 3;
        src = q == 0 ? g->Fdx[f] : (q == 1 ? g->Fdy[f] : g->Fds[f]); n = 2 * v;
    } else if (which == 19) {
        if (out) for (long q = 0; q < c; ++q) out[q] = g->ign[q];
        return c;
    }
    if (out && src) memcpy(out, src, n * sizeof(double));
    return n;
}

static int iclamp(int n, int lo, int hi) { return n < lo ? lo : (n > hi ? hi : n); }

/* square_vals / cube_vals (:368-384) */
static void sq(const or_grid* g, double* C, const double* d, int i, int j, int tid) {
    const int nx = g->nx, ny = g->ny;
    C[0] = d[tid * nx * ny + i * ny + j];
    C[1] = d[tid * nx * ny + i * ny + iclamp(j + 1, 0, ny - 1)];
    C[2] = d[tid * nx * ny + iclamp(i + 1, 0, nx - 1) * ny + j];
    C[3] = d[tid * nx * ny + iclamp(i + 1, 0, nx - 1) * ny + iclamp(j + 1, 0, ny - 1)];
}
static void cube(const or_grid* g, double* C, const double* d, int i, int j, int l, int tid) {
    const long nx = g->nx, ny = g->ny, nl = g->nl;
    const long base = tid * nx * ny * nl;
    const long i1 = iclamp(i + 1, 0, (int)nx - 1), j1 = iclamp(j + 1, 0, (int)ny - 1), l1 = iclamp(l + 1, 0, (int)nl - 1);
    C[0] = d[base + i * ny * nl + j * nl + l];
    C[1] = d[base + i * ny * nl + j1 * nl + l];
    C[2] = d[base + i1 * ny * nl + j * nl + l];
    C[3] = d[base + i1 * ny * nl + j1 * nl + l];
    C[4] = d[base + i * ny * nl + j * nl + l1];
    C[5] = d[base + i * ny * nl + j1 * nl + l1];
    C[6] = d[base + i1 * ny * nl + j * nl + l1];
    C[7] = d[base + i1 * ny * nl + j1 * nl + l1];
}

/* cubic Hermite on [0,1]: f0(1-u) + f1 u + u(1-u)(a(1-u) + b u), written as in :441-448 */
static double herm(double f0, double f1, double a, double b, double u) {
    return f0 * (1. - u) + f1 * u + u * (1 - u) * (a * (1 - u) + b * u);
}
static double lin(double f0, double f1, double u) { return f0 * (1. - u) + f1 * u; }

/* spline_interp_velo (:386-502) */
static double spline_velo(const or_grid* g, int f, int i, int j, int l0, int l1,
                          double xd, double yd, double s0, double s1, double td, int tid1, int tid2) {
    double C[8], D[8], Ct[8], Dt[8], Cx[8], Dx[8], Cy[8], Dy[8], Cs[8], Ds[8];
    cube(g, C, g->F[f], i, j, l0, tid1);     cube(g, D, g->F[f], i, j, l1, tid2);
    cube(g, Ct, g->Fdt[f], i, j, l0, tid1);  cube(g, Dt, g->Fdt[f], i, j, l1, tid2);
    cube(g, Cx, g->Fdx[f], i, j, l0, 0);     cube(g, Dx, g->Fdx[f], i, j, l1, 1);
    cube(g, Cy, g->Fdy[f], i, j, l0, 0);     cube(g, Dy, g->Fdy[f], i, j, l1, 1);
    cube(g, Cs, g->Fds[f], i, j, l0, 0);     cube(g, Ds, g->Fds[f], i, j, l1, 1);
    const double dx = g->dx, dy = g->dy, ds = g->ds, dt = g->dt;
    /* x reduction: corners (y,s) = 00,10,01,11 map to indices (0,2),(1,3),(4,6),(5,7) */
    const int lo[4] = {0, 1, 4, 5}, hi[4] = {2, 3, 6, 7};
    double Cxr[4], Dxr[4], Cyr[4], Dyr[4], Ctr[4], Dtr[4], Csr[4], Dsr[4];
    for (int q = 0; q < 4; ++q) {
        const double ac = dx * Cx[lo[q]] - (C[hi[q]] - C[lo[q]]);
        const double bc = -dx * Cx[hi[q]] + (C[hi[q]] - C[lo[q]]);
        const double ad = dx * Dx[lo[q]] - (D[hi[q]] - D[lo[q]]);
        const double bd = -dx * Dx[hi[q]] + (D[hi[q]] - D[lo[q]]);
        Cxr[q] = herm(C[lo[q]], C[hi[q]], ac, bc, xd);
        Dxr[q] = herm(D[lo[q]], D[hi[q]], ad, bd, xd);
        Cyr[q] = lin(Cy[lo[q]], Cy[hi[q]], xd); Dyr[q] = lin(Dy[lo[q]], Dy[hi[q]], xd);
        Ctr[q] = lin(Ct[lo[q]], Ct[hi[q]], xd); Dtr[q] = lin(Dt[lo[q]], Dt[hi[q]], xd);
        Csr[q] = lin(Cs[lo[q]], Cs[hi[q]], xd); Dsr[q] = lin(Ds[lo[q]], Ds[hi[q]], xd);
    }
    /* y reduction: (s=0: q0,q1), (s=1: q2,q3) */
    double Cyv[2], Dyv[2], Csv[2], Dsv[2], Ctv[2], Dtv[2];
    for (int q = 0; q < 2; ++q) {
        const int a0 = 2 * q, a1 = 2 * q + 1;
        const double ac = dy * Cyr[a0] - (Cxr[a1] - Cxr[a0]);
        const double bc = -dy * Cyr[a1] + (Cxr[a1] - Cxr[a0]);
        const double ad = dy * Dyr[a0] - (Dxr[a1] - Dxr[a0]);
        const double bd = -dy * Dyr[a1] + (Dxr[a1] - Dxr[a0]);
        Cyv[q] = herm(Cxr[a0], Cxr[a1], ac, bc, yd);
        Dyv[q] = herm(Dxr[a0], Dxr[a1], ad, bd, yd);
        Csv[q] = lin(Csr[a0], Csr[a1], yd); Dsv[q] = lin(Dsr[a0], Dsr[a1], yd);
        Ctv[q] = lin(Ctr[a0], Ctr[a1], yd); Dtv[q] = lin(Dtr[a0], Dtr[a1], yd);
    }
    /* s reduction */
    const double a0 = ds * Csv[0] - (Cyv[1] - Cyv[0]);
    const double b0 = -ds * Csv[1] + (Cyv[1] - Cyv[0]);
    const double a1 = ds * Dsv[0] - (Dyv[1] - Dyv[0]);
    const double b1 = -ds * Dsv[1] + (Dyv[1] - Dyv[0]);
    const double CC = herm(Cyv[0], Cyv[1], a0, b0, s0);
    const double DD = herm(Dyv[0], Dyv[1], a1, b1, s1);
    const double CCt = lin(Ctv[0], Ctv[1], s0);
    const double DDt = lin(Dtv[0], Dtv[1], s1);
    /* t reduction */
    const double a = dt * CCt - (DD - CC);
    const double b = -dt * DDt + (DD - CC);
    return herm(CC, DD, a, b, td);
}

/* get_kinematics_at_point (:505-605).  The cell fractions go through
 * std::modf(double, float*), which resolves to the float overload: the
 * argument is rounded to float first, so the fraction has float precision. */
static void query1(const or_grid* g, double tpt, double xpt, double ypt, double zpt, double* o) {
    const int tid1 = (g->tstep + 1) % 4, tid2 = (g->tstep + 2) % 4;
    float nxpf, nypf;
    double cx = (xpt - g->dom[0]) / g->dx; cx = dmax(0., dmin(cx, g->nx - 1.));
    double cy = (ypt - g->dom[2]) / g->dy; cy = dmax(0., dmin(cy, g->ny - 1.));
    const double xd = modff((float)cx, &nxpf);
    const double yd = modff((float)cy, &nypf);
    const double td = dmin(1., dmax(0., (tpt - g->t0) / g->dt));
    const int i = (int)nxpf, j = (int)nypf;
    double C[4], D[4], Ct[4], Dt[4], Cx[4], Dx[4], Cy[4], Dy[4];
    sq(g, C, g->ETA, i, j, tid1);    sq(g, D, g->ETA, i, j, tid2);
    sq(g, Ct, g->ETAdt, i, j, tid1); sq(g, Dt, g->ETAdt, i, j, tid2);
    sq(g, Cx, g->ETAdx, i, j, 0);    sq(g, Dx, g->ETAdx, i, j, 1);
    sq(g, Cy, g->ETAdy, i, j, 0);    sq(g, Dy, g->ETAdy, i, j, 1);
    const double dx = g->dx, dy = g->dy, dt = g->dt;
    const double a00 = dx * Cx[0] - (C[2] - C[0]), a01 = dx * Cx[1] - (C[3] - C[1]);
    const double b00 = -dx * Cx[2] + (C[2] - C[0]), b01 = -dx * Cx[3] + (C[3] - C[1]);
    const double a10 = dx * Dx[0] - (D[2] - D[0]), a11 = dx * Dx[1] - (D[3] - D[1]);
    const double b10 = -dx * Dx[2] + (D[2] - D[0]), b11 = -dx * Dx[3] + (D[3] - D[1]);
    const double C0 = herm(C[0], C[2], a00, b00, xd), C1 = herm(C[1], C[3], a01, b01, xd);
    const double D0 = herm(D[0], D[2], a10, b10, xd), D1 = herm(D[1], D[3], a11, b11, xd);
    const double C0y = lin(Cy[0], Cy[2], xd), C1y = lin(Cy[1], Cy[3], xd);
    const double D0y = lin(Dy[0], Dy[2], xd), D1y = lin(Dy[1], Dy[3], xd);
    const double C0t = lin(Ct[0], Ct[2], xd), C1t = lin(Ct[1], Ct[3], xd);
    const double D0t = lin(Dt[0], Dt[2], xd), D1t = lin(Dt[1], Dt[3], xd);
    const double a0 = dy * C0y - (C1 - C0), b0 = -dy * C1y + (C1 - C0);
    const double a1 = dy * D0y - (D1 - D0), b1 = -dy * D1y + (D1 - D0);
    const double e0 = herm(C0, C1, a0, b0, yd), e1 = herm(D0, D1, a1, b1, yd);
    const double CCt = lin(C0t, C1t, yd), DDt = lin(D0t, D1t, yd);
    const double a = dt * CCt - (e1 - e0), b = -dt * DDt + (e1 - e0);
    const double eta = herm(e0, e1, a, b, td);
    const double sp0 = or_cart2sigmaS(zpt, e0, g->depth, g->swl, g->a, g->b);
    const double sp1 = or_cart2sigmaS(zpt, e1, g->depth, g->swl, g->a, g->b);
    float s0f, s1f;
    const double sd0 = modff((float)(sp0 / g->ds), &s0f);
    const double sd1 = modff((float)(sp1 / g->ds), &s1f);
    const int l0 = (int)s0f, l1 = (int)s1f;
    o[0] = eta;
    for (int f = 0; f < 3; ++f) o[1 + f] = spline_velo(g, f, i, j, l0, l1, xd, yd, sd0, sd1, td, tid1, tid2);
}

void or_grid_query(const or_grid* g, long n, const double* t, const double* x, const double* y,
                   const double* z, double* out4) {
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < n; ++p) query1(g, t[p], x[p], y[p], z[p], out4 + 4 * p);
}

double or_column_fill(const or_sea* s, long ncol, const double* x, const double* y, double t,
                      int nl, double tan_a, double tan_b, double water_depth, double swl,
                      double* eta, double* ux, double* uy, double* uz) {
    const double ds = 1. / dmax(1., (double)(nl - 1));
    const double t_start = omp_get_wtime();
    #pragma omp parallel for schedule(dynamic, 1)
    for (long c = 0; c < ncol; ++c)
        column_body(s, t, x[c], y[c], nl, ds, tan_a, tan_b, water_depth, swl,
                    &eta[c], &ux[c * nl], &uy[c * nl], &uz[c * nl]);
    return omp_get_wtime() - t_start;
}

/* ============================================================ legacy lsGrid */
/* Two time slots (slot 0 at t0, slot 1 at t0 + dt), filled by the same column
 * body as the spline grid, and queried by tri/bilinear interpolation
 * (src/lsgrid.h:15-114, src/lsgrid.cpp).  With init_only, slot 1 is the same
 * storage as slot 0 (lsgrid.cpp:466-470, 665-667: the arrays alias). */
struct or_lsgrid {
    const or_sea* sea;
    int nx, ny, nl, init_only;
    double dom[4], domi[4], t0, dt, a, b, depth, swl, dx, dy, ds;
    int* ign;
    double *ETA0, *ETA1, *F0[3], *F1[3];
};

/* lsGrid::s2tan / tan2s / z2s / s2z (lsgrid.cpp:82-101): the swl-free maps */
static double ls_s2tan(const or_lsgrid* g, double s) { return -pow(tan(-s * g->a), g->b) / pow(tan(g->a), g->b); }
static double ls_tan2s(const or_lsgrid* g, double t) { return -atan(pow(-t * pow(tan(g->a), g->b), 1. / g->b)) / g->a; }
static double ls_z2s(double z, double we, double d) { return (z - we) / (d + we); }
static double ls_s2z(double s, double we, double d) { return we + s * (d + we); }

/* the level loop of initialize_kinematics (lsgrid.cpp:445-459) for one column
 * at time t whose surface slot holds eta (incl. swl) */
static void ls_levels(const or_lsgrid* g, double t, double x, double y, double eta, double* ux, double* uy, double* uz) {
    const or_sea* s = g->sea;
    const double et = eta - g->swl;
    const double gx = phi1_dxdz(s, t, x, y), gy = phi1_dydz(s, t, x, y), gz = phi1_dzdz(s, t, x, y);
    for (int m = 0; m < g->nl; ++m) {
        const double spt = ls_s2tan(g, -1. + g->ds * m);
        const double z = ls_s2z(spt, et, g->depth);
        const double zmax = dmax(0., z - g->swl);
        const double zmin = dmin(0., z);
        double o4[4];
        uvw2(s, t, x, y, zmin, o4);
        ux[m] = u1(s, t, x, y, zmin) + o4[0] + gx * zmax;
        uy[m] = v1(s, t, x, y, zmin) + o4[1] + gy * zmax;
        uz[m] = w1(s, t, x, y, zmin) + o4[2] + gz * zmax;
    }
}

/* initialize_surface_elevation[_with_ignore] (lsgrid.cpp:622-747) */
static void ls_surface(or_lsgrid* g, double t_target, int use_ignore) {
    g->t0 = t_target;
    #pragma omp parallel for collapse(2)
    for (int i = 0; i < g->nx; ++i)
        for (int j = 0; j < g->ny; ++j) {
            const long c = (long)i * g->ny + j;
            const double x = g->dom[0] + g->dx * i, y = g->dom[2] + g->dy * j;
            if (use_ignore && g->ign[c]) { g->ETA0[c] = g->swl; continue; }
            g->ETA0[c] = eta1(g->sea, g->t0, x, y) + eta2(g->sea, g->t0, x, y) + g->swl;
        }
    if (g->init_only) return;
    #pragma omp parallel for collapse(2)
    for (int i = 0; i < g->nx; ++i)
        for (int j = 0; j < g->ny; ++j) {
            const long c = (long)i * g->ny + j;
            const double x = g->dom[0] + g->dx * i, y = g->dom[2] + g->dy * j;
            if (use_ignore && g->ign[c]) { g->ETA1[c] = g->swl; continue; }
            g->ETA1[c] = eta1(g->sea, g->t0 + g->dt, x, y) + eta2(g->sea, g->t0 + g->dt, x, y) + g->swl;
        }
}

/* initialize_kinematics[_with_ignore] (lsgrid.cpp:419-620) */
static void ls_kinematics(or_lsgrid* g, int use_ignore) {
    const int nl = g->nl;
    for (int slot = 0; slot < (g->init_only ? 1 : 2); ++slot) {
        const double t = slot ? g->t0 + g->dt : g->t0;
        double* E = slot ? g->ETA1 : g->ETA0;
        double** F = slot ? g->F1 : g->F0;
        #pragma omp parallel for schedule(dynamic, 1) collapse(2)
        for (int i = 0; i < g->nx; ++i)
            for (int j = 0; j < g->ny; ++j) {
                const long c = (long)i * g->ny + j;
                if (use_ignore && g->ign[c]) continue;
                const double x = g->dom[0] + g->dx * i, y = g->dom[2] + g->dy * j;
                ls_levels(g, t, x, y, E[c], &F[0][c * nl], &F[1][c * nl], &F[2][c * nl]);
            }
    }
}

or_lsgrid* or_lsgrid_create(const or_sea* s, const double* domain, int nx, int ny, int nl,
                            double t0, double dt, double tan_a, double tan_b,
                            double water_depth, double swl, const double* domain_ignore,
                            int ignore_at_init, int init_only) {
    or_lsgrid* g = (or_lsgrid*)calloc(1, sizeof(or_lsgrid));
    g->sea = s; g->nx = nx; g->ny = ny; g->nl = nl; g->dt = dt; g->init_only = init_only != 0;
    g->a = tan_a; g->b = tan_b; g->depth = water_depth; g->swl = swl;
    for (int q = 0; q < 4; ++q) { g->dom[q] = domain[q]; g->domi[q] = domain_ignore[q]; }
    const long c = (long)nx * ny, v = c * nl;
    /* allocate (lsgrid.cpp:23-56): every array zeroed */
    g->ETA0 = (double*)calloc(c, sizeof(double));
    g->ETA1 = g->init_only ? g->ETA0 : (double*)calloc(c, sizeof(double));
    for (int f = 0; f < 3; ++f) {
        g->F0[f] = (double*)calloc(v, sizeof(double));
        g->F1[f] = g->init_only ? g->F0[f] : (double*)calloc(v, sizeof(double));
    }
    g->ign = (int*)calloc(c, sizeof(int));
    /* set_ignore (lsgrid.cpp:1554-1573) */
    g->dx = (g->dom[1] - g->dom[0]) / dmax(1., (double)(nx - 1));
    g->dy = (g->dom[3] - g->dom[2]) / dmax(1., (double)(ny - 1));
    for (int i = 0; i < nx; ++i) {
        const double xp = g->dom[0] + g->dx * i;
        for (int j = 0; j < ny; ++j) {
            const double yp = g->dom[2] + g->dy * j;
            g->ign[i * ny + j] = (xp >= g->domi[0] && xp <= g->domi[1] && yp >= g->domi[2] && yp <= g->domi[3]) ? 1 : 0;
        }
    }
    g->ds = 1. / dmax(1., (double)(nl - 1));
    ls_surface(g, t0, ignore_at_init);
    ls_kinematics(g, ignore_at_init);
    return g;
}

/* A grid holding given slot data (no sea): for checking the interpolation on
 * fields filled elsewhere.  eta [2][nx][ny], u/v/w [2][nx][ny][nl]. */
or_lsgrid* or_lsgrid_from_fields(const double* domain, int nx, int ny, int nl, double t0, double dt,
                                 double tan_a, double tan_b, double water_depth, double swl,
                                 const double* eta, const double* u, const double* v, const double* w) {
    or_lsgrid* g = (or_lsgrid*)calloc(1, sizeof(or_lsgrid));
    g->nx = nx; g->ny = ny; g->nl = nl; g->t0 = t0; g->dt = dt;
    g->a = tan_a; g->b = tan_b; g->depth = water_depth; g->swl = swl;
    for (int q = 0; q < 4; ++q) g->dom[q] = domain[q];
    g->dx = (g->dom[1] - g->dom[0]) / dmax(1., (double)(nx - 1));
    g->dy = (g->dom[3] - g->dom[2]) / dmax(1., (double)(ny - 1));
    g->ds = 1. / dmax(1., (double)(nl - 1));
    const long c = (long)nx * ny, vv = c * nl;
    g->ETA0 = (double*)malloc(c * sizeof(double)); g->ETA1 = (double*)malloc(c * sizeof(double));
    memcpy(g->ETA0, eta, c * sizeof(double)); memcpy(g->ETA1, eta + c, c * sizeof(double));
    const double* src[3] = {u, v, w};
    for (int f = 0; f < 3; ++f) {
        g->F0[f] = (double*)malloc(vv * sizeof(double)); g->F1[f] = (double*)malloc(vv * sizeof(double));
        memcpy(g->F0[f], src[f], vv * sizeof(double)); memcpy(g->F1[f], src[f] + vv, vv * sizeof(double));
    }
    g->ign = (int*)calloc(c, sizeof(int));
    return g;
}

void or_lsgrid_destroy(or_lsgrid* g) {
    if (!g) return;
    if (!g->init_only) { free(g->ETA1); for (int f = 0; f < 3; ++f) free(g->F1[f]); }
    free(g->ETA0); for (int f = 0; f < 3; ++f) free(g->F0[f]);
    free(g->ign); free(g);
}

/* lsGrid::CheckTime (lsgrid.cpp:378-386): 1 while tpt <= t0 + dt */
int or_lsgrid_check_time(const or_lsgrid* g, double t) { return t > g->t0 + g->dt ? 0 : 1; }

/* lsGrid::update (lsgrid.cpp:749-827): a jump of more than one step
 * re-initialises every column at dt * floor(t_target / dt) (returns 1);
 * otherwise one step forward, slot 1 -> slot 0 and a new slot 1, for the
 * columns outside the ignore region (returns 0) */
int or_lsgrid_update(or_lsgrid* g, double t_target) {
    const double dt = g->dt;
    if ((t_target / dt - (g->t0 + 2 * dt) / dt) > 0.) {
        ls_surface(g, dt * floor(t_target / dt), 0);
        ls_kinematics(g, 0);
        return 1;
    }
    g->t0 += dt;
    const int nl = g->nl;
    #pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int j = 0; j < g->ny; ++j)
        for (int i = 0; i < g->nx; ++i) {
            const long c = (long)i * g->ny + j;
            if (g->ign[c]) continue;
            const double x = g->dom[0] + g->dx * i, y = g->dom[2] + g->dy * j;
            g->ETA0[c] = g->ETA1[c];
            g->ETA1[c] = eta1(g->sea, g->t0 + dt, x, y) + eta2(g->sea, g->t0 + dt, x, y) + g->swl;
            for (int f = 0; f < 3; ++f)
                for (int m = 0; m < nl; ++m) g->F0[f][c * nl + m] = g->F1[f][c * nl + m];
            ls_levels(g, g->t0 + dt, x, y, g->ETA1[c], &g->F1[0][c * nl], &g->F1[1][c * nl], &g->F1[2][c * nl]);
        }
    return 0;
}

double or_lsgrid_t0(const or_lsgrid* g) { return g->t0; }

/* which: 0 ETA, 4 UX, 5 UY, 6 UZ; slot 0 then slot 1 */
long or_lsgrid_export(const or_lsgrid* g, int which, double* out) {
    const long c = (long)g->nx * g->ny, v = c * g->nl;
    const double *s0, *s1; long n;
    if (which == 0) { s0 = g->ETA0; s1 = g->ETA1; n = c; }
    else if (which >= 4 && which <= 6) { s0 = g->F0[which - 4]; s1 = g->F1[which - 4]; n = v; }
    else return 0;
    if (out) { memcpy(out, s0, n * sizeof(double)); memcpy(out + n, s1, n * sizeof(double)); }
    return 2 * n;
}

/* clip (lsgrid.cpp:187-190): std::max(lower, std::min(n, upper)) on ints */
static int ls_clipi(int n, int lo, int hi) { const int m = (hi < n) ? hi : n; return (lo < m) ? m : lo; }

/* bilinear_interpolation2 (lsgrid.cpp:322-357) */
static double ls_bilinear2(const or_lsgrid* g, const double* V0, const double* V1, double tpt, double xpt, double ypt) {
    const int nx = g->nx, ny = g->ny;
    float nxf, nyf;
    const double xd = modff((float)dmax(0., dmin((xpt - g->dom[0]) / g->dx, nx - 1.)), &nxf);
    const double yd = modff((float)dmax(0., dmin((ypt - g->dom[2]) / g->dy, ny - 1.)), &nyf);
    const int i = (int)nxf, j = (int)nyf;
    const int i1 = ls_clipi(i + 1, 0, nx - 1), j1 = ls_clipi(j + 1, 0, ny - 1);
    const double C00 = V0[i * ny + j], C01 = V0[i * ny + j1], C10 = V0[i1 * ny + j], C11 = V0[i1 * ny + j1];
    const double D00 = V1[i * ny + j], D01 = V1[i * ny + j1], D10 = V1[i1 * ny + j], D11 = V1[i1 * ny + j1];
    const double td = dmin(1., dmax(0., (tpt - g->t0) / g->dt));
    const double C0 = C00 * (1. - xd) + C10 * xd, C1 = C01 * (1. - xd) + C11 * xd;
    const double D0 = D00 * (1. - xd) + D10 * xd, D1 = D01 * (1. - xd) + D11 * xd;
    return (C0 * (1. - yd) + C1 * yd) * (1. - td) + (D0 * (1. - yd) + D1 * yd) * td;
}

/* trilinear_interpolation2 (lsgrid.cpp:192-287) */
static double ls_trilinear2(const or_lsgrid* g, const double* V0, const double* V1, double tpt, double xpt,
                            double ypt, double zpt) {
    const long nx = g->nx, ny = g->ny, nl = g->nl;
    float nxf, nyf;
    const double xd = modff((float)dmax(0., dmin((xpt - g->dom[0]) / g->dx, nx - 1.)), &nxf);
    const double yd = modff((float)dmax(0., dmin((ypt - g->dom[2]) / g->dy, ny - 1.)), &nyf);
    const long i = (int)nxf, j = (int)nyf;
    const long i1 = ls_clipi((int)i + 1, 0, (int)nx - 1), j1 = ls_clipi((int)j + 1, 0, (int)ny - 1);
    const double C00 = g->ETA0[i * ny + j], C01 = g->ETA0[i * ny + j1];
    const double C10 = g->ETA0[i1 * ny + j], C11 = g->ETA0[i1 * ny + j1];
    const double D00 = g->ETA1[i * ny + j], D01 = g->ETA1[i * ny + j1];
    const double D10 = g->ETA1[i1 * ny + j], D11 = g->ETA1[i1 * ny + j1];
    const double td = dmin(1., dmax(0., (tpt - g->t0) / g->dt));
    double C0 = C00 * (1. - xd) + C10 * xd, C1 = C01 * (1. - xd) + C11 * xd;
    double D0 = D00 * (1. - xd) + D10 * xd, D1 = D01 * (1. - xd) + D11 * xd;
    const double we0 = C0 * (1. - yd) + C1 * yd, we1 = D0 * (1. - yd) + D1 * yd;
    const double swl = g->swl, d = g->depth;
    const double spt0 = ls_tan2s(g, dmax(ls_z2s(dmin(zpt - swl, we0 - swl), we0 - swl, d), -1.));
    const double spt1 = ls_tan2s(g, dmax(ls_z2s(dmin(zpt - swl, we1 - swl), we1 - swl, d), -1.));
    float s0f, s1f;
    const double sd0 = modff((float)((spt0 + 1.) / g->ds), &s0f);
    const double sd1 = modff((float)((spt1 + 1.) / g->ds), &s1f);
    const long l0 = (int)s0f, l1 = (int)s1f;
    const long l0p = ls_clipi((int)l0 + 1, 0, (int)nl - 1), l1p = ls_clipi((int)l1 + 1, 0, (int)nl - 1);
    const double C000 = V0[i * ny * nl + j * nl + l0], C001 = V0[i * ny * nl + j * nl + l0p];
    const double C010 = V0[i * ny * nl + j1 * nl + l0], C011 = V0[i * ny * nl + j1 * nl + l0p];
    const double C100 = V0[i1 * ny * nl + j * nl + l0], C101 = V0[i1 * ny * nl + j * nl + l0p];
    const double C110 = V0[i1 * ny * nl + j1 * nl + l0], C111 = V0[i1 * ny * nl + j1 * nl + l0p];
    const double D000 = V1[i * ny * nl + j * nl + l1], D001 = V1[i * ny * nl + j * nl + l1p];
    const double D010 = V1[i * ny * nl + j1 * nl + l1], D011 = V1[i * ny * nl + j1 * nl + l1p];
    const double D100 = V1[i1 * ny * nl + j * nl + l1], D101 = V1[i1 * ny * nl + j * nl + l1p];
    const double D110 = V1[i1 * ny * nl + j1 * nl + l1], D111 = V1[i1 * ny * nl + j1 * nl + l1p];
    const double c00 = C000 * (1. - xd) + C100 * xd, c01 = C001 * (1. - xd) + C101 * xd;
    const double c10 = C010 * (1. - xd) + C110 * xd, c11 = C011 * (1. - xd) + C111 * xd;
    const double d00 = D000 * (1. - xd) + D100 * xd, d01 = D001 * (1. - xd) + D101 * xd;
    const double d10 = D010 * (1. - xd) + D110 * xd, d11 = D011 * (1. - xd) + D111 * xd;
    C0 = c00 * (1. - yd) + c10 * yd; C1 = c01 * (1. - yd) + c11 * yd;
    D0 = d00 * (1. - yd) + d10 * yd; D1 = d01 * (1. - yd) + d11 * yd;
    return (C0 * (1. - sd0) + C1 * sd0) * (1. - td) + (D0 * (1. - sd1) + D1 * sd1) * td;
}

/* lsGrid::eta / u / v / w (lsgrid.cpp:360-376) -> out4 = {eta, u, v, w} */
void or_lsgrid_query(const or_lsgrid* g, long n, const double* t, const double* x, const double* y,
                     const double* z, double* out4) {
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < n; ++p) {
        out4[4 * p] = ls_bilinear2(g, g->ETA0, g->ETA1, t[p], x[p], y[p]);
        for (int f = 0; f < 3; ++f) out4[4 * p + 1 + f] = ls_trilinear2(g, g->F0[f], g->F1[f], t[p], x[p], y[p], z[p]);
    }
}
