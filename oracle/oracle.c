/*
 * oracle.c -- CPU restatement of gpu_cross's FP64 cross-section path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP library,
 * never part of the product.  Built by oracle/Makefile as liboracle.so with
 * -O2 -ffp-contract=off and no fast-math, so that every FP64 operation rounds
 * exactly where the reference's compiled C/C++ rounds.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Humlicek (1982, JQSRT 27, 437) W4 rational approximation of the      */
/* Faddeeva function, real part.  Region tests, coefficients and the    */
/* evaluation order follow profiles.cpp:9-42 exactly; the coefficient   */
/* table is the published one (no other values are valid).              */
/* ------------------------------------------------------------------ */

/* region III numerator / denominator, highest power first (profiles.cpp:28-30) */
static const double W4_III_NUM[5] = {0.5642236, 3.778987, 11.96482, 20.20933, 16.4955};
static const double W4_III_DEN[6] = {1.0, 6.699398, 21.69274, 39.27121, 38.82363, 16.4955};
/* region IV, P(u) = a0 - u(a1 - u(a2 - ...)), profiles.cpp:37-39 */
static const double W4_IV_NUM[7] = {0.56419, 1.320522, 35.76683, 219.0313, 1540.787, 3321.9905, 36183.31};
static const double W4_IV_DEN[8] = {1.0, 1.841439, 61.57037, 364.2191, 2186.181, 9022.228, 24322.84, 32066.6};

double orc_humlic(double x, double y)
{
    const double complex t = CMPLX(y, -x);
    const double s = fabs(x) + y;
    if (s >= 15.0) {                                    /* region I   (profiles.cpp:14-19) */
        double complex r = t * 0.5641896 / (0.5 + t * t);
        return creal(r);
    }
    if (s >= 5.5) {                                     /* region II  (profiles.cpp:20-25) */
        double complex u = t * t;
        double complex r = t * (1.410474 + u * .5641896) / (.75 + u * (3. + u));
        return creal(r);
    }
    if (y >= 0.195 * fabs(x) - 0.176) {                 /* region III (profiles.cpp:26-32) */
        /* innermost term is t * c (complex times real), then Horner outwards */
        double complex num = t * W4_III_NUM[0];
        for (int k = 1; k < 4; ++k) num = t * (W4_III_NUM[k] + num);
        num = W4_III_NUM[4] + num;
        /* denominator: innermost (6.699398 + T), then c_k + T*d outwards */
        double complex den = W4_III_DEN[1] + t;
        for (int k = 2; k < 6; ++k) den = W4_III_DEN[k] + t * den;
        return creal(num / den);
    }
    /* region IV (profiles.cpp:34-41): exp(u) - t * P(u) / Q(u) */
    double complex u = t * t;
    double complex p = u * W4_IV_NUM[0];                   /* innermost: U*.56419 */
    p = W4_IV_NUM[1] - p;
    for (int k = 2; k < 7; ++k) p = W4_IV_NUM[k] - u * p;
    double complex q = W4_IV_DEN[1] - u;                /* (1.841439 - U) */
    for (int k = 2; k < 8; ++k) q = W4_IV_DEN[k] - u * q;
    double complex r = cexp(u) - t * p / q;
    return creal(r);
}

/* ------------------------------------------------------------------ */
/* Constants and per-line quantities                                    */
/* ------------------------------------------------------------------ */

double orc_partition(size_t n, const double *energy, const int *gtot, double temperature)
{
    /* exomol_functions.cpp:54-66: local constants equal to defines.h, serial sum */
    const double beta = 6.6260693e-27 * 2.99792458e10 / (1.380658e-16 * temperature);
    double q = 0.0;
    for (size_t i = 0; i < n; ++i) q += gtot[i] * exp(-beta * energy[i]);
    return q;
}

double orc_beta(double temperature)
{
    return ORC_PLANCK * ORC_VELLGT / (ORC_BOLTZ * temperature);   /* OpenMPManager.cpp:66 */
}

double orc_cmcoef(int convention)
{
    if (convention == ORC_CONV_GPU) return 1.0 / (8.0 * acos(-1.0) * 2.99792458e10); /* GPUManager.cpp:31,39 */
    return 1.0 / (8.0 * ORC_PI * ORC_VELLGT);                                         /* OpenMPManager.cpp:65 */
}

double orc_dpwcoef(double temperature, double mean_mass, int convention)
{
    if (convention == ORC_CONV_GPU)   /* 1/e half-width per unit nu, cross_kernal.cu:938 */
        return sqrt(2.0 * ORC_BOLTZ * temperature * ORC_NA / (mean_mass)) / ORC_VELLGT;
    /* HWHM per unit nu, OpenMPManager.cpp:191 */
    return sqrt(2.0 * ORC_LN2 * ORC_BOLTZ * temperature * ORC_NA / (mean_mass)) / ORC_VELLGT;
}

double orc_reference_cutoff(int profile, double pressure)
{
    if (profile == ORC_DOPPLER) return ORC_DOPPLER_CUTOFF;          /* defines.h:20 */
    return fmin(25.0 * pressure, 100.0);                           /* OpenMPManager.cpp:192 */
}

double orc_line_intensity(const orc_consts *c, double elow, double nu, double aif, int gns)
{
    const double beta = orc_beta(c->temperature);
    const double cm = orc_cmcoef(c->convention);
    double nu_if = nu;
    if (c->profile == ORC_DOPPLER) {
        if (c->convention == ORC_CONV_GPU) {             /* cross_kernal.cu:645-647 */
            double s = cm * aif * gns * exp(-beta * elow) * (1.0 - exp(-beta * nu_if)) /
                       (nu_if * nu_if * c->partition);
            return nu_if == 0 ? 0.0 : s;
        }
        if (nu_if == 0) nu_if = 1e-6;                    /* OpenMPManager.cpp:153-157 */
        return cm * aif * gns * exp(-beta * elow) * (1.0 - exp(-beta * nu_if)) /
               (nu_if * nu_if * c->partition);
    }
    if (nu_if == 0) nu_if = 1e-6;                        /* OpenMPManager.cpp:198-202 */
    double s = cm * aif * gns * ORC_ISQRTPI * exp(-beta * elow) * (1.0 - exp(-beta * nu_if)) /
               (nu_if * nu_if * c->partition);
    if (c->convention == ORC_CONV_GPU && gns == -1) s = aif;   /* cross_kernal.cu:777 */
    return s;
}

double orc_line_intensity_plain(const orc_consts *c, double elow, double nu, double aif, int gns)
{
    /* cross_kernal.cu:471-496 (Gaussian) and :731-756 (Voigt without gamma): the CUDA manager's
     * cmcoef (acos(-1), GPUManager.cpp:31,39), no ISQRTPI, nu == 0 gives S = 0 */
    const double beta = orc_beta(c->temperature);
    const double cm = orc_cmcoef(ORC_CONV_GPU);
    const double g = (double)gns;
    double s = cm * aif * g * exp(-beta * elow) * (1.0 - exp(-beta * nu)) / (nu * nu * c->partition);
    return nu == 0 ? 0.0 : s;
}

/* erfcx(y) = exp(y^2) erfc(y) for y >= 0.  Below 26: y^2 = h + l exactly (Dekker's product
 * with fma), exp(y^2) = exp(h) (1 + l) to first order in |l| <= ulp(h)/2, times libm erfc
 * (erfc(26) ~ 5e-296 is still a normal number).  From 26 up: Laplace's continued fraction
 * sqrt(pi) erfcx(y) = 1/(y + (1/2)/(y + 1/(y + (3/2)/(y + 2/(y + ...))))) (Abramowitz &
 * Stegun 7.1.14), evaluated from 40 levels down (converged to double precision there). */
double orc_erfcx(double y)
{
    if (y < 26.0) {
        const double h = y * y;
        const double l = fma(y, y, -h);
        return exp(h) * (1.0 + l) * erfc(y);
    }
    double t = y;
    for (int k = 40; k >= 1; --k) t = y + (0.5 * k) / t;
    return 1.0 / (sqrt(acos(-1.0)) * t);
}

/* Sigma series of Zaghloul & Ali, Alg. 916, truncated as cross_kernal.cu:61-111 truncates it:
 * n = 1..4, the third series centred on n0 = ceil(x/a) */
static void orc_sigma916(double x, double y, double *s1, double *s2, double *s3, double a)
{
    double f, f3p, f3n, an, an3p, an3n;
    const double yy = y * y;
    const int n0 = (int)(ceil(x / a));
    *s1 = 0.0; *s2 = 0.0; *s3 = 0.0;
    for (int n = 1; n < 5; ++n) {
        const int n3p = n0 + n - 1, n3n = n0 - n;
        an = a * n;
        an3p = a * n3p;
        an3n = a * n3n;
        f = 1.0 / (an * an + yy);
        f3p = 1.0 / (an3p * an3p + yy);
        f3n = 1.0 / (an3n * an3n + yy);
        *s1 += f * exp(-(an * an + x * x));
        *s2 += f * exp(-(an + x) * (an + x));
        *s3 += f3p * exp(-(an3p - x) * (an3p - x));
        if (n3n >= 1) *s3 += f3n * exp(-(an3n - x) * (an3n - x));
    }
}

double orc_voigt916(double x, double y, double a)
{
    /* cross_kernal.cu:121-147 (equation 13 of Alg. 916) */
    const double ex2 = exp(-x * x);
    if (x == 0) return orc_erfcx(y);
    if (y == 0) return ex2;
    double s1, s2, s3;
    orc_sigma916(x, y, &s1, &s2, &s3, a);
    const double xy = x * y;
    const double a2ipi = 2.0 * a / M_PI;
    const double cos2xy = cos(2.0 * xy);
    const double sinxy = sin(xy);
    double t1 = ex2 * orc_erfcx(y) * cos2xy;
    t1 += a2ipi * x * sinxy * ex2 * sinxy / xy;
    t1 += a2ipi * y * (-cos2xy * s1 + 0.5 * (s2 + s3));
    return t1;
}

double orc_gh3(double x, double y, double xxyy)
{
    /* cross_kernal.cu:183-186, constants and order as written (PI the defines.h literal) */
    return 1.1181635900 * y / (ORC_PI * xxyy) + 2.0 * 0.2954089751 * (xxyy + 1.5) /
           ((xxyy + 1.5) * (xxyy + 1.5) - 4 * x * x * 1.5);
}

double orc_gamma_t(const orc_consts *c, double gamma, double nexp)
{
    return gamma * pow(c->ref_temp / c->temperature, nexp);   /* OpenMPManager.cpp:204 */
}

double orc_line_cutoff(const orc_consts *c, double nu)
{
    if (c->profile == ORC_DOPPLER && c->doppler_widths > 0.0) {
        double nu_if = nu == 0 ? 1e-6 : nu;
        return c->doppler_widths * (nu_if * orc_dpwcoef(c->temperature, c->mean_mass, c->convention)) + c->dfreq;
    }
    return c->cutoff;
}

/* first j in [lo, hi) with pred(j) true, for a predicate that is false then true */
static int first_true_le(double nu, double bound, const double *freq, int lo, int hi)
{   /* pred: fl(nu - f_j) <= bound; monotone since f_j increases */
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (nu - freq[mid] <= bound) hi = mid; else lo = mid + 1;
    }
    return lo;
}
static int first_true_lt(double nu, double bound, const double *freq, int lo, int hi)
{   /* pred: fl(nu - f_j) < bound */
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (nu - freq[mid] < bound) hi = mid; else lo = mid + 1;
    }
    return lo;
}

long orc_window(const orc_consts *c, double nu, double cutoff_i, const double *freq, int npoints_total,
                int points, int start_idx, int *ib, int *ie)
{
    (void)npoints_total;
    if (c->convention == ORC_CONV_GPU || c->profile == ORC_GAUSSIAN) {
        /* cross_kernal.cu:954-957 / :686-687: bin b of the chunk takes the line iff -C <= nu - f_b <= C */
        int s = start_idx, e = start_idx + points;
        *ib = first_true_le(nu, cutoff_i, freq, s, e);
        *ie = first_true_lt(nu, -cutoff_i, freq, s, e);
        if (*ie < *ib) *ie = *ib;
        return (long)(*ie - *ib);
    }
    /* OpenMPManager.cpp:205-206 (nu==0 is replaced by 1e-6 before, :198) */
    double nu_if = nu == 0 ? 1e-6 : nu;
    const double start_nu = freq[0];
    *ib = (int)fmax(round((nu_if - cutoff_i - start_nu) / c->dfreq), (double)start_idx);
    *ie = (int)fmin(round((nu_if + cutoff_i - start_nu) / c->dfreq), (double)(points + start_idx));
    return (long)(*ie - *ib);
}

void orc_chunk_range(double min_nu, double max_nu, double hw, double start_nu, double dfreq,
                     int npoints, int *ib, int *ie)
{
    *ib = (int)fmax(round((min_nu - hw - start_nu) / dfreq), 0.0);            /* BaseProfile.cpp:55 */
    *ie = (int)fmin(round((max_nu + hw - start_nu) / dfreq), (double)npoints); /* BaseProfile.cpp:56 */
}

long long orc_count_evals(const orc_consts *c, size_t nlines, const double *nu, const double *freq,
                          int npoints_total, int points, int start_idx)
{
    long long total = 0;
    for (size_t i = 0; i < nlines; ++i) {
        int ib, ie;
        long w = orc_window(c, nu[i], orc_line_cutoff(c, nu[i]), freq, npoints_total, points, start_idx, &ib, &ie);
        if (w > 0) total += w;
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* Accumulation: one thread slice (OpenMPManager.cpp:138-224)           */
/* ------------------------------------------------------------------ */

typedef struct slice_job {
    const orc_consts *c;
    const double *elow, *nu, *aif, *gamma, *nexp, *freq;
    const int *gns;
    int npoints_total, points, start_idx;
    size_t begin, end;
    double *grid;    /* private grid indexed by absolute bin - start_idx */
} slice_job;

/* The CUDA manager's Gaussian (cross_kernal.cu:567-625) and no-gamma Voigt (:860-922) sums of
 * one line, bin by bin in the line's per-bin window (|nu - f| <= C, inclusive both ends).
 * Gaussian: ln2 is the defines.h:4 literal (the reference never sets cross_constants.ln2,
 * GPUManager.cpp:24-47) and every line in the window is summed (the reference's sorted-order
 * early exit tests f - nu > hw, which stops at the first line BELOW the window).  No-gamma
 * Voigt: gamma_L = 0.05, window 500 gamma_L, Doppler width nu sqrt(2 ln2 k T / m) / c as
 * written there (no Avogadro factor), Alg. 916 where x^2 + y^2 < 100 (at |x|), else the 3-point
 * form.  With the mean mass in amu, as the reference's drivers pass it, the Doppler width is
 * ~1e-16 cm^-1 and every bin takes the 3-point form. */
static void plain_line(const slice_job *jb, size_t i)
{
    const orc_consts *c = jb->c;
    double *grid = jb->grid - jb->start_idx;
    const double nu = jb->nu[i];
    const double abscoef = orc_line_intensity_plain(c, jb->elow[i], nu, jb->aif[i], jb->gns[i]);
    if (abscoef == 0.0) return;
    const double cut = c->cutoff;
    const int s = jb->start_idx, e = jb->start_idx + jb->points;
    const int ib = first_true_le(nu, cut, jb->freq, s, e);
    int ie = first_true_lt(nu, -cut, jb->freq, s, e);
    if (ie < ib) ie = ib;
    if (c->profile == ORC_GAUSSIAN) {
        const double x_ = sqrt(ORC_LN2) / c->half_width;          /* cross_kernal.cu:582 */
        const double x0 = x_ * c->dfreq * 0.5;
        for (int j = ib; j < ie; ++j) {
            const double dfreq_ = jb->freq[j] - nu;                    /* :594 */
            const double xp = x_ * (dfreq_) + x0;
            const double xm = x_ * (dfreq_) - x0;
            grid[j] += abscoef * (erf(xp) - erf(xm));
        }
        return;
    }
    const double gammaL = 0.05;
    const double dpwcoeff = sqrt(2.0 * ORC_LN2 * ORC_BOLTZ * c->temperature / (c->mean_mass)) / ORC_VELLGT;  /* :872 */
    for (int j = ib; j < ie; ++j) {
        const double dfreq_ = nu - jb->freq[j];                        /* :884 */
        const double gammaG = nu * dpwcoeff;
        const double x = ORC_SQRTLN2 * dfreq_ / gammaG;
        const double y = ORC_SQRTLN2 * gammaL / gammaG;
        const double xxyy = x * x + y * y;
        /* the series at |x|: the reference passes the signed offset with its x < 0 guard
         * commented out (cross_kernal.cu:77), which adds n <= 0 terms for bins above the centre */
        if (xxyy < 100.0)
            grid[j] += abscoef * ORC_SQRTLN2PI / (gammaG) * y * orc_voigt916(fabs(x), y, 1.0);
        else
            grid[j] += abscoef * (ORC_SQRTLN2PI / gammaG) * orc_gh3(x, y, xxyy);
    }
}

static void *run_slice(void *arg)
{
    slice_job *jb = (slice_job *)arg;
    const orc_consts *c = jb->c;
    const double dpwcoef = orc_dpwcoef(c->temperature, c->mean_mass, c->convention);
    const double x0 = c->dfreq * 0.5;                                    /* OpenMPManager.cpp:148 */
    double *grid = jb->grid - jb->start_idx;
    for (size_t i = jb->begin; i < jb->end; ++i) {
        double nu_if = jb->nu[i];
        if (c->profile == ORC_GAUSSIAN || (c->profile == ORC_VOIGT && !jb->gamma)) {
            plain_line(jb, i);
            continue;
        }
        if (c->profile == ORC_VOIGT || c->convention == ORC_CONV_OPENMP)
            if (nu_if == 0) nu_if = 1e-6;
        const double abscoef = orc_line_intensity(c, jb->elow[i], jb->nu[i], jb->aif[i], jb->gns[i]);
        int ib, ie;
        orc_window(c, jb->nu[i], orc_line_cutoff(c, jb->nu[i]), jb->freq, jb->npoints_total, jb->points,
                   jb->start_idx, &ib, &ie);
        if (abscoef == 0.0) continue;                                    /* OpenMPManager.cpp:163,208 */
        if (c->profile == ORC_VOIGT) {
            const double gammaL = orc_gamma_t(c, jb->gamma[i], jb->nexp[i]);
            for (int j = ib; j < ie; ++j) {                              /* OpenMPManager.cpp:209-216 */
                double gammaG = 1.0 / (nu_if * dpwcoef);
                double dfreq_ = nu_if - jb->freq[j];
                double x = fabs(dfreq_) * gammaG;
                double y = gammaL * gammaG;
                grid[j] += abscoef * orc_humlic(x, y) * gammaG;
            }
        } else {
            for (int j = ib; j < ie; ++j) {                              /* OpenMPManager.cpp:164-173 */
                double gammaG = 1.0 / (nu_if * dpwcoef);
                double dfreq_ = fabs(nu_if - jb->freq[j]);
                double xp = gammaG * ((dfreq_) + x0);
                double xm = gammaG * ((dfreq_) - x0);
                double de = erf(xp) - erf(xm);
                grid[j] += abscoef * de;
            }
        }
    }
    return NULL;
}

int orc_cross_section(const orc_consts *c, size_t nlines, const double *elow, const double *nu,
                      const double *aif, const int *gns, const double *gamma, const double *nexp,
                      const double *freq, int npoints_total, int points, int start_idx,
                      int nthreads, double *out)
{
    if (!c || nthreads < 1 || points < 0 || start_idx < 0 || start_idx + points > npoints_total) return 1;
    if (c->profile == ORC_VOIGT && (!gamma != !nexp)) return 1;      /* both or neither */
    if (c->profile == ORC_GAUSSIAN && !(c->half_width > 0)) return 1;
    if (nlines > 0x7fffffff) return 1;
    const int nener = (int)nlines;
    /* OpenMPManager.cpp:186-188: slice size computed in float */
    const int for_block = (int)ceilf((float)nener / (float)nthreads);
    if ((long long)for_block * nthreads < nener) return 2;   /* reference would drop lines */
    slice_job *jobs = (slice_job *)calloc((size_t)nthreads, sizeof(slice_job));
    pthread_t *tids = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
    double *grids = (double *)calloc((size_t)nthreads * (size_t)(points > 0 ? points : 1), sizeof(double));
    if (!jobs || !tids || !grids) { free(jobs); free(tids); free(grids); return 3; }
    for (int t = 0; t < nthreads; ++t) {
        slice_job *jb = &jobs[t];
        jb->c = c; jb->elow = elow; jb->nu = nu; jb->aif = aif; jb->gns = gns; jb->gamma = gamma;
        jb->nexp = nexp; jb->freq = freq; jb->npoints_total = npoints_total; jb->points = points;
        jb->start_idx = start_idx;
        long long b = (long long)t * for_block, e = b + for_block;
        if (b > nener) b = nener;
        if (e > nener) e = nener;
        jb->begin = (size_t)b; jb->end = (size_t)e;
        jb->grid = grids + (size_t)t * (size_t)points;
    }
    int spawned = 0;
    for (int t = 1; t < nthreads; ++t)
        if (pthread_create(&tids[t], NULL, run_slice, &jobs[t]) == 0) ++spawned; else run_slice(&jobs[t]), tids[t] = 0;
    run_slice(&jobs[0]);
    for (int t = 1; t < nthreads; ++t) if (tids[t]) pthread_join(tids[t], NULL);
    (void)spawned;
    /* OpenMPManager.cpp:123-127: g_intens[j] += t_intens[t][j] in thread order */
    for (int j = 0; j < points; ++j) {
        double acc = 0.0;
        for (int t = 0; t < nthreads; ++t) acc += grids[(size_t)t * (size_t)points + j];
        out[start_idx + j] += acc;
    }
    free(jobs); free(tids); free(grids);
    return 0;
}

/* vectorised helpers for the Python test side */
void orc_humlic_vec(size_t n, const double *x, const double *y, double *out)
{
    for (size_t i = 0; i < n; ++i) out[i] = orc_humlic(x[i], y[i]);
}

void orc_erf_vec(size_t n, const double *x, double *out)
{
    for (size_t i = 0; i < n; ++i) out[i] = erf(x[i]);
}

void orc_voigt916_vec(size_t n, const double *x, const double *y, double a, double *out)
{
    for (size_t i = 0; i < n; ++i) out[i] = orc_voigt916(x[i], y[i], a);
}

void orc_erfcx_vec(size_t n, const double *x, double *out)
{
    for (size_t i = 0; i < n; ++i) out[i] = orc_erfcx(x[i]);
}
