/*
 * Host geometry for the coherent search (see core.h).
 *
 * Reference semantics followed (file:line):
 *   Julia-compatible sinpi/cospi/eps            src/interp.c:17-43
 *   params_defaults / decimation_set            src/interp.c:48-75
 *   continued-fraction grid step p/q            src/interp.c:80-118
 *   per-harmonic phase tables (W, A, carry)     src/interp.c:219-281
 *   exact harmonic position and fit predicate   src/interp.c:349-405
 *   change points + chunk starts                src/interp.c:866-966
 *   boxcar width banks, ladder                  src/boxcar.c:78-133
 *   analytic sigma, per-width shape             src/boxcar.c:169-198
 * The V=32 group tables of the CPU filler (src/interp.c:147-217) are not
 * built: the device kernels read the phase tables directly.
 */
#include "core.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.141592653589793238462643383279502884
#endif

double csx_sinpi(double x)
{
    if (!isfinite(x)) return x - x;
    double r = fmod(x, 2.0);
    if (r < 0) r += 2.0;
    if (r == 0.0 || r == 1.0) return 0.0 * r;
    if (r == 0.5) return 1.0;
    if (r == 1.5) return -1.0;
    return sin(M_PI * r);
}

double csx_cospi(double x)
{
    if (!isfinite(x)) return x - x;
    double r = fmod(x, 2.0);
    if (r < 0) r += 2.0;
    if (r == 0.0) return 1.0;
    if (r == 0.5 || r == 1.5) return 0.0;
    if (r == 1.0) return -1.0;
    return cos(M_PI * r);
}

double csx_eps(double x)
{
    x = fabs(x);
    if (x == 0.0) return 4.9406564584124654e-324;
    return nextafter(x, INFINITY) - x;
}

int csx_decimation_set(int nharms, int maxdecim, int *out, int outcap)
{
    if (maxdecim < 1) maxdecim = 1;
    int n = 0;
    for (int k = 1; k <= maxdecim; k++) {
        if (nharms / k < 2) continue;
        if (n < outcap) out[n] = k;
        n++;
    }
    return n < outcap ? n : outcap;
}

void csx_params_defaults(SearchParams *p)
{
    memset(p, 0, sizeof(*p));
    p->nharms = 60;
    p->m = 16;
    p->numbetween = 16;
    p->hidr = 0.5;
    p->threshold = 8.0;
    p->boxcar_fsp = 1.5;
    p->boxcar_maxfrac = 0.3;
    p->boxcar_gatemargin = 0.01;
    p->precision = CS_PREC_F32;
    p->sigma = CS_SIGMA_ANALYTIC;
    p->ndecim = csx_decimation_set(p->nharms, 6, p->decimations, 64);
}

/* Continued fractions of lodr until p/q is within 1e-13 (q <= 2^16). */
int csx_trial_grid_rational(double lodr, int *pnum_out, int *q_out)
{
    if (!(lodr > 0.0)) return 0;
    const double tol = 1e-13;
    long long hprev = 0, kprev = 1, hcur = 1, kcur = 0, p = 0, q = 1;
    double y = lodr;
    int ok = 0;
    for (int it = 0; it < 200 && !ok; it++) {
        double fl = floor(y);
        if (fl > 2.0e9) break;
        long long a = (long long)fl;
        __int128 hn = (__int128)a * hcur + hprev;
        __int128 kn = (__int128)a * kcur + kprev;
        if (hn > (__int128)1e18 || kn > (__int128)1e18) break;
        p = (long long)hn;
        q = (long long)kn;
        hprev = hcur; kprev = kcur; hcur = p; kcur = q;
        if (q > CS_DIRECT_QMAX) break;
        if (fabs((double)p / (double)q - lodr) <= tol) { ok = 1; break; }
        double frac = y - fl;
        if (frac <= 0.0) break;
        y = 1.0 / frac;
        if (!isfinite(y)) break;
    }
    if (!ok && q > 0 && q <= CS_DIRECT_QMAX && p >= 1 &&
        fabs((double)p / (double)q - lodr) <= tol)
        ok = 1;
    if (!ok || q <= 0 || q > CS_DIRECT_QMAX || p < 1) return 0;
    if (fabs((double)p / (double)q - lodr) > 8.0 * csx_eps(lodr)) return 0;
    *pnum_out = (int)p;
    *q_out = (int)q;
    return 1;
}

void csx_plan_free(DirectPlan *dp)
{
    free(dp->row); free(dp->W); free(dp->A_re); free(dp->A_im); free(dp->carry);
    free(dp->grow); free(dp->goff); free(dp->gnj); free(dp->gW);
    free(dp->gA_re); free(dp->gA_im);
    memset(dp, 0, sizeof(*dp));
}

int csx_plan_build(DirectPlan *dp, int h, int m, double r_lo, int has_rational,
                   int pnum, int q)
{
    memset(dp, 0, sizeof(*dp));
    const int m2 = m / 2;
    dp->h = h;
    dp->m = m;
    dp->pnum = pnum;
    dp->q = q;
    const double hr = (double)h * r_lo;
    dp->rfloor0 = (int)floor(hr);
    dp->dr0 = hr - (double)dp->rfloor0;
    if (!has_rational) return 0;          /* P = 0: per-trial coefficients */

    const long long step = (long long)h * pnum;
    dp->s = (int)(step % q);
    dp->base_adv = (int)(step / q);
    long long g = step, a = q;
    while (a) { long long t = g % a; g = a; a = t; }
    const int P = (int)(q / g);
    dp->P = P;
    dp->row = (int32_t *)calloc((size_t)q, sizeof(int32_t));
    dp->W = (float *)malloc((size_t)m * P * sizeof(float));
    dp->A_re = (float *)malloc((size_t)P * sizeof(float));
    dp->A_im = (float *)malloc((size_t)P * sizeof(float));
    dp->carry = (uint8_t *)malloc((size_t)P);
    if (!dp->row || !dp->W || !dp->A_re || !dp->A_im || !dp->carry) return -1;

    /* Phase p visits residue res = p*s mod q; fractional offset
     * u = dr0 + res/q wraps into [0,1) with a carry into the bin index. */
    int res = 0;
    for (int ph = 0; ph < P; ph++) {
        dp->row[res] = ph + 1;
        const double u = dp->dr0 + (double)res / (double)q;
        const int cy = (u >= 1.0);
        const double dr = cy ? u - 1.0 : u;
        dp->carry[ph] = (uint8_t)cy;
        float *w = dp->W + (size_t)ph * m;
        if (dr == 0.0) {
            /* on-bin trial: identity tap at i = m/2 - 1 */
            dp->A_re[ph] = 1.0f;
            dp->A_im[ph] = 0.0f;
            for (int i = 0; i < m; i++) w[i] = (i == m2 - 1) ? 1.0f : 0.0f;
        } else {
            const double sn = csx_sinpi(dr), cs = csx_cospi(dr);
            dp->A_re[ph] = (float)(sn * cs / M_PI);
            dp->A_im[ph] = (float)(-sn * sn / M_PI);
            for (int i = 0; i < m; i++)
                w[i] = (float)(1.0 / (dr - (double)(i + 1 - m2)));
        }
        res += dp->s;
        if (res >= q) res -= q;
    }
    return 0;
}

int csx_plans_build(DirectPlanSet *set, const SearchParams *p, double r_lo)
{
    memset(set, 0, sizeof(*set));
    if (p->m <= 0 || (p->m & 1)) {
        fprintf(stderr, "Error: m must be a positive even integer\n");
        return -1;
    }
    const double lodr = p->hidr / (double)p->nharms;
    int pnum = 0, q = 0;
    const int rat = csx_trial_grid_rational(lodr, &pnum, &q);
    set->n = p->nharms;
    set->plans = (DirectPlan *)calloc((size_t)set->n, sizeof(DirectPlan));
    if (!set->plans) return -1;
    for (int h = 1; h <= p->nharms; h++)
        if (csx_plan_build(&set->plans[h - 1], h, p->m, r_lo, rat, pnum, q) != 0) {
            csx_plans_free(set);
            return -1;
        }
    return 0;
}

void csx_plans_free(DirectPlanSet *set)
{
    if (!set) return;
    for (int i = 0; i < set->n; i++) csx_plan_free(&set->plans[i]);
    free(set->plans);
    set->plans = NULL;
    set->n = 0;
}

void csx_chunk_state(const DirectPlan *dp, long long t, int *res, long long *qint)
{
    const __int128 prod = (__int128)t * (long long)(dp->h * dp->pnum);
    long long rem = (long long)(prod % dp->q);
    if (rem < 0) rem += dp->q;
    *res = (int)rem;
    *qint = (long long)(prod / dp->q);
}

/* floor(h*r_t) with the carry; P == 0 falls back to FP64 positions. */
long long csx_plan_base(const DirectPlan *dp, double lodr, long long t)
{
    if (dp->P == 0) {
        double r = (double)(dp->rfloor0 + dp->dr0) + (double)t * ((double)dp->h * lodr);
        return (long long)floor(r + 1e-15);
    }
    int res; long long qint;
    csx_chunk_state(dp, t, &res, &qint);
    return (long long)dp->rfloor0 + qint + (dp->carry[dp->row[res] - 1] ? 1 : 0);
}

long long csx_plan_nyq(const DirectPlan *dp, double lodr, long long t)
{
    if (dp->P == 0) {
        double r = (double)(dp->rfloor0 + dp->dr0) + (double)t * ((double)dp->h * lodr);
        return (long long)floor(r);
    }
    int res; long long qint;
    csx_chunk_state(dp, t, &res, &qint);
    return (long long)dp->rfloor0 + qint;
}

int csx_harmonic_fits(const DirectPlan *dp, double lodr, long long namps,
                      long long N, long long t)
{
    const int m2 = dp->m / 2;
    const long long b = csx_plan_base(dp, lodr, t);
    return (b + 2 - m2) >= 1 && (b + 1 + m2) <= namps && csx_plan_nyq(dp, lodr, t) < N / 2;
}

static int fit_lo(const DirectPlan *dp, double lodr, long long t)
{
    return csx_plan_base(dp, lodr, t) + 2 - dp->m / 2 >= 1;
}
static int fit_hi(const DirectPlan *dp, double lodr, long long namps, long long N,
                  long long t)
{
    return csx_plan_base(dp, lodr, t) + 1 + dp->m / 2 <= namps &&
           csx_plan_nyq(dp, lodr, t) < N / 2;
}

long long csx_first_fit(const DirectPlan *dp, double lodr, long long namps,
                        long long N, long long total)
{
    (void)namps; (void)N;
    if (total < 1) return total;
    if (fit_lo(dp, lodr, 0)) return 0;
    if (!fit_lo(dp, lodr, total - 1)) return total;
    long long lo = 0, hi = total - 1;
    while (hi - lo > 1) {
        long long mid = (lo + hi) >> 1;
        if (fit_lo(dp, lodr, mid)) hi = mid; else lo = mid;
    }
    return hi;
}

long long csx_first_unfit(const DirectPlan *dp, double lodr, long long namps,
                          long long N, long long total)
{
    if (total < 1) return total;
    if (!fit_hi(dp, lodr, namps, N, 0)) return 0;
    if (fit_hi(dp, lodr, namps, N, total - 1)) return total;
    long long lo = 0, hi = total - 1;
    while (hi - lo > 1) {
        long long mid = (lo + hi) >> 1;
        if (fit_hi(dp, lodr, namps, N, mid)) lo = mid; else hi = mid;
    }
    return hi;
}

static int cmp_ll(const void *a, const void *b)
{
    const long long x = *(const long long *)a, y = *(const long long *)b;
    return (x > y) - (x < y);
}

int csx_chunk_starts(const DirectPlanSet *set, double lodr, long long namps,
                     long long N, long long total, int Nprof,
                     long long **out, long long *nout)
{
    *out = NULL;
    *nout = 0;
    if (total < 1) return 0;
    const long long nreg = (total + Nprof - 1) / Nprof;
    long long *s = (long long *)malloc((size_t)(nreg + 2 * set->n + 1) * sizeof(long long));
    if (!s) return -1;
    long long n = 0;
    for (long long i = 0; i < total; i += Nprof) s[n++] = i;
    if (total > 1)
        for (int i = 0; i < set->n; i++) {
            const DirectPlan *dp = &set->plans[i];
            const long long a = csx_first_fit(dp, lodr, namps, N, total);
            const long long c = csx_first_unfit(dp, lodr, namps, N, total);
            if (a > 0 && a < total) s[n++] = a;
            if (c > 0 && c < total) s[n++] = c;
        }
    qsort(s, (size_t)n, sizeof(long long), cmp_ll);
    long long u = 0;
    for (long long i = 0; i < n; i++)
        if (u == 0 || s[i] != s[u - 1]) s[u++] = s[i];
    *out = s;
    *nout = u;
    return 0;
}

int csx_boxcar_widths(int nbins, double fsp, double maxfrac, int *out, int cap)
{
    if (nbins < 2) return 0;
    int wmax = (int)floor(maxfrac * (double)nbins);
    if (wmax < 1) wmax = 1;
    if (wmax > nbins - 1) wmax = nbins - 1;
    int n = 0;
    for (int w = 1; w <= wmax;) {
        if (n < cap) out[n] = w;
        n++;
        const int grown = (int)floor(fsp * (double)w);
        w = grown > w + 1 ? grown : w + 1;
    }
    return n < cap ? n : cap;
}

static int cmp_int(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* Ladder: when consecutive decimation factors differ by at most 3x, the
 * lowest k keeps width 1, only the highest keeps widths above 6. */
int csx_ladder_widths(int nbins, int k, const SearchParams *p, int *out, int cap)
{
    int full[256];
    const int nf = csx_boxcar_widths(nbins, p->boxcar_fsp, p->boxcar_maxfrac, full, 256);
    int ks[64];
    const int nk = p->ndecim < 64 ? p->ndecim : 64;
    int ladder = nk > 1;
    if (ladder) {
        memcpy(ks, p->decimations, (size_t)nk * sizeof(int));
        qsort(ks, (size_t)nk, sizeof(int), cmp_int);
        for (int i = 0; i + 1 < nk; i++)
            if (ks[i + 1] > (CS_LADDER_WMAX / 2) * ks[i]) ladder = 0;
    }
    int n = 0;
    if (ladder) {
        const int kmin = ks[0], kmax = ks[nk - 1];
        for (int i = 0; i < nf; i++) {
            const int w = full[i];
            if ((w >= 2 || k == kmin) && (w <= CS_LADDER_WMAX || k == kmax)) {
                if (n < cap) out[n] = w;
                n++;
            }
        }
    }
    if (n == 0) {
        n = nf < cap ? nf : cap;
        memcpy(out, full, (size_t)n * sizeof(int));
    }
    return n < cap ? n : cap;
}

double csx_analytic_sigma(const uint8_t *filled, int k, int Hk)
{
    int nlow = 0;
    for (int j = 1; j <= Hk - 1; j++) nlow += filled[j * k - 1] ? 1 : 0;
    const int nnyq = filled[Hk * k - 1] ? 1 : 0;
    if (nlow + nnyq < 1) return 0.0;
    return sqrt(2.0 * nlow + 0.5 * nnyq);
}

/* shape_w = sigma / sqrt(Var(boxsum_w)) for unit-power white harmonics. */
void csx_boxcar_shape(double *shape, int nbins, const int *widths, int nw,
                      const uint8_t *filled, int k, int Hk)
{
    const double sigma = csx_analytic_sigma(filled, k, Hk);
    for (int i = 0; i < nw; i++) {
        const int w = widths[i];
        double v = 0.0;
        for (int j = 1; j <= Hk - 1; j++) {
            if (!filled[j * k - 1]) continue;
            const double d = csx_sinpi((double)(j * w) / (double)nbins) /
                             csx_sinpi((double)j / (double)nbins);
            v += 2.0 * d * d;
        }
        if (filled[Hk * k - 1]) {
            const double d = csx_sinpi(0.5 * (double)w);
            v += 0.5 * d * d;
        }
        shape[i] = (sigma > 0.0 && v > 0.0) ? sigma / sqrt(v) : 0.0;
    }
}

int csx_nslots(const SearchParams *p)
{
    int n = 0;
    for (int i = 0; i < p->ndecim; i++) n += p->decimations[i] > 1;
    return n;
}

/* ------------------------------------------------------------------ */
/* Zaplist intervals (src/zap.c:94-187)                                */
/* ------------------------------------------------------------------ */
typedef struct {
    double lo, hi;
} ZapIv;

/* the median window of count bins from start, clipped to [0, namps) */
static void zap_window(long long start, int count, long long namps, long long *s, int *n)
{
    if (start < 0) {
        count += (int)start;
        start = 0;
    }
    if (start + count > namps) count = (int)(namps - start);
    *s = start;
    *n = count;
}

int csx_zap_ops(const ZapList *z, double T, long long N, long long namps, double baryv,
                CsZapOp **out)
{
    *out = NULL;
    if (!z || z->n <= 0) return 0;
    ZapIv *iv = (ZapIv *)malloc((size_t)z->n * sizeof(ZapIv));
    CsZapOp *ops = (CsZapOp *)malloc((size_t)z->n * sizeof(CsZapOp));
    if (!iv || !ops) {
        free(iv);
        free(ops);
        return -1;
    }
    for (int i = 0; i < z->n; i++) {
        const double scale = z->v[i].bary ? 1.0 : 1.0 + baryv;
        /* fused as the reference's -ffp-contract=fast build of zap.c fuses
         * (f -/+ 0.5 w), src/Makefile:135-138 */
        double lo = fma(-0.5, z->v[i].width, z->v[i].freq) * T * scale;
        double hi = fma(0.5, z->v[i].width, z->v[i].freq) * T * scale;
        if (hi - lo < 1.0) {
            const double mid = 0.5 * (hi + lo);
            lo = mid - 0.5;
            hi = mid + 0.5;
        }
        iv[i].lo = lo;
        iv[i].hi = hi;
    }
    /* insertion sort by the low edge: stable, and lists are short */
    for (int i = 1; i < z->n; i++) {
        const ZapIv x = iv[i];
        int j = i - 1;
        while (j >= 0 && iv[j].lo > x.lo) {
            iv[j + 1] = iv[j];
            j--;
        }
        iv[j + 1] = x;
    }
    long long lim = N / 2;
    if (lim > namps) lim = namps;
    int n = 0;
    for (int i = 0; i < z->n; i++) {
        const double lo = iv[i].lo;
        double hi = iv[i].hi;
        if (lo >= (double)lim) break;
        if (hi >= (double)lim) hi = (double)(lim - 1);
        long long ilo = (long long)floor(lo), ihi = (long long)ceil(hi);
        if (ilo < 1) ilo = 1;
        if (ihi > namps) ihi = namps;
        if (ihi - ilo <= 0) continue;
        if (!(lo - 1.5 * CS_ZAP_MEDIANBINS > 1.0)) continue;
        CsZapOp *o = &ops[n++];
        o->ilo = ilo;
        o->ihi = ihi;
        zap_window((long long)(lo - 3 * CS_ZAP_MEDIANBINS / 2), CS_ZAP_MEDIANBINS, namps,
                   &o->s_lo, &o->n_lo);
        zap_window((long long)(hi - CS_ZAP_MEDIANBINS / 2), CS_ZAP_MEDIANBINS, namps,
                   &o->s_hi, &o->n_hi);
    }
    free(iv);
    if (n == 0) {
        free(ops);
        return 0;
    }
    *out = ops;
    return n;
}
