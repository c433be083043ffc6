/*
 * ORACLE -- test infrastructure only (see oracle.h).  CPU restatement of the
 * reference's CPU search path, used only as the checker by tests/, by the
 * smoke check and by the benchmark's CPU-baseline leg.
 *
 * Build: oracle/Makefile (gcc -O3 -march=native -ffp-contract=off -fopenmp),
 * the same FP-contraction policy as the reference (src/Makefile:43-44).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.141592653589793238462643383279502884
#endif

#define ORC_V 32
#define ORC_QMAX (1 << 16)
#define ORC_SIGCAP 8192
#define ORC_LADDER_WMAX 6
#define ORC_SUB 256            /* trials per FFT/gate sub-block */

/* ------------------------------------------------------------------ */
/* Small helpers: interp.c:17-43                                      */
/* ------------------------------------------------------------------ */
double orc_sinpi(double x)
{
    if (!isfinite(x)) return x - x;
    double r = fmod(x, 2.0);
    if (r < 0) r += 2.0;
    if (r == 0.0 || r == 1.0) return 0.0 * r;
    if (r == 0.5) return 1.0;
    if (r == 1.5) return -1.0;
    return sin(M_PI * r);
}
double orc_cospi(double x)
{
    if (!isfinite(x)) return x - x;
    double r = fmod(x, 2.0);
    if (r < 0) r += 2.0;
    if (r == 0.0) return 1.0;
    if (r == 0.5 || r == 1.5) return 0.0;
    if (r == 1.0) return -1.0;
    return cos(M_PI * r);
}
static double orc_eps(double x)
{
    x = fabs(x);
    if (x == 0.0) return 4.9406564584124654e-324;
    return nextafter(x, INFINITY) - x;
}

void orc_config_defaults(OrcConfig *c)
{
    memset(c, 0, sizeof(*c));
    c->nharms = 60;
    c->m = 16;
    c->hidr = 0.5;
    c->precision = 0;
    c->sigma = 0;
    c->fsp = 1.5;
    c->maxfrac = 0.3;
    c->lofreq = 0.1;
    c->hifreq = 125.0;
    c->lobin = 100;
    c->blocksize = 0;
    c->ndecim = orc_decimation_set(c->nharms, 6, c->decimations, 64);
}

/* interp.c:64-75 */
int orc_decimation_set(int nharms, int maxdecim, int *out, int outcap)
{
    if (maxdecim < 1) maxdecim = 1;
    int n = 0;
    for (int k = 1; k <= maxdecim; k++)
        if (nharms / k >= 2) {
            if (n < outcap) out[n] = k;
            n++;
        }
    return n < outcap ? n : outcap;
}

/* ------------------------------------------------------------------ */
/* Rational trial grid: interp.c:80-118 (continued fractions, 1e-13)  */
/* ------------------------------------------------------------------ */
int orc_trial_grid_rational(double lodr, int *pout, int *qout)
{
    if (!(lodr > 0.0)) return 0;
    const double tol = 1e-13;
    long long p0 = 0, q0 = 1, p1 = 1, q1 = 0, p = 0, q = 1;
    double y = lodr;
    int found = 0;
    for (int it = 0; it < 200; it++) {
        double fl = floor(y);
        if (fl > 2.0e9) break;
        long long a = (long long)fl;
        __int128 pn = (__int128)a * p1 + p0;
        __int128 qn = (__int128)a * q1 + q0;
        if (pn > (__int128)1e18 || qn > (__int128)1e18) break;
        p = (long long)pn;
        q = (long long)qn;
        p0 = p1; q0 = q1; p1 = p; q1 = q;
        if (q > ORC_QMAX) break;
        if (fabs((double)p / (double)q - lodr) <= tol) { found = 1; break; }
        double frac = y - fl;
        if (frac <= 0.0) break;
        y = 1.0 / frac;
        if (!isfinite(y)) break;
    }
    if (!found && q > 0 && q <= ORC_QMAX && p >= 1 &&
        fabs((double)p / (double)q - lodr) <= tol)
        found = 1;
    if (!found || q <= 0 || q > ORC_QMAX || p < 1) return 0;
    if (fabs((double)p / (double)q - lodr) > 8.0 * orc_eps(lodr)) return 0;
    *pout = (int)p;
    *qout = (int)q;
    return 1;
}

/* ------------------------------------------------------------------ */
/* Plans: interp.c:147-281                                            */
/* ------------------------------------------------------------------ */
typedef struct {
    int h, m, q, s, base_adv, pnum, rfloor0, P;
    double dr0;
    double lodr;        /* P = 0: the trial step, positions in FP64 per trial */
    int32_t *row;       /* q, 1-based or 0 */
    float *W;           /* P*m */
    float *Are, *Aim;   /* P */
    uint8_t *carry;     /* P */
    int ngrp;
    int32_t *grow, *goff, *gnj;
    float *gW, *gAre, *gAim;
} OPlan;

static int gcd_i(int a, int b)
{
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) { int t = a % b; a = b; b = t; }
    return a;
}

static void oplan_free(OPlan *p)
{
    free(p->row); free(p->W); free(p->Are); free(p->Aim); free(p->carry);
    free(p->grow); free(p->goff); free(p->gnj); free(p->gW);
    free(p->gAre); free(p->gAim);
    memset(p, 0, sizeof(*p));
}

static int oplan_groups(OPlan *dp)
{
    const int V = ORC_V, m = dp->m, q = dp->q, s = dp->s;
    int gstep = (int)(((long long)V * s) % q);
    int ngrp = (gstep == 0) ? 1 : q / gcd_i(gstep, q);
    dp->ngrp = ngrp;
    dp->grow = calloc((size_t)q, sizeof(int32_t));
    dp->goff = malloc((size_t)(ngrp + 1) * sizeof(int32_t));
    dp->gnj = malloc((size_t)ngrp * sizeof(int32_t));
    dp->gAre = malloc((size_t)V * ngrp * sizeof(float));
    dp->gAim = malloc((size_t)V * ngrp * sizeof(float));
    if (!dp->grow || !dp->goff || !dp->gnj || !dp->gAre || !dp->gAim) return -1;
    int ps[ORC_V], ds[ORC_V];
    size_t total = 0;
    for (int pass = 0; pass < 2; pass++) {
        int res0 = 0;
        for (int g = 0; g < ngrp; g++) {
            if (pass == 0) dp->grow[res0] = g + 1;
            int res = res0, B = 0, B0 = 0;
            for (int k = 0; k < V; k++) {
                int p = dp->row[res] - 1;
                int Bk = B + (dp->carry[p] ? 1 : 0);
                if (k == 0) B0 = Bk;
                ps[k] = p;
                ds[k] = Bk - B0;
                if (pass == 0) {
                    dp->gAre[(size_t)g * V + k] = dp->Are[p];
                    dp->gAim[(size_t)g * V + k] = dp->Aim[p];
                }
                res += s;
                B += dp->base_adv;
                if (res >= q) { res -= q; B += 1; }
            }
            if (pass == 0) {
                int nj = m + ds[V - 1];
                dp->gnj[g] = nj;
                dp->goff[g] = (int32_t)total;
                total += (size_t)V * nj;
            } else {
                size_t o = (size_t)dp->goff[g];
                for (int k = 0; k < V; k++)
                    for (int i = 0; i < m; i++)
                        dp->gW[o + (size_t)(ds[k] + i) * V + k] = dp->W[(size_t)ps[k] * m + i];
            }
            res0 = (res0 + gstep) % q;
        }
        if (pass == 0) {
            dp->goff[ngrp] = (int32_t)total;
            dp->gW = calloc(total ? total : 1, sizeof(float));
            if (!dp->gW) return -1;
        }
    }
    return 0;
}

static int oplan_build(OPlan *dp, int h, int m, double r_lo, int pnum, int q, double lodr,
                       int rational)
{
    memset(dp, 0, sizeof(*dp));
    const int m2 = m / 2;
    dp->h = h; dp->m = m; dp->pnum = pnum; dp->q = q;
    dp->lodr = lodr;
    double hr = (double)h * r_lo;
    dp->rfloor0 = (int)floor(hr);
    dp->dr0 = hr - (double)dp->rfloor0;
    if (!rational) return 0;            /* P = 0: per-trial coefficients (interp.c:155-160) */
    long long step = (long long)h * pnum;
    dp->s = (int)(step % q);
    dp->base_adv = (int)(step / q);
    {
        long long g = step, a = q;
        while (a) { long long t = g % a; g = a; a = t; }
        dp->P = (int)(q / g);
    }
    int P = dp->P;
    dp->row = calloc((size_t)q, sizeof(int32_t));
    dp->W = malloc((size_t)m * P * sizeof(float));
    dp->Are = malloc((size_t)P * sizeof(float));
    dp->Aim = malloc((size_t)P * sizeof(float));
    dp->carry = malloc((size_t)P);
    if (!dp->row || !dp->W || !dp->Are || !dp->Aim || !dp->carry) return -1;
    int res = 0;
    for (int pp = 0; pp < P; pp++) {
        dp->row[res] = pp + 1;
        double u = dp->dr0 + (double)res / (double)q;
        dp->carry[pp] = (u >= 1.0);
        double dr = dp->carry[pp] ? u - 1.0 : u;
        if (dr == 0.0) {
            dp->Are[pp] = 1.0f;
            dp->Aim[pp] = 0.0f;
            for (int i = 0; i < m; i++)
                dp->W[(size_t)pp * m + i] = (i == m2 - 1) ? 1.0f : 0.0f;
        } else {
            double s_ = orc_sinpi(dr), c_ = orc_cospi(dr);
            dp->Are[pp] = (float)(s_ * c_ / M_PI);
            dp->Aim[pp] = (float)(-s_ * s_ / M_PI);
            for (int i = 0; i < m; i++)
                dp->W[(size_t)pp * m + i] = (float)(1.0 / (dr - (double)(i + 1 - m2)));
        }
        res += dp->s;
        if (res >= q) res -= q;
    }
    return oplan_groups(dp);
}

/* ------------------------------------------------------------------ */
/* Exact positions / availability: interp.c:349-405                   */
/* ------------------------------------------------------------------ */
static inline void ostate(const OPlan *dp, long long t, int *res, long long *qint)
{
    __int128 prod = (__int128)t * (long long)(dp->h * dp->pnum);
    long long q = dp->q;
    long long rem = (long long)(prod % q);
    if (rem < 0) rem += q;
    *res = (int)rem;
    *qint = (long long)(prod / q);
}
/* P = 0: the trial's FP64 position r = (rfloor0 + dr0) + t * (h * lodr)
 * (interp.c:359-381) */
static inline double opos_slow(const OPlan *dp, long long t)
{
    return (double)(dp->rfloor0 + dp->dr0) + (double)t * ((double)dp->h * dp->lodr);
}
static inline long long obase(const OPlan *dp, long long t)
{
    if (dp->P == 0) return (long long)floor(opos_slow(dp, t) + 1e-15);
    int res; long long qint;
    ostate(dp, t, &res, &qint);
    return (long long)dp->rfloor0 + qint + (dp->carry[dp->row[res] - 1] ? 1 : 0);
}
static inline long long onyq(const OPlan *dp, long long t)
{
    if (dp->P == 0) return (long long)floor(opos_slow(dp, t));
    int res; long long qint;
    ostate(dp, t, &res, &qint);
    return (long long)dp->rfloor0 + qint;
}
static inline int ofits(const OPlan *dp, long long namps, long long Nhalf, long long t)
{
    const int m2 = dp->m / 2;
    long long b = obase(dp, t);
    return (b + 2 - m2) >= 1 && (b + 1 + m2) <= namps && onyq(dp, t) < Nhalf;
}

/* ------------------------------------------------------------------ */
/* Context                                                            */
/* ------------------------------------------------------------------ */
struct OrcCtx {
    OrcConfig cfg;
    const float *amps;
    long long namps, N;
    double T, r_lo, r_hi, lodr;
    long long total;
    int Nprof;
    int pnum, q, rational;
    OPlan *plans;           /* nharms */
    long long *cstarts;
    long long nchunks;
    int ndec;               /* decimations with k > 1 */
    int dk[64], dHk[64], dfold[64];
    int nw[65];             /* per slot */
    int widths[65][256];
};

static int ll_cmp(const void *a, const void *b)
{
    long long x = *(const long long *)a, y = *(const long long *)b;
    return (x > y) - (x < y);
}

/* interp.c:866-942 bisection for first fits / first unfits */
static long long ofirst_lo(const OrcCtx *c, const OPlan *dp)
{
    long long total = c->total;
    const int m2 = dp->m / 2;
#define PLO(t) (obase(dp, (t)) + 2 - m2 >= 1)
    if (total < 1) return total;
    if (PLO(0)) return 0;
    if (!PLO(total - 1)) return total;
    long long lo = 0, hi = total - 1;
    while (hi - lo > 1) {
        long long mid = (lo + hi) >> 1;
        if (PLO(mid)) hi = mid; else lo = mid;
    }
#undef PLO
    return hi;
}
static long long ofirst_hi_false(const OrcCtx *c, const OPlan *dp)
{
    long long total = c->total;
    const int m2 = dp->m / 2;
#define PHI(t) (obase(dp, (t)) + 1 + m2 <= c->namps && onyq(dp, (t)) < c->N / 2)
    if (total < 1) return total;
    if (!PHI(0)) return 0;
    if (PHI(total - 1)) return total;
    long long lo = 0, hi = total - 1;
    while (hi - lo > 1) {
        long long mid = (lo + hi) >> 1;
        if (PHI(mid)) lo = mid; else hi = mid;
    }
#undef PHI
    return hi;
}

int orc_boxcar_widths(int nbins, double fsp, double maxfrac, int *out, int cap)
{
    if (nbins < 2) return 0;
    int wmax = (int)floor(maxfrac * (double)nbins);
    if (wmax < 1) wmax = 1;
    if (wmax > nbins - 1) wmax = nbins - 1;
    int n = 0, w = 1;
    while (w <= wmax) {
        if (n < cap) out[n] = w;
        n++;
        int nw = (int)floor(fsp * (double)w);
        w = nw > w + 1 ? nw : w + 1;
    }
    return n < cap ? n : cap;
}

static int int_cmp(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* boxcar.c:100-133 */
int orc_ladder_widths(int nbins, int k, const OrcConfig *cfg, int *out, int cap)
{
    int full[256];
    int nf = orc_boxcar_widths(nbins, cfg->fsp, cfg->maxfrac, full, 256);
    if (cfg->ndecim <= 1) {
        int n = nf < cap ? nf : cap;
        memcpy(out, full, (size_t)n * sizeof(int));
        return n;
    }
    int ks[64];
    int nk = cfg->ndecim < 64 ? cfg->ndecim : 64;
    memcpy(ks, cfg->decimations, (size_t)nk * sizeof(int));
    qsort(ks, (size_t)nk, sizeof(int), int_cmp);
    for (int i = 0; i + 1 < nk; i++)
        if (ks[i + 1] > (ORC_LADDER_WMAX / 2) * ks[i]) {
            int n = nf < cap ? nf : cap;
            memcpy(out, full, (size_t)n * sizeof(int));
            return n;
        }
    int kmin = ks[0], kmax = ks[nk - 1];
    int n = 0;
    for (int i = 0; i < nf; i++) {
        int w = full[i];
        if ((w >= 2 || k == kmin) && (w <= ORC_LADDER_WMAX || k == kmax)) {
            if (n < cap) out[n] = w;
            n++;
        }
    }
    if (n == 0) {
        n = nf < cap ? nf : cap;
        memcpy(out, full, (size_t)n * sizeof(int));
    }
    return n < cap ? n : cap;
}

OrcCtx *orc_ctx_create(const float *amps, long long namps, long long N, double T,
                       const OrcConfig *cfg)
{
    if (cfg->m <= 0 || (cfg->m & 1) || cfg->nharms < 1) return NULL;
    OrcCtx *c = calloc(1, sizeof(OrcCtx));
    if (!c) return NULL;
    c->cfg = *cfg;
    c->amps = amps;
    c->namps = namps;
    c->N = N;
    c->T = T;
    const int nh = cfg->nharms;
    c->lodr = cfg->hidr / (double)nh;
    c->r_lo = cfg->lofreq * T;
    if (cfg->lobin != 100) c->r_lo = (double)cfg->lobin;
    c->r_hi = cfg->hifreq * T;
    long long total = (long long)floor((c->r_hi - c->r_lo) / c->lodr) + 1;
    if (total < 0) total = 0;
    c->total = total;
    c->Nprof = cfg->blocksize == 0 ? 2048 : cfg->blocksize;
    if (c->Nprof < 1) c->Nprof = 1;
    c->rational = orc_trial_grid_rational(c->lodr, &c->pnum, &c->q);
    /* a non-rational step in double precision keeps a double stack in the
     * reference (interp.c:464-467); only its f32 form is restated */
    if (!c->rational && cfg->precision == 1) { free(c); return NULL; }
    if (!c->rational) { c->pnum = 0; c->q = 1; }
    c->plans = calloc((size_t)nh, sizeof(OPlan));
    for (int h = 1; h <= nh; h++)
        if (oplan_build(&c->plans[h - 1], h, cfg->m, c->r_lo, c->pnum, c->q, c->lodr,
                        c->rational) != 0) {
            orc_ctx_free(c);
            return NULL;
        }
    /* chunk_starts (interp.c:944-966) */
    long long nreg = total > 0 ? (total + c->Nprof - 1) / c->Nprof : 0;
    long long *b = malloc((size_t)(2 * nh + nreg + 1) * sizeof(long long));
    long long nb = 0;
    if (total > 1) {
        for (int i = 0; i < nh; i++) {
            long long a = ofirst_lo(c, &c->plans[i]);
            long long e = ofirst_hi_false(c, &c->plans[i]);
            if (a > 0 && a < total) b[nb++] = a;
            if (e > 0 && e < total) b[nb++] = e;
        }
    }
    for (long long i = 0; i < total; i += c->Nprof) b[nb++] = i;
    qsort(b, (size_t)nb, sizeof(long long), ll_cmp);
    long long u = 0;
    for (long long i = 0; i < nb; i++)
        if (u == 0 || b[i] != b[u - 1]) b[u++] = b[i];
    c->cstarts = b;
    c->nchunks = u;
    /* decimation slots and width banks */
    c->ndec = 0;
    c->nw[0] = orc_ladder_widths(2 * nh, 1, cfg, c->widths[0], 256);
    for (int i = 0; i < cfg->ndecim; i++) {
        int k = cfg->decimations[i];
        if (k <= 1) continue;
        int d = c->ndec++;
        c->dk[d] = k;
        c->dHk[d] = nh / k;
        c->dfold[d] = (nh % k == 0) && (2 * (nh / k) * k == 2 * nh);
        c->nw[d + 1] = orc_ladder_widths(2 * c->dHk[d], k, cfg, c->widths[d + 1], 256);
    }
    return c;
}

void orc_ctx_free(OrcCtx *c)
{
    if (!c) return;
    if (c->plans)
        for (int i = 0; i < c->cfg.nharms; i++) oplan_free(&c->plans[i]);
    free(c->plans);
    free(c->cstarts);
    free(c);
}

long long orc_ctx_total(const OrcCtx *c) { return c->total; }
long long orc_ctx_nchunks(const OrcCtx *c) { return c->nchunks; }
double orc_ctx_r_lo(const OrcCtx *c) { return c->r_lo; }
double orc_ctx_lodr(const OrcCtx *c) { return c->lodr; }
int orc_ctx_rational(const OrcCtx *c, int *p, int *q)
{
    *p = c->pnum; *q = c->q;
    return c->rational;
}
int orc_ctx_chunk(const OrcCtx *c, long long ci, long long *i0, int *n, double *rstart)
{
    if (ci < 0 || ci >= c->nchunks) return -1;
    long long a = c->cstarts[ci];
    long long e = ci + 1 < c->nchunks ? c->cstarts[ci + 1] : c->total;
    *i0 = a;
    *n = (int)(e - a);
    *rstart = c->r_lo + (double)a * c->lodr;
    return 0;
}
int orc_ctx_ndec(const OrcCtx *c, int *ks, int *hks)
{
    for (int d = 0; d < c->ndec; d++) {
        if (ks) ks[d] = c->dk[d];
        if (hks) hks[d] = c->dHk[d];
    }
    return c->ndec;
}
long long orc_ctx_harmonic_trials(const OrcCtx *c)
{
    long long s = 0;
    for (int i = 0; i < c->cfg.nharms; i++) {
        long long a = ofirst_lo(c, &c->plans[i]);
        long long e = ofirst_hi_false(c, &c->plans[i]);
        /* fits on [a, e) when a < e (monotone predicates) */
        if (e > a) s += e - a;
    }
    return s;
}
int orc_plan_info(const OrcCtx *c, int h, int *rfloor0, double *dr0, int *P,
                  int *s, int *base_adv)
{
    if (h < 1 || h > c->cfg.nharms) return -1;
    const OPlan *p = &c->plans[h - 1];
    *rfloor0 = p->rfloor0; *dr0 = p->dr0; *P = p->P; *s = p->s;
    *base_adv = p->base_adv;
    return 0;
}
int orc_plan_tables(const OrcCtx *c, int h, float *W, float *A, uint8_t *carry,
                    int32_t *row)
{
    if (h < 1 || h > c->cfg.nharms) return -1;
    const OPlan *p = &c->plans[h - 1];
    if (W) memcpy(W, p->W, (size_t)p->P * p->m * sizeof(float));
    if (A) for (int i = 0; i < p->P; i++) { A[2 * i] = p->Are[i]; A[2 * i + 1] = p->Aim[i]; }
    if (carry) memcpy(carry, p->carry, (size_t)p->P);
    if (row) memcpy(row, p->row, (size_t)p->q * sizeof(int32_t));
    return p->P;
}
int orc_harmonic_fits(const OrcCtx *c, int h, long long t)
{
    if (h < 1 || h > c->cfg.nharms) return 0;
    return ofits(&c->plans[h - 1], c->namps, c->N / 2, t);
}

/* ------------------------------------------------------------------ */
/* Project c2r.                                                       */
/*                                                                    */
/* Definition (both the oracle and the HIP kernels follow it op for   */
/* op in IEEE arithmetic, with fused multiply-adds exactly where it    */
/* writes fma and no other contraction, so their profiles are         */
/* bit-identical; the reference uses FFTW/cuFFT instead):             */
/*  tw(j,N) = ((T)cos(a), (T)sin(a)), a = 2*pi*(double)j/N, cos and   */
/*            sin separate libm calls (not a merged sincos)           */
/*  cmul(y,w) = (fma(y.re, w.re, -(y.im*w.im)),                       */
/*               fma(y.re, w.im,   y.im*w.re))                        */
/*  c2r(X[0..H]) -> x[0..2H-1]:                                        */
/*    Z_0 = (X0.re + XH.re, X0.re - XH.re)                             */
/*    k=1..H-1: a=X_k, b=X_{H-k}; e=(a.re+b.re, a.im-b.im);            */
/*              d=(a.re-b.re, a.im+b.im); o=cmul(d,tw(k,2H))           */
/*              Z_k = (e.re - o.im, e.im + o.re)                       */
/*    z = IDFT_H(Z); x[2s] = z_s.re; x[2s+1] = z_s.im                  */
/*  IDFT_N (e^{+i}) of x[0..N-1]:                                      */
/*   N = 1: copy.                                                      */
/*   p = smallest prime factor, N1 = the power of p dividing N.        */
/*   N1 < N (prime-factor split, no twiddles), N2 = N/N1:              */
/*     T[n2] = IDFT_N1(x[(N2*n1 + N1*n2) mod N], n1 = 0..N1-1);       */
/*     for k1 < N1: o = IDFT_N2(T[n2][k1], n2 = 0..N2-1);             */
/*       X[(k1*E1 + k2*E2) mod N] = o[k2], E1 = N2*(N2^-1 mod N1),     */
/*       E2 = N1*(N1^-1 mod N2)  (the CRT map).                        */
/*   N1 = N (a prime power), decimation in time: R = 4 if 4|N, else   */
/*     2 if 2|N, else p; M = N/R; Y_r = IDFT_M(x_r, x_{r+R}, ...);    */
/*     for k<M: u_0 = Y_0[k],                                          */
/*     u_r = Y_r[k] if r*k==0 else cmul(Y_r[k], tw(r*k, N));           */
/*     X[k + M*s] = butterfly_R(u)_s, with                             */
/*    R=2: (u0+u1, u0-u1)                                             */
/*    R=4: t0=u0+u2 t1=u0-u2 t2=u1+u3 t3=u1-u3;                       */
/*         (t0+t2, (t1.re-t3.im, t1.im+t3.re), t0-t2,                 */
/*          (t1.re+t3.im, t1.im-t3.re))                               */
/*    R odd (p): sp_r = u_r+u_{p-r}, sm_r = u_r-u_{p-r}, r=1..(p-1)/2;*/
/*         X_0 = ((u0+sp_1)+sp_2)+...;                                 */
/*         s=1..(p-1)/2, c,sn = tw(r*s mod p, p):                     */
/*         A = u0; A = fma(sp_r, c, A) for r = 1, 2, ... (per part);  */
/*         B = sm_1*sn; B = fma(sm_r, sn, B) for r = 2, 3, ...;       */
/*         X_s = (A.re-B.im, A.im+B.re), X_{p-s} = (A.re+B.im, A.im-B.re) */
/* ------------------------------------------------------------------ */
/* cos and sin as two separate libm calls, through pointers the compiler
 * cannot merge into one sincos call (which differs from them in the last
 * bit for some angles): the twiddle definition is libm cos and sin. */
static double (*volatile orc_cos_fn)(double) = cos;
static double (*volatile orc_sin_fn)(double) = sin;
static inline void otw(int j, int N, double *c, double *s)
{
    double a = 2.0 * M_PI * (double)j / (double)N;
    *c = orc_cos_fn(a);
    *s = orc_sin_fn(a);
}

static int ospf(int N)
{
    if (N % 2 == 0) return 2;
    for (int p = 3; p * p <= N; p += 2)
        if (N % p == 0) return p;
    return N;
}

static int opfa_n1(int N)
{
    const int p = ospf(N);
    int n1 = 1;
    while (N % (n1 * p) == 0) n1 *= p;
    return n1;
}

static int oradix(int N)
{
    if (N % 4 == 0) return 4;
    if (N % 2 == 0) return 2;
    return ospf(N);
}

static long omodinv(long a, long m)
{
    a %= m;
    for (long x = 1; x < m; x++)
        if (a * x % m == 1) return x;
    return 0;
}

/* dft_SUF: output k (B lanes at or_/oi + ooff[k]) = IDFT_N of the inputs
 * n (B lanes at ir/ii + ioff[n]); the index lists carry the prime-factor
 * and decimation-in-time reorderings, which involve no arithmetic. */
#define DEFINE_DFT(T, SUF, CVT, FMA)                                           \
static void dft_##SUF(int N, const T *ir, const T *ii, const long *ioff,      \
                      T *or_, T *oi, const long *ooff, int B)                 \
{                                                                             \
    if (N == 1) {                                                             \
        for (int b = 0; b < B; b++) {                                         \
            or_[ooff[0] + b] = ir[ioff[0] + b];                               \
            oi[ooff[0] + b] = ii[ioff[0] + b];                                \
        }                                                                     \
        return;                                                               \
    }                                                                         \
    const int N1 = opfa_n1(N);                                                \
    T *tr = malloc(sizeof(T) * (size_t)N * B), *ti = malloc(sizeof(T) * (size_t)N * B); \
    long *si = malloc(sizeof(long) * (size_t)N), *so = malloc(sizeof(long) * (size_t)N); \
    if (N1 < N) {                                                             \
        const int N2 = N / N1;                                                \
        for (int n2 = 0; n2 < N2; n2++) {                                     \
            for (int n1 = 0; n1 < N1; n1++) {                                 \
                si[n1] = ioff[((long)N2 * n1 + (long)N1 * n2) % N];           \
                so[n1] = ((long)n2 * N1 + n1) * B;                            \
            }                                                                 \
            dft_##SUF(N1, ir, ii, si, tr, ti, so, B);                         \
        }                                                                     \
        const long E1 = N2 * omodinv(N2, N1), E2 = N1 * omodinv(N1, N2);      \
        for (int k1 = 0; k1 < N1; k1++) {                                     \
            for (int n2 = 0; n2 < N2; n2++) {                                 \
                si[n2] = ((long)n2 * N1 + k1) * B;                            \
                so[n2] = ooff[(k1 * E1 + n2 * E2) % N];                       \
            }                                                                 \
            dft_##SUF(N2, tr, ti, si, or_, oi, so, B);                        \
        }                                                                     \
        free(tr); free(ti); free(si); free(so);                               \
        return;                                                               \
    }                                                                         \
    const int R = oradix(N), M = N / R;                                       \
    /* Y_r[j] at tr/ti + (r*M + j)*B */                                       \
    for (int r = 0; r < R; r++) {                                             \
        for (int j = 0; j < M; j++) {                                         \
            si[j] = ioff[r + (long)j * R];                                    \
            so[j] = ((long)r * M + j) * B;                                    \
        }                                                                     \
        dft_##SUF(M, ir, ii, si, tr, ti, so, B);                              \
    }                                                                         \
    T *ur = malloc(sizeof(T) * (size_t)R * B);                                \
    T *ui = malloc(sizeof(T) * (size_t)R * B);                                \
    T *sp = malloc(sizeof(T) * (size_t)(R + 1) * 2 * B);                      \
    T *sm = malloc(sizeof(T) * (size_t)(R + 1) * 2 * B);                      \
    T *twc = malloc(sizeof(T) * (size_t)R), *tws = malloc(sizeof(T) * (size_t)R); \
    T *Ar = malloc(sizeof(T) * B), *Ai = malloc(sizeof(T) * B);               \
    T *Br = malloc(sizeof(T) * B), *Bi = malloc(sizeof(T) * B);               \
    for (int j = 0; j < R; j++) {                                             \
        double c_, s_; otw(j, R, &c_, &s_);                                   \
        twc[j] = CVT(c_); tws[j] = CVT(s_);                                   \
    }                                                                         \
    for (int k = 0; k < M; k++) {                                             \
        for (int r = 0; r < R; r++) {                                         \
            const T *yr = tr + ((long)r * M + k) * B;                         \
            const T *yi = ti + ((long)r * M + k) * B;                         \
            T *dr = ur + (size_t)r * B, *di = ui + (size_t)r * B;            \
            if (r * k == 0) {                                                 \
                for (int b = 0; b < B; b++) { dr[b] = yr[b]; di[b] = yi[b]; } \
            } else {                                                          \
                double c_, s_; otw(r * k, N, &c_, &s_);                       \
                const T wc = CVT(c_), ws = CVT(s_);                           \
                for (int b = 0; b < B; b++) {                                 \
                    const T bb = yi[b] * ws, dd = yi[b] * wc;                 \
                    dr[b] = FMA(yr[b], wc, -bb);                              \
                    di[b] = FMA(yr[b], ws, dd);                               \
                }                                                             \
            }                                                                 \
        }                                                                     \
        T *o0r = or_ + ooff[k], *o0i = oi + ooff[k];                          \
        if (R == 2) {                                                         \
            T *o1r = or_ + ooff[k + M], *o1i = oi + ooff[k + M];              \
            for (int b = 0; b < B; b++) {                                     \
                T a0 = ur[b], a1 = ui[b], b0 = ur[B + b], b1 = ui[B + b];     \
                o0r[b] = a0 + b0; o0i[b] = a1 + b1;                           \
                o1r[b] = a0 - b0; o1i[b] = a1 - b1;                           \
            }                                                                 \
        } else if (R == 4) {                                                  \
            T *p1r = or_ + ooff[k + M], *p1i = oi + ooff[k + M];              \
            T *p2r = or_ + ooff[k + 2 * M], *p2i = oi + ooff[k + 2 * M];      \
            T *p3r = or_ + ooff[k + 3 * M], *p3i = oi + ooff[k + 3 * M];      \
            for (int b = 0; b < B; b++) {                                     \
                T u0r = ur[b], u0i = ui[b], u1r = ur[B + b], u1i = ui[B + b]; \
                T u2r = ur[2 * B + b], u2i = ui[2 * B + b];                   \
                T u3r = ur[3 * B + b], u3i = ui[3 * B + b];                   \
                T t0r = u0r + u2r, t0i = u0i + u2i;                           \
                T t1r = u0r - u2r, t1i = u0i - u2i;                           \
                T t2r = u1r + u3r, t2i = u1i + u3i;                           \
                T t3r = u1r - u3r, t3i = u1i - u3i;                           \
                o0r[b] = t0r + t2r; o0i[b] = t0i + t2i;                       \
                p2r[b] = t0r - t2r; p2i[b] = t0i - t2i;                       \
                p1r[b] = t1r - t3i; p1i[b] = t1i + t3r;                       \
                p3r[b] = t1r + t3i; p3i[b] = t1i - t3r;                       \
            }                                                                 \
        } else {                                                              \
            const int hh = (R - 1) / 2;                                       \
            /* sp/sm planes: [2r + part][b], r = 1..hh */                     \
            for (int r = 1; r <= hh; r++) {                                   \
                const T *arr = ur + (size_t)r * B, *aii = ui + (size_t)r * B; \
                const T *brr = ur + (size_t)(R - r) * B, *bii = ui + (size_t)(R - r) * B; \
                T *p0 = sp + (size_t)(2 * r) * B, *p1 = sp + (size_t)(2 * r + 1) * B; \
                T *m0 = sm + (size_t)(2 * r) * B, *m1 = sm + (size_t)(2 * r + 1) * B; \
                for (int b = 0; b < B; b++) {                                 \
                    p0[b] = arr[b] + brr[b]; p1[b] = aii[b] + bii[b];         \
                    m0[b] = arr[b] - brr[b]; m1[b] = aii[b] - bii[b];         \
                }                                                             \
            }                                                                 \
            for (int s = 1; s <= hh; s++) {                                   \
                for (int b = 0; b < B; b++) { Ar[b] = ur[b]; Ai[b] = ui[b]; } \
                for (int r = 1; r <= hh; r++) {                               \
                    const int j = (r * s) % R;                                \
                    const T c = twc[j], sn = tws[j];                          \
                    const T *p0 = sp + (size_t)(2 * r) * B, *p1 = sp + (size_t)(2 * r + 1) * B; \
                    const T *m0 = sm + (size_t)(2 * r) * B, *m1 = sm + (size_t)(2 * r + 1) * B; \
                    for (int b = 0; b < B; b++) {                             \
                        Ar[b] = FMA(p0[b], c, Ar[b]);                         \
                        Ai[b] = FMA(p1[b], c, Ai[b]);                         \
                        if (r == 1) { Br[b] = m0[b] * sn; Bi[b] = m1[b] * sn; } \
                        else { Br[b] = FMA(m0[b], sn, Br[b]);                 \
                               Bi[b] = FMA(m1[b], sn, Bi[b]); }               \
                    }                                                         \
                }                                                             \
                T *osr = or_ + ooff[k + M * s], *osi = oi + ooff[k + M * s];  \
                T *omr = or_ + ooff[k + M * (R - s)], *omi = oi + ooff[k + M * (R - s)]; \
                for (int b = 0; b < B; b++) {                                 \
                    osr[b] = Ar[b] - Bi[b]; osi[b] = Ai[b] + Br[b];           \
                    omr[b] = Ar[b] + Bi[b]; omi[b] = Ai[b] - Br[b];           \
                }                                                             \
            }                                                                 \
            /* X_0 last: with an in-place output list it must not clobber */ \
            /* u0 before the s loop reads it (ur is a copy, so it cannot) */ \
            for (int b = 0; b < B; b++) {                                     \
                T acr = ur[b], aci = ui[b];                                   \
                for (int r = 1; r <= hh; r++) {                               \
                    acr = acr + sp[(size_t)(2 * r) * B + b];                  \
                    aci = aci + sp[(size_t)(2 * r + 1) * B + b];              \
                }                                                             \
                o0r[b] = acr; o0i[b] = aci;                                   \
            }                                                                 \
        }                                                                     \
    }                                                                         \
    free(ur); free(ui); free(sp); free(sm); free(twc); free(tws);             \
    free(Ar); free(Ai); free(Br); free(Bi);                                   \
    free(tr); free(ti); free(si); free(so);                                   \
}                                                                             \
/* Batched c2r: X bins h=0..H at (xr,xi)[h*xs + b]; out bins x[i*lo + b]. */ \
static void c2r_##SUF(int H, const T *xr, const T *xi, long xs, T *out,      \
                      long lo, int B, T *wk)                                  \
{                                                                             \
    if (H < 1 || B < 1) return;                                               \
    T *zr = wk, *zi = wk + (size_t)H * B;                                     \
    T *yr = wk + (size_t)2 * H * B, *yi = wk + (size_t)3 * H * B;             \
    for (int b = 0; b < B; b++) {                                             \
        T x0 = xr[b], xh = xr[(long)H * xs + b];                              \
        zr[b] = x0 + xh;                                                      \
        zi[b] = x0 - xh;                                                      \
    }                                                                         \
    for (int k = 1; k < H; k++) {                                             \
        double c_, s_; otw(k, 2 * H, &c_, &s_);                               \
        const T wc = CVT(c_), ws = CVT(s_);                                   \
        const T *ar = xr + (long)k * xs, *ai = xi + (long)k * xs;             \
        const T *br = xr + (long)(H - k) * xs, *bi = xi + (long)(H - k) * xs; \
        for (int b = 0; b < B; b++) {                                         \
            T er = ar[b] + br[b], ei = ai[b] - bi[b];                         \
            T dr = ar[b] - br[b], di = ai[b] + bi[b];                         \
            T p2 = di * ws, p4 = di * wc;                                     \
            T or_ = FMA(dr, wc, -p2), oi = FMA(dr, ws, p4);                   \
            zr[(size_t)k * B + b] = er - oi;                                  \
            zi[(size_t)k * B + b] = ei + or_;                                 \
        }                                                                     \
    }                                                                         \
    long *io = calloc((size_t)H, sizeof(long));                                \
    for (int s = 0; s < H; s++) io[s] = (long)s * B;                          \
    dft_##SUF(H, zr, zi, io, yr, yi, io, B);                                  \
    free(io);                                                                 \
    for (int s = 0; s < H; s++)                                               \
        for (int b = 0; b < B; b++) {                                         \
            out[(long)(2 * s) * lo + b] = yr[(size_t)s * B + b];              \
            out[(long)(2 * s + 1) * lo + b] = yi[(size_t)s * B + b];          \
        }                                                                     \
}

#define CVT_F(x) ((float)(x))
#define CVT_D(x) ((double)(x))
DEFINE_DFT(float, f32, CVT_F, fmaf)
DEFINE_DFT(double, f64, CVT_D, fma)

void orc_c2r_f32(int H, const float *X, float *x)
{
    float *xr = malloc(sizeof(float) * (H + 1)), *xi = malloc(sizeof(float) * (H + 1));
    float *wk = malloc(sizeof(float) * 4 * (H > 0 ? H : 1));
    if (H >= 1) {
        for (int h = 0; h <= H; h++) { xr[h] = X[2 * h]; xi[h] = X[2 * h + 1]; }
        c2r_f32(H, xr, xi, 1, x, 1, 1, wk);
    }
    free(xr); free(xi); free(wk);
}
/* IDFT_N of N interleaved complex values (the project definition above). */
void orc_idft_f64(int N, const double *in, double *out)
{
    if (N <= 0) return;
    double *ir = malloc(sizeof(double) * N), *ii = malloc(sizeof(double) * N);
    double *or_ = malloc(sizeof(double) * N), *oi = malloc(sizeof(double) * N);
    long *io = calloc((size_t)N, sizeof(long));
    for (int k = 0; k < N; k++) { ir[k] = in[2 * k]; ii[k] = in[2 * k + 1]; io[k] = k; }
    dft_f64(N, ir, ii, io, or_, oi, io, 1);
    for (int k = 0; k < N; k++) { out[2 * k] = or_[k]; out[2 * k + 1] = oi[k]; }
    free(ir); free(ii); free(or_); free(oi); free(io);
}
void orc_idft_f32(int N, const float *in, float *out)
{
    if (N <= 0) return;
    float *ir = malloc(sizeof(float) * N), *ii = malloc(sizeof(float) * N);
    float *or_ = malloc(sizeof(float) * N), *oi = malloc(sizeof(float) * N);
    long *io = calloc((size_t)N, sizeof(long));
    for (int k = 0; k < N; k++) { ir[k] = in[2 * k]; ii[k] = in[2 * k + 1]; io[k] = k; }
    dft_f32(N, ir, ii, io, or_, oi, io, 1);
    for (int k = 0; k < N; k++) { out[2 * k] = or_[k]; out[2 * k + 1] = oi[k]; }
    free(ir); free(ii); free(or_); free(oi); free(io);
}

void orc_c2r_f64(int H, const double *X, double *x)
{
    double *xr = malloc(sizeof(double) * (H + 1)), *xi = malloc(sizeof(double) * (H + 1));
    double *wk = malloc(sizeof(double) * 4 * (H > 0 ? H : 1));
    if (H >= 1) {
        for (int h = 0; h <= H; h++) { xr[h] = X[2 * h]; xi[h] = X[2 * h + 1]; }
        c2r_f64(H, xr, xi, 1, x, 1, 1, wk);
    }
    free(xr); free(xi); free(wk);
}

/* ------------------------------------------------------------------ */
/* Spectrum stack fill: interp.c:608-861                              */
/* st_re/st_im: (nh+1) rows of n, h-major; filled[h-1].               */
/* ------------------------------------------------------------------ */
static void fill_group(const OrcCtx *c, const OPlan *dp, long long t0, int n,
                       float *rowr, float *rowi, float *re, float *im)
{
    const int V = ORC_V, m2 = dp->m / 2;
    long long T_first = (t0 / V) * V;
    long long T_last = ((t0 + n - 1) / V) * V + (V - 1);
    int res; long long qint;
    ostate(dp, T_first, &res, &qint);
    long long lo1 = (long long)dp->rfloor0 + qint +
                    (dp->carry[dp->row[res] - 1] ? 1 : 0) + 2 - m2;
    long long hi1 = obase(dp, T_last) + 1 + m2;
    long long nw = hi1 - lo1 + 1;
    /* planes, zero padded outside [1, namps] (1-based) */
    for (long long i = 0; i < nw; i++) {
        long long idx = lo1 + i - 1;
        if (idx >= 0 && idx < c->namps) {
            re[i] = c->amps[2 * idx];
            im[i] = c->amps[2 * idx + 1];
        } else {
            re[i] = 0.0f;
            im[i] = 0.0f;
        }
    }
    long long T = T_first;
    while (T <= T_last) {
        int g = dp->grow[res] - 1;
        int p = dp->row[res] - 1;
        long long b0 = (long long)dp->rfloor0 + qint + (dp->carry[p] ? 1 : 0) + 2 - m2 - lo1;
        size_t o = (size_t)dp->goff[g];
        int nj = dp->gnj[g];
        float ar[ORC_V], ai[ORC_V];
        for (int k = 0; k < V; k++) { ar[k] = 0.0f; ai[k] = 0.0f; }
        for (int j = 0; j < nj; j++) {
            const float rv = re[b0 + j], iv = im[b0 + j];
            const float *w = dp->gW + o + (size_t)j * V;
            for (int k = 0; k < V; k++) {
                ar[k] = fmaf(w[k], rv, ar[k]);
                ai[k] = fmaf(w[k], iv, ai[k]);
            }
        }
        for (int k = 0; k < V; k++) {
            long long j = T + k - t0;
            if (j < 0 || j >= n) continue;
            float gr = dp->gAre[(size_t)g * V + k], gi = dp->gAim[(size_t)g * V + k];
            float p1 = gr * ar[k], p2 = gi * ai[k], p3 = gr * ai[k], p4 = gi * ar[k];
            rowr[j] = p1 - p2;
            rowi[j] = p3 + p4;
        }
        long long tot = (long long)res + (long long)V * dp->s;
        qint += (long long)V * dp->base_adv + tot / dp->q;
        res = (int)(tot % dp->q);
        T += V;
    }
}

static void fill_direct(const OrcCtx *c, const OPlan *dp, long long t0, int n,
                        float *rowr, float *rowi)
{
    const int m = dp->m, m2 = m / 2;
    for (int j = 0; j < n; j++) {
        int res; long long qint;
        ostate(dp, t0 + j, &res, &qint);
        int p = dp->row[res] - 1;
        long long b = (long long)dp->rfloor0 + qint + (dp->carry[p] ? 1 : 0) + 1 - m2;
        float ar = 0.0f, ai = 0.0f;
        for (int i = 0; i < m; i++) {
            float w = dp->W[(size_t)p * m + i];
            ar = fmaf(w, c->amps[2 * (b + i)], ar);
            ai = fmaf(w, c->amps[2 * (b + i) + 1], ai);
        }
        float gr = dp->Are[p], gi = dp->Aim[p];
        float p1 = gr * ar, p2 = gi * ai, p3 = gr * ai, p4 = gi * ar;
        rowr[j] = p1 - p2;
        rowi[j] = p3 + p4;
    }
}

/* P = 0 (a trial step that is no simple rational): every trial's position
 * in FP64, the m-tap weights 1/(dr - (i - m/2)) and the rotation from
 * sinpi/cospi of the fractional offset, summed in double and rounded to the
 * float stack (interp.c:427-475, fill_row_slow, the f32 workspace) */
static void fill_slow(const OrcCtx *c, const OPlan *dp, long long t0, int n,
                      float *rowr, float *rowi)
{
    const int m = dp->m, m2 = m / 2;
    const long long namps = c->namps;
    const double hr_lo = (double)(dp->rfloor0 + dp->dr0);
    const double dh = (double)dp->h * dp->lodr;
    for (int k = 0; k < n; k++) {
        const double r = hr_lo + (double)(t0 + k) * dh;
        const long long rint0 = (long long)floor(r + 1e-15) + 1;
        double dr = fmod(r, 1.0);
        if (dr < 0) dr += 1.0;
        const long long base = rint0 - m2;     /* 1-based */
        double xr, xi;
        if (dr == 0.0) {
            const long long idx = base + m2 - 1;
            xr = (idx >= 0 && idx < namps) ? (double)c->amps[2 * idx] : 0.0;
            xi = (idx >= 0 && idx < namps) ? (double)c->amps[2 * idx + 1] : 0.0;
        } else {
            double sre = 0.0, sim = 0.0;
            for (int i = 1; i <= m; i++) {
                const double w = 1.0 / (dr - (double)(i - m2));
                const long long idx = base + i - 1;
                const double ar = (idx >= 0 && idx < namps) ? (double)c->amps[2 * idx] : 0.0;
                const double ai = (idx >= 0 && idx < namps) ? (double)c->amps[2 * idx + 1] : 0.0;
                sre += w * ar;
                sim += w * ai;
            }
            const double s_ = orc_sinpi(dr), c_ = orc_cospi(dr);
            const double pr = s_ * c_ / M_PI;
            const double pi_ = -s_ * s_ / M_PI;
            xr = pr * sre - pi_ * sim;
            xi = pr * sim + pi_ * sre;
        }
        rowr[k] = (float)xr;
        rowi[k] = (float)xi;
    }
}

static void chunk_filled(const OrcCtx *c, long long t0, int n, uint8_t *filled)
{
    for (int h = 1; h <= c->cfg.nharms; h++) {
        const OPlan *dp = &c->plans[h - 1];
        filled[h - 1] = (uint8_t)(ofits(dp, c->namps, c->N / 2, t0) &&
                                  ofits(dp, c->namps, c->N / 2, t0 + n - 1));
    }
}

static void fill_stack(const OrcCtx *c, long long t0, int n, int form,
                       float *st_re, float *st_im, uint8_t *filled,
                       float *re, float *im)
{
    const int nh = c->cfg.nharms;
    chunk_filled(c, t0, n, filled);
    memset(st_re, 0, (size_t)n * sizeof(float));
    memset(st_im, 0, (size_t)n * sizeof(float));
    for (int h = 1; h <= nh; h++) {
        float *rr = st_re + (size_t)h * n, *ri = st_im + (size_t)h * n;
        if (!filled[h - 1]) {
            memset(rr, 0, (size_t)n * sizeof(float));
            memset(ri, 0, (size_t)n * sizeof(float));
            continue;
        }
        if (c->plans[h - 1].P == 0) fill_slow(c, &c->plans[h - 1], t0, n, rr, ri);
        else if (form == 0) fill_group(c, &c->plans[h - 1], t0, n, rr, ri, re, im);
        else fill_direct(c, &c->plans[h - 1], t0, n, rr, ri);
    }
}

static size_t window_cap(const OrcCtx *c, int n)
{
    return (size_t)ceil((double)(n - 1 + 2 * ORC_V) * c->cfg.hidr) + c->cfg.m + 8;
}

int orc_chunk_stack(const OrcCtx *c, long long ci, int form, float *out,
                    uint8_t *filled)
{
    long long i0; int n; double rs;
    if (orc_ctx_chunk(c, ci, &i0, &n, &rs) != 0) return -1;
    const int nh = c->cfg.nharms;
    float *sr = malloc(sizeof(float) * (size_t)(nh + 1) * n);
    float *si = malloc(sizeof(float) * (size_t)(nh + 1) * n);
    size_t wc = window_cap(c, n);
    float *re = malloc(sizeof(float) * wc), *im = malloc(sizeof(float) * wc);
    fill_stack(c, i0, n, form, sr, si, filled, re, im);
    for (size_t i = 0; i < (size_t)(nh + 1) * n; i++) {
        out[2 * i] = sr[i];
        out[2 * i + 1] = si[i];
    }
    free(sr); free(si); free(re); free(im);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Noise scales: boxcar.c:18-73, 138-198                              */
/* ------------------------------------------------------------------ */
double orc_analytic_sigma(const uint8_t *filled, int k, int Hk)
{
    int nlow = 0;
    for (int j = 1; j <= Hk - 1; j++) if (filled[j * k - 1]) nlow++;
    int nnyq = filled[Hk * k - 1] ? 1 : 0;
    if (nlow + nnyq < 1) return 0.0;
    return sqrt(2.0 * nlow + 0.5 * nnyq);
}

void orc_boxcar_shape(double *shape, int nbins, const int *widths, int nw,
                      const uint8_t *filled, int k, int Hk)
{
    double sigma = orc_analytic_sigma(filled, k, Hk);
    for (int i = 0; i < nw; i++) {
        int w = widths[i];
        double v = 0.0;
        for (int j = 1; j <= Hk - 1; j++) {
            if (!filled[j * k - 1]) continue;
            double d = orc_sinpi((double)(j * w) / (double)nbins) /
                       orc_sinpi((double)j / (double)nbins);
            v += 2.0 * d * d;
        }
        if (filled[Hk * k - 1]) {
            double d = orc_sinpi(0.5 * (double)w);
            v += 0.5 * d * d;
        }
        shape[i] = (sigma > 0.0 && v > 0.0) ? sigma / sqrt(v) : 0.0;
    }
}

static int fcmp(const void *a, const void *b)
{
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}
static int dcmp(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* Exact median (order statistics are selection-order independent). */
static float median_f(float *v, int n)
{
    qsort(v, (size_t)n, sizeof(float), fcmp);
    int half = n >> 1;
    if (n & 1) return v[half];
    return 0.5f * (v[half - 1] + v[half]);
}
static double median_d(double *v, int n)
{
    qsort(v, (size_t)n, sizeof(double), dcmp);
    int half = n >> 1;
    if (n & 1) return v[half];
    return 0.5 * (v[half - 1] + v[half]);
}

/* Sample selection of block_sigma (boxcar.c:138-164); get(j, i) reads
 * profile value (trial j, bin i). */
#define SIGMA_SAMPLE(T, GET, nbins, n, cap, buf, ns)                          \
    do {                                                                      \
        long long N_ = (long long)(nbins) * (n);                              \
        ns = 0;                                                               \
        if (N_ <= (cap)) {                                                    \
            for (int j_ = 0; j_ < (n); j_++)                                  \
                for (int i_ = 0; i_ < (nbins); i_++) buf[ns++] = GET(j_, i_); \
        } else {                                                              \
            int ncol_ = (int)((cap) / (nbins));                               \
            if (ncol_ < 1) ncol_ = 1;                                         \
            if (ncol_ > (n)) ncol_ = (n);                                     \
            for (int c_ = 0; c_ < ncol_; c_++) {                              \
                long long j0_ = ((long long)c_ * (n)) / ncol_;                \
                int take_ = (nbins);                                          \
                if (take_ > (cap) - ns) take_ = (cap) - ns;                   \
                if (take_ == 0) break;                                        \
                for (int i_ = 0; i_ < take_; i_++)                            \
                    buf[ns++] = GET((int)j0_, i_);                            \
            }                                                                 \
        }                                                                     \
    } while (0)

float orc_block_sigma_f32(const float *M, int nbins, int n, int cap)
{
    if ((long long)nbins * n == 0) return 0.0f;
    float *buf = malloc(sizeof(float) * (size_t)cap);
    int ns;
#define GETROW(j, i) M[(size_t)(j) * nbins + (i)]
    SIGMA_SAMPLE(float, GETROW, nbins, n, cap, buf, ns);
#undef GETROW
    for (int i = 0; i < ns; i++) buf[i] = (float)fabs((double)buf[i]);
    float med = median_f(buf, ns);
    free(buf);
    return (float)1.4826 * med;
}

/* ------------------------------------------------------------------ */
/* Scalar boxcar metric: boxcar.c:203-252                             */
/* ------------------------------------------------------------------ */
float orc_profile_boxcar_f32(const float *prof, int nbins, const int *widths,
                             int nw, float invsigma, const double *shape)
{
    if (!(invsigma > 0.0f)) return 0.0f;
    int wmax = widths[nw - 1];
    float *ps = malloc(sizeof(float) * (size_t)(nbins + wmax + 1));
    ps[0] = 0.0f;
    for (int i = 1; i <= nbins + wmax; i++) {
        int idx = i > nbins ? i - nbins : i;
        ps[i] = ps[i - 1] + (prof[idx - 1] - 0.0f);
    }
    float best = -INFINITY;
    float stot = ps[nbins] - ps[0];
    for (int i = 0; i < nw; i++) {
        int w = widths[i];
        float duty = (float)w / (float)nbins;
        float invsw = (float)((double)invsigma * shape[i]);
        float m = ps[w] - ps[0];
        for (int p = 2; p <= nbins; p++) {
            float d = ps[p + w - 1] - ps[p - 1];
            if (d > m) m = d;
        }
        float cand = (m - duty * stot) * invsw;
        if (cand > best) best = cand;
    }
    free(ps);
    return best;
}

double orc_profile_boxcar_f64(const double *prof, int nbins, const int *widths,
                              int nw, double invsigma, const double *shape)
{
    if (!(invsigma > 0.0)) return 0.0;
    int wmax = widths[nw - 1];
    double *ps = malloc(sizeof(double) * (size_t)(nbins + wmax + 1));
    ps[0] = 0.0;
    for (int i = 1; i <= nbins + wmax; i++) {
        int idx = i > nbins ? i - nbins : i;
        ps[i] = ps[i - 1] + (prof[idx - 1] - 0.0);
    }
    double best = -INFINITY;
    double stot = ps[nbins] - ps[0];
    for (int i = 0; i < nw; i++) {
        int w = widths[i];
        double duty = (double)w / (double)nbins;
        double invsw = (double)((double)invsigma * shape[i]);
        double m = ps[w] - ps[0];
        for (int p = 2; p <= nbins; p++) {
            double d = ps[p + w - 1] - ps[p - 1];
            if (d > m) m = d;
        }
        double cand = (m - duty * stot) * invsw;
        if (cand > best) best = cand;
    }
    free(ps);
    return best;
}

/* ------------------------------------------------------------------ */
/* Chunk engine (per-thread workspace)                                */
/* ------------------------------------------------------------------ */
typedef struct {
    int n_cap;
    float *st_re, *st_im;       /* (nh+1) x Nprof */
    void *prof[65];             /* per slot, bin-major [i * n + j] */
    float *re, *im;             /* window planes */
    size_t wcap;
    void *wk;                   /* c2r work: 4*H*SUB */
    void *ps;                   /* gate prefix: (nb+wmax+1)*SUB */
    double *mvals;              /* (1+ndec) * Nprof */
    uint8_t *filled;
    /* count-keyed shape cache (search.c:36-46) */
    int shapeH[65];
    double shape[65][256];
    uint8_t shape_src[65][256]; /* unused: kept for debugging symmetry */
} OWork;

static int slot_nb(const OrcCtx *c, int s)
{
    return s == 0 ? 2 * c->cfg.nharms : 2 * c->dHk[s - 1];
}
static int slot_k(const OrcCtx *c, int s) { return s == 0 ? 1 : c->dk[s - 1]; }
static int slot_Hk(const OrcCtx *c, int s)
{
    return s == 0 ? c->cfg.nharms : c->dHk[s - 1];
}

static OWork *owork_create(const OrcCtx *c)
{
    OWork *w = calloc(1, sizeof(OWork));
    const int nh = c->cfg.nharms, Np = c->Nprof;
    const size_t es = c->cfg.precision ? sizeof(double) : sizeof(float);
    w->n_cap = Np;
    w->st_re = malloc(sizeof(float) * (size_t)(nh + 1) * Np);
    w->st_im = malloc(sizeof(float) * (size_t)(nh + 1) * Np);
    for (int s = 0; s <= c->ndec; s++)
        w->prof[s] = malloc(es * (size_t)slot_nb(c, s) * Np);
    w->wcap = window_cap(c, Np);
    w->re = malloc(sizeof(float) * w->wcap);
    w->im = malloc(sizeof(float) * w->wcap);
    w->wk = malloc(es * 4 * (size_t)nh * ORC_SUB + 64);
    int maxps = 0;
    for (int s = 0; s <= c->ndec; s++) {
        int v = slot_nb(c, s) + c->widths[s][c->nw[s] - 1] + 1;
        if (v > maxps) maxps = v;
    }
    w->ps = malloc(es * (size_t)maxps * ORC_SUB);
    w->mvals = malloc(sizeof(double) * (size_t)(c->ndec + 1) * Np);
    w->filled = malloc((size_t)nh);
    for (int s = 0; s <= 64; s++) w->shapeH[s] = -1;
    return w;
}

static void owork_free(OWork *w, const OrcCtx *c)
{
    if (!w) return;
    free(w->st_re); free(w->st_im);
    for (int s = 0; s <= c->ndec; s++) free(w->prof[s]);
    free(w->re); free(w->im); free(w->wk); free(w->ps); free(w->mvals);
    free(w->filled);
    free(w);
}

static int nvalid_of(const OrcCtx *c, int k, double rstart, int n)
{
    double nyq = (double)c->N / 2.0;
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = lo + ((hi - lo) >> 1);
        if ((double)k * (rstart + (double)mid * c->lodr) < nyq) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* Batched gate over bin-major profiles prof[i*ldp + j], trials j0..j0+B-1.
 * Same adds and compares per trial as orc_profile_boxcar_*. */
static void gate_f32(const float *prof, long ldp, int j0, int B, int nb,
                     const int *widths, int nw, float invsigma,
                     const double *shape, float *ps, double *mout)
{
    if (!(invsigma > 0.0f)) {
        for (int b = 0; b < B; b++) mout[b] = 0.0;
        return;
    }
    const int wmax = widths[nw - 1];
    for (int b = 0; b < B; b++) ps[b] = 0.0f;
    for (int i = 1; i <= nb + wmax; i++) {
        int idx = (i > nb ? i - nb : i) - 1;
        const float *src = prof + (long)idx * ldp + j0;
        float *d = ps + (size_t)i * B, *s = ps + (size_t)(i - 1) * B;
        for (int b = 0; b < B; b++) d[b] = s[b] + (src[b] - 0.0f);
    }
    float best[ORC_SUB], mx[ORC_SUB];
    for (int b = 0; b < B; b++) best[b] = -INFINITY;
    for (int wi = 0; wi < nw; wi++) {
        int w = widths[wi];
        float duty = (float)w / (float)nb;
        float invsw = (float)((double)invsigma * shape[wi]);
        for (int b = 0; b < B; b++) mx[b] = ps[(size_t)w * B + b] - ps[b];
        for (int p = 2; p <= nb; p++) {
            const float *hi = ps + (size_t)(p + w - 1) * B, *lo = ps + (size_t)(p - 1) * B;
            for (int b = 0; b < B; b++) {
                float d = hi[b] - lo[b];
                mx[b] = d > mx[b] ? d : mx[b];
            }
        }
        for (int b = 0; b < B; b++) {
            float st = ps[(size_t)nb * B + b] - ps[b];
            float cand = (mx[b] - duty * st) * invsw;
            best[b] = cand > best[b] ? cand : best[b];
        }
    }
    for (int b = 0; b < B; b++) mout[b] = (double)best[b];
}

static void gate_f64(const double *prof, long ldp, int j0, int B, int nb,
                     const int *widths, int nw, double invsigma,
                     const double *shape, double *ps, double *mout)
{
    if (!(invsigma > 0.0)) {
        for (int b = 0; b < B; b++) mout[b] = 0.0;
        return;
    }
    const int wmax = widths[nw - 1];
    for (int b = 0; b < B; b++) ps[b] = 0.0;
    for (int i = 1; i <= nb + wmax; i++) {
        int idx = (i > nb ? i - nb : i) - 1;
        const double *src = prof + (long)idx * ldp + j0;
        double *d = ps + (size_t)i * B, *s = ps + (size_t)(i - 1) * B;
        for (int b = 0; b < B; b++) d[b] = s[b] + (src[b] - 0.0);
    }
    double best[ORC_SUB], mx[ORC_SUB];
    for (int b = 0; b < B; b++) best[b] = -INFINITY;
    for (int wi = 0; wi < nw; wi++) {
        int w = widths[wi];
        double duty = (double)w / (double)nb;
        double invsw = invsigma * shape[wi];
        for (int b = 0; b < B; b++) mx[b] = ps[(size_t)w * B + b] - ps[b];
        for (int p = 2; p <= nb; p++) {
            const double *hi = ps + (size_t)(p + w - 1) * B, *lo = ps + (size_t)(p - 1) * B;
            for (int b = 0; b < B; b++) {
                double d = hi[b] - lo[b];
                mx[b] = d > mx[b] ? d : mx[b];
            }
        }
        for (int b = 0; b < B; b++) {
            double st = ps[(size_t)nb * B + b] - ps[b];
            double cand = (mx[b] - duty * st) * invsw;
            best[b] = cand > best[b] ? cand : best[b];
        }
    }
    for (int b = 0; b < B; b++) mout[b] = best[b];
}

/* Profiles of every slot for one chunk into w->prof (bin-major, ld = n). */
static void chunk_profiles(const OrcCtx *c, OWork *w, long long i0, int n)
{
    const int nh = c->cfg.nharms;
    fill_stack(c, i0, n, 0, w->st_re, w->st_im, w->filled, w->re, w->im);
    const int f64 = c->cfg.precision == 1;
    for (int j0 = 0; j0 < n; j0 += ORC_SUB) {
        int B = n - j0 < ORC_SUB ? n - j0 : ORC_SUB;
        if (!f64) {
            c2r_f32(nh, w->st_re + j0, w->st_im + j0, n, (float *)w->prof[0] + j0,
                    n, B, (float *)w->wk);
        } else {
            /* f64: the f32 interpolated values widened to double */
            double *xr = malloc(sizeof(double) * (size_t)(nh + 1) * B);
            double *xi = malloc(sizeof(double) * (size_t)(nh + 1) * B);
            for (int h = 0; h <= nh; h++)
                for (int b = 0; b < B; b++) {
                    xr[(size_t)h * B + b] = (double)w->st_re[(size_t)h * n + j0 + b];
                    xi[(size_t)h * B + b] = (double)w->st_im[(size_t)h * n + j0 + b];
                }
            c2r_f64(nh, xr, xi, B, (double *)w->prof[0] + j0, n, B, (double *)w->wk);
            for (int d = 0; d < c->ndec; d++) {
                if (c->dfold[d]) continue;
                int k = c->dk[d];
                c2r_f64(c->dHk[d], xr, xi, (long)k * B, (double *)w->prof[d + 1] + j0,
                        n, B, (double *)w->wk);
            }
            free(xr); free(xi);
        }
        for (int d = 0; d < c->ndec; d++) {
            const int k = c->dk[d], nk = 2 * c->dHk[d];
            if (c->dfold[d]) {
                if (!f64) {
                    const float *s = (const float *)w->prof[0];
                    float *o = (float *)w->prof[d + 1];
                    const float invk = 1.0f / (float)k;
                    for (int jj = 0; jj < nk; jj++)
                        for (int b = 0; b < B; b++) {
                            float a = s[(long)jj * n + j0 + b];
                            for (int qq = 1; qq < k; qq++)
                                a += s[(long)(jj + qq * nk) * n + j0 + b];
                            o[(long)jj * n + j0 + b] = a * invk;
                        }
                } else {
                    const double *s = (const double *)w->prof[0];
                    double *o = (double *)w->prof[d + 1];
                    const double invk = 1.0 / (double)k;
                    for (int jj = 0; jj < nk; jj++)
                        for (int b = 0; b < B; b++) {
                            double a = s[(long)jj * n + j0 + b];
                            for (int qq = 1; qq < k; qq++)
                                a += s[(long)(jj + qq * nk) * n + j0 + b];
                            o[(long)jj * n + j0 + b] = a * invk;
                        }
                }
            } else if (!f64) {
                c2r_f32(c->dHk[d], w->st_re + j0, w->st_im + j0, (long)k * n,
                        (float *)w->prof[d + 1] + j0, n, B, (float *)w->wk);
            }
        }
    }
}

/* Metrics of every slot of one chunk into w->mvals[slot*n + j] (NaN where a
 * decimation's trial is above its Nyquist).  Updates the shape cache. */
static void chunk_metrics(const OrcCtx *c, OWork *w, long long i0, int n,
                          double rstart)
{
    chunk_profiles(c, w, i0, n);
    const int f64 = c->cfg.precision == 1;
    for (int s = 0; s <= c->ndec; s++) {
        const int k = slot_k(c, s), Hk = slot_Hk(c, s), nb = slot_nb(c, s);
        const int nv = s == 0 ? n : nvalid_of(c, k, rstart, n);
        double *mv = w->mvals + (size_t)s * n;
        for (int j = 0; j < n; j++) mv[j] = NAN;
        /* refresh_shape (search.c:36-46) */
        int nf = 0;
        for (int j = 1; j <= Hk; j++) if (w->filled[j * k - 1]) nf++;
        if (w->shapeH[s] != nf) {
            orc_boxcar_shape(w->shape[s], nb, c->widths[s], c->nw[s], w->filled, k, Hk);
            w->shapeH[s] = nf;
        }
        if (nv <= 0) continue;
        if (!f64) {
            float sig;
            if (c->cfg.sigma == 0) sig = (float)orc_analytic_sigma(w->filled, k, Hk);
            else {
                const float *P = (const float *)w->prof[s];
                float *buf = malloc(sizeof(float) * ORC_SIGCAP);
                int ns;
#define GETBM(j, i) P[(size_t)(i) * n + (j)]
                if ((long long)nb * nv == 0) { sig = 0.0f; }
                else {
                    SIGMA_SAMPLE(float, GETBM, nb, nv, ORC_SIGCAP, buf, ns);
                    for (int i = 0; i < ns; i++) buf[i] = (float)fabs((double)buf[i]);
                    sig = (float)1.4826 * median_f(buf, ns);
                }
#undef GETBM
                free(buf);
            }
            float invsigma = sig > 0.0f ? 1.0f / sig : 0.0f;
            for (int j0 = 0; j0 < nv; j0 += ORC_SUB) {
                int B = nv - j0 < ORC_SUB ? nv - j0 : ORC_SUB;
                gate_f32((const float *)w->prof[s], n, j0, B, nb, c->widths[s],
                         c->nw[s], invsigma, w->shape[s], (float *)w->ps, mv + j0);
            }
        } else {
            double sig;
            if (c->cfg.sigma == 0) sig = orc_analytic_sigma(w->filled, k, Hk);
            else {
                const double *P = (const double *)w->prof[s];
                double *buf = malloc(sizeof(double) * ORC_SIGCAP);
                int ns;
#define GETBMD(j, i) P[(size_t)(i) * n + (j)]
                if ((long long)nb * nv == 0) { sig = 0.0; }
                else {
                    SIGMA_SAMPLE(double, GETBMD, nb, nv, ORC_SIGCAP, buf, ns);
                    for (int i = 0; i < ns; i++) buf[i] = fabs(buf[i]);
                    sig = 1.4826 * median_d(buf, ns);
                }
#undef GETBMD
                free(buf);
            }
            double invsigma = sig > 0.0 ? 1.0 / sig : 0.0;
            for (int j0 = 0; j0 < nv; j0 += ORC_SUB) {
                int B = nv - j0 < ORC_SUB ? nv - j0 : ORC_SUB;
                gate_f64((const double *)w->prof[s], n, j0, B, nb, c->widths[s],
                         c->nw[s], invsigma, w->shape[s], (double *)w->ps, mv + j0);
            }
        }
    }
}

int orc_chunk_profiles(const OrcCtx *c, long long ci, int slot, void *out)
{
    long long i0; int n; double rs;
    if (orc_ctx_chunk(c, ci, &i0, &n, &rs) != 0 || slot < 0 || slot > c->ndec) return -1;
    OWork *w = owork_create(c);
    chunk_profiles(c, w, i0, n);
    const int nb = slot_nb(c, slot);
    if (c->cfg.precision == 1) {
        const double *P = (const double *)w->prof[slot];
        double *o = (double *)out;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < nb; i++) o[(size_t)j * nb + i] = P[(size_t)i * n + j];
    } else {
        const float *P = (const float *)w->prof[slot];
        float *o = (float *)out;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < nb; i++) o[(size_t)j * nb + i] = P[(size_t)i * n + j];
    }
    owork_free(w, c);
    return nb;
}

/* Replay the single-thread shape cache up to (not including) chunk ci. */
static void replay_shape_cache(const OrcCtx *c, OWork *w, long long ci)
{
    uint8_t *f = malloc((size_t)c->cfg.nharms);
    for (long long cc = 0; cc < ci; cc++) {
        long long i0; int n; double rs;
        orc_ctx_chunk(c, cc, &i0, &n, &rs);
        chunk_filled(c, i0, n, f);
        for (int s = 0; s <= c->ndec; s++) {
            const int k = slot_k(c, s), Hk = slot_Hk(c, s);
            int nf = 0;
            for (int j = 1; j <= Hk; j++) if (f[j * k - 1]) nf++;
            if (w->shapeH[s] != nf) {
                orc_boxcar_shape(w->shape[s], slot_nb(c, s), c->widths[s], c->nw[s],
                                 f, k, Hk);
                w->shapeH[s] = nf;
            }
        }
    }
    free(f);
}

int orc_chunk_metrics(const OrcCtx *c, long long ci, double *out)
{
    long long i0; int n; double rs;
    if (orc_ctx_chunk(c, ci, &i0, &n, &rs) != 0) return -1;
    OWork *w = owork_create(c);
    replay_shape_cache(c, w, ci);
    chunk_metrics(c, w, i0, n, rs);
    memcpy(out, w->mvals, sizeof(double) * (size_t)(c->ndec + 1) * n);
    owork_free(w, c);
    return n;
}

/* ------------------------------------------------------------------ */
/* Search: search.c:947-1095 (OpenMP round-robin chunks)              */
/* ------------------------------------------------------------------ */
typedef struct { OrcCand *v; size_t n, cap; } OList;
static void olist_push(OList *l, OrcCand c)
{
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 256;
        l->v = realloc(l->v, l->cap * sizeof(OrcCand));
    }
    l->v[l->n++] = c;
}

long long orc_search(const OrcCtx *c, double threshold, int nthreads,
                     long long chunk_lo, long long chunk_hi, OrcCand **out)
{
    long long nch = c->nchunks;
    if (chunk_hi < 0 || chunk_hi > nch) chunk_hi = nch;
    if (chunk_lo < 0) chunk_lo = 0;
    long long span = chunk_hi - chunk_lo;
    int nt = nthreads < 1 ? 1 : nthreads;
    if (nt > span) nt = (int)(span > 0 ? span : 1);
    OList *tl = calloc((size_t)nt, sizeof(OList));
    #pragma omp parallel num_threads(nt)
    {
        int tid = omp_get_thread_num();
        OWork *w = owork_create(c);
        for (long long ci = chunk_lo + tid; ci < chunk_hi; ci += nt) {
            long long i0; int n; double rstart;
            orc_ctx_chunk(c, ci, &i0, &n, &rstart);
            chunk_metrics(c, w, i0, n, rstart);
            for (int s = 0; s <= c->ndec; s++) {
                const int k = slot_k(c, s), Hk = slot_Hk(c, s);
                const double *mv = w->mvals + (size_t)s * n;
                for (int j = 0; j < n; j++) {
                    double m = mv[j];
                    if (!(m == m)) continue;
                    double rf = rstart + (double)j * c->lodr;
                    double r = s == 0 ? rf : (double)k * (rstart + (double)j * c->lodr);
                    if (s != 0 && !(r < (double)c->N / 2.0)) continue;
                    if (m > threshold) {
                        OrcCand cd;
                        cd.freq = r / c->T;
                        cd.metric = m;
                        cd.r = r;
                        cd.nharm = Hk;
                        cd.ducy = NAN;
                        olist_push(&tl[tid], cd);
                    }
                }
            }
        }
        owork_free(w, c);
    }
    size_t tot = 0;
    for (int t = 0; t < nt; t++) tot += tl[t].n;
    OrcCand *v = malloc(sizeof(OrcCand) * (tot ? tot : 1));
    size_t o = 0;
    for (int t = 0; t < nt; t++) {
        if (tl[t].n) memcpy(v + o, tl[t].v, tl[t].n * sizeof(OrcCand));
        o += tl[t].n;
        free(tl[t].v);
    }
    free(tl);
    *out = v;
    return (long long)tot;
}

/* ------------------------------------------------------------------ */
/* FP64 oracle                                                        */
/* ------------------------------------------------------------------ */
void orc_fourier_interp_complex(double r, const float *amps, long long namps,
                                int m, double *out_re, double *out_im)
{
    const int m2 = m / 2;
    double dr = fmod(r, 1.0);
    if (dr < 0) dr += 1.0;
    long long rint0 = (long long)floor(r + 1e-15) + 1;
    long long lo0 = rint0 - m2;
    double sre = 0.0, sim = 0.0;
    for (int j = 0; j < m; j++) {
        double off = dr - (double)(-m2 + 1 + j);
        double sinc = off == 0.0 ? 1.0 : sin(M_PI * off) / (M_PI * off);
        double cc = cos(M_PI * off), ss = sin(M_PI * off);
        double cr = sinc * cc, ci = -sinc * ss;
        long long idx = lo0 + j;
        double bre = 0.0, bim = 0.0;
        if (idx >= 0 && idx < namps) {
            bre = (double)amps[2 * idx];
            bim = (double)amps[2 * idx + 1];
        }
        sre += cr * bre - ci * bim;
        sim += cr * bim + ci * bre;
    }
    *out_re = sre;
    *out_im = sim;
}

/* x[t] = X0 + (-1)^t XH + 2 sum_{h=1}^{H-1} Re(X_h e^{2 pi i h t / n}) */
static void dft_direct_d(int H, const double *Xr, const double *Xi, int hs,
                         double *x)
{
    const int n = 2 * H;
    for (int t = 0; t < n; t++) {
        double s = Xr[0] + ((t & 1) ? -Xr[(long)H * hs] : Xr[(long)H * hs]);
        double acc = 0.0;
        for (int h = 1; h < H; h++) {
            long j = ((long)h * t) % n;
            double a = 2.0 * M_PI * (double)j / (double)n;
            acc += Xr[(long)h * hs] * cos(a) - Xi[(long)h * hs] * sin(a);
        }
        x[t] = s + 2.0 * acc;
    }
}

/* FP64 metrics of rows [j0, j0+nr) of chunk ci (NaN elsewhere).  Measured
 * sigma still sees the whole chunk: the rows its sample reads are marked by a
 * dry run of the sampler and computed in the first pass. */
static int chunk_metrics_fp64_rows(const OrcCtx *c, long long ci, int j0, int nr,
                                   double *out)
{
    long long i0; int n; double rstart;
    if (orc_ctx_chunk(c, ci, &i0, &n, &rstart) != 0) return -1;
    if (j0 < 0 || nr < 0 || j0 + nr > n) return -1;
    const int nh = c->cfg.nharms;
    uint8_t *f = malloc((size_t)nh);
    chunk_filled(c, i0, n, f);
    uint8_t *need = calloc((size_t)n, 1);
    for (int j = j0; j < j0 + nr; j++) need[j] = 1;
    if (c->cfg.sigma == 1) {
        double *dry = malloc(sizeof(double) * ORC_SIGCAP);
        for (int s = 0; s <= c->ndec; s++) {
            const int k = slot_k(c, s), nb = slot_nb(c, s);
            int nv = s == 0 ? n : nvalid_of(c, k, rstart, n);
            int ns;
#define GETMARK(j, i) (need[(j)] = 1, 0.0)
            if ((long long)nb * nv > 0) SIGMA_SAMPLE(double, GETMARK, nb, nv, ORC_SIGCAP, dry, ns);
#undef GETMARK
            (void)ns;
        }
        free(dry);
    }
    double *Xr = calloc((size_t)nh + 1, sizeof(double));
    double *Xi = calloc((size_t)nh + 1, sizeof(double));
    double *x = malloc(sizeof(double) * 2 * (nh + 1));
    double shape[256];
    /* measured sigma needs the whole chunk's double profiles */
    double **profs = NULL;
    if (c->cfg.sigma == 1) {
        profs = calloc(c->ndec >= 0 ? (size_t)c->ndec + 1 : 1, sizeof(double *));
        for (int s = 0; s <= c->ndec; s++)
            profs[s] = malloc(sizeof(double) * (size_t)slot_nb(c, s) * n);
    }
    double *mv = malloc(sizeof(double) * (size_t)(c->ndec + 1) * n);
    for (int pass = 0; pass < (c->cfg.sigma == 1 ? 2 : 1); pass++) {
        double sig[65];
        if (pass == 1 || c->cfg.sigma == 0) {
            for (int s = 0; s <= c->ndec; s++) {
                const int k = slot_k(c, s), Hk = slot_Hk(c, s), nb = slot_nb(c, s);
                if (c->cfg.sigma == 0) sig[s] = orc_analytic_sigma(f, k, Hk);
                else {
                    int nv = s == 0 ? n : nvalid_of(c, k, rstart, n);
                    double *buf = malloc(sizeof(double) * ORC_SIGCAP);
                    int ns = 0;
                    const double *P = profs[s];
#define GETRD(j, i) P[(size_t)(j) * nb + (i)]
                    if ((long long)nb * nv > 0) {
                        SIGMA_SAMPLE(double, GETRD, nb, nv, ORC_SIGCAP, buf, ns);
                        for (int i = 0; i < ns; i++) buf[i] = fabs(buf[i]);
                        sig[s] = 1.4826 * median_d(buf, ns);
                    } else sig[s] = 0.0;
#undef GETRD
                    free(buf);
                }
            }
        }
        for (int j = 0; j < n; j++) {
            if (!need[j]) {
                for (int s = 0; s <= c->ndec; s++) mv[(size_t)s * n + j] = NAN;
                continue;
            }
            if (pass == 1 && (j < j0 || j >= j0 + nr)) continue;
            long long t = i0 + j;
            for (int h = 1; h <= nh; h++) {
                Xr[h] = Xi[h] = 0.0;
                if (!f[h - 1]) continue;
                const OPlan *dp = &c->plans[h - 1];
                double pos;
                if (dp->P == 0) {
                    pos = opos_slow(dp, t);
                } else {
                    int res; long long qint;
                    ostate(dp, t, &res, &qint);
                    pos = (double)((long long)dp->rfloor0 + qint) +
                          (dp->dr0 + (double)res / (double)dp->q);
                }
                orc_fourier_interp_complex(pos, c->amps, c->namps, c->cfg.m, &Xr[h], &Xi[h]);
            }
            for (int s = 0; s <= c->ndec; s++) {
                const int k = slot_k(c, s), Hk = slot_Hk(c, s), nb = slot_nb(c, s);
                dft_direct_d(Hk, Xr, Xi, k, x);
                if (pass == 0 && c->cfg.sigma == 1) {
                    memcpy(profs[s] + (size_t)j * nb, x, sizeof(double) * nb);
                    continue;
                }
                int nv = s == 0 ? n : nvalid_of(c, k, rstart, n);
                if (j >= nv) { mv[(size_t)s * n + j] = NAN; continue; }
                orc_boxcar_shape(shape, nb, c->widths[s], c->nw[s], f, k, Hk);
                double invs = sig[s] > 0.0 ? 1.0 / sig[s] : 0.0;
                mv[(size_t)s * n + j] =
                    orc_profile_boxcar_f64(x, nb, c->widths[s], c->nw[s], invs, shape);
            }
        }
    }
    for (int s = 0; s <= c->ndec; s++)
        for (int j = 0; j < n; j++)
            if (j < j0 || j >= j0 + nr) mv[(size_t)s * n + j] = NAN;
    memcpy(out, mv, sizeof(double) * (size_t)(c->ndec + 1) * n);
    if (profs) {
        for (int s = 0; s <= c->ndec; s++) free(profs[s]);
        free(profs);
    }
    free(mv); free(f); free(Xr); free(Xi); free(x); free(need);
    return n;
}

int orc_chunk_metrics_fp64(const OrcCtx *c, long long ci, double *out)
{
    long long i0; int n; double rstart;
    if (orc_ctx_chunk(c, ci, &i0, &n, &rstart) != 0) return -1;
    return chunk_metrics_fp64_rows(c, ci, 0, n, out);
}

int orc_chunk_metrics_fp64_rows(const OrcCtx *c, long long ci, int j0, int nr, double *out)
{
    return chunk_metrics_fp64_rows(c, ci, j0, nr, out);
}

/* ------------------------------------------------------------------ */
/* Candidate post-processing: candidate.c:43-161                      */
/* ------------------------------------------------------------------ */
typedef int (*ocmp_fn)(const OrcCand *, const OrcCand *);
static int ocmp_metric_desc(const OrcCand *a, const OrcCand *b)
{
    return (b->metric > a->metric) - (b->metric < a->metric);
}
static int ocmp_freq(const OrcCand *a, const OrcCand *b)
{
    return (a->freq > b->freq) - (a->freq < b->freq);
}
static int ocmp_r(const OrcCand *a, const OrcCand *b)
{
    return (a->r > b->r) - (a->r < b->r);
}
/* stable bottom-up merge sort */
static void osort(OrcCand *v, size_t n, ocmp_fn cmp)
{
    if (n < 2) return;
    OrcCand *tmp = malloc(n * sizeof(OrcCand));
    for (size_t width = 1; width < n; width *= 2) {
        for (size_t i = 0; i < n; i += 2 * width) {
            size_t mid = i + width < n ? i + width : n;
            size_t end = i + 2 * width < n ? i + 2 * width : n;
            size_t a = i, b = mid, k = i;
            while (a < mid && b < end) {
                if (cmp(&v[b], &v[a]) < 0) tmp[k++] = v[b++];
                else tmp[k++] = v[a++];
            }
            while (a < mid) tmp[k++] = v[a++];
            while (b < end) tmp[k++] = v[b++];
        }
        memcpy(v, tmp, n * sizeof(OrcCand));
    }
    free(tmp);
}
void orc_sort_metric_desc(OrcCand *v, size_t n) { osort(v, n, ocmp_metric_desc); }
void orc_sort_freq(OrcCand *v, size_t n) { osort(v, n, ocmp_freq); }

size_t orc_remove_duplicates(OrcCand *v, size_t n, double dr_tol)
{
    if (n == 0) return 0;
    osort(v, n, ocmp_r);
    OrcCand *o = malloc(n * sizeof(OrcCand));
    size_t no = 0;
    OrcCand best = v[0];
    double prev = best.r;
    for (size_t i = 1; i < n; i++) {
        OrcCand c = v[i];
        if (c.r - prev <= dr_tol) {
            if (c.metric > best.metric) best = c;
        } else {
            o[no++] = best;
            best = c;
        }
        prev = c.r;
    }
    o[no++] = best;
    osort(o, no, ocmp_freq);
    memcpy(v, o, no * sizeof(OrcCand));
    free(o);
    return no;
}

int orc_harmonically_related(double r1, double r2, int numharm, double tol)
{
    double lo = r1 < r2 ? r1 : r2, hi = r1 < r2 ? r2 : r1;
    if (!(lo > 0.0)) return 0;
    for (int m = 1; m <= numharm; m++) {
        int nn = (int)nearbyint((double)m * hi / lo);
        if (nn < 1 || nn > numharm) continue;
        if (fabs((double)m * hi - (double)nn * lo) <= tol * (double)m) return 1;
    }
    return 0;
}

size_t orc_remove_harmonics(OrcCand *v, size_t n, int numharm, double tol)
{
    if (n == 0) return 0;
    osort(v, n, ocmp_metric_desc);
    OrcCand *o = malloc(n * sizeof(OrcCand));
    size_t no = 0;
    for (size_t i = 0; i < n; i++) {
        int rel = 0;
        for (size_t j = 0; j < no; j++)
            if (orc_harmonically_related(v[i].r, o[j].r, numharm, tol)) { rel = 1; break; }
        if (!rel) o[no++] = v[i];
    }
    osort(o, no, ocmp_freq);
    memcpy(v, o, no * sizeof(OrcCand));
    free(o);
    return no;
}

/* boxcar.c:736-785 */
int orc_boxcar_best_width(const double *prof, int nbins, double fsp,
                          double maxfrac, int *out_w, double *out_ducy)
{
    if (nbins < 1) return -1;
    int widths[256];
    int nw = orc_boxcar_widths(nbins, fsp, maxfrac, widths, 256);
    if (nw <= 0) { *out_w = 1; *out_ducy = 1.0 / nbins; return 0; }
    int wmax = widths[nw - 1];
    double *ps = malloc(sizeof(double) * (size_t)(nbins + wmax + 1));
    double sum = 0.0;
    for (int i = 0; i < nbins; i++) sum += prof[i];
    double mean = sum / (double)nbins;
    ps[0] = 0.0;
    for (int i = 1; i <= nbins + wmax; i++) {
        int idx = i > nbins ? i - nbins : i;
        ps[i] = ps[i - 1] + (prof[idx - 1] - mean);
    }
    double stot = ps[nbins] - ps[0];
    int Hk = nbins / 2;
    uint8_t *filled = malloc((size_t)(Hk > 0 ? Hk : 1));
    for (int j = 0; j < Hk; j++) filled[j] = 1;
    double shape[256];
    if (Hk >= 1) orc_boxcar_shape(shape, nbins, widths, nw, filled, 1, Hk);
    else for (int i = 0; i < nw; i++) shape[i] = 0.0;
    double best = -INFINITY;
    int bestw = widths[0];
    for (int i = 0; i < nw; i++) {
        int w = widths[i];
        double duty = (double)w / (double)nbins;
        double m = ps[w] - ps[0];
        for (int p = 2; p <= nbins; p++) {
            double d = ps[p + w - 1] - ps[p - 1];
            if (d > m) m = d;
        }
        double cand = (m - duty * stot) * shape[i];
        if (cand > best) { best = cand; bestw = w; }
    }
    *out_w = bestw;
    *out_ducy = (double)bestw / (double)nbins;
    free(filled);
    free(ps);
    return 0;
}

/* candidate.c:208-329 (c2r by direct double DFT, normalised by 1/n) */
void orc_measure_ducy(const float *amps, long long namps, long long N,
                      OrcCand *v, size_t n, double fsp, double maxfrac)
{
    const int m = 64, m2 = 32;
    for (size_t ci = 0; ci < n; ci++) {
        OrcCand *c = &v[ci];
        c->ducy = NAN;
        if (!(c->nharm >= 1 && c->r > 0.0)) continue;
        int nharm = c->nharm;
        double *Xr = calloc((size_t)nharm + 1, sizeof(double));
        double *Xi = calloc((size_t)nharm + 1, sizeof(double));
        int H = 0;
        for (int h = 1; h <= nharm; h++) {
            double rh = c->r * (double)h;
            long long rint0 = (long long)floor(rh + 1e-15) + 1;
            if (!(rint0 - m2 >= 0 && rint0 + m2 <= namps && rh < (double)(N / 2))) break;
            orc_fourier_interp_complex(rh, amps, namps, m, &Xr[h], &Xi[h]);
            H++;
        }
        if (H < 1) {
            double z[2] = {0.0, 0.0};
            int w; double d;
            if (orc_boxcar_best_width(z, 2, fsp, maxfrac, &w, &d) == 0) c->ducy = d;
            free(Xr); free(Xi);
            continue;
        }
        int nb = 2 * H;
        double *x = malloc(sizeof(double) * nb);
        dft_direct_d(H, Xr, Xi, 1, x);
        for (int i = 0; i < nb; i++) x[i] *= 1.0 / (double)nb;
        int w; double d;
        if (orc_boxcar_best_width(x, nb, fsp, maxfrac, &w, &d) == 0) c->ducy = d;
        free(x); free(Xr); free(Xi);
    }
}

void orc_format_double(char *buf, size_t n, double x)
{
    if (isnan(x)) { snprintf(buf, n, "NaN"); return; }
    if (isinf(x)) { snprintf(buf, n, x > 0 ? "Inf" : "-Inf"); return; }
    for (int prec = 1; prec <= 17; prec++) {
        snprintf(buf, n, "%.*g", prec, x);
        char *end = NULL;
        if (strtod(buf, &end) == x && end && *end == '\0') break;
    }
    if (!strpbrk(buf, ".eE") && !strstr(buf, "inf") && !strstr(buf, "nan")) {
        size_t l = strlen(buf);
        if (l + 3 < n) { buf[l] = '.'; buf[l + 1] = '0'; buf[l + 2] = '\0'; }
    }
}

size_t orc_format_candidates(const OrcCand *v, size_t n, char *buf, size_t cap)
{
    char *s = NULL;
    size_t len = 0;
    FILE *f = open_memstream(&s, &len);
    fprintf(f, "#Num    'S/N'      Frequency (Hz)        Period (ms)    #Harm  Ducy(%%)\n");
    for (size_t i = 0; i < n; i++) {
        char d[32];
        if (isnan(v[i].ducy)) snprintf(d, sizeof(d), "-");
        else snprintf(d, sizeof(d), "%.2f", 100.0 * v[i].ducy);
        fprintf(f, "%-4zu  %8.2f  %18.12f  %18.12f   %3d   %6s\n", i + 1,
                v[i].metric, v[i].freq, 1000.0 / v[i].freq, v[i].nharm, d);
    }
    fclose(f);
    if (buf && cap > 0) {
        size_t c = len < cap - 1 ? len : cap - 1;
        memcpy(buf, s, c);
        buf[c] = '\0';
    }
    free(s);
    return len + 1;
}

/* ------------------------------------------------------------------ */
/* realfft + rednoise: datpipe.c:52-253 with a double FFT             */
/* ------------------------------------------------------------------ */
/* In-place iterative radix-2 complex FFT (sign -1), n power of two. */
static void fft_pow2_d(double *re, double *im, long n, int sign)
{
    for (long i = 1, j = 0; i < n; i++) {
        long bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (long len = 2; len <= n; len <<= 1) {
        long half = len >> 1;
        for (long k = 0; k < half; k++) {
            double a = sign * 2.0 * M_PI * (double)k / (double)len;
            double wr = cos(a), wi = sin(a);
            for (long i = k; i < n; i += len) {
                long j = i + half;
                double xr = re[j] * wr - im[j] * wi;
                double xi = re[j] * wi + im[j] * wr;
                re[j] = re[i] - xr; im[j] = im[i] - xi;
                re[i] += xr; im[i] += xi;
            }
        }
    }
}

/* Complex forward DFT of any length (Bluestein for non powers of two). */
static int fft_any_d(double *re, double *im, long n)
{
    if (n <= 0) return -1;
    if ((n & (n - 1)) == 0) { fft_pow2_d(re, im, n, -1); return 0; }
    long m = 1;
    while (m < 2 * n - 1) m <<= 1;
    double *ar = calloc((size_t)m, sizeof(double)), *ai = calloc((size_t)m, sizeof(double));
    double *br = calloc((size_t)m, sizeof(double)), *bi = calloc((size_t)m, sizeof(double));
    double *wr = malloc(sizeof(double) * n), *wi = malloc(sizeof(double) * n);
    if (!ar || !ai || !br || !bi || !wr || !wi) return -1;
    for (long k = 0; k < n; k++) {
        long kk = (long)(((__int128)k * k) % (2 * n));
        double a = M_PI * (double)kk / (double)n;
        wr[k] = cos(a); wi[k] = -sin(a);     /* e^{-i pi k^2 / n} */
    }
    for (long k = 0; k < n; k++) {
        ar[k] = re[k] * wr[k] - im[k] * wi[k];
        ai[k] = re[k] * wi[k] + im[k] * wr[k];
    }
    br[0] = wr[0]; bi[0] = -wi[0];
    for (long k = 1; k < n; k++) {
        br[k] = br[m - k] = wr[k];
        bi[k] = bi[m - k] = -wi[k];
    }
    fft_pow2_d(ar, ai, m, -1);
    fft_pow2_d(br, bi, m, -1);
    for (long k = 0; k < m; k++) {
        double xr = ar[k] * br[k] - ai[k] * bi[k];
        double xi = ar[k] * bi[k] + ai[k] * br[k];
        ar[k] = xr; ai[k] = xi;
    }
    fft_pow2_d(ar, ai, m, 1);
    for (long k = 0; k < n; k++) {
        double xr = ar[k] / (double)m, xi = ai[k] / (double)m;
        re[k] = xr * wr[k] - xi * wi[k];
        im[k] = xr * wi[k] + xi * wr[k];
    }
    free(ar); free(ai); free(br); free(bi); free(wr); free(wi);
    return 0;
}

/* FP contraction of the staging and zap code.  The reference compiles
 * datpipe.c and zap.c with -ffp-contract=fast (src/Makefile:129-138), so gcc
 * fuses, where FMA exists, exactly these expressions of theirs:
 *   r*r + im*im                          -> fmaf(r, r, im*im)
 *   mean_old + dslope*(float)ii          -> fmaf(dslope, (float)ii, mean_old)
 *   y*(1.5f - 0.5f*x*y*y)                -> y*fmaf(-(0.5f*x*y), y, 1.5f)
 *   h1r + a - f, a = h2r*wr, f = h2i*wi  -> fma(-h2i, wi, fma(h2r, wr, h1r)), and the
 *                                           other three unpack outputs alike
 *   tmp1*wpr - wi*wpi + wr               -> fma(tmp1, wpr, -(wi*wpi)) + wr
 *   wi*wpr + tmp1*wpi + wi               -> fma(wi, wpr, tmp1*wpi) + wi
 *   (f -/+ 0.5*w)*T*scale                -> fma(-/+0.5, w, f)*T*scale
 * This restatement writes the fused forms out and is compiled without
 * contraction like the rest.  With ORC_SOURCE_FORM defined it keeps the
 * reference's unfused source expressions instead, for a build with
 * -ffp-contract=fast whose outputs must equal the fused forms bit for bit. */
#ifdef ORC_SOURCE_FORM
#define ORC_POW(r, i) ((r) * (r) + (i) * (i))
#define ORC_LEVEL(m, d, ii) ((m) + (d) * (float)(ii))
#define ORC_NEWTON(x, y) ((y) * (1.5f - 0.5f * (x) * (y) * (y)))
#define ORC_PQM(p, a, b, c, d) ((p) + (a) * (b) - (c) * (d))
#define ORC_MQP(p, a, b, c, d) ((p) - (a) * (b) + (c) * (d))
#define ORC_PPP(p, a, b, c, d) ((p) + (a) * (b) + (c) * (d))
#define ORC_WR(t, wpr, wi, wpi, wr) ((t) * (wpr) - (wi) * (wpi) + (wr))
#define ORC_WI(wi, wpr, t, wpi) ((wi) * (wpr) + (t) * (wpi) + (wi))
#define ORC_EDGE(f, h, w) (((f) + (h) * (w)))
#else
#define ORC_POW(r, i) fmaf((r), (r), (i) * (i))
#define ORC_LEVEL(m, d, ii) fmaf((d), (float)(ii), (m))
#define ORC_NEWTON(x, y) ((y) * fmaf(-(0.5f * (x) * (y)), (y), 1.5f))
#define ORC_PQM(p, a, b, c, d) fma(-(c), (d), fma((a), (b), (p)))
#define ORC_MQP(p, a, b, c, d) fma((c), (d), fma(-(a), (b), (p)))
#define ORC_PPP(p, a, b, c, d) fma((c), (d), fma((a), (b), (p)))
#define ORC_WR(t, wpr, wi, wpi, wr) (fma((t), (wpr), -((wi) * (wpi))) + (wr))
#define ORC_WI(wi, wpr, t, wpi) (fma((wi), (wpr), (t) * (wpi)) + (wi))
#define ORC_EDGE(f, h, w) fma((h), (w), (f))
#endif

int orc_realfft(const float *samples, long long N, float *out)
{
    if (N <= 0 || (N & 1)) return -1;
    long n2 = (long)(N / 2);
    double *re = malloc(sizeof(double) * n2), *im = malloc(sizeof(double) * n2);
    if (!re || !im) return -1;
    for (long i = 0; i < n2; i++) { re[i] = samples[2 * i]; im[i] = samples[2 * i + 1]; }
    if (fft_any_d(re, im, n2) != 0) return -1;
    /* single-precision storage of the complex FFT, as fftwf leaves it */
    float *d = out;
    for (long i = 0; i < n2; i++) { d[2 * i] = (float)re[i]; d[2 * i + 1] = (float)im[i]; }
    free(re); free(im);
    return orc_nr_unpack(d, N);
}

/* Numerical Recipes unpack of the N/2-point complex FFT in d (N floats, in
 * place), double trig and recurrence (datpipe.c:101-138) */
int orc_nr_unpack(float *d, long long N)
{
    if (N <= 0 || (N & 1)) return -1;
    const long n2 = (long)(N / 2);
    const double cc = -0.5, theta = -2.0 * M_PI / (double)N;
    double wr = cos(theta), wi = sin(theta), tmp1 = sin(0.5 * theta);
    const double wpr = -2.0 * tmp1 * tmp1, wpi = wi;
    long il = 1, ih = n2 - il;
    for (; il <= (long)(N >> 2); il++, ih--) {
        double h1r = 0.5 * (d[2 * il] + d[2 * ih]);
        double h1i = 0.5 * (d[2 * il + 1] - d[2 * ih + 1]);
        double h2r = -cc * (d[2 * il + 1] + d[2 * ih + 1]);
        double h2i = cc * (d[2 * il] - d[2 * ih]);
        d[2 * il] = (float)ORC_PQM(h1r, h2r, wr, h2i, wi);
        d[2 * il + 1] = (float)ORC_PPP(h1i, h2i, wr, h2r, wi);
        d[2 * ih] = (float)ORC_MQP(h1r, h2r, wr, h2i, wi);
        d[2 * ih + 1] = (float)ORC_PPP(-h1i, h2i, wr, h2r, wi);
        tmp1 = wr;
        wr = ORC_WR(tmp1, wpr, wi, wpi, wr);
        wi = ORC_WI(wi, wpr, tmp1, wpi);
    }
    tmp1 = d[0];
    d[0] = (float)(tmp1 + d[1]);
    d[1] = (float)(tmp1 - d[1]);
    return 0;
}


static float presto_invsqrtf(float x)
{
    union { float f; int32_t i; } t;
    t.f = x;
    t.i = 0x5f3759df - (t.i >> 1);
    float y = t.f;
    return ORC_NEWTON(x, y);
}

#define ESWAP(a, b) do { float t_ = (a); (a) = (b); (b) = t_; } while (0)
static float presto_median(float *arr, int n)
{
    int low = 0, high = n - 1, median = (low + high) / 2, middle, ll, hh;
    for (;;) {
        if (high <= low) return arr[median];
        if (high == low + 1) {
            if (arr[low] > arr[high]) ESWAP(arr[low], arr[high]);
            return arr[median];
        }
        middle = (low + high) / 2;
        if (arr[middle] > arr[high]) ESWAP(arr[middle], arr[high]);
        if (arr[low] > arr[high]) ESWAP(arr[low], arr[high]);
        if (arr[middle] > arr[low]) ESWAP(arr[middle], arr[low]);
        ESWAP(arr[middle], arr[low + 1]);
        ll = low + 1;
        hh = high;
        for (;;) {
            do ll++; while (arr[low] > arr[ll]);
            do hh--; while (arr[hh] > arr[low]);
            if (hh < ll) break;
            ESWAP(arr[ll], arr[hh]);
        }
        ESWAP(arr[low], arr[hh]);
        if (hh <= median) low = ll;
        if (hh >= median) high = hh - 1;
    }
}

/* the two rednoise helpers on their own, for the pins */
float orc_presto_invsqrtf(float x) { return presto_invsqrtf(x); }
float orc_presto_median(const float *x, int n)
{
    float buf[1024];
    if (n < 1 || n > 1024) return NAN;
    memcpy(buf, x, sizeof(float) * (size_t)n);
    return presto_median(buf, n);
}

/* rednoise.c restatement (datpipe.c:143-253), startwidth 6, endwidth 100,
 * endfreq 6 Hz.  in/out: ncomplex interleaved complex. */
static int rednoise_o(const float *in, float *out, long long nc, long long N, double dt)
{
    const int startwidth = 6, endwidth = 100;
    const double endfreq = 6.0;
    float powbuf[100], oldb[200], newb[200], outb[200];
    long long binnum = 1, numwrote = 0, pos = 1;
    int bufflen, nold, nnew, mid_old, mid_new, ii, ind;
    float mean_old, mean_new, dslope = 1.0f, norm;
    const float T = (float)((double)N * dt);
    out[0] = 1.0f; out[1] = 0.0f;
    numwrote = 1;
    bufflen = startwidth;
    if (nc - 1 < bufflen) return -1;
    nold = bufflen;
    memcpy(oldb, in + 2 * pos, sizeof(float) * 2 * nold);
    pos += nold;
    mid_old = nold / 2;
    for (ii = 0; ii < nold; ii++) {
        float r = oldb[2 * ii], im = oldb[2 * ii + 1];
        powbuf[ii] = ORC_POW(r, im);
    }
    mean_old = (float)(presto_median(powbuf, nold) / log(2.0));
    norm = presto_invsqrtf(mean_old);
    for (ii = 0; ii < mid_old; ii++) {
        outb[2 * ii] = oldb[2 * ii] * norm;
        outb[2 * ii + 1] = oldb[2 * ii + 1] * norm;
    }
    memcpy(out + 2 * numwrote, outb, sizeof(float) * 2 * mid_old);
    numwrote += mid_old;
    binnum += nold;
    bufflen = (int)(startwidth * log((double)binnum));
    for (;;) {
        long long avail = nc - pos;
        if (avail <= 0) break;
        nnew = avail < bufflen ? (int)avail : bufflen;
        memcpy(newb, in + 2 * pos, sizeof(float) * 2 * nnew);
        pos += nnew;
        mid_new = nnew / 2;
        for (ii = 0; ii < nnew; ii++) {
            float r = newb[2 * ii], im = newb[2 * ii + 1];
            powbuf[ii] = ORC_POW(r, im);
        }
        mean_new = (float)(presto_median(powbuf, nnew) / log(2.0));
        dslope = (float)((mean_new - mean_old) / (0.5 * (nold + nnew)));
        ii = 0;
        for (ind = mid_old; ind < nold; ii++, ind++) {
            norm = presto_invsqrtf(ORC_LEVEL(mean_old, dslope, ii));
            outb[2 * ii] = oldb[2 * ind] * norm;
            outb[2 * ii + 1] = oldb[2 * ind + 1] * norm;
        }
        for (ind = 0; ind < mid_new; ii++, ind++) {
            norm = presto_invsqrtf(ORC_LEVEL(mean_old, dslope, ii));
            outb[2 * ii] = newb[2 * ind] * norm;
            outb[2 * ii + 1] = newb[2 * ind + 1] * norm;
        }
        memcpy(out + 2 * numwrote, outb, sizeof(float) * 2 * ii);
        numwrote += ii;
        binnum += nnew;
        if ((float)binnum / T < endfreq) bufflen = (int)(startwidth * log((double)binnum));
        else bufflen = endwidth;
        if (bufflen > endwidth) return -1;
        memcpy(oldb, newb, sizeof(float) * 2 * nnew);
        nold = nnew;
        mean_old = mean_new;
        mid_old = mid_new;
    }
    ii = 0;
    for (ind = mid_old; ind < nold; ii++, ind++) {
        norm = presto_invsqrtf(ORC_LEVEL(mean_old, dslope, ii));
        outb[2 * ii] = oldb[2 * ind] * norm;
        outb[2 * ii + 1] = oldb[2 * ind + 1] * norm;
    }
    memcpy(out + 2 * numwrote, outb, sizeof(float) * 2 * ii);
    numwrote += ii;
    return numwrote == nc ? 0 : -1;
}

int orc_rednoise(const float *in, long long N, double dt, float *out)
{
    return rednoise_o(in, out, N / 2, N, dt);
}

int orc_realfft_rednoise(const float *samples, long long N, double dt, float *out)
{
    float *tmp = malloc(sizeof(float) * (size_t)N);
    if (!tmp) return -1;
    if (orc_realfft(samples, N, tmp) != 0) { free(tmp); return -1; }
    int rc = rednoise_o(tmp, out, N / 2, N, dt);
    free(tmp);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Zaplist: src/zap.c:77-118 (median_power, zap_one), 155-187 (apply) */
/* ------------------------------------------------------------------ */
typedef struct { double lo, hi; int idx; } OrcZapBin;

static int orc_zapbin_cmp(const void *a, const void *b)
{
    const OrcZapBin *x = a, *y = b;
    if (x->lo != y->lo) return (x->lo > y->lo) - (x->lo < y->lo);
    return (x->idx > y->idx) - (x->idx < y->idx);   /* input order among equal edges */
}

static int orc_float_cmp(const void *a, const void *b)
{
    const float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* lower median of the float powers of `count` bins from `start`, the window
 * clipped to the spectrum (zap.c:77-92) */
static float orc_zap_median(const float *amps, long long namps, long long start, int count)
{
    if (start < 0) { count += (int)start; start = 0; }
    if (start + count > namps) count = (int)(namps - start);
    if (count <= 0) return 0.0f;
    float buf[200];
    for (int i = 0; i < count; i++) {
        const float re = amps[2 * (start + i)], im = amps[2 * (start + i) + 1];
        buf[i] = ORC_POW(re, im);
    }
    qsort(buf, (size_t)count, sizeof(float), orc_float_cmp);
    return buf[(count - 1) / 2];
}

int orc_zap(float *amps, long long namps, long long N, double T, const double *freq_hz,
            const double *width_hz, const int *bary, int nbirds, double baryv)
{
    if (nbirds <= 0) return 0;
    OrcZapBin *b = malloc(sizeof(OrcZapBin) * (size_t)nbirds);
    for (int i = 0; i < nbirds; i++) {
        const double scale = bary[i] ? 1.0 : (1.0 + baryv);
        double lo = ORC_EDGE(freq_hz[i], -0.5, width_hz[i]) * T * scale;
        double hi = ORC_EDGE(freq_hz[i], 0.5, width_hz[i]) * T * scale;
        if (hi - lo < 1.0) {
            const double mid = 0.5 * (hi + lo);
            lo = mid - 0.5;
            hi = mid + 0.5;
        }
        b[i].lo = lo;
        b[i].hi = hi;
        b[i].idx = i;
    }
    qsort(b, (size_t)nbirds, sizeof(OrcZapBin), orc_zapbin_cmp);
    long long limit = N / 2 < namps ? N / 2 : namps;
    int applied = 0;
    for (int i = 0; i < nbirds; i++) {
        if (b[i].lo >= (double)limit) break;
        if (b[i].hi >= (double)limit) b[i].hi = (double)(limit - 1);
        const double lobin = b[i].lo, hibin = b[i].hi;
        long long ilo = (long long)floor(lobin), ihi = (long long)ceil(hibin);
        if (ilo < 1) ilo = 1;
        if (ihi > namps) ihi = namps;
        if (ihi - ilo <= 0) continue;
        if (!(lobin - 300.0 > 1.0)) continue;
        const float mlo = orc_zap_median(amps, namps, (long long)(lobin - 300), 200);
        const float mhi = orc_zap_median(amps, namps, (long long)(hibin - 100), 200);
        const float avg = (float)sqrt(0.5 * (mlo + mhi) / -log(0.5));
        for (long long k = ilo; k < ihi; k++) {
            amps[2 * k] = avg;
            amps[2 * k + 1] = 0.0f;
        }
        applied++;
    }
    free(b);
    return applied;
}

/* ------------------------------------------------------------------ */
/* PRESTO prepsubband de-dispersion as the prep front end runs it      */
/* (nobary, no clipping, masks or zero-DM): the delay tables of        */
/* src/prep_dedisp.c:320-331, 362-367, 405-470 with                    */
/* third_party/presto/src/dispersion.c:100-215 and misc_utils.c:325-373,*/
/* then the write loop of src/prep_dedisp.c:485-530, 745-790 over      */
/* get_data's CPU branch (src/prep_dedisp.c:1104-1290), read_subbands  */
/* and prep_subbands (backend_common.c:936-1000, 1684-1738),           */
/* dedisp_subbands and float_dedisp (dispersion.c:216-263, 300-345).   */
/* ------------------------------------------------------------------ */
static long orc_nearest_long(double x)
{
    return (long)(x < 0 ? ceil(x - 0.5) : floor(x + 0.5));
}

double orc_delay_from_dm(double dm, double f)
{
    return f == 0.0 ? 0.0 : dm / (0.000241 * f * f);
}

/* per-channel delays from the lowest channel up (the doppler factor is
 * 1 + v/c with v/c = 0, i.e. exactly 1) */
static void orc_chan_delays(int n, double dm, double lofreq, double width, double *out)
{
    for (int i = 0; i < n; i++) out[i] = orc_delay_from_dm(dm, lofreq + i * width);
}

/* delays of every subband's highest channel */
static void orc_subband_delays(int numchan, int nsub, double dm, double lofreq, double cw,
                               double *out)
{
    const double sbw = cw * (numchan / nsub);
    orc_chan_delays(nsub, dm, lofreq + sbw - cw, sbw, out);
}

static int orc_smooth11(long long n)
{
    for (int f = 2; f <= 11; f++)
        while (n % f == 0) n /= f;
    return n == 1;
}

static int orc_pow10(long long n)
{
    while (n > 9 && n % 10 == 0) n /= 10;
    return n == 1;
}

/* The length of the good-factor table is 114: 1000 and every multiple of 16
 * in (1000, 10000] without a prime factor above 11, ascending. */
long long orc_choose_good_N(long long n)
{
    if (n <= 0) return 0;
    char digits[32];
    snprintf(digits, sizeof(digits), "%lld", n);
    digits[4] = '\0';
    const long long first4 = atoll(digits);
    long long g = 0;
    for (long long v = 1000; v <= 10000; v++) {
        if (v != 1000 && (v % 16 || !orc_smooth11(v))) continue;
        g = v;
        if (v == first4 && n % v == 0 && orc_pow10(n / v)) break;
        if (v > first4) break;
    }
    while (g < n) g *= 10;
    long long two = 2;
    while (two < n) two *= 2;
    return two < g ? two : g;
}

int orc_dedisp_tables(const OrcDedispConfig *c, long long nspec, double *dms, int *idispdt,
                      int *offsets, long long *numout)
{
    const int nch = c->numchan, nsub = c->nsub, nd = c->numdms, ds = c->downsamp;
    if (nch <= 0 || nsub <= 0 || nsub > nch || nch % nsub || nd <= 0 || ds <= 0 || c->P <= 0 ||
        c->P % ds || !(c->dt > 0.0))
        return -1;
    double avgdm = 0.0;
    for (int i = 0; i < nd; i++) {
        dms[i] = c->lodm + i * c->dmstep;
        avgdm += dms[i];
    }
    avgdm /= nd;
    const double maxdm = dms[nd - 1], dsdt = ds * c->dt;
    const double bw = orc_delay_from_dm(maxdm, c->lofreq) -
                      orc_delay_from_dm(maxdm, c->lofreq + (nch - 1) * c->chan_wid);
    const int bpr = (int)(bw / c->dt) / c->P + 1;
    *numout = c->numout > 0 ? c->numout : orc_choose_good_N(nspec / ds);
    double *cd = (double *)malloc(sizeof(double) * (size_t)nch);
    double *sd = (double *)malloc(sizeof(double) * (size_t)nsub);
    const int cps = nch / nsub;
    orc_subband_delays(nch, nsub, avgdm, c->lofreq, c->chan_wid, sd);
    orc_chan_delays(nch, avgdm, c->lofreq, c->chan_wid, cd);
    for (int s = 0; s < nsub; s++)
        for (int j = 0; j < cps; j++) cd[s * cps + j] -= sd[s];
    int bad = 0;
    for (int ch = 0; ch < nch; ch++) {
        idispdt[ch] = (int)orc_nearest_long(cd[ch] / c->dt);
        /* the reference refuses more dispersion across a subband than one
         * raw block (backend_common.c:1706-1712); past it, or below zero,
         * its loops would index outside the two blocks */
        if (idispdt[ch] < 0 || idispdt[ch] > c->P) bad = 1;
    }
    const int dsW = c->P * bpr / ds;
    for (int i = 0; i < nd; i++) {
        orc_subband_delays(nch, nsub, dms[i], c->lofreq, c->chan_wid, sd);
        const double top = sd[nsub - 1];
        for (int s = 0; s < nsub; s++) {
            const int o = (int)orc_nearest_long((sd[s] - top) / dsdt);
            offsets[i * nsub + s] = o;
            if (o < 0 || o > dsW) bad = 1;
        }
    }
    free(cd);
    free(sd);
    return bad ? -1 : bpr;
}

/* AS 52 running mean and variance, in double over float samples
 * (misc_utils.c:641-675) */
void orc_avg_var(const float *x, int n, double *mean, double *var)
{
    double m = (double)x[0], v = 0.0, an1 = 0.0;
    for (int i = 1; i < n; i++) {
        const double an = (double)(i + 1);
        an1 = (double)i;
        const double dx = (x[i] - m) / an;
        v += an * an1 * dx * dx;
        m += dx;
    }
    if (n > 1) v /= an1;
    *mean = m;
    *var = v;
}

static int orc_fcmp(const void *a, const void *b)
{
    const float x = *(const float *)a, y = *(const float *)b;
    return x < y ? -1 : x > y;
}

/* the element of rank (n - 1) / 2, which PRESTO's quickselect median
 * returns (median.c:12-75) */
float orc_lower_median(const float *x, int n)
{
    float *t = (float *)malloc(sizeof(float) * (size_t)n);
    memcpy(t, x, sizeof(float) * (size_t)n);
    qsort(t, (size_t)n, sizeof(float), orc_fcmp);
    const float m = t[(n - 1) / 2];
    free(t);
    return m;
}

typedef struct {
    float running_avg, running_std;
    float *chan_running_avg;
    int blocksread;
} OrcClip;

/* clip_times on one raw block [P][nch] in place (clipping.c:41-215, without
 * the CLIPBINSFILE list); levels receives the block's channel levels */
static int orc_clip_times(OrcClip *st, float *raw, int P, int nch, float clip_sigma,
                          float *levels)
{
    float *zdm = (float *)malloc(sizeof(float) * (size_t)P);
    float *good = (float *)malloc(sizeof(float) * (size_t)P);
    double *cavg = (double *)calloc((size_t)nch, sizeof(double));
    for (int i = 0; i < P; i++) {
        float z = 0.0f;
        for (int c = 0; c < nch; c++) z += raw[(long long)i * nch + c];
        zdm[i] = z;
    }
    double cur_avg, cur_std;
    orc_avg_var(zdm, P, &cur_avg, &cur_std);
    cur_std = sqrt(cur_std);
    const float med = orc_lower_median(zdm, P);
    const float lo = med - 3.0 * cur_std, hi = med + 3.0 * cur_std;
    int ngood = 0;
    for (int i = 0; i < P; i++)
        if (zdm[i] > lo && zdm[i] < hi) {
            good[ngood++] = zdm[i];
            for (int c = 0; c < nch; c++) cavg[c] += raw[(long long)i * nch + c];
        }
    if (ngood < 1) {
        cur_avg = st->running_avg;
        cur_std = st->running_std;
        for (int c = 0; c < nch; c++) cavg[c] = st->chan_running_avg[c];
    } else {
        orc_avg_var(good, ngood, &cur_avg, &cur_std);
        cur_std = sqrt(cur_std);
        for (int c = 0; c < nch; c++) cavg[c] /= ngood;
    }
    if (st->blocksread) {
        st->running_avg = 0.9 * st->running_avg + 0.1 * cur_avg;
        st->running_std = 0.9 * st->running_std + 0.1 * cur_std;
        for (int c = 0; c < nch; c++)
            st->chan_running_avg[c] = 0.9 * st->chan_running_avg[c] + 0.1 * cavg[c];
    } else {
        st->running_avg = cur_avg;
        st->running_std = cur_std;
        for (int c = 0; c < nch; c++) st->chan_running_avg[c] = cavg[c];
    }
    const float trigger = clip_sigma * st->running_std;
    if (levels) memcpy(levels, st->chan_running_avg, sizeof(float) * (size_t)nch);
    int clipped = 0;
    for (int i = 0; i < P; i++)
        if (fabsf(zdm[i] - st->running_avg) > trigger) {
            memcpy(raw + (long long)i * nch, st->chan_running_avg, sizeof(float) * (size_t)nch);
            clipped++;
        }
    st->blocksread++;
    free(zdm); free(good); free(cavg);
    return clipped;
}

long long orc_clip_all(const OrcDedispConfig *c, float *raw, long long nspec, float *levels)
{
    const int P = c->P, nch = c->numchan;
    if (P <= 0 || nch <= 0 || nspec % P) return -1;
    OrcClip st = {0.0f, 0.0f, (float *)calloc((size_t)nch, sizeof(float)), 0};
    long long n = 0;
    for (long long b = 0; b < nspec / P; b++)
        n += orc_clip_times(&st, raw + b * (long long)P * nch, P, nch, c->clip_sigma,
                            levels ? levels + b * (long long)nch : NULL);
    free(st.chan_running_avg);
    return n;
}

typedef struct {
    const float *raw;
    long long nblk, next;   /* whole raw blocks, the next one to read */
    int first;
    float *last, *cur;      /* raw blocks, [P][numchan] */
    OrcClip clip;           /* clip_times' state across blocks */
} OrcReader;

/* read_subbands + prep_subbands + dedisp_subbands into sub block `blk` of
 * fdata ([nsub][W]); P on success, 0 when the data has ended */
static int orc_read_subbands(OrcReader *r, const OrcDedispConfig *c, const int *idispdt,
                             float *fdata, int W, int blk)
{
    const int P = c->P, nch = c->numchan, cps = nch / c->nsub;
    const size_t bytes = sizeof(float) * (size_t)P * nch;
    /* prep_subbands clips every block it copies in, the first included
     * (backend_common.c:936-1000) */
    if (r->first) {
        if (r->next >= r->nblk) return -1;
        memcpy(r->last, r->raw + r->next * (long long)P * nch, bytes);
        if (c->clip_sigma > 0.0f) orc_clip_times(&r->clip, r->last, P, nch, c->clip_sigma, NULL);
        r->next++;
        r->first = 0;
    }
    if (r->next >= r->nblk) return 0;
    memcpy(r->cur, r->raw + r->next * (long long)P * nch, bytes);
    if (c->clip_sigma > 0.0f) orc_clip_times(&r->clip, r->cur, P, nch, c->clip_sigma, NULL);
    r->next++;
    for (int s = 0; s < c->nsub; s++)
        for (int j = 0; j < P; j++) fdata[(long long)s * W + (long long)blk * P + j] = 0.0f;
    for (int ch = 0; ch < nch; ch++) {
        float *sub = fdata + (long long)(ch / cps) * W + (long long)blk * P;
        const int d = idispdt[ch];
        int j = 0;
        for (; j < P - d; j++) sub[j] += r->last[(long long)(j + d) * nch + ch];
        for (; j < P; j++) sub[j] += r->cur[(long long)(j + d - P) * nch + ch];
    }
    float *t = r->last;
    r->last = r->cur;
    r->cur = t;
    return P;
}

long long orc_dedisp(const OrcDedispConfig *c, const float *raw, long long nspec, float *out,
                     float *pad)
{
    const int nch = c->numchan, nsub = c->nsub, nd = c->numdms, ds = c->downsamp, P = c->P;
    if (nch <= 0 || nsub <= 0 || P <= 0 || nspec % P) return -1;
    double *dms = (double *)malloc(sizeof(double) * (size_t)nd);
    int *idispdt = (int *)malloc(sizeof(int) * (size_t)nch);
    int *offs = (int *)malloc(sizeof(int) * (size_t)nd * nsub);
    long long numout = 0;
    const int bpr = orc_dedisp_tables(c, nspec, dms, idispdt, offs, &numout);
    if (bpr < 0) {
        free(dms); free(idispdt); free(offs);
        return -1;
    }
    const int W = P * bpr, dsW = W / ds;
    OrcReader rd = {raw, nspec / P, 0, 1, NULL, NULL,
                    {0.0f, 0.0f, (float *)calloc((size_t)nch, sizeof(float)), 0}};
    rd.last = (float *)malloc(sizeof(float) * (size_t)P * nch);
    rd.cur = (float *)malloc(sizeof(float) * (size_t)P * nch);
    float *data[2], *dsdata[2];
    for (int i = 0; i < 2; i++) {
        data[i] = (float *)calloc((size_t)nsub * W, sizeof(float));
        dsdata[i] = ds > 1 ? (float *)calloc((size_t)nsub * dsW, sizeof(float)) : data[i];
    }
    float *res = (float *)malloc(sizeof(float) * (size_t)nd * dsW);
    int cur = 0, firsttime = 1, failed = 0;
    long long totwrote = 0, statnum = 0;
    double avg = 0.0;
    for (;;) {
        /* get_data: one work block read ahead of the one de-dispersed */
        int totnumread = 0;
        for (;;) {
            for (int b = 0; b < bpr; b++) {
                const int n = orc_read_subbands(&rd, c, idispdt, data[cur], W, b);
                if (n < 0) { failed = 1; break; }
                if (n != P)
                    for (int s = 0; s < nsub; s++)
                        memset(data[cur] + (long long)s * W + (long long)b * P, 0,
                               sizeof(float) * (size_t)P);
                if (!firsttime) totnumread += n;
            }
            if (failed) break;
            if (ds > 1)
                for (int s = 0; s < nsub; s++)
                    for (int u = 0; u < dsW; u++) {
                        const float *x = data[cur] + (long long)s * W + (long long)u * ds;
                        float acc = 0.0f;
                        for (int k = 0; k < ds; k++) acc += x[k];
                        dsdata[cur][(long long)s * dsW + u] = acc / ds;
                    }
            if (firsttime) {
                cur ^= 1;
                firsttime = 0;
            } else {
                break;
            }
        }
        if (failed) break;
        /* float_dedisp of (last, current) for every DM */
        const float *lastd = dsdata[cur ^ 1], *curd = dsdata[cur];
        for (int i = 0; i < nd; i++) {
            float *r = res + (long long)i * dsW;
            for (int u = 0; u < dsW; u++) r[u] = -0.0f;
            for (int s = 0; s < nsub; s++) {
                const int o = offs[i * nsub + s];
                int u = 0;
                for (; u < dsW - o; u++) r[u] += lastd[(long long)s * dsW + o + u];
                for (; u < dsW; u++) r[u] += curd[(long long)s * dsW + u - (dsW - o)];
            }
        }
        cur ^= 1;
        if (totnumread != W) break;
        long long nw = dsW;
        if (totwrote + nw > numout) nw = numout - totwrote;
        for (int i = 0; i < nd; i++)
            memcpy(out + (long long)i * numout + totwrote, res + (long long)i * dsW,
                   sizeof(float) * (size_t)nw);
        for (long long u = 0; u < nw; u++) {
            /* update_stats: the running mean of the first DM's samples */
            const float x = res[u];
            const double dev = x - avg;
            avg += dev / ((double)(statnum + u) + 1.0);
        }
        statnum += nw;
        totwrote += nw;
        if (totwrote == numout) break;
    }
    const float padv = (float)avg;
    for (int i = 0; i < nd && !failed; i++)
        for (long long u = totwrote; u < numout; u++) out[(long long)i * numout + u] = padv;
    if (pad) *pad = padv;
    for (int i = 0; i < 2; i++) {
        if (ds > 1) free(dsdata[i]);
        free(data[i]);
    }
    free(res); free(rd.last); free(rd.cur); free(rd.clip.chan_running_avg);
    free(dms); free(idispdt); free(offs);
    return failed ? -1 : totwrote;
}

int orc_dedisp_direct(const OrcDedispConfig *c, const float *raw, long long nspec,
                      long long ntime, float *out)
{
    const int nch = c->numchan, nsub = c->nsub, nd = c->numdms, ds = c->downsamp;
    double *dms = (double *)malloc(sizeof(double) * (size_t)nd);
    int *idispdt = (int *)malloc(sizeof(int) * (size_t)nch);
    int *offs = (int *)malloc(sizeof(int) * (size_t)nd * nsub);
    long long numout = 0;
    int rc = orc_dedisp_tables(c, nspec, dms, idispdt, offs, &numout) < 0 ? -1 : 0;
    const int cps = rc == 0 ? nch / nsub : 1;
    int maxo = 0, maxd = 0;
    for (int i = 0; rc == 0 && i < nd * nsub; i++) maxo = offs[i] > maxo ? offs[i] : maxo;
    for (int i = 0; rc == 0 && i < nch; i++) maxd = idispdt[i] > maxd ? idispdt[i] : maxd;
    const long long nsubt = ntime + maxo;          /* downsampled subband samples needed */
    if (rc == 0 && nsubt * ds + maxd > nspec) rc = -1;
    float *sub = rc == 0 ? (float *)malloc(sizeof(float) * (size_t)nsub * nsubt) : NULL;
    for (int s = 0; rc == 0 && s < nsub; s++)
        for (long long u = 0; u < nsubt; u++) {
            float acc = 0.0f;
            for (int k = 0; k < ds; k++) {
                const long long t = u * ds + k;
                float v = 0.0f;
                for (int j = 0; j < cps; j++) {
                    const int ch = s * cps + j;
                    v += raw[(t + idispdt[ch]) * nch + ch];
                }
                acc = ds > 1 ? acc + v : v;
            }
            sub[(long long)s * nsubt + u] = ds > 1 ? acc / ds : acc;
        }
    for (int i = 0; rc == 0 && i < nd; i++)
        for (long long u = 0; u < ntime; u++) {
            float acc = -0.0f;
            for (int s = 0; s < nsub; s++) acc += sub[(long long)s * nsubt + u + offs[i * nsub + s]];
            out[(long long)i * ntime + u] = acc;
        }
    free(sub); free(dms); free(idispdt); free(offs);
    return rc;
}
