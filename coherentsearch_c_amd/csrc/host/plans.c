/* Reference-named plan / grid / chunk API over the shared host geometry
 * (src/interp.c:17-118, 283-416, 915-966). */
#include "host_internal.h"

void params_defaults(SearchParams *p) { csx_params_defaults(p); }
int decimation_set(int nharms, int maxdecim, int *out, int outcap)
{
    return csx_decimation_set(nharms, maxdecim, out, outcap);
}
int trial_grid_rational(double lodr, int *pnum, int *q)
{
    return csx_trial_grid_rational(lodr, pnum, q);
}
int direct_plans_build(DirectPlanSet *set, const SearchParams *p, double r_lo)
{
    return csx_plans_build(set, p, r_lo);
}
void direct_plans_free(DirectPlanSet *set) { csx_plans_free(set); }
void cs_chunk_state(const DirectPlan *dp, long long t, int *res, long long *qint)
{
    csx_chunk_state(dp, t, res, qint);
}
int cs_harmonic_fits(const DirectPlan *dp, const FFTFile *ft, const SearchParams *p,
                     long long t)
{
    return csx_harmonic_fits(dp, p->hidr / (double)p->nharms, ft->namps, ft->N, t);
}
double cs_sinpi(double x) { return csx_sinpi(x); }
double cs_cospi(double x) { return csx_cospi(x); }
double cs_eps(double x) { return csx_eps(x); }

static int cmp_ll(const void *a, const void *b)
{
    const long long x = *(const long long *)a, y = *(const long long *)b;
    return (x > y) - (x < y);
}

int harmonic_change_trials(const DirectPlanSet *set, const FFTFile *ft,
                           const SearchParams *p, long long total, long long **out)
{
    *out = NULL;
    if (total <= 1) return 0;
    const double lodr = p->hidr / (double)p->nharms;
    long long *b = (long long *)malloc((size_t)(2 * set->n + 1) * sizeof(long long));
    if (!b) return 0;
    int n = 0;
    for (int i = 0; i < set->n; i++) {
        const long long a = csx_first_fit(&set->plans[i], lodr, ft->namps, ft->N, total);
        const long long c = csx_first_unfit(&set->plans[i], lodr, ft->namps, ft->N, total);
        if (a > 0 && a < total) b[n++] = a;
        if (c > 0 && c < total) b[n++] = c;
    }
    qsort(b, (size_t)n, sizeof(long long), cmp_ll);
    int u = 0;
    for (int i = 0; i < n; i++)
        if (u == 0 || b[i] != b[u - 1]) b[u++] = b[i];
    *out = b;
    return u;
}

int chunk_starts(const DirectPlanSet *set, const FFTFile *ft, const SearchParams *p,
                 long long total, int Nprof, long long **out, long long *nout)
{
    return csx_chunk_starts(set, p->hidr / (double)p->nharms, ft->namps, ft->N, total,
                            Nprof, out, nout);
}

int boxcar_widths(int nbins, double fsp, double maxfrac, int *out, int cap)
{
    return csx_boxcar_widths(nbins, fsp, maxfrac, out, cap);
}
int ladder_boxcar_widths(int nbins, int k, const SearchParams *p, int *out, int cap)
{
    return csx_ladder_widths(nbins, k, p, out, cap);
}
double analytic_sigma(const uint8_t *filled, int k, int Hk)
{
    return csx_analytic_sigma(filled, k, Hk);
}
void boxcar_shape(double *shape, int nbins, const int *widths, int nw,
                  const uint8_t *filled, int k, int Hk)
{
    csx_boxcar_shape(shape, nbins, widths, nw, filled, k, Hk);
}
