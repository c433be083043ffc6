/*
 * Host-side geometry shared by the HIP library and the host C library:
 * trial grid, per-harmonic interpolation plans, fit predicates, chunk table,
 * boxcar width banks and analytic noise scales.  Pure C, no device code.
 * The csx_ names are internal; the host library re-exports them under the
 * reference's names and the HIP library keeps them hidden.
 */
#ifndef CS_CORE_H
#define CS_CORE_H

#include "cs_types.h"

#ifdef __cplusplus
extern "C" {
#endif

double csx_sinpi(double x);
double csx_cospi(double x);
double csx_eps(double x);

void csx_params_defaults(SearchParams *p);
int  csx_decimation_set(int nharms, int maxdecim, int *out, int outcap);
int  csx_trial_grid_rational(double lodr, int *pnum, int *q);

/* One plan without the CPU group tables.  P == 0 for a non-rational grid. */
int  csx_plan_build(DirectPlan *dp, int h, int m, double r_lo, int has_rational,
                    int pnum, int q);
void csx_plan_free(DirectPlan *dp);
int  csx_plans_build(DirectPlanSet *set, const SearchParams *p, double r_lo);
void csx_plans_free(DirectPlanSet *set);

void csx_chunk_state(const DirectPlan *dp, long long t, int *res, long long *qint);
long long csx_plan_base(const DirectPlan *dp, double lodr, long long t);
long long csx_plan_nyq(const DirectPlan *dp, double lodr, long long t);
int  csx_harmonic_fits(const DirectPlan *dp, double lodr, long long namps,
                       long long N, long long t);
/* first trial where the low-edge predicate holds / first where the high-edge
 * predicate fails, by bisection over [0, total) */
long long csx_first_fit(const DirectPlan *dp, double lodr, long long namps,
                        long long N, long long total);
long long csx_first_unfit(const DirectPlan *dp, double lodr, long long namps,
                          long long N, long long total);
int  csx_chunk_starts(const DirectPlanSet *set, double lodr, long long namps,
                      long long N, long long total, int Nprof,
                      long long **out, long long *nout);

int    csx_boxcar_widths(int nbins, double fsp, double maxfrac, int *out, int cap);
int    csx_ladder_widths(int nbins, int k, const SearchParams *p, int *out, int cap);
double csx_analytic_sigma(const uint8_t *filled, int k, int Hk);
void   csx_boxcar_shape(double *shape, int nbins, const int *widths, int nw,
                        const uint8_t *filled, int k, int Hk);

/* Decimation count with k > 1 (slots after the base). */
int csx_nslots(const SearchParams *p);

/* One zaplist interval to apply: bins [ilo, ihi) become (avgamp, 0), where
 * avgamp = f32(sqrt(0.5 * f32(med_lo + med_hi) / ln 2)) and med_lo / med_hi
 * are the lower medians of the powers re^2 + im^2 (float) of n_lo bins from
 * s_lo and n_hi bins from s_hi (0 when the count is not positive), read after
 * the earlier intervals have been applied. */
typedef struct {
    long long ilo, ihi;
    long long s_lo, s_hi;
    int n_lo, n_hi;
} CsZapOp;

#define CS_ZAP_MEDIANBINS 200

/* The zaplist's intervals for a spectrum of namps bins over T seconds (N
 * samples), in application order: bins (f -/+ w/2) T scale with scale 1 or
 * 1 + baryv, widened to one bin, sorted by their low edge (stable), cut at
 * min(N/2, namps), dropping those that zap nothing or lie within 301 bins of
 * DC (src/zap.c:94-187).  *out (malloc'd, NULL when none) receives the ops;
 * returns their count, -1 on allocation failure. */
int csx_zap_ops(const ZapList *z, double T, long long N, long long namps, double baryv,
                CsZapOp **out);

#ifdef __cplusplus
}
#endif
#endif
