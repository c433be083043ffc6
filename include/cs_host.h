/*
 * libcs_host.so -- the host side of the coherent search, in C, mirroring the
 * reference's host interface for this path (include/coherent_search.h) with
 * the same names, argument meaning and error behaviour.  It drives the GPU
 * backend (cs_gpu.h) and has no CPU search path of its own: a search needs a
 * gfx950 device.
 */
#ifndef CS_HOST_H
#define CS_HOST_H

#include <stdio.h>

#include "cs_gpu.h"
#include "cs_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- parameters / plans / chunks: src/interp.c ------------------------ */
CS_API void params_defaults(SearchParams *p);                               /* interp.c:48-62 */
CS_API int  decimation_set(int nharms, int maxdecim, int *out, int outcap); /* interp.c:64-75 */
CS_API int  trial_grid_rational(double lodr, int *pnum, int *q);           /* interp.c:80-118 */
CS_API int  direct_plans_build(DirectPlanSet *set, const SearchParams *p,
                               double r_lo);                                /* interp.c:283-303 */
CS_API void direct_plans_free(DirectPlanSet *set);                         /* interp.c:305-312 */
CS_API void cs_chunk_state(const DirectPlan *dp, long long t, int *res,
                           long long *qint);                                /* interp.c:407-410 */
CS_API int  cs_harmonic_fits(const DirectPlan *dp, const FFTFile *ft,
                             const SearchParams *p, long long t);          /* interp.c:412-416 */
CS_API int  harmonic_change_trials(const DirectPlanSet *set, const FFTFile *ft,
                                   const SearchParams *p, long long total,
                                   long long **out_bnds);                  /* interp.c:915-942 */
CS_API int  chunk_starts(const DirectPlanSet *set, const FFTFile *ft,
                         const SearchParams *p, long long total, int Nprof,
                         long long **out, long long *nout);                /* interp.c:944-966 */
CS_API double cs_sinpi(double x);                                          /* interp.c:17-26 */
CS_API double cs_cospi(double x);                                          /* interp.c:27-36 */
CS_API double cs_eps(double x);                                            /* interp.c:38-43 */
CS_API void fourier_interp_complex(double r, const float *amps, long long namps,
                                   int m, double *out_re, double *out_im); /* interp.c:317-344 */

/* ---- boxcar widths and noise scales: src/boxcar.c --------------------- */
CS_API int    boxcar_widths(int nbins, double fsp, double maxfrac, int *out, int cap);
CS_API int    ladder_boxcar_widths(int nbins, int k, const SearchParams *p, int *out,
                                   int cap);
CS_API double analytic_sigma(const uint8_t *filled, int k, int Hk);
CS_API void   boxcar_shape(double *shape, int nbins, const int *widths, int nw,
                           const uint8_t *filled, int k, int Hk);
CS_API int    boxcar_best_width(const double *prof, int nbins, double fsp,
                                double maxfrac, int *out_w, double *out_ducy);

/* ---- candidates: src/candidate.c -------------------------------------- */
CS_API void candlist_init(CandidateList *l);
CS_API void candlist_free(CandidateList *l);
CS_API void candlist_push(CandidateList *l, Candidate c);
CS_API void candlist_append(CandidateList *dst, const CandidateList *src);
CS_API void candlist_sort_metric_desc(CandidateList *l);
CS_API void candlist_sort_freq(CandidateList *l);
CS_API void candlist_sort_r(CandidateList *l);
CS_API void candlist_sort_metric_desc_stable(CandidateList *l);
CS_API void candlist_sort_freq_stable(CandidateList *l);
CS_API void candlist_sort_r_stable(CandidateList *l);
CS_API CandidateList remove_duplicates(const CandidateList *in, double dr_tol);
CS_API CandidateList remove_harmonics(const CandidateList *in, int numharm, double tol);
CS_API int  harmonically_related(double r1, double r2, int numharm, double tol);
CS_API double *candidate_profile(const FFTFile *ft, double r, int nharm, int m,
                                 int *out_n);
CS_API CandidateList measure_ducy(const FFTFile *ft, const CandidateList *in,
                                  const SearchParams *p);
CS_API int  write_candidates(const CandidateList *cands, const char *outfile,
                             double threshold);
CS_API void cs_format_double(char *buf, size_t n, double x);
CS_API CandidateList read_candidates_file(const char *path, double T);

/* ---- inputs: src/fileio.c, src/datpipe.c ------------------------------ */
CS_API int  inf_read(const char *path, SimpleInf *inf);
CS_API void inf_free(SimpleInf *inf);
CS_API int  zaplist_read(const char *path, ZapList *z);                   /* zap.c:120-147 */
CS_API void zaplist_free(ZapList *z);                                      /* zap.c:149-153 */
CS_API int  zaplist_apply(const ZapList *z, FFTFile *ft, double baryv);    /* zap.c:155-187, 94-118 */
CS_API int  fftfile_open(FFTFile *ft, const char *path);
CS_API void fftfile_close(FFTFile *ft);
typedef struct DatPipe DatPipe;
CS_API DatPipe *datpipe_create(int startwidth, int endwidth, double endfreq);
CS_API void datpipe_free(DatPipe *p);
/* No counterpart in the reference (its pipeline runs a single-threaded FFTW
 * plan): OpenMP threads of the pipeline's complex transform, 0 for the OpenMP
 * default.  Callers that stage several files at once give each pipeline a
 * share of the cores. */
CS_API void datpipe_set_threads(DatPipe *p, int threads);
/* No counterpart in the reference: stage on GPU `device` (realfft + rednoise
 * by cs_gpu_stage_*, the spectrum copied back to the pipeline's buffer) for
 * lengths whose half has no prime factor above 7, the host transform for the
 * others; -1 (the default) stages every length on the host. */
CS_API void datpipe_set_device(DatPipe *p, int device);
CS_API int  datpipe_redfft(DatPipe *p, const char *datpath, FFTFile *ft);
CS_API int  datpipe_redfft_mem(DatPipe *p, const float *samples, long long N,
                               double dt, const char *path, FFTFile *ft);
CS_API int  cs_open_input(const char *path, DatPipe *pipe, FFTFile *ft, int *is_dat);
/* Writers for PRESTO-style files (no reference counterpart; tests and
 * synthetic data).  0 on success. */
CS_API int  cs_write_inf(const char *path, long long N, double dt, double DM,
                         const char *object);
CS_API int  cs_write_floats(const char *path, const float *v, long long n);

/* ---- search driver: src/search.c -------------------------------------- */
typedef struct {
    int Nprof;
    int plan_valid;
    const SearchParams *params;
    double plan_T;
    long long plan_N, plan_namps;
    double plan_lofreq, plan_hifreq;
    int plan_lobin, plan_Nprof;
    long long plan_total;
    DirectPlanSet dplans;
    long long *cstarts;
    long long nchunks;
    void *gpu_session;   /* GpuSession* */
} SearchCache;

/* ---- metric statistics and normalisation: src/search.c:409-822 ---------- */
/* coherent_search.h:262-267 */
typedef struct {
    int block, k, Hk, nbins;
    double ngoodbins, flo, fhi;
    long long n;
    double min, median, mean, std, max;
} BlockMetricStats;

/* coherent_search.h:269-276 */
typedef struct {
    int k, Hk, nbins, win;
    double flo, fhi, lo, hi, invw;
    long long *counts;
    int nb;
    long long under, over, total;
    double sum, sumsq, vmin, vmax;
} MetricHistogram;

/* coherent_search.h:278-287 */
typedef struct {
    double hist_lo, hist_hi;
    int hist_nb, nwin;
    BlockMetricStats *blocks;
    size_t nblocks, capblocks;
    MetricHistogram *hists;   /* one per decimation k */
    size_t nhists, caphists;
    MetricHistogram *whists;  /* one per (k, frequency window) */
    size_t nwhists, capwhists;
} MetricStats;

/* coherent_search.h:391-402 */
typedef struct {
    int k;
    double *edges;   /* nwin + 1 */
    double *loc;
    double *scale;
    int nwin;
} NormK;

typedef struct {
    NormK *ks;
    int n;
} MetricNorm;

CS_API void metricstats_init(MetricStats *ms, double lo, double hi, int nb, int nwin);  /* search.c:493-498 */
CS_API void metricstats_free(MetricStats *ms);                                         /* search.c:499-508 */
CS_API void metricstats_append_block(MetricStats *ms, BlockMetricStats s);             /* search.c:509-517 */
CS_API void metricstats_append_hist(MetricStats *ms, MetricHistogram h);               /* search.c:518-526 */
CS_API void metricstats_append_whist(MetricStats *ms, MetricHistogram h);              /* search.c:527-535 */
CS_API void hist_init(MetricHistogram *h, int k, int Hk, int win, double flo, double fhi,
                      double lo, double hi, int nb);                                   /* search.c:412-423 */
CS_API void hist_free(MetricHistogram *h);                                             /* search.c:424-429 */
CS_API void hist_push(MetricHistogram *h, double x);                                   /* search.c:430-445 */
CS_API void hist_merge(MetricHistogram *a, const MetricHistogram *b);                  /* search.c:446-453 */
CS_API double hist_quantile(const MetricHistogram *h, double q);                       /* search.c:454-470 */
CS_API void log_edges(double a, double b, int nwin, double *edges);                    /* search.c:471-478 */
CS_API int window_index(const double *edges, int nwin, double f);                      /* search.c:479-492 */
CS_API MetricNorm *build_metricnorm(const MetricStats *ms);                            /* search.c:752-801 */
CS_API void metricnorm_free(MetricNorm *n);                                            /* search.c:802-813 */
CS_API double metric_normalize(const MetricNorm *n, int k, double f, double M);        /* search.c:814-823 */
/* cli.c:170-303: the stderr tables and <base>_metricstats.txt, _metricfap.txt,
 * _metrichist.txt */
CS_API void report_metricstats(const MetricStats *ms, const char *base, const SearchParams *p,
                               double threshold);

/* search.c:1247-1486 with the GPU path only.  gpu_device < 0 is an error
 * (this library has no CPU search).  With metricstats or normalize the GPU
 * keeps every trial's metric (cs_gpu_session_metric_stats): metricstats
 * receives the per-block statistics and the per-(k, window) histograms, and
 * normalize scores the candidates against the histograms' robust location
 * and scale per frequency window, as the reference's two-pass CPU search
 * does.  The wisdom arguments are accepted and ignored. */
CS_API CandidateList search_file(const FFTFile *ft, const SearchParams *p,
                                 double lofreq, double hifreq, int lobin, int blocksize,
                                 double threshold, int remove, double dr_tol,
                                 int harm_remove, int numharm, MetricStats *metricstats,
                                 int normalize, int verbose, int wisdom,
                                 const char *wisdom_file, SearchCache *cache,
                                 double harm_tol, int gpu_device);
CS_API void search_cache_free(SearchCache *cache);                          /* search.c:1488-1506 */

/* ---- CLI: src/cli.c ----------------------------------------------------- */
CS_API int cs_main(int argc, char **argv);        /* cli.c:1115 */
CS_API int cs_main_list(int argc, char **argv);   /* cli.c:1116 */

#ifdef __cplusplus
}
#endif
#endif
