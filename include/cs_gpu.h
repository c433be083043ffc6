/*
 * libcs_hip.so -- the MI355X (gfx950) backend of the coherent harmonic search.
 *
 * Drop-in for the reference's CUDA backend object (src/cuda.cu): the first
 * six entry points have the reference's exact signatures and semantics, so a
 * host built against include/coherent_search.h links this library in place
 * of cuda.o / gpu_stub.o.  All structs are the reference's layouts
 * (cs_types.h).  No torch types, no C++ in the signatures.
 *
 * Differences a caller can observe:
 *   - cs_gpu_session_create returns NULL (with a message on stderr) for a
 *     configuration this build does not implement, instead of exit(1);
 *   - the analytic/measured noise scale and f32/f64 precision follow the
 *     reference's CPU path (the reference GPU path forbids measured sigma).
 */
#ifndef CS_GPU_H
#define CS_GPU_H

#include "cs_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define CS_API __attribute__((visibility("default")))
#else
#define CS_API
#endif

typedef struct GpuSession GpuSession;

/* Replaces src/cuda.cu:488-528 (declared include/coherent_search.h:438).
 * 1: device present (errbuf gets "name (device d, gfx950)"), -1: no device /
 * bad index / runtime error.  This build always has GPU support, so it never
 * returns 0 (the reference's gpu_stub.c:7-14 case). */
CS_API int cs_gpu_available(int device, char *errbuf, size_t errlen);

/* Replaces src/cuda.cu:554-792 (coherent_search.h:442-446).  Builds the
 * device-resident plan arena, chunk and tile tables for one observation
 * geometry.  dplans must be rational (every P > 0); p is borrowed for the
 * session's lifetime and compared by pointer in _matches, like the
 * reference.  Any nharms 1-512, even m <= 64, f32 or f64, decimations up to
 * 16 and the fsp 1.5 width bank up to width 211: register kernels serve
 * nharms <= 64 (m <= 32) and 65-128 (m <= 16, f32) when the tile's windows
 * fit their LDS buffer, the general kernel everything else.  Returns NULL on
 * failure or outside that range. */
CS_API GpuSession *cs_gpu_session_create(const FFTFile *ft, const SearchParams *p,
                                         const DirectPlanSet *dplans,
                                         const long long *cstarts, long long total,
                                         int Nprof, long long nchunks,
                                         double r_lo, double lodr, int device);

/* Replaces src/cuda.cu:1257-1266 (coherent_search.h:447-449). */
CS_API int cs_gpu_session_matches(const GpuSession *s, const FFTFile *ft,
                                  const SearchParams *p, int Nprof, long long total,
                                  double r_lo, double lodr, int device);

/* Replaces src/cuda.cu:793-1176 (coherent_search.h:450-451).  Uploads
 * ft->amps and returns every trial with metric > threshold: chunk order, then
 * base before decimations in SearchParams order, ascending trial; ducy NaN.
 * The list's v is malloc'ed; free it with free() / candlist_free().  An
 * N/namps mismatch returns an empty list, as in the reference. */
CS_API CandidateList cs_gpu_session_run(GpuSession *s, const FFTFile *ft,
                                        double threshold);

/* Replaces src/cuda.cu:1178-1235 (coherent_search.h:452). */
CS_API void cs_gpu_session_free(GpuSession *s);

/* Replaces src/cuda.cu:1238-1255 (coherent_search.h:454-460); r_hi unused. */
CS_API CandidateList cs_gpu_search_region(const FFTFile *ft, const SearchParams *p,
                                          const DirectPlanSet *dplans,
                                          const long long *cstarts, long long total,
                                          int Nprof, long long nchunks,
                                          double r_lo, double r_hi, double lodr,
                                          double threshold, int device);

/* ---- extensions (no counterpart in the reference) ---------------------- */

/* Interpolated harmonic stack of trials [t0, t0+n) after the last upload of
 * ft->amps: out[(t - t0)*(nharms+1) + h] as interleaved float pairs, h = 0
 * (DC, zero) .. nharms, zero for harmonics that do not fit the trial's chunk.
 * For parity tests against the CPU restatement. 0 on success. */
CS_API int cs_gpu_session_dump_stack(GpuSession *s, const FFTFile *ft,
                                     long long t0, long long n, float *out);

/* Upload ft->amps into the session's resident spectrum (the first half of
 * cs_gpu_session_run).  0 on success, -1 on an N/namps mismatch or error. */
CS_API int cs_gpu_session_upload(GpuSession *s, const FFTFile *ft);

/* Search the resident spectrum (the second half of cs_gpu_session_run):
 * the same candidate list without the host-to-device copy. */
CS_API CandidateList cs_gpu_session_search_resident(GpuSession *s, double threshold);

/* No counterpart in the reference: a search of the resident spectrum that
 * allocates nothing, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same candidates as cs_gpu_session_search_resident, written
 * to out[0..return value) in its order; the records and their counter live in
 * a fixed per-GPU device array of the library (at most 262144 records), so
 * the call returns -1 with a message when the search finds more records than
 * that or than `cap`, and never allocates, frees or builds a graph.  Calls on
 * one GPU share that array and must not overlap (they would race); each call
 * clears its counter first.  Returns the number of candidates, -1 on
 * error. */
CS_API long long cs_gpu_session_search_noalloc(GpuSession *s, double threshold,
                                               Candidate *out, long long cap);

/* Replay the search's launch sequence (counter reset, measured-sigma passes,
 * search kernel) from a captured hipGraph (enable = 1, the default for
 * sessions with their own stream) or launch it eagerly (0).  The graph is
 * re-captured when the threshold or the record buffer changes; results are
 * identical either way.  0 on success. */
CS_API int cs_gpu_session_set_graphs(GpuSession *s, int enable);

/* Apply a zaplist to the session's resident spectrum, as zaplist_apply
 * (src/zap.c:155-187) does to a host spectrum: each interval's bins become
 * (avgamp, 0), avgamp from the lower medians of the powers of 200 bins on
 * either side, intervals in order of their low edge.  amps_host (optional,
 * 2*namps floats) receives the spectrum afterwards.  Returns the number of
 * intervals applied, -1 on failure. */
CS_API int cs_gpu_session_zap(GpuSession *s, const ZapList *z, double baryv, float *amps_host);

/* The final candidate list of a search, from the session's resident spectrum:
 * cands (a collapsed list, e.g. search_file's) sorted by S/N (stable), cut to
 * the first ncands (all when ncands < 0), with duty cycles measured on the
 * device -- the steps src/cli.c:1059-1061 run on the host (sort, truncate,
 * measure_ducy of src/candidate.c:305-329: exact 64-tap interpolation of each
 * harmonic that fits, a 2H-bin profile, boxcar_best_width).  The spectrum does
 * not come back to the host.  cands is left as it is; the result is a new list
 * (empty, with a message on stderr, on failure or for nharm > 1024). */
CS_API CandidateList cs_gpu_session_finalize(GpuSession *s, const CandidateList *cands,
                                             int ncands);

/* Device time of the last cs_gpu_session_run's search kernel(s), in ms,
 * measured with HIP events on the session's stream. */
CS_API double cs_gpu_session_last_kernel_ms(const GpuSession *s);

/* Work of one session: filled harmonic-trials (sum over harmonics of the
 * trials where the harmonic fits), trials, and the algorithmic HBM bytes of
 * the reference GPU dataflow, per trial
 *   16(nh+1) + 8 nh (2 + [any divisor decimation])
 *   + sum over non-divisor decimations k of (8(Hk+1) + 16 Hk)
 * plus 8*namps for the spectrum. */
CS_API int cs_gpu_session_work(const GpuSession *s, long long *harmonic_trials,
                               long long *trials, double *alg_bytes);

/* ------------------------------------------------------------------ */
/* Device staging of raw series (no reference counterpart on the GPU: the
 * reference stages on the host, src/datpipe.c:52-253 -- realfft_forward and
 * rednoise_apply).  A stage plan holds the FFT plan and the rednoise block
 * layout of one series length N at sample time dt, with datpipe_create's
 * startwidth/endwidth/endfreq (0 selects 6, 100, 6 Hz). */
typedef struct CsStage CsStage;

/* NULL (with a message) if N is odd, N/2 has a prime factor above 7, or the
 * rednoise widths exceed the device limit. */
CS_API CsStage *cs_gpu_stage_create(long long N, double dt, int startwidth, int endwidth,
                                    double endfreq, int device);
CS_API void cs_gpu_stage_free(CsStage *g);

/* samples (host, N floats) -> rednoised spectrum amps (host, N floats = N/2
 * complex, bin 0 = (1, 0)); packed (optional, N floats) receives the realfft
 * output before rednoise.  0 on success. */
CS_API int cs_gpu_stage_run_host(CsStage *g, const float *samples, float *amps, float *packed);

/* Stage samples (host, N floats) straight into the session's resident
 * spectrum (the device-side equivalent of datpipe_redfft_mem followed by
 * cs_gpu_session_upload); amps_host (optional, N floats) receives a copy for
 * host post-processing.  -1 if the plan's N does not match the session. */
CS_API int cs_gpu_session_stage(GpuSession *s, CsStage *g, const float *samples,
                                float *amps_host);

/* ------------------------------------------------------------------ */
/* Autotuner (no reference counterpart).  For one kernel class (device,
 * nharms, padded taps 16/32, precision, sigma mode) it times the search on a
 * synthetic 2^22-sample noise spectrum for every candidate launch shape --
 * the lane order inside a tile (trials in residue classes or in natural
 * order), the mapping of tiles to the 8 XCDs (contiguous ranges or stripes),
 * and, for analytic sigma, the chunk length (blocksize 16384, 65536 or
 * 262144) -- and keeps the fastest.  Sessions created afterwards for the class
 * use its launch shape, which changes no result.  The chunk length is advice:
 * engine requests with blocksize < 0 take it (an analytic-sigma metric does
 * not depend on it, but the raw list's chunk -> decimation -> trial order
 * does, so it is never applied silently). */
typedef struct {
    int nharms, m_padded, precision, sigma;   /* the class */
    int lane_order;        /* 0 residue classes, 1 natural order */
    int xcd_stripe;        /* 0 contiguous range per XCD, S: stripes of S tiles */
    int blocksize;         /* chosen chunk length (the default for measured sigma) */
    int ntried;            /* launch shapes timed */
    double default_ms;     /* search kernel time of the default shape */
    double best_ms;        /* ... of the chosen one */
} CsTuneResult;

/* tunes the class of p on `device` (or returns the cached result); 0 on
 * success, -1 if no GPU kernel serves p */
CS_API int cs_gpu_autotune(int device, const SearchParams *p, CsTuneResult *out);
/* the tuned classes as text (one line each) / read back; 0 on success */
CS_API int cs_gpu_tune_save(const char *path);
CS_API int cs_gpu_tune_load(const char *path);

/* ------------------------------------------------------------------ */
/* Metric statistics and normalisation on the device (no counterpart in the
 * reference's GPU library, whose caller refuses --metricstats/--normalize
 * with --gpu, src/cli.c:911-915; they replace its CPU collection in
 * chunk_after_fill/decim_pass and the second, normalised search_region of
 * src/search.c:537-565, 605-747, 830-945, 1428-1440).
 *
 * Slot s of a session is the base (k = 1) for s = 0, then the decimations
 * k > 1 of SearchParams.decimations in order; trial j of chunk c belongs to
 * slot s when j < the chunk's length (base) or j < its count of trials below
 * the decimated Nyquist frequency (src/search.c:619-629). */
#define CS_MAXWIN 64        /* frequency windows per decimation (the CLI uses 16) */

typedef struct {
    long long n;            /* trials with a metric */
    double min, max;
    double sum, sumsq;      /* sums of x and x*x (a fixed-order tree on the device) */
    double ss;              /* sum of (x - sum/n)^2 */
    double median;          /* median_f64 (src/boxcar.c:62-71): the middle value, or the
                               mean of the two middle values for even n */
    long long under, over;  /* metrics below hist_lo / at or above hist_hi */
} CsMetricBlock;

/* One search of ft (NULL: the resident spectrum) keeping every trial's metric
 * on the device (the measured-sigma passes first with --sigma measured).  For
 * every (chunk c, slot s) it fills blocks[c*nslot + s] and adds the chunk's
 * metrics in [hist_lo, hist_hi) to counts[(s*nwin + w-1)*hist_nb + bin] for
 * w = win[c*nslot + s] (1..nwin; 0 adds nothing), bin = floor((x - hist_lo)
 * * hist_nb/(hist_hi - hist_lo)) clamped to [0, hist_nb) -- hist_push's rule
 * (src/search.c:430-445).  counts must be zeroed by the caller (it is added
 * to).  hist_nb <= 4096.  0 on success. */
CS_API int cs_gpu_session_metric_stats(GpuSession *s, const FFTFile *ft, const int *win,
                                       int nwin, double hist_lo, double hist_hi, int hist_nb,
                                       CsMetricBlock *blocks, long long *counts);
/* After cs_gpu_session_metric_stats: the candidates of the kept metrics in
 * the order cs_gpu_session_run gives, each scored (M - loc)/scale with the
 * loc and scale of the window holding its frequency -- window_index's rule
 * over edges[s*(nwin+1) ...] (src/search.c:479-492, 814-823) -- or M itself
 * when has_norm[s] = 0; only scores above threshold.  loc and scale hold
 * nwin values per slot. */
CS_API CandidateList cs_gpu_session_metric_candidates(GpuSession *s, double threshold,
                                                      int nwin, const int *has_norm,
                                                      const double *edges, const double *loc,
                                                      const double *scale);

/* ------------------------------------------------------------------ */
/* Search engine: continuous batching of whole searches (no reference
 * counterpart: the reference searches one file at a time behind prefetch
 * threads, src/cli.c:590-617).  `nslots` worker streams; a request is
 * admitted, oldest first, when its HBM footprint fits `hbm_budget` next to the
 * running ones (a request larger than the budget runs alone).  Each request
 * stages its raw series on the device (cs_gpu_stage_*) and searches it as
 * search_file would (src/search.c:1247-1406) -- on a resident plan when an
 * earlier request had the same geometry -- returning the raw candidate
 * list of cs_gpu_session_run, or the final list with duty cycles when the
 * request asks for it (the spectrum then never leaves the device). */
typedef struct CsEngine CsEngine;

typedef struct {
    const float *samples;   /* N floats in host memory or in the engine device's memory
                               (e.g. a de-dispersed series, cs_gpu_dedisp_*); must stay
                               valid until the request is done */
    long long N;            /* even */
    double dt;
    const SearchParams *p;  /* copied at submit */
    double lofreq, hifreq;  /* search band (Hz); lobin != 100 overrides lofreq, as search_file */
    int lobin;
    int blocksize;          /* 0: CS_GPU_DEFAULT_BLOCKSIZE; < 0: the autotuner's chunk
                               length for the class (analytic sigma; same metrics and,
                               after the collapse, the same list; the raw list's
                               chunk order differs) */
    double threshold;
    float *amps_out;        /* optional host buffer of N floats: the staged spectrum */
    /* ncands > 0: the final list instead of the raw one -- collapsed as
     * search_file does (remove_duplicates within dr_tol, or a stable frequency
     * sort when remove is 0; then remove_harmonics up to numharm within
     * harm_tol when harm_remove), then cs_gpu_session_finalize (sort by S/N,
     * first ncands, duty cycles on the device) */
    int ncands;
    int numharm;
    int remove, harm_remove;
    double dr_tol, harm_tol;
    /* optional zaplist applied to the staged spectrum before the search
     * (cs_gpu_session_zap; copied at submit), non-'B' lines scaled by
     * 1 + baryv; amps_out then receives the zapped spectrum */
    const ZapList *zaplist;
    double baryv;
} CsSearchRequest;

typedef struct {
    int status;             /* 0 ok, -1 failed (message on stderr) */
    int reused;             /* 1: ran on a resident plan of the same geometry */
    long long harmonic_trials, trials, footprint;
    double queue_ms;        /* submit -> admitted */
    double run_ms;          /* admitted -> done */
    double latency_ms;      /* submit -> done */
    double kernel_ms;       /* search kernel, HIP events */
    double plan_ms;         /* host: plans, chunk starts, session and stage tables */
    double stage_ms;        /* upload + device realfft/rednoise (+ spectrum copy back) */
    double search_ms;       /* search launch to candidate list on the host */
    double post_ms;         /* collapse and device duty cycles (ncands > 0) */
} CsSearchInfo;

typedef struct {
    long long budget, in_use, peak_in_use;
    int nslots, running, peak_running, pad;
    long long queued, completed;
    long long plan_hits;        /* requests that ran on a resident plan */
    long long plans_resident;   /* plans kept now */
} CsEngineStats;

/* hbm_budget <= 0: 80 % of the device's free memory; nslots <= 0: 4 */
CS_API CsEngine *cs_engine_create(int device, long long hbm_budget, int nslots);
/* ticket >= 0, or -1 for an invalid request */
CS_API long long cs_engine_submit(CsEngine *e, const CsSearchRequest *r);
/* 1 done, 0 pending, -1 unknown ticket */
CS_API int cs_engine_poll(CsEngine *e, long long ticket);
/* the ticket of the next request to finish (completion order), -1 on timeout
 * (timeout_ms < 0 waits indefinitely) */
CS_API long long cs_engine_next_done(CsEngine *e, double timeout_ms);
/* waits for the request, hands over its list (caller frees with candlist_free
 * or free(out->v)) and releases the ticket; returns the request's status */
CS_API int cs_engine_wait(CsEngine *e, long long ticket, CandidateList *out,
                          CsSearchInfo *info);
CS_API void cs_engine_stats(CsEngine *e, CsEngineStats *s);
/* finishes every submitted request, then releases the engine */
CS_API void cs_engine_free(CsEngine *e);
/* resident plans kept between requests (default 8; 0 frees every plan when
 * its request ends).  A request whose geometry -- N, dt, search options,
 * band, chunk length -- matches an idle plan reuses it: no planning,
 * allocation or table upload, staging and search replayed from graphs. */
CS_API void cs_engine_set_plan_cache(CsEngine *e, int max_plans);

/* pinned host memory for series and spectra (hipHostMalloc) */
CS_API void *cs_host_alloc(long long bytes);
CS_API void cs_host_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
