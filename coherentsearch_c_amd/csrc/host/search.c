/*
 * search_file: the trial grid, plan/chunk cache and GPU dispatch of
 * src/search.c:1247-1486, followed by the same collapse (remove_duplicates or
 * a stable frequency sort, then remove_harmonics).  The plan cache is reused
 * across files while T, N, namps, band, blocksize and total match
 * (src/search.c:1292-1297); the GPU session while cs_gpu_session_matches.
 * There is no CPU search here: gpu_device < 0, --metricstats and --normalize
 * are rejected with a message and an empty list.
 */
#include "host_internal.h"

static void cache_drop_plans(SearchCache *c)
{
    if (c->plan_valid) {
        direct_plans_free(&c->dplans);
        free(c->cstarts);
        c->cstarts = NULL;
        c->plan_valid = 0;
    }
}

CandidateList search_file(const FFTFile *ft, const SearchParams *p, double lofreq,
                          double hifreq, int lobin, int blocksize, double threshold,
                          int remove, double dr_tol, int harm_remove, int numharm,
                          void *metricstats, int normalize, int verbose, int wisdom,
                          const char *wisdom_file, SearchCache *cache, double harm_tol,
                          int gpu_device)
{
    (void)wisdom;
    (void)wisdom_file;
    CandidateList empty;
    candlist_init(&empty);
    if (metricstats || normalize) {
        fprintf(stderr, "Error: --metricstats / --normalize are not supported by the GPU "
                        "search\n");
        return empty;
    }
    if (gpu_device < 0) {
        fprintf(stderr, "Error: this build searches on the GPU only (use --gpu N)\n");
        return empty;
    }
    const double lodr = p->hidr / (double)p->nharms;
    double r_lo = lofreq * ft->T;
    if (lobin != 100) r_lo = (double)lobin;   /* src/search.c:1259-1260 */
    const double r_hi = hifreq * ft->T;
    long long total = (long long)floor((r_hi - r_lo) / lodr) + 1;
    if (total < 0) total = 0;
    if (total == 0) return empty;
    int Nprof = blocksize == 0 ? CS_GPU_DEFAULT_BLOCKSIZE : blocksize;
    if (Nprof < 1) Nprof = 1;

    DirectPlanSet local;
    memset(&local, 0, sizeof(local));
    DirectPlanSet *dplans = NULL;
    long long *cstarts = NULL, nchunks = 0;
    int own = 0;
    const int hit = cache && cache->plan_valid && cache->params == p &&
                    cache->plan_T == ft->T && cache->plan_N == ft->N &&
                    cache->plan_namps == ft->namps && cache->plan_lofreq == lofreq &&
                    cache->plan_hifreq == hifreq && cache->plan_lobin == lobin &&
                    cache->plan_Nprof == Nprof && cache->plan_total == total;
    if (hit) {
        dplans = &cache->dplans;
        cstarts = cache->cstarts;
        nchunks = cache->nchunks;
    } else {
        DirectPlanSet *target = &local;
        if (cache) {
            cache_drop_plans(cache);
            target = &cache->dplans;
        } else {
            own = 1;
        }
        if (direct_plans_build(target, p, r_lo) != 0) return empty;
        if (chunk_starts(target, ft, p, total, Nprof, &cstarts, &nchunks) != 0) {
            direct_plans_free(target);
            return empty;
        }
        dplans = target;
        if (cache) {
            cache->cstarts = cstarts;
            cache->nchunks = nchunks;
            cache->params = p;
            cache->plan_T = ft->T;
            cache->plan_N = ft->N;
            cache->plan_namps = ft->namps;
            cache->plan_lofreq = lofreq;
            cache->plan_hifreq = hifreq;
            cache->plan_lobin = lobin;
            cache->plan_Nprof = Nprof;
            cache->plan_total = total;
            cache->plan_valid = 1;
        }
    }
    if (verbose)
        fprintf(stderr,
                "Search: %lld trials in %lld chunks of %d over r = %.1f ... %.1f bins "
                "(%.4f ... %.4f Hz), GPU\n",
                total, nchunks, Nprof, r_lo, r_hi, r_lo / ft->T, r_hi / ft->T);
    for (int i = 0; i < dplans->n; i++)
        if (dplans->plans[i].P == 0) {
            fprintf(stderr, "Error: --hidr is not a simple rational; the GPU search needs "
                            "the tabulated kernel\n");
            if (own) { free(cstarts); direct_plans_free(dplans); }
            return empty;
        }

    CandidateList cands;
    candlist_init(&cands);
    GpuSession *gs = cache ? (GpuSession *)cache->gpu_session : NULL;
    int own_session = 0;
    if (!gs || !cs_gpu_session_matches(gs, ft, p, Nprof, total, r_lo, lodr, gpu_device)) {
        if (gs) cs_gpu_session_free(gs);
        if (cache) cache->gpu_session = NULL;
        gs = cs_gpu_session_create(ft, p, dplans, cstarts, total, Nprof, nchunks, r_lo, lodr,
                                   gpu_device);
        if (!gs) {
            fprintf(stderr, "Error: failed to create the GPU session\n");
            if (own) { free(cstarts); direct_plans_free(dplans); }
            return empty;
        }
        if (cache) cache->gpu_session = gs;
        else own_session = 1;
    }
    cands = cs_gpu_session_run(gs, ft, threshold);
    if (own_session) cs_gpu_session_free(gs);

    if (remove) {
        CandidateList kept = remove_duplicates(&cands, dr_tol);
        candlist_free(&cands);
        cands = kept;
    } else {
        candlist_sort_freq_stable(&cands);
    }
    if (harm_remove) {
        CandidateList kept = remove_harmonics(&cands, numharm, harm_tol);
        candlist_free(&cands);
        cands = kept;
    }
    if (own) {
        free(cstarts);
        direct_plans_free(dplans);
    }
    return cands;
}

void search_cache_free(SearchCache *cache)
{
    if (!cache) return;
    if (cache->gpu_session) {
        cs_gpu_session_free((GpuSession *)cache->gpu_session);
        cache->gpu_session = NULL;
    }
    cache_drop_plans(cache);
    cache->params = NULL;
}
