/*
 * search_file: the trial grid, plan/chunk cache and GPU dispatch of
 * src/search.c:1247-1486, followed by the same collapse (remove_duplicates or
 * a stable frequency sort, then remove_harmonics).  The plan cache is reused
 * across files while T, N, namps, band, blocksize and total match
 * (src/search.c:1292-1297); the GPU session while cs_gpu_session_matches.
 * With --metricstats / --normalize the GPU keeps every trial's metric and
 * summarises it per (chunk, decimation) (cs_gpu_session_metric_stats); the
 * reference's statistics are assembled from those summaries here and the
 * candidates come from the kept metrics, normalised per frequency window
 * with --normalize (src/search.c:989-1095, 1427-1444).  There is no CPU
 * search: gpu_device < 0 is rejected with a message and an empty list.
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

static int int_asc(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static int block_order(const void *a, const void *b)
{
    const BlockMetricStats *x = (const BlockMetricStats *)a, *y = (const BlockMetricStats *)b;
    if (x->block != y->block) return (x->block > y->block) - (x->block < y->block);
    return (x->k > y->k) - (x->k < y->k);
}

/* The metric-statistics search of src/search.c:989-1095 on the GPU session
 * gs: one search keeping every trial's metric, then
 *   - per (chunk, decimation): the block statistics (src/search.c:537-565),
 *     and the chunk's metrics added to the histogram of the frequency window
 *     holding the chunk's mean frequency (src/search.c:864-868, 924-929),
 *     accumulated over chunks in order;
 *   - per decimation: the band-wide histogram, the windows merged in order;
 *   - the candidates: the kept metrics above threshold or, with normalize,
 *     their normalised scores (src/search.c:1428-1440: the reference's
 *     second search recomputes the same metrics, so it is not repeated).
 * The statistics are appended to ms. */
static CandidateList metric_search(GpuSession *gs, const FFTFile *ft, const SearchParams *p,
                                   MetricStats *ms, int normalize, double threshold,
                                   double r_lo, double r_hi, double lodr, long long total,
                                   const long long *cstarts, long long nchunks)
{
    CandidateList out;
    candlist_init(&out);
    const int nh = p->nharms, nwin = ms->nwin, nb = ms->hist_nb;
    /* slots: the base, then the decimations k > 1 in SearchParams order */
    int slot_k[64], nslot = 0;
    slot_k[nslot++] = 1;
    for (int i = 0; i < p->ndecim && nslot < 64; i++)
        if (p->decimations[i] > 1) slot_k[nslot++] = p->decimations[i];
    /* the decimations with statistics, ascending (src/search.c:1004-1020) */
    int statk[64], nstat = 0;
    for (int s = 0; s < nslot; s++) {
        int dup = 0;
        for (int j = 0; j < nstat; j++) dup |= statk[j] == slot_k[s];
        if (!dup) statk[nstat++] = slot_k[s];
    }
    qsort(statk, (size_t)nstat, sizeof(int), int_asc);
    int slot_si[64];
    for (int s = 0; s < nslot; s++) {
        int si = 0;
        while (si < nstat - 1 && statk[si] < slot_k[s]) si++;
        slot_si[s] = si;
    }
    const size_t nbs = (size_t)nchunks * nslot;
    double *wedges = (double *)malloc((size_t)nstat * (nwin + 1) * sizeof(double));
    int *win = (int *)malloc((nbs ? nbs : 1) * sizeof(int));
    CsMetricBlock *blk = (CsMetricBlock *)malloc((nbs ? nbs : 1) * sizeof(CsMetricBlock));
    long long *counts = (long long *)calloc((size_t)nslot * nwin * nb, sizeof(long long));
    MetricHistogram *wh = (MetricHistogram *)calloc((size_t)nstat * nwin, sizeof(MetricHistogram));
    int *has = (int *)calloc((size_t)nslot, sizeof(int));
    double *ed = (double *)calloc((size_t)nslot * (nwin + 1), sizeof(double));
    double *loc = (double *)calloc((size_t)nslot * nwin, sizeof(double));
    double *scl = (double *)calloc((size_t)nslot * nwin, sizeof(double));
    MetricNorm *norm = NULL;
    if (!wedges || !win || !blk || !counts || !wh || !has || !ed || !loc || !scl) goto done;
    for (int si = 0; si < nstat; si++)
        log_edges((double)statk[si] * r_lo / ft->T, (double)statk[si] * r_hi / ft->T, nwin,
                  wedges + (size_t)si * (nwin + 1));
    for (long long c = 0; c < nchunks; c++) {
        const long long i0 = cstarts[c];
        const long long n = (c + 1 < nchunks ? cstarts[c + 1] : total) - i0;
        const double rmean = r_lo + (double)i0 * lodr + (double)(n - 1) * lodr / 2.0;
        for (int s = 0; s < nslot; s++)
            win[c * nslot + s] = window_index(wedges + (size_t)slot_si[s] * (nwin + 1), nwin,
                                              (double)slot_k[s] * rmean / ft->T);
    }
    if (cs_gpu_session_metric_stats(gs, ft, win, nwin, ms->hist_lo, ms->hist_hi, nb, blk,
                                    counts) != 0) {
        fprintf(stderr, "Error: GPU metric statistics failed\n");
        goto done;
    }
    for (int si = 0; si < nstat; si++)
        for (int w = 0; w < nwin; w++) {
            const double *e = wedges + (size_t)si * (nwin + 1);
            hist_init(&wh[si * nwin + w], statk[si], nh / statk[si], w + 1, e[w], e[w + 1],
                      ms->hist_lo, ms->hist_hi, nb);
        }
    for (int s = 0; s < nslot; s++)
        for (int w = 0; w < nwin; w++) {
            MetricHistogram *h = &wh[slot_si[s] * nwin + w];
            const long long *src = counts + ((size_t)s * nwin + w) * nb;
            for (int b = 0; b < nb; b++) h->counts[b] += src[b];
        }
    for (long long c = 0; c < nchunks; c++) {
        const long long i0 = cstarts[c];
        const long long n = (c + 1 < nchunks ? cstarts[c + 1] : total) - i0;
        const double rstart = r_lo + (double)i0 * lodr;
        const double rmean = rstart + (double)(n - 1) * lodr / 2.0;
        for (int s = 0; s < nslot; s++) {
            const CsMetricBlock *b = &blk[c * nslot + s];
            if (b->n <= 0) continue;
            const int k = slot_k[s], Hk = nh / k;
            MetricHistogram *h = &wh[slot_si[s] * nwin + win[c * nslot + s] - 1];
            h->total += b->n;
            h->under += b->under;
            h->over += b->over;
            h->sum += b->sum;
            h->sumsq += b->sumsq;
            if (b->min < h->vmin) h->vmin = b->min;
            if (b->max > h->vmax) h->vmax = b->max;
            BlockMetricStats st;
            memset(&st, 0, sizeof(st));
            st.block = (int)(c + 1);
            st.k = k;
            st.Hk = Hk;
            st.nbins = 2 * Hk;
            st.ngoodbins = (double)ft->N / 2.0 / ((double)k * rmean);
            if (st.ngoodbins > (double)Hk) st.ngoodbins = (double)Hk;
            st.flo = (double)k * rstart / ft->T;
            st.fhi = (double)k * (rstart + (double)(b->n - 1) * lodr) / ft->T;
            if (k == 1) {
                st.ngoodbins = (double)ft->N / 2.0 / rmean;
                if (st.ngoodbins > (double)nh) st.ngoodbins = (double)nh;
                st.flo = rstart / ft->T;
                st.fhi = (rstart + (double)(b->n - 1) * lodr) / ft->T;
            }
            st.n = b->n;
            st.min = b->min;
            st.median = b->median;
            st.mean = b->sum / (double)b->n;
            st.std = b->n > 1 ? sqrt(b->ss / (double)(b->n - 1)) : 0.0;
            st.max = b->max;
            metricstats_append_block(ms, st);
        }
    }
    for (int si = 0; si < nstat; si++) {
        const double *e = wedges + (size_t)si * (nwin + 1);
        MetricHistogram g;
        hist_init(&g, statk[si], nh / statk[si], 0, e[0], e[nwin], ms->hist_lo, ms->hist_hi, nb);
        for (int w = 0; w < nwin; w++) {
            hist_merge(&g, &wh[si * nwin + w]);
            metricstats_append_whist(ms, wh[si * nwin + w]);
            wh[si * nwin + w].counts = NULL;       /* owned by ms now */
        }
        metricstats_append_hist(ms, g);
    }
    qsort(ms->blocks, ms->nblocks, sizeof(BlockMetricStats), block_order);
    if (normalize) {
        norm = build_metricnorm(ms);
        if (!norm) goto done;
        for (int s = 0; s < nslot; s++)
            for (int i = 0; i < norm->n; i++) {
                const NormK *nk = &norm->ks[i];
                if (nk->k != slot_k[s] || nk->nwin != nwin) continue;
                has[s] = 1;
                memcpy(ed + (size_t)s * (nwin + 1), nk->edges, (size_t)(nwin + 1) * sizeof(double));
                memcpy(loc + (size_t)s * nwin, nk->loc, (size_t)nwin * sizeof(double));
                memcpy(scl + (size_t)s * nwin, nk->scale, (size_t)nwin * sizeof(double));
                break;
            }
    }
    out = cs_gpu_session_metric_candidates(gs, threshold, nwin, has, ed, loc, scl);
done:
    if (wh)
        for (int i = 0; i < nstat * nwin; i++) free(wh[i].counts);
    free(wh);
    metricnorm_free(norm);
    free(wedges);
    free(win);
    free(blk);
    free(counts);
    free(has);
    free(ed);
    free(loc);
    free(scl);
    return out;
}

CandidateList search_file(const FFTFile *ft, const SearchParams *p, double lofreq,
                          double hifreq, int lobin, int blocksize, double threshold,
                          int remove, double dr_tol, int harm_remove, int numharm,
                          MetricStats *metricstats, int normalize, int verbose, int wisdom,
                          const char *wisdom_file, SearchCache *cache, double harm_tol,
                          int gpu_device)
{
    (void)wisdom;
    (void)wisdom_file;
    CandidateList empty;
    candlist_init(&empty);
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
    if (metricstats || normalize) {
        MetricStats local, *ms = metricstats;
        if (!ms) {
            metricstats_init(&local, 0.0, 64.0, 3200, 16);
            ms = &local;
        }
        cands = metric_search(gs, ft, p, ms, normalize, threshold, r_lo, r_hi, lodr, total,
                              cstarts, nchunks);
        if (!metricstats) metricstats_free(&local);
    } else {
        cands = cs_gpu_session_run(gs, ft, threshold);
    }
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
