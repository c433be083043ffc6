/*
 * Metric statistics and the per-(decimation, frequency) normalisation of
 * --metricstats / --normalize (src/search.c:409-535, 752-823) and their
 * report (src/cli.c:157-303).  The GPU search supplies the numbers
 * (cs_gpu_session_metric_stats); this file holds the histogram arithmetic,
 * the normalisation model built from it and the text output, with the
 * reference's rules:
 *   - a histogram of nb equal bins over [lo, hi) with underflow/overflow
 *     counts and the running sum, sum of squares, minimum and maximum;
 *   - quantiles interpolate linearly inside the bin that crosses q*total;
 *   - windows are log-spaced in frequency, nwin per decimation, and a
 *     frequency belongs to the window of the last edge at or below it;
 *   - the normalisation of decimation k in window w is (M - loc)/scale with
 *     loc the window's median and scale its 84.13 % quantile minus the
 *     median, the band-wide values standing in for windows with fewer than
 *     CS_MIN_WIN_TRIALS trials or no spread.
 */
#include "host_internal.h"

#include <float.h>

#define CS_MIN_WIN_TRIALS 200     /* include/coherent_search.h:43 */

void hist_init(MetricHistogram *h, int k, int Hk, int win, double flo, double fhi, double lo,
               double hi, int nb)
{
    memset(h, 0, sizeof(*h));
    h->k = k;
    h->Hk = Hk;
    h->nbins = 2 * Hk;
    h->win = win;
    h->flo = flo;
    h->fhi = fhi;
    h->lo = lo;
    h->hi = hi;
    h->nb = nb;
    h->invw = (double)nb / (hi - lo);
    h->counts = (long long *)calloc((size_t)(nb > 0 ? nb : 1), sizeof(long long));
    h->vmin = INFINITY;
    h->vmax = -INFINITY;
}

void hist_free(MetricHistogram *h)
{
    if (!h) return;
    free(h->counts);
    memset(h, 0, sizeof(*h));
}

void hist_push(MetricHistogram *h, double x)
{
    h->total++;
    h->sum += x;
    h->sumsq += x * x;
    if (x < h->vmin) h->vmin = x;
    if (x > h->vmax) h->vmax = x;
    if (x < h->lo) {
        h->under++;
        return;
    }
    if (x >= h->hi) {
        h->over++;
        return;
    }
    int b = (int)floor((x - h->lo) * h->invw);
    if (b < 0) b = 0;
    if (b > h->nb - 1) b = h->nb - 1;
    h->counts[b]++;
}

void hist_merge(MetricHistogram *a, const MetricHistogram *b)
{
    for (int i = 0; i < a->nb; i++) a->counts[i] += b->counts[i];
    a->under += b->under;
    a->over += b->over;
    a->total += b->total;
    a->sum += b->sum;
    a->sumsq += b->sumsq;
    if (b->vmin < a->vmin) a->vmin = b->vmin;
    if (b->vmax > a->vmax) a->vmax = b->vmax;
}

double hist_quantile(const MetricHistogram *h, double q)
{
    if (h->total == 0) return NAN;
    const double target = q * (double)h->total;
    long long below = h->under;
    if ((double)below >= target) return h->lo;
    const double width = 1.0 / h->invw;
    for (int i = 0; i < h->nb; i++) {
        const long long c = h->counts[i];
        if ((double)(below + c) >= target) {
            const double left = h->lo + (double)i * width;
            if (c == 0) return left;
            return left + (target - (double)below) / (double)c * width;
        }
        below += c;
    }
    return h->hi;
}

void log_edges(double a, double b, int nwin, double *edges)
{
    if (a < DBL_MIN) a = DBL_MIN;
    if (b < a * (1.0 + 1e-9)) b = a * (1.0 + 1e-9);
    const double la = log10(a), lb = log10(b);
    for (int i = 0; i <= nwin; i++)
        edges[i] = pow(10.0, la + (lb - la) * (double)i / (double)nwin);
}

int window_index(const double *edges, int nwin, double f)
{
    /* count of edges at or below f, clamped to 1..nwin */
    int lo = 0, hi = nwin + 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (edges[mid] <= f) lo = mid + 1;
        else hi = mid;
    }
    return lo < 1 ? 1 : (lo > nwin ? nwin : lo);
}

void metricstats_init(MetricStats *ms, double lo, double hi, int nb, int nwin)
{
    memset(ms, 0, sizeof(*ms));
    ms->hist_lo = lo;
    ms->hist_hi = hi;
    ms->hist_nb = nb;
    ms->nwin = nwin < 1 ? 1 : nwin;
}

void metricstats_free(MetricStats *ms)
{
    free(ms->blocks);
    for (size_t i = 0; i < ms->nhists; i++) hist_free(&ms->hists[i]);
    free(ms->hists);
    for (size_t i = 0; i < ms->nwhists; i++) hist_free(&ms->whists[i]);
    free(ms->whists);
    memset(ms, 0, sizeof(*ms));
}

/* append one element to a growable array of elements of size sz */
static int grow_push(void **v, size_t *n, size_t *cap, size_t first, const void *x, size_t sz)
{
    if (*n == *cap) {
        const size_t nc = *cap ? 2 * *cap : first;
        void *nv = realloc(*v, nc * sz);
        if (!nv) return -1;
        *v = nv;
        *cap = nc;
    }
    memcpy((char *)*v + *n * sz, x, sz);
    (*n)++;
    return 0;
}

void metricstats_append_block(MetricStats *ms, BlockMetricStats s)
{
    grow_push((void **)&ms->blocks, &ms->nblocks, &ms->capblocks, 256, &s, sizeof(s));
}

void metricstats_append_hist(MetricStats *ms, MetricHistogram h)
{
    grow_push((void **)&ms->hists, &ms->nhists, &ms->caphists, 16, &h, sizeof(h));
}

void metricstats_append_whist(MetricStats *ms, MetricHistogram h)
{
    grow_push((void **)&ms->whists, &ms->nwhists, &ms->capwhists, 64, &h, sizeof(h));
}

/* median and the 84.13 % quantile minus it (NaN when not positive) */
static void robust_loc_scale(const MetricHistogram *h, double *loc, double *scale)
{
    *loc = hist_quantile(h, 0.5);
    const double up = hist_quantile(h, 0.8413);
    *scale = up - *loc;
    if (!(*scale > 0.0)) *scale = NAN;
}

MetricNorm *build_metricnorm(const MetricStats *ms)
{
    MetricNorm *nm = (MetricNorm *)calloc(1, sizeof(MetricNorm));
    if (!nm) return NULL;
    nm->ks = (NormK *)calloc(ms->nhists ? ms->nhists : 1, sizeof(NormK));
    if (!nm->ks) {
        free(nm);
        return NULL;
    }
    for (size_t g = 0; g < ms->nhists; g++) {
        const int k = ms->hists[g].k;
        double gloc, gscale;
        robust_loc_scale(&ms->hists[g], &gloc, &gscale);
        if (isnan(gscale)) gscale = 1.0;
        int nwin = 0;
        for (size_t w = 0; w < ms->nwhists; w++) nwin += ms->whists[w].k == k;
        if (nwin == 0) continue;
        NormK *nk = &nm->ks[nm->n];
        nk->k = k;
        nk->nwin = nwin;
        nk->edges = (double *)malloc((size_t)(nwin + 1) * sizeof(double));
        nk->loc = (double *)malloc((size_t)nwin * sizeof(double));
        nk->scale = (double *)malloc((size_t)nwin * sizeof(double));
        if (!nk->edges || !nk->loc || !nk->scale) {
            free(nk->edges);
            free(nk->loc);
            free(nk->scale);
            metricnorm_free(nm);
            return NULL;
        }
        int i = 0;
        double top = 0.0;
        for (size_t w = 0; w < ms->nwhists; w++) {
            const MetricHistogram *h = &ms->whists[w];
            if (h->k != k) continue;
            nk->edges[i] = h->flo;
            top = h->fhi;
            double l, s;
            robust_loc_scale(h, &l, &s);
            const int thin = h->total < CS_MIN_WIN_TRIALS || isnan(s);
            nk->loc[i] = thin ? gloc : l;
            nk->scale[i] = thin ? gscale : s;
            i++;
        }
        nk->edges[nwin] = top;
        nm->n++;
    }
    return nm;
}

void metricnorm_free(MetricNorm *n)
{
    if (!n) return;
    for (int i = 0; i < n->n; i++) {
        free(n->ks[i].edges);
        free(n->ks[i].loc);
        free(n->ks[i].scale);
    }
    free(n->ks);
    free(n);
}

double metric_normalize(const MetricNorm *n, int k, double f, double M)
{
    for (int i = 0; i < n->n; i++) {
        if (n->ks[i].k != k) continue;
        const int w = window_index(n->ks[i].edges, n->ks[i].nwin, f) - 1;
        return (M - n->ks[i].loc[w]) / n->ks[i].scale[w];
    }
    return M;
}

/* ---- the report (src/cli.c:157-303) ------------------------------------ */

/* mean, sample standard deviation from the running sums, median, the metric
 * at each single-trial false-alarm probability, and the overflow fraction */
static void hist_summary(const MetricHistogram *h, const double *faps, int nfaps, double *mean,
                         double *std, double *median, double *fapv, double *overflow)
{
    const long long N = h->total;
    *mean = N > 0 ? h->sum / (double)N : NAN;
    *std = N > 1 ? sqrt(fmax(0.0, (h->sumsq - (double)N * (*mean) * (*mean)) / (double)(N - 1)))
                 : 0.0;
    *median = hist_quantile(h, 0.5);
    for (int i = 0; i < nfaps; i++) fapv[i] = hist_quantile(h, 1.0 - faps[i]);
    *overflow = N > 0 ? (double)h->over / (double)N : 0.0;
}

void report_metricstats(const MetricStats *ms, const char *base, const SearchParams *p,
                        double threshold)
{
    int any = 0;
    for (size_t i = 0; i < ms->nhists; i++) any |= ms->hists[i].total > 0;
    if (!any) {
        fprintf(stderr, "Warning: No metric statistics collected (no blocks searched)\n");
        return;
    }
    static const double faps[5] = {0.1, 0.01, 1e-3, 1e-4, 1e-5};
    static const char *const flab[5] = {"1e-1", "1e-2", "1e-3", "1e-4", "1e-5"};
    char thr[64];
    cs_format_double(thr, sizeof(thr), threshold);

    /* per-decimation table on stderr */
    fprintf(stderr, "\nMetric statistics by decimation  (peak boxcar S/N, threshold=%s)\n", thr);
    fprintf(stderr, "  Pure-noise floor scales ~sqrt(nbins); a fixed threshold picks a "
                    "DIFFERENT false-alarm rate per k.\n");
    fprintf(stderr, "  %-3s %4s %5s %10s %7s %7s %7s | metric at single-trial FAP =\n", "k",
            "Hk", "nbins", "ntrials", "mean", "std", "max");
    fprintf(stderr, "  %-3s %4s %5s %10s %7s %7s %7s | %8s %8s %8s %8s %8s\n", "", "", "", "",
            "", "", "", flab[0], flab[1], flab[2], flab[3], flab[4]);
    for (size_t i = 0; i < ms->nhists; i++) {
        const MetricHistogram *h = &ms->hists[i];
        double mean, std, med, fv[5], ov;
        hist_summary(h, faps, 5, &mean, &std, &med, fv, &ov);
        fprintf(stderr, "  %-3d %4d %5d %10lld %7.3f %7.3f %7.3f |", h->k, h->Hk, h->nbins,
                h->total, mean, std, h->total > 0 ? h->vmax : NAN);
        for (int q = 0; q < 5; q++) fprintf(stderr, " %8.3f", fv[q]);
        if (ov > faps[4]) fprintf(stderr, "  (overflow %.1e!)", ov);
        fprintf(stderr, "\n");
    }
    fprintf(stderr, "  A single --threshold gives these per-k FAPs; for comparable FAP, "
                    "threshold each k separately\n");
    fprintf(stderr, "  (or normalise the metric per (k, frequency)).  FAP columns are "
                    "single-trial, single-decimation.\n");

    /* the 1e-4 threshold's range over the windows of each k, in order of
     * first appearance */
    fprintf(stderr, "\n  FAP=1e-4 threshold vs frequency (%d log-spaced windows per k; drift = "
                    "red noise / Nyquist):\n", ms->nwin);
    fprintf(stderr, "  %-3s %5s %21s %9s %9s\n", "k", "nbins", "freq range (Hz)", "min", "max");
    int seen[256], nseen = 0;
    for (size_t i = 0; i < ms->nwhists; i++) {
        int dup = 0;
        for (int j = 0; j < nseen; j++) dup |= seen[j] == ms->whists[i].k;
        if (!dup && nseen < 256) seen[nseen++] = ms->whists[i].k;
    }
    for (int j = 0; j < nseen; j++) {
        double vmin = INFINITY, vmax = -INFINITY, fl = INFINITY, fh = -INFINITY;
        int nbins = 0, first = 1;
        for (size_t i = 0; i < ms->nwhists; i++) {
            const MetricHistogram *h = &ms->whists[i];
            if (h->k != seen[j]) continue;
            const double v = hist_quantile(h, 1.0 - 1e-4);
            if (v < vmin) vmin = v;
            if (v > vmax) vmax = v;
            if (h->flo < fl) fl = h->flo;
            if (h->fhi > fh) fh = h->fhi;
            if (first) {
                nbins = h->nbins;
                first = 0;
            }
        }
        fprintf(stderr, "  %-3d %5d %9.3f -%9.3f %9.3f %9.3f\n", seen[j], nbins, fl, fh, vmin,
                vmax);
    }
    fprintf(stderr, "  Full per-window thresholds -> %s_metricfap.txt\n\n", base);

    char path[4096];
    FILE *io;
    snprintf(path, sizeof(path), "%s_metricstats.txt", base);
    if ((io = fopen(path, "w")) != NULL) {
        fprintf(io, "# Per-block, per-decimation metric statistics\n");
        fprintf(io, "# metric=boxcar nharms=%d threshold=%s\n", p->nharms, thr);
        fprintf(io, "# nbins = 2*Hk; the pure-noise metric floor scales ~sqrt(nbins).\n");
        fprintf(io, "#%-7s %3s %4s %5s %10s %14s %14s %8s %9s %9s %9s %9s %9s\n", "block", "k",
                "Hk", "nbins", "ngoodbins", "f_lo(Hz)", "f_hi(Hz)", "n", "min", "median", "mean",
                "std", "max");
        for (size_t i = 0; i < ms->nblocks; i++) {
            const BlockMetricStats *b = &ms->blocks[i];
            fprintf(io, "%-8d %3d %4d %5d %10.3f %14.8f %14.8f %8lld %9.3f %9.3f %9.3f %9.3f "
                        "%9.3f\n", b->block, b->k, b->Hk, b->nbins, b->ngoodbins, b->flo, b->fhi,
                    b->n, b->min, b->median, b->mean, b->std, b->max);
        }
        fclose(io);
    }

    snprintf(path, sizeof(path), "%s_metricfap.txt", base);
    if ((io = fopen(path, "w")) != NULL) {
        fprintf(io, "# Per-(decimation, frequency-window) empirical false-alarm thresholds\n");
        fprintf(io, "# metric=boxcar nharms=%d\n", p->nharms);
        fprintf(io, "# fap_X = metric value whose single-trial, single-decimation false-alarm "
                    "prob is X (in-window)\n");
        fprintf(io, "#%-3s %4s %5s %4s %13s %13s %10s %8s %8s %8s | %s\n", "k", "Hk", "nbins",
                "win", "f_lo(Hz)", "f_hi(Hz)", "ntrials", "mean", "std", "median",
                "fap1e-1 fap1e-2 fap1e-3 fap1e-4 fap1e-5");
        for (size_t i = 0; i < ms->nwhists; i++) {
            const MetricHistogram *h = &ms->whists[i];
            double mean, std, med, fv[5], ov;
            hist_summary(h, faps, 5, &mean, &std, &med, fv, &ov);
            fprintf(io, "%-4d %4d %5d %4d %13.7f %13.7f %10lld %8.3f %8.3f %8.3f |", h->k, h->Hk,
                    h->nbins, h->win, h->flo, h->fhi, h->total, mean, std, med);
            for (int q = 0; q < 5; q++) fprintf(io, " %7.3f", fv[q]);
            if (ov > faps[4]) fprintf(io, "  # overflow %.1e", ov);
            fprintf(io, "\n");
        }
        fclose(io);
    }

    snprintf(path, sizeof(path), "%s_metrichist.txt", base);
    if ((io = fopen(path, "w")) != NULL) {
        const double bw = (ms->hist_hi - ms->hist_lo) / (double)ms->hist_nb;
        fprintf(io, "# Per-(decimation, frequency-window) metric histograms (one streaming "
                    "pass)\n");
        fprintf(io, "# metric=boxcar nharms=%d\n", p->nharms);
        fprintf(io, "# bin = left edge of a bin of width %g; count = trials in [bin, bin+width)\n",
                bw);
        fprintf(io, "#%-3s %4s %4s %13s %13s %12s %14s\n", "k", "Hk", "win", "f_lo(Hz)",
                "f_hi(Hz)", "bin", "count");
        for (size_t i = 0; i < ms->nwhists; i++) {
            const MetricHistogram *h = &ms->whists[i];
            int b0 = -1, b1 = -1;
            for (int b = 0; b < h->nb; b++)
                if (h->counts[b] > 0) {
                    if (b0 < 0) b0 = b;
                    b1 = b;
                }
            if (b0 < 0) continue;
            for (int b = b0; b <= b1; b++)
                fprintf(io, "%-4d %4d %4d %13.7f %13.7f %12.5f %14lld\n", h->k, h->Hk, h->win,
                        h->flo, h->fhi, ms->hist_lo + (double)b * bw, h->counts[b]);
        }
        fclose(io);
    }
}
