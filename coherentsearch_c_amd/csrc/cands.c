/*
 * Candidate lists and the collapse steps, shared by the HIP library (the
 * engine's post-processing) and the host C library (which re-exports them
 * under the reference's names): stable bottom-up merge sorts,
 * remove_duplicates (src/candidate.c:98-123), harmonically_related and
 * remove_harmonics (src/candidate.c:125-161).  Pure C, no device code.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "cands.h"

void csx_candlist_init(CandidateList *l) { l->v = NULL; l->n = 0; l->cap = 0; }

void csx_candlist_free(CandidateList *l)
{
    free(l->v);
    csx_candlist_init(l);
}

static int candlist_reserve(CandidateList *l, size_t need)
{
    if (need <= l->cap) return 0;
    size_t nc = l->cap ? l->cap : 256;
    while (nc < need) nc *= 2;
    Candidate *v = (Candidate *)realloc(l->v, nc * sizeof(Candidate));
    if (!v) return -1;
    l->v = v;
    l->cap = nc;
    return 0;
}

void csx_candlist_push(CandidateList *l, Candidate c)
{
    if (candlist_reserve(l, l->n + 1) != 0) return;
    l->v[l->n++] = c;
}

void csx_candlist_append(CandidateList *dst, const CandidateList *src)
{
    if (src->n == 0 || candlist_reserve(dst, dst->n + src->n) != 0) return;
    memcpy(dst->v + dst->n, src->v, src->n * sizeof(Candidate));
    dst->n += src->n;
}

typedef int (*cmp_fn)(const Candidate *, const Candidate *);

/* stable: on ties the element from the left run goes first */
static void stable_sort(Candidate *v, size_t n, cmp_fn cmp)
{
    if (n < 2) return;
    Candidate *tmp = (Candidate *)malloc(n * sizeof(Candidate));
    if (!tmp) return;
    for (size_t w = 1; w < n; w *= 2) {
        for (size_t i = 0; i < n; i += 2 * w) {
            const size_t mid = i + w < n ? i + w : n;
            const size_t end = i + 2 * w < n ? i + 2 * w : n;
            size_t a = i, b = mid, k = i;
            while (a < mid && b < end) tmp[k++] = cmp(&v[b], &v[a]) < 0 ? v[b++] : v[a++];
            while (a < mid) tmp[k++] = v[a++];
            while (b < end) tmp[k++] = v[b++];
        }
        memcpy(v, tmp, n * sizeof(Candidate));
    }
    free(tmp);
}

static int by_metric_desc(const Candidate *a, const Candidate *b)
{
    return (b->metric > a->metric) - (b->metric < a->metric);
}
static int by_freq(const Candidate *a, const Candidate *b)
{
    return (a->freq > b->freq) - (a->freq < b->freq);
}
static int by_r(const Candidate *a, const Candidate *b)
{
    return (a->r > b->r) - (a->r < b->r);
}

void csx_candlist_sort_metric_desc(CandidateList *l) { stable_sort(l->v, l->n, by_metric_desc); }
void csx_candlist_sort_freq(CandidateList *l) { stable_sort(l->v, l->n, by_freq); }
void csx_candlist_sort_r(CandidateList *l) { stable_sort(l->v, l->n, by_r); }

static CandidateList copy_list(const CandidateList *in)
{
    CandidateList t;
    csx_candlist_init(&t);
    if (candlist_reserve(&t, in->n) == 0) {
        memcpy(t.v, in->v, in->n * sizeof(Candidate));
        t.n = in->n;
    }
    return t;
}

/* Runs of candidates within dr_tol bins of their predecessor keep their
 * strongest member; result sorted by frequency. */
CandidateList csx_remove_duplicates(const CandidateList *in, double dr_tol)
{
    CandidateList out;
    csx_candlist_init(&out);
    if (in->n == 0) return out;
    CandidateList t = copy_list(in);
    csx_candlist_sort_r(&t);
    Candidate best = t.v[0];
    double prev = best.r;
    for (size_t i = 1; i < t.n; i++) {
        const Candidate c = t.v[i];
        if (c.r - prev <= dr_tol) {
            if (c.metric > best.metric) best = c;
        } else {
            csx_candlist_push(&out, best);
            best = c;
        }
        prev = c.r;
    }
    csx_candlist_push(&out, best);
    csx_candlist_free(&t);
    csx_candlist_sort_freq(&out);
    return out;
}

int csx_harmonically_related(double r1, double r2, int numharm, double tol)
{
    const double lo = r1 < r2 ? r1 : r2, hi = r1 < r2 ? r2 : r1;
    if (!(lo > 0.0)) return 0;
    for (int m = 1; m <= numharm; m++) {
        const int n = (int)nearbyint((double)m * hi / lo);
        if (n < 1 || n > numharm) continue;
        if (fabs((double)m * hi - (double)n * lo) <= tol * (double)m) return 1;
    }
    return 0;
}

/* Greedy by descending metric: keep a candidate unless it is m/n-related
 * to one already kept; result sorted by frequency. */
CandidateList csx_remove_harmonics(const CandidateList *in, int numharm, double tol)
{
    CandidateList out;
    csx_candlist_init(&out);
    if (in->n == 0) return out;
    CandidateList t = copy_list(in);
    csx_candlist_sort_metric_desc(&t);
    for (size_t i = 0; i < t.n; i++) {
        int rel = 0;
        for (size_t j = 0; j < out.n && !rel; j++)
            rel = csx_harmonically_related(t.v[i].r, out.v[j].r, numharm, tol);
        if (!rel) csx_candlist_push(&out, t.v[i]);
    }
    csx_candlist_free(&t);
    csx_candlist_sort_freq(&out);
    return out;
}
