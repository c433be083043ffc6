/*
 * Zaplists (src/zap.c, PRESTO accelsearch -zaplist semantics): the file
 * reader and the in-memory zap of a spectrum.  The intervals come from
 * csx_zap_ops (shared with the device path); each is applied in turn, so the
 * median windows of a later interval see the bins an earlier one set.
 */
#include "host_internal.h"

int zaplist_read(const char *path, ZapList *z)
{
    memset(z, 0, sizeof(*z));
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char line[1024];
    int cap = 0;
    while (fgets(line, sizeof(line), f)) {
        const char *p = line;
        if (*p == '#' || *p == '\n' || *p == '\r' || *p == '\0') continue;
        int bary = 0;
        if (*p == 'B') {
            bary = 1;
            p++;
        }
        double freq = 0.0, width = 0.0;
        if (sscanf(p, "%lf %lf", &freq, &width) != 2) continue;
        if (!(freq > 0.0) || !(width > 0.0)) continue;
        if (z->n == cap) {
            const int nc = cap ? 2 * cap : 32;
            ZapBird *v = (ZapBird *)realloc(z->v, (size_t)nc * sizeof(ZapBird));
            if (!v) {
                fclose(f);
                zaplist_free(z);
                return -1;
            }
            z->v = v;
            cap = nc;
        }
        z->v[z->n].freq = freq;
        z->v[z->n].width = width;
        z->v[z->n].bary = bary;
        z->n++;
    }
    fclose(f);
    return 0;
}

void zaplist_free(ZapList *z)
{
    if (!z) return;
    free(z->v);
    z->v = NULL;
    z->n = 0;
}

/* lower median (element (n-1)/2 in ascending order) of the powers of n bins */
static float median_power(const float *amps, long long s, int n)
{
    if (n <= 0) return 0.0f;
    float v[CS_ZAP_MEDIANBINS];
    for (int i = 0; i < n; i++) {
        const float re = amps[2 * (s + i)], im = amps[2 * (s + i) + 1];
        v[i] = fmaf(re, re, im * im);     /* the reference's contracted r*r + im*im */
    }
    for (int i = 1; i < n; i++) {
        const float x = v[i];
        int j = i - 1;
        while (j >= 0 && v[j] > x) {
            v[j + 1] = v[j];
            j--;
        }
        v[j + 1] = x;
    }
    return v[(n - 1) / 2];
}

int zaplist_apply(const ZapList *z, FFTFile *ft, double baryv)
{
    if (!z || z->n <= 0) return 0;
    CsZapOp *ops = NULL;
    const int n = csx_zap_ops(z, ft->T, ft->N, ft->namps, baryv, &ops);
    if (n < 0) return -1;
    const double ln2 = -log(0.5);
    for (int k = 0; k < n; k++) {
        const CsZapOp *o = &ops[k];
        const float lo = median_power(ft->amps, o->s_lo, o->n_lo);
        const float hi = median_power(ft->amps, o->s_hi, o->n_hi);
        const float avg = (float)sqrt(0.5 * (lo + hi) / ln2);
        for (long long i = o->ilo; i < o->ihi; i++) {
            ft->amps[2 * i] = avg;
            ft->amps[2 * i + 1] = 0.0f;
        }
    }
    free(ops);
    return n;
}
