/*
 * Candidate lists, collapsing, duty cycles and the .cohout writer.
 * Semantics of src/candidate.c:11-379 (stable bottom-up merge sorts,
 * remove_duplicates :98-123, remove_harmonics :138-161, candidate_profile
 * :208-271, measure_ducy :305-329, write_candidates :350-379) and of
 * boxcar_best_width (src/boxcar.c:736-785) and fourier_interp_complex
 * (src/interp.c:317-344).  The duty-cycle profile's c2r uses this
 * library's own FFT in place of FFTW.
 */
#include "host_internal.h"

#ifndef M_PI
#define M_PI 3.141592653589793238462643383279502884
#endif

void candlist_init(CandidateList *l) { l->v = NULL; l->n = 0; l->cap = 0; }

void candlist_free(CandidateList *l)
{
    free(l->v);
    candlist_init(l);
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

void candlist_push(CandidateList *l, Candidate c)
{
    if (candlist_reserve(l, l->n + 1) != 0) return;
    l->v[l->n++] = c;
}

void candlist_append(CandidateList *dst, const CandidateList *src)
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

void candlist_sort_metric_desc(CandidateList *l) { stable_sort(l->v, l->n, by_metric_desc); }
void candlist_sort_freq(CandidateList *l) { stable_sort(l->v, l->n, by_freq); }
void candlist_sort_r(CandidateList *l) { stable_sort(l->v, l->n, by_r); }
void candlist_sort_metric_desc_stable(CandidateList *l) { candlist_sort_metric_desc(l); }
void candlist_sort_freq_stable(CandidateList *l) { candlist_sort_freq(l); }
void candlist_sort_r_stable(CandidateList *l) { candlist_sort_r(l); }

static CandidateList copy_list(const CandidateList *in)
{
    CandidateList t;
    candlist_init(&t);
    if (candlist_reserve(&t, in->n) == 0) {
        memcpy(t.v, in->v, in->n * sizeof(Candidate));
        t.n = in->n;
    }
    return t;
}

/* Runs of candidates within dr_tol bins of their predecessor keep their
 * strongest member; result sorted by frequency. */
CandidateList remove_duplicates(const CandidateList *in, double dr_tol)
{
    CandidateList out;
    candlist_init(&out);
    if (in->n == 0) return out;
    CandidateList t = copy_list(in);
    candlist_sort_r(&t);
    Candidate best = t.v[0];
    double prev = best.r;
    for (size_t i = 1; i < t.n; i++) {
        const Candidate c = t.v[i];
        if (c.r - prev <= dr_tol) {
            if (c.metric > best.metric) best = c;
        } else {
            candlist_push(&out, best);
            best = c;
        }
        prev = c.r;
    }
    candlist_push(&out, best);
    candlist_free(&t);
    candlist_sort_freq(&out);
    return out;
}

int harmonically_related(double r1, double r2, int numharm, double tol)
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
CandidateList remove_harmonics(const CandidateList *in, int numharm, double tol)
{
    CandidateList out;
    candlist_init(&out);
    if (in->n == 0) return out;
    CandidateList t = copy_list(in);
    candlist_sort_metric_desc(&t);
    for (size_t i = 0; i < t.n; i++) {
        int rel = 0;
        for (size_t j = 0; j < out.n && !rel; j++)
            rel = harmonically_related(t.v[i].r, out.v[j].r, numharm, tol);
        if (!rel) candlist_push(&out, t.v[i]);
    }
    candlist_free(&t);
    candlist_sort_freq(&out);
    return out;
}

void fourier_interp_complex(double r, const float *amps, long long namps, int m,
                            double *out_re, double *out_im)
{
    const int m2 = m / 2;
    double dr = fmod(r, 1.0);
    if (dr < 0) dr += 1.0;
    const long long first = (long long)floor(r + 1e-15) + 1 - m2;
    double sre = 0.0, sim = 0.0;
    for (int j = 0; j < m; j++) {
        const double off = dr - (double)(-m2 + 1 + j);
        const double sinc = off == 0.0 ? 1.0 : sin(M_PI * off) / (M_PI * off);
        const double cr = sinc * cos(M_PI * off), ci = -sinc * sin(M_PI * off);
        const long long idx = first + j;
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

/* Normalised pulse profile from the first harmonics that fit (exact m-tap
 * interpolation at h*r, then a length-2H real inverse transform / 2H). */
double *candidate_profile(const FFTFile *ft, double r, int nharm, int m, int *out_n)
{
    *out_n = 0;
    if (nharm < 1) return NULL;
    const int m2 = m / 2;
    const long long Nhalf = ft->N / 2;
    double *amps = (double *)malloc((size_t)nharm * 2 * sizeof(double));
    if (!amps) return NULL;
    int H = 0;
    for (int h = 1; h <= nharm; h++) {
        const double rh = r * (double)h;
        const long long rint0 = (long long)floor(rh + 1e-15) + 1;
        if (!(rint0 - m2 >= 0 && rint0 + m2 <= ft->namps && rh < (double)Nhalf)) break;
        fourier_interp_complex(rh, ft->amps, ft->namps, m, &amps[2 * H], &amps[2 * H + 1]);
        H++;
    }
    if (H < 1) {
        free(amps);
        double *z = (double *)calloc(2, sizeof(double));
        if (z) *out_n = 2;
        return z;
    }
    const int n = 2 * H;
    double *z = (double *)calloc((size_t)n * 2, sizeof(double));
    double *out = (double *)malloc((size_t)n * sizeof(double));
    if (!z || !out) { free(amps); free(z); free(out); return NULL; }
    /* Hermitian spectrum: bin h (1..H-1) and n-h conj, Nyquist real */
    for (int h = 1; h <= H; h++) {
        const double re = amps[2 * (h - 1)], im = amps[2 * (h - 1) + 1];
        if (h == H) {
            z[2 * h] = re;
            continue;
        }
        z[2 * h] = re;
        z[2 * h + 1] = im;
        z[2 * (n - h)] = re;
        z[2 * (n - h) + 1] = -im;
    }
    cs_fft_any(z, n, +1);
    const double inv = 1.0 / (double)n;
    for (int i = 0; i < n; i++) out[i] = z[2 * i] * inv;
    free(z);
    free(amps);
    *out_n = n;
    return out;
}

int boxcar_best_width(const double *prof, int nbins, double fsp, double maxfrac,
                      int *out_w, double *out_ducy)
{
    if (nbins < 1) return -1;
    int widths[256];
    const int nw = csx_boxcar_widths(nbins, fsp, maxfrac, widths, 256);
    if (nw <= 0) {
        *out_w = 1;
        *out_ducy = 1.0 / nbins;
        return 0;
    }
    const int wmax = widths[nw - 1];
    double *ps = (double *)malloc((size_t)(nbins + wmax + 1) * sizeof(double));
    const int Hk = nbins / 2;
    uint8_t *filled = (uint8_t *)malloc((size_t)(Hk > 0 ? Hk : 1));
    double *shape = (double *)malloc((size_t)nw * sizeof(double));
    if (!ps || !filled || !shape) { free(ps); free(filled); free(shape); return -1; }
    double sum = 0.0;
    for (int i = 0; i < nbins; i++) sum += prof[i];
    const double mean = sum / (double)nbins;
    ps[0] = 0.0;
    for (int i = 1; i <= nbins + wmax; i++) {
        const int idx = i > nbins ? i - nbins : i;
        ps[i] = ps[i - 1] + (prof[idx - 1] - mean);
    }
    const double stot = ps[nbins] - ps[0];
    for (int j = 0; j < Hk; j++) filled[j] = 1;
    if (Hk >= 1) csx_boxcar_shape(shape, nbins, widths, nw, filled, 1, Hk);
    else for (int i = 0; i < nw; i++) shape[i] = 0.0;
    double best = -INFINITY;
    int bestw = widths[0];
    for (int i = 0; i < nw; i++) {
        const int w = widths[i];
        const double duty = (double)w / (double)nbins;
        double mx = ps[w] - ps[0];
        for (int q = 2; q <= nbins; q++) {
            const double d = ps[q + w - 1] - ps[q - 1];
            if (d > mx) mx = d;
        }
        const double cand = (mx - duty * stot) * shape[i];
        if (cand > best) { best = cand; bestw = w; }
    }
    *out_w = bestw;
    *out_ducy = (double)bestw / (double)nbins;
    free(shape);
    free(filled);
    free(ps);
    return 0;
}

CandidateList measure_ducy(const FFTFile *ft, const CandidateList *in, const SearchParams *p)
{
    CandidateList out;
    candlist_init(&out);
    for (size_t i = 0; i < in->n; i++) {
        Candidate c = in->v[i];
        c.ducy = NAN;
        if (c.nharm >= 1 && c.r > 0.0) {
            int n = 0;
            double *prof = candidate_profile(ft, c.r, c.nharm, 64, &n);
            if (prof) {
                int w = 1;
                double d = NAN;
                if (n >= 2 && boxcar_best_width(prof, n, p->boxcar_fsp, p->boxcar_maxfrac,
                                                &w, &d) == 0)
                    c.ducy = d;
                free(prof);
            }
        }
        candlist_push(&out, c);
    }
    return out;
}

void cs_format_double(char *buf, size_t n, double x)
{
    if (isnan(x)) { snprintf(buf, n, "NaN"); return; }
    if (isinf(x)) { snprintf(buf, n, x > 0 ? "Inf" : "-Inf"); return; }
    for (int prec = 1; prec <= 17; prec++) {
        snprintf(buf, n, "%.*g", prec, x);
        char *end = NULL;
        if (strtod(buf, &end) == x && end && *end == '\0') break;
    }
    if (!strpbrk(buf, ".eE") && !strstr(buf, "inf") && !strstr(buf, "nan")) {
        const size_t l = strlen(buf);
        if (l + 3 < n) { buf[l] = '.'; buf[l + 1] = '0'; buf[l + 2] = '\0'; }
    }
}

int write_candidates(const CandidateList *cands, const char *outfile, double threshold)
{
    (void)threshold;
    const int to_stdout = outfile == NULL || *outfile == '\0';
    FILE *f = to_stdout ? stdout : fopen(outfile, "w");
    if (!f) {
        fprintf(stderr, "Error: cannot write candidates to '%s'\n", outfile);
        return -1;
    }
    fprintf(f, "#Num    'S/N'      Frequency (Hz)        Period (ms)    #Harm  Ducy(%%)\n");
    for (size_t i = 0; i < cands->n; i++) {
        const Candidate *c = &cands->v[i];
        char d[32];
        if (isnan(c->ducy)) snprintf(d, sizeof(d), "-");
        else snprintf(d, sizeof(d), "%.2f", 100.0 * c->ducy);
        fprintf(f, "%-4zu  %8.2f  %18.12f  %18.12f   %3d   %6s\n", i + 1, c->metric,
                c->freq, 1000.0 / c->freq, c->nharm, d);
    }
    if (to_stdout) fflush(stdout);
    else fclose(f);
    return 0;
}

CandidateList read_candidates_file(const char *path, double T)
{
    CandidateList out;
    candlist_init(&out);
    FILE *f = fopen(path, "r");
    if (!f) {
        fprintf(stderr, "Error: cannot open candidate file '%s'\n", path);
        return out;
    }
    char line[4096];
    while (fgets(line, sizeof(line), f)) {
        char *s = line;
        while (*s == ' ' || *s == '\t') s++;
        if (*s == '#' || *s == '\n' || *s == '\r' || *s == '\0') continue;
        int rank = 0, nharm = 0;
        double snr = 0.0, freq = 0.0, period = 0.0;
        if (sscanf(s, "%d %lf %lf %lf %d", &rank, &snr, &freq, &period, &nharm) >= 5 &&
            freq > 0.0) {
            Candidate c = {freq, snr, freq * T, nharm, NAN};
            candlist_push(&out, c);
        }
    }
    fclose(f);
    return out;
}
