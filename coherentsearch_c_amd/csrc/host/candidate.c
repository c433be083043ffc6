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

void candlist_init(CandidateList *l) { csx_candlist_init(l); }
void candlist_free(CandidateList *l) { csx_candlist_free(l); }
void candlist_push(CandidateList *l, Candidate c) { csx_candlist_push(l, c); }
void candlist_append(CandidateList *dst, const CandidateList *src)
{
    csx_candlist_append(dst, src);
}
void candlist_sort_metric_desc(CandidateList *l) { csx_candlist_sort_metric_desc(l); }
void candlist_sort_freq(CandidateList *l) { csx_candlist_sort_freq(l); }
void candlist_sort_r(CandidateList *l) { csx_candlist_sort_r(l); }
void candlist_sort_metric_desc_stable(CandidateList *l) { csx_candlist_sort_metric_desc(l); }
void candlist_sort_freq_stable(CandidateList *l) { csx_candlist_sort_freq(l); }
void candlist_sort_r_stable(CandidateList *l) { csx_candlist_sort_r(l); }
CandidateList remove_duplicates(const CandidateList *in, double dr_tol)
{
    return csx_remove_duplicates(in, dr_tol);
}
int harmonically_related(double r1, double r2, int numharm, double tol)
{
    return csx_harmonically_related(r1, r2, numharm, tol);
}
CandidateList remove_harmonics(const CandidateList *in, int numharm, double tol)
{
    return csx_remove_harmonics(in, numharm, tol);
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
