/*
 * In-memory .dat -> realfft -> rednoise staging (src/datpipe.c:52-448):
 * PRESTO's realfft (the N reals as N/2 complex, forward complex transform,
 * Numerical-Recipes unpack with double trigonometry and float storage, DC and
 * Nyquist packed into bin 0) followed by PRESTO's rednoise with startwidth 6,
 * endwidth 100, endfreq 6 Hz (running-median power normalisation with the
 * fast inverse square root and the lower-median quickselect).  The complex
 * transform is this library's double-precision FFT rounded to float where the
 * reference uses FFTW in single precision.
 */
#include "host_internal.h"

#ifndef M_PI
#define M_PI 3.141592653589793238462643383279502884
#endif

struct DatPipe {
    long long N;
    float *red;          /* N floats = N/2 complex, rednoise output */
    double *work;        /* N doubles, complex FFT scratch */
    float *packed;       /* N floats, realfft output */
    int startwidth, endwidth;
    double endfreq;
};

static float fast_invsqrtf(float x)
{
    union { float f; int32_t i; } t;
    t.f = x;
    t.i = 0x5f3759df - (t.i >> 1);
    const float y = t.f;
    return y * (1.5f - 0.5f * x * y * y);
}

/* Numerical Recipes quickselect; returns element (n-1)/2; reorders arr. */
static float nr_median(float *arr, int n)
{
    int low = 0, high = n - 1;
    const int median = (low + high) / 2;
    for (;;) {
        if (high <= low) return arr[median];
        if (high == low + 1) {
            if (arr[low] > arr[high]) { float t = arr[low]; arr[low] = arr[high]; arr[high] = t; }
            return arr[median];
        }
        const int middle = (low + high) / 2;
#define SWAPF(a, b) do { float t_ = (a); (a) = (b); (b) = t_; } while (0)
        if (arr[middle] > arr[high]) SWAPF(arr[middle], arr[high]);
        if (arr[low] > arr[high]) SWAPF(arr[low], arr[high]);
        if (arr[middle] > arr[low]) SWAPF(arr[middle], arr[low]);
        SWAPF(arr[middle], arr[low + 1]);
        int ll = low + 1, hh = high;
        for (;;) {
            do ll++; while (arr[low] > arr[ll]);
            do hh--; while (arr[hh] > arr[low]);
            if (hh < ll) break;
            SWAPF(arr[ll], arr[hh]);
        }
        SWAPF(arr[low], arr[hh]);
#undef SWAPF
        if (hh <= median) low = ll;
        if (hh >= median) high = hh - 1;
    }
}

static int realfft_forward(const float *x, long long N, double *work, float *d)
{
    const long long n2 = N / 2;
    for (long long i = 0; i < N; i++) work[i] = (double)x[i];
    if (cs_fft_any(work, n2, -1) != 0) return -1;
    for (long long i = 0; i < N; i++) d[i] = (float)work[i];
    const double cc = -0.5, theta = -2.0 * M_PI / (double)N;
    double wr = cos(theta), wi = sin(theta), tmp = sin(0.5 * theta);
    const double wpr = -2.0 * tmp * tmp, wpi = wi;
    for (long long il = 1, ih = n2 - 1; il <= (N >> 2); il++, ih--) {
        const double h1r = 0.5 * (d[2 * il] + d[2 * ih]);
        const double h1i = 0.5 * (d[2 * il + 1] - d[2 * ih + 1]);
        const double h2r = -cc * (d[2 * il + 1] + d[2 * ih + 1]);
        const double h2i = cc * (d[2 * il] - d[2 * ih]);
        const double a = h2r * wr, b = h2r * wi, e = h2i * wr, f = h2i * wi;
        d[2 * il] = (float)(h1r + a - f);
        d[2 * il + 1] = (float)(h1i + e + b);
        d[2 * ih] = (float)(h1r - a + f);
        d[2 * ih + 1] = (float)(-h1i + e + b);
        tmp = wr;
        wr = tmp * wpr - wi * wpi + wr;
        wi = wi * wpr + tmp * wpi + wi;
    }
    tmp = d[0];
    d[0] = (float)(tmp + d[1]);
    d[1] = (float)(tmp - d[1]);
    return 0;
}

static int rednoise(const float *in, float *out, long long nc, long long N, double dt,
                    int startwidth, int endwidth, double endfreq)
{
    float powbuf[1024], oldb[2048], newb[2048], outb[2048];
    if (endwidth > 1024) return -1;
    const float T = (float)((double)N * dt);
    long long binnum = 1, numwrote = 1, pos = 1;
    out[0] = 1.0f;
    out[1] = 0.0f;
    int bufflen = startwidth;
    if (nc - 1 < bufflen) {
        fprintf(stderr, "Error: rednoise: too few Fourier bins (%lld)\n", nc);
        return -1;
    }
    int nold = bufflen, nnew, mid_new, ii, ind;
    memcpy(oldb, in + 2 * pos, sizeof(float) * 2 * (size_t)nold);
    pos += nold;
    int mid_old = nold / 2;
    for (ii = 0; ii < nold; ii++)
        powbuf[ii] = oldb[2 * ii] * oldb[2 * ii] + oldb[2 * ii + 1] * oldb[2 * ii + 1];
    float mean_old = (float)(nr_median(powbuf, nold) / log(2.0)), mean_new;
    float dslope = 1.0f, norm = fast_invsqrtf(mean_old);
    for (ii = 0; ii < mid_old; ii++) {
        outb[2 * ii] = oldb[2 * ii] * norm;
        outb[2 * ii + 1] = oldb[2 * ii + 1] * norm;
    }
    memcpy(out + 2 * numwrote, outb, sizeof(float) * 2 * (size_t)mid_old);
    numwrote += mid_old;
    binnum += nold;
    bufflen = (int)(startwidth * log((double)binnum));
    for (;;) {
        const long long avail = nc - pos;
        if (avail <= 0) break;
        nnew = avail < bufflen ? (int)avail : bufflen;
        memcpy(newb, in + 2 * pos, sizeof(float) * 2 * (size_t)nnew);
        pos += nnew;
        mid_new = nnew / 2;
        for (ii = 0; ii < nnew; ii++)
            powbuf[ii] = newb[2 * ii] * newb[2 * ii] + newb[2 * ii + 1] * newb[2 * ii + 1];
        mean_new = (float)(nr_median(powbuf, nnew) / log(2.0));
        dslope = (float)((mean_new - mean_old) / (0.5 * (nold + nnew)));
        ii = 0;
        for (ind = mid_old; ind < nold; ii++, ind++) {
            norm = fast_invsqrtf(mean_old + dslope * (float)ii);
            outb[2 * ii] = oldb[2 * ind] * norm;
            outb[2 * ii + 1] = oldb[2 * ind + 1] * norm;
        }
        for (ind = 0; ind < mid_new; ii++, ind++) {
            norm = fast_invsqrtf(mean_old + dslope * (float)ii);
            outb[2 * ii] = newb[2 * ind] * norm;
            outb[2 * ii + 1] = newb[2 * ind + 1] * norm;
        }
        memcpy(out + 2 * numwrote, outb, sizeof(float) * 2 * (size_t)ii);
        numwrote += ii;
        binnum += nnew;
        if ((float)binnum / T < endfreq) bufflen = (int)(startwidth * log((double)binnum));
        else bufflen = endwidth;
        if (bufflen > endwidth) {
            fprintf(stderr, "Error: rednoise block length %d exceeds buffer %d\n", bufflen,
                    endwidth);
            return -1;
        }
        memcpy(oldb, newb, sizeof(float) * 2 * (size_t)nnew);
        nold = nnew;
        mean_old = mean_new;
        mid_old = mid_new;
    }
    ii = 0;
    for (ind = mid_old; ind < nold; ii++, ind++) {
        norm = fast_invsqrtf(mean_old + dslope * (float)ii);
        outb[2 * ii] = oldb[2 * ind] * norm;
        outb[2 * ii + 1] = oldb[2 * ind + 1] * norm;
    }
    memcpy(out + 2 * numwrote, outb, sizeof(float) * 2 * (size_t)ii);
    numwrote += ii;
    if (numwrote != nc) {
        fprintf(stderr, "Error: rednoise wrote %lld of %lld bins\n", numwrote, nc);
        return -1;
    }
    return 0;
}

DatPipe *datpipe_create(int startwidth, int endwidth, double endfreq)
{
    DatPipe *p = (DatPipe *)calloc(1, sizeof(DatPipe));
    if (!p) return NULL;
    p->startwidth = startwidth > 0 ? startwidth : 6;
    p->endwidth = endwidth > 0 ? endwidth : 100;
    p->endfreq = endfreq > 0.0 ? endfreq : 6.0;
    return p;
}

void datpipe_free(DatPipe *p)
{
    if (!p) return;
    free(p->red);
    free(p->work);
    free(p->packed);
    free(p);
}

static int prepare(DatPipe *p, long long N)
{
    if (p->N == N && p->red) return 0;
    free(p->red); free(p->work); free(p->packed);
    p->red = NULL; p->work = NULL; p->packed = NULL;
    p->N = 0;
    if (N <= 0 || (N & 1)) return -1;
    p->red = (float *)malloc((size_t)N * sizeof(float));
    p->work = (double *)malloc((size_t)N * sizeof(double));
    p->packed = (float *)malloc((size_t)N * sizeof(float));
    if (!p->red || !p->work || !p->packed) return -1;
    p->N = N;
    return 0;
}

static int stage(DatPipe *p, const float *samples, long long N, double dt, FFTFile *ft)
{
    if (realfft_forward(samples, N, p->work, p->packed) != 0) return -1;
    if (rednoise(p->packed, p->red, N / 2, N, dt, p->startwidth, p->endwidth, p->endfreq) != 0)
        return -1;
    ft->N = N;
    ft->T = (double)N * dt;
    ft->df = 1.0 / ft->T;
    ft->dereddened = 1;
    ft->detrended = 1;
    ft->amps = p->red;
    ft->namps = N / 2;
    ft->map = NULL;
    ft->mapsize = 0;
    ft->owns_amps = 0;
    ft->DC = ft->amps[0];
    ft->Nyquist = ft->amps[1];
    return 0;
}

int datpipe_redfft(DatPipe *p, const char *datpath, FFTFile *ft)
{
    memset(ft, 0, sizeof(*ft));
    ft->path = strdup(datpath);
    const size_t plen = strlen(datpath), keep = plen > 4 ? plen - 4 : 0;
    char *ip = (char *)malloc(keep + 5);
    if (!ft->path || !ip) { free(ip); return -1; }
    memcpy(ip, datpath, keep);
    memcpy(ip + keep, ".inf", 5);
    if (inf_read(ip, &ft->inf) != 0) {
        fprintf(stderr, "Error: The .inf file '%s' was not found.\n", ip);
        free(ip);
        return -1;
    }
    free(ip);
    if (!ft->inf.has_N || !ft->inf.has_dt) {
        fprintf(stderr, "Error: missing N or dt in %s\n", ft->inf.path);
        return -1;
    }
    const long long N = ft->inf.N;
    if (prepare(p, N) != 0) {
        fprintf(stderr, "Error: cannot prepare realfft/rednoise for N=%lld\n", N);
        return -1;
    }
    FILE *f = fopen(datpath, "rb");
    if (!f) {
        fprintf(stderr, "Error: cannot open '%s'\n", datpath);
        return -1;
    }
    float *x = (float *)malloc((size_t)N * sizeof(float));
    const size_t got = x ? fread(x, sizeof(float), (size_t)N, f) : 0;
    fclose(f);
    if (got != (size_t)N) {
        fprintf(stderr, "Error: '%s' holds %zu floats, expected %lld\n", datpath, got, N);
        free(x);
        return -1;
    }
    const int rc = stage(p, x, N, ft->inf.dt, ft);
    free(x);
    return rc;
}

int datpipe_redfft_mem(DatPipe *p, const float *samples, long long N, double dt,
                       const char *path, FFTFile *ft)
{
    memset(ft, 0, sizeof(*ft));
    ft->path = strdup(path ? path : "memory");
    if (!ft->path || prepare(p, N) != 0) return -1;
    if (stage(p, samples, N, dt, ft) != 0) return -1;
    ft->inf.has_N = 1;
    ft->inf.N = N;
    ft->inf.has_dt = 1;
    ft->inf.dt = dt;
    return 0;
}

int cs_open_input(const char *path, DatPipe *pipe, FFTFile *ft, int *is_dat)
{
    *is_dat = 0;
    const size_t n = strlen(path);
    if (n >= 4 && strcmp(path + n - 4, ".dat") == 0) {
        *is_dat = 1;
        if (!pipe) {
            fprintf(stderr, "Error: no .dat pipeline available\n");
            return -1;
        }
        return datpipe_redfft(pipe, path, ft);
    }
    if (n >= 4 && strcmp(path + n - 4, ".fft") == 0) return fftfile_open(ft, path);
    fprintf(stderr, "Error: input '%s' must end in .dat or .fft\n", path);
    return -1;
}
