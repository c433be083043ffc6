/*
 * In-memory .dat -> realfft -> rednoise staging (src/datpipe.c:52-448):
 * PRESTO's realfft (the N reals as N/2 complex, forward complex transform,
 * Numerical-Recipes unpack with double trigonometry and float storage, DC and
 * Nyquist packed into bin 0) followed by PRESTO's rednoise with startwidth 6,
 * endwidth 100, endfreq 6 Hz (running-median power normalisation with the
 * fast inverse square root and the lower-median quickselect).  The complex
 * transform is this library's double-precision FFT rounded to float where the
 * reference uses FFTW in single precision.  With datpipe_set_device the same
 * staging runs on the GPU (cs_gpu_stage_*) for any even length.
 *
 * The reference compiles datpipe.c with -ffp-contract=fast
 * (src/Makefile:129-133); the products gcc fuses there are fused here with
 * explicit fma/fmaf (this file is built without contraction): the powers
 * fmaf(r, r, im*im), the rednoise level fmaf(dslope, ii, mean_old), the
 * inverse square root's Newton step y*fmaf(-(0.5x*y), y, 1.5), the unpack's
 * outputs and trigonometric recurrence.
 */
#include "host_internal.h"

#include <omp.h>

#ifndef M_PI
#define M_PI 3.141592653589793238462643383279502884
#endif

struct DatPipe {
    long long N;
    float *red;          /* N floats = N/2 complex, rednoise output */
    double *work;        /* N doubles, the complex transform */
    double *scratch;     /* N doubles, the long transform's second buffer */
    int threads;         /* OpenMP threads of the transform (0: the default) */
    float *packed;       /* N floats, realfft output */
    int startwidth, endwidth;
    double endfreq;
    int device;          /* GPU for the staging, -1 for the host transform */
    CsStage *gst;        /* device stage plan for (gN, gdt) */
    long long gN;
    double gdt;
    int red_pinned;      /* red came from cs_host_alloc */
    float *xin;          /* pinned sample buffer of the device staging, grow-only */
    long long xcap;
};

static float fast_invsqrtf(float x)
{
    union { float f; int32_t i; } t;
    t.f = x;
    t.i = 0x5f3759df - (t.i >> 1);
    const float y = t.f;
    return y * fmaf(-(0.5f * x * y), y, 1.5f);
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

static int realfft_forward(const float *x, long long N, double *work, double *scratch,
                           int threads, float *d)
{
    const long long n2 = N / 2;
    const int nt = threads > 0 ? threads : omp_get_max_threads();
    #pragma omp parallel for num_threads(nt) schedule(static)
    for (long long i = 0; i < N; i++) work[i] = (double)x[i];
    if (cs_fft_work(work, scratch, n2, -1, threads) != 0) return -1;
    #pragma omp parallel for num_threads(nt) schedule(static)
    for (long long i = 0; i < N; i++) d[i] = (float)work[i];
    const double cc = -0.5, theta = -2.0 * M_PI / (double)N;
    double wr = cos(theta), wi = sin(theta), tmp = sin(0.5 * theta);
    const double wpr = -2.0 * tmp * tmp, wpi = wi;
    for (long long il = 1, ih = n2 - 1; il <= (N >> 2); il++, ih--) {
        const double h1r = 0.5 * (d[2 * il] + d[2 * ih]);
        const double h1i = 0.5 * (d[2 * il + 1] - d[2 * ih + 1]);
        const double h2r = -cc * (d[2 * il + 1] + d[2 * ih + 1]);
        const double h2i = cc * (d[2 * il] - d[2 * ih]);
        d[2 * il] = (float)fma(-h2i, wi, fma(h2r, wr, h1r));
        d[2 * il + 1] = (float)fma(h2r, wi, fma(h2i, wr, h1i));
        d[2 * ih] = (float)fma(h2i, wi, fma(-h2r, wr, h1r));
        d[2 * ih + 1] = (float)fma(h2r, wi, fma(h2i, wr, -h1i));
        tmp = wr;
        wr = fma(tmp, wpr, -(wi * wpi)) + wr;
        wi = fma(wi, wpr, tmp * wpi) + wi;
    }
    tmp = d[0];
    d[0] = (float)(tmp + d[1]);
    d[1] = (float)(tmp - d[1]);
    return 0;
}

/* rednoise (src/datpipe.c:143-253) in three passes with the same arithmetic as
 * the reference's single sweep: the block layout (it depends only on N, dt and
 * the widths), every block's median power over threads, then every output
 * segment -- from the middle of one block to the middle of the next,
 * normalised by invsqrtf(mean_old + dslope*ii) -- over threads. */
static int rednoise(const float *in, float *out, long long nc, long long N, double dt,
                    int startwidth, int endwidth, double endfreq, int threads)
{
    if (endwidth > 1024) return -1;
    const float T = (float)((double)N * dt);
    int bufflen = startwidth;
    if (nc - 1 < bufflen) {
        fprintf(stderr, "Error: rednoise: too few Fourier bins (%lld)\n", nc);
        return -1;
    }
    long long cap = 1024, nb = 0;
    long long *bs = (long long *)malloc((size_t)cap * sizeof(long long));
    int *bl = (int *)malloc((size_t)cap * sizeof(int));
    float *mean = NULL;
    int rc = -1;
    if (!bs || !bl) goto done;
    {
        long long pos = 1, binnum = 1;
        int len = bufflen;                 /* block 0 is startwidth bins */
        for (;;) {
            if (nb == cap) {
                cap *= 2;
                long long *nbs = (long long *)realloc(bs, (size_t)cap * sizeof(long long));
                if (!nbs) goto done;
                bs = nbs;
                int *nbl = (int *)realloc(bl, (size_t)cap * sizeof(int));
                if (!nbl) goto done;
                bl = nbl;
            }
            bs[nb] = pos;
            bl[nb] = len;
            nb++;
            pos += len;
            binnum += len;
            if (nb == 1 || (float)binnum / T < endfreq)
                bufflen = (int)(startwidth * log((double)binnum));
            else
                bufflen = endwidth;
            if (nb > 1 && bufflen > endwidth) {
                fprintf(stderr, "Error: rednoise block length %d exceeds buffer %d\n", bufflen,
                        endwidth);
                goto done;
            }
            const long long avail = nc - pos;
            if (avail <= 0) break;
            len = avail < bufflen ? (int)avail : bufflen;
            if (len < 1 || len > 1024) {
                fprintf(stderr, "Error: rednoise block length %d out of range\n", len);
                goto done;
            }
        }
    }
    mean = (float *)malloc((size_t)nb * sizeof(float));
    if (!mean) goto done;
    const int nt = threads > 0 ? threads : omp_get_max_threads();
    #pragma omp parallel for num_threads(nt) schedule(static)
    for (long long b = 0; b < nb; b++) {
        float powbuf[1024];
        const float *v = in + 2 * bs[b];
        for (int ii = 0; ii < bl[b]; ii++)
            powbuf[ii] = fmaf(v[2 * ii], v[2 * ii], v[2 * ii + 1] * v[2 * ii + 1]);
        mean[b] = (float)(nr_median(powbuf, bl[b]) / log(2.0));
    }
    out[0] = 1.0f;
    out[1] = 0.0f;
    {
        const float norm = fast_invsqrtf(mean[0]);
        for (int ii = 0; ii < bl[0] / 2; ii++) {
            out[2 * (1 + ii)] = in[2 * (1 + ii)] * norm;
            out[2 * (1 + ii) + 1] = in[2 * (1 + ii) + 1] * norm;
        }
    }
    /* segment k runs from the middle of block k-1 to the middle of block k
     * (k < nb), or to the end of the last block (k = nb) */
    #pragma omp parallel for num_threads(nt) schedule(static)
    for (long long k = 1; k <= nb; k++) {
        const long long o = k - 1;
        float dslope = 1.0f;
        if (k < nb)
            dslope = (float)((mean[k] - mean[o]) / (0.5 * (bl[o] + bl[k])));
        else if (nb >= 2)
            dslope = (float)((mean[o] - mean[o - 1]) / (0.5 * (bl[o - 1] + bl[o])));
        const long long s0 = bs[o] + bl[o] / 2;
        const int n = bl[o] - bl[o] / 2 + (k < nb ? bl[k] / 2 : 0);
        for (int ii = 0; ii < n; ii++) {
            const float norm = fast_invsqrtf(fmaf(dslope, (float)ii, mean[o]));
            out[2 * (s0 + ii)] = in[2 * (s0 + ii)] * norm;
            out[2 * (s0 + ii) + 1] = in[2 * (s0 + ii) + 1] * norm;
        }
    }
    rc = 0;
done:
    free(bs);
    free(bl);
    free(mean);
    return rc;
}

DatPipe *datpipe_create(int startwidth, int endwidth, double endfreq)
{
    DatPipe *p = (DatPipe *)calloc(1, sizeof(DatPipe));
    if (!p) return NULL;
    p->startwidth = startwidth > 0 ? startwidth : 6;
    p->endwidth = endwidth > 0 ? endwidth : 100;
    p->endfreq = endfreq > 0.0 ? endfreq : 6.0;
    p->device = -1;
    return p;
}

void datpipe_set_device(DatPipe *p, int device)
{
    if (p) p->device = device >= 0 ? device : -1;
}

void datpipe_set_threads(DatPipe *p, int threads)
{
    if (p) p->threads = threads > 0 ? threads : 0;
}

void datpipe_free(DatPipe *p)
{
    if (!p) return;
    if (p->gst) cs_gpu_stage_free(p->gst);
    if (p->red_pinned) cs_host_free(p->red); else free(p->red);
    if (p->xin) cs_host_free(p->xin);
    free(p->work);
    free(p->scratch);
    free(p->packed);
    free(p);
}

/* the output buffer for N samples; the host transform's buffers are
 * allocated on first use (host_buffers) */
static int prepare(DatPipe *p, long long N)
{
    const int pin = p->device >= 0;
    if (p->N == N && p->red && p->red_pinned == pin) return 0;
    if (p->red_pinned) cs_host_free(p->red); else free(p->red);
    free(p->work); free(p->scratch); free(p->packed);
    p->red = NULL; p->work = NULL; p->scratch = NULL; p->packed = NULL;
    p->N = 0;
    if (N <= 0 || (N & 1)) return -1;
    /* the device staging copies the spectrum back into pinned memory: a
     * pageable destination would be copied through a bounce buffer */
    p->red = pin ? (float *)cs_host_alloc(N * (long long)sizeof(float))
                 : (float *)malloc((size_t)N * sizeof(float));
    if (!p->red) return -1;
    p->red_pinned = pin;
    p->N = N;
    return 0;
}

/* room for N samples: pinned (grow-only) for the device staging, else malloc */
static float *sample_buffer(DatPipe *p, long long N)
{
    if (p->device < 0) return (float *)malloc((size_t)N * sizeof(float));
    if (p->xcap < N) {
        if (p->xin) cs_host_free(p->xin);
        p->xin = (float *)cs_host_alloc(N * (long long)sizeof(float));
        p->xcap = p->xin ? N : 0;
    }
    return p->xin;
}

static void sample_release(DatPipe *p, float *x)
{
    if (x != p->xin) free(x);
}

static int host_buffers(DatPipe *p)
{
    const long long N = p->N;
    if (!p->work) p->work = (double *)malloc((size_t)N * sizeof(double));
    if (!p->scratch) p->scratch = (double *)malloc((size_t)N * sizeof(double));
    if (!p->packed) p->packed = (float *)malloc((size_t)N * sizeof(float));
    return p->work && p->scratch && p->packed ? 0 : -1;
}

/* realfft + rednoise on the GPU into p->red (cs_gpu_stage_*, the stage plan
 * kept while N and dt repeat; any even N, chirp-z on the device when N/2 has
 * a prime factor above 7): 0 done, -1 on a device error */
static int stage_device(DatPipe *p, const float *samples, long long N, double dt)
{
    if (!p->gst || p->gN != N || p->gdt != dt) {
        if (p->gst) cs_gpu_stage_free(p->gst);
        p->gst = cs_gpu_stage_create(N, dt, p->startwidth, p->endwidth, p->endfreq, p->device);
        p->gN = N;
        p->gdt = dt;
        if (!p->gst) {
            fprintf(stderr, "Error: no device stage plan for N=%lld on GPU %d\n", N, p->device);
            return -1;
        }
    }
    if (cs_gpu_stage_run_host(p->gst, samples, p->red, NULL) != 0) {
        fprintf(stderr, "Error: device staging failed for N=%lld on GPU %d\n", N, p->device);
        return -1;
    }
    return 0;
}

static int stage(DatPipe *p, const float *samples, long long N, double dt, FFTFile *ft)
{
    const int dev = p->device >= 0 ? stage_device(p, samples, N, dt) : 1;
    if (dev < 0) return -1;
    if (dev > 0) {
        if (host_buffers(p) != 0) return -1;
        if (realfft_forward(samples, N, p->work, p->scratch, p->threads, p->packed) != 0)
            return -1;
        if (rednoise(p->packed, p->red, N / 2, N, dt, p->startwidth, p->endwidth, p->endfreq,
                     p->threads) != 0)
            return -1;
    }
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
    float *x = sample_buffer(p, N);
    const size_t got = x ? fread(x, sizeof(float), (size_t)N, f) : 0;
    fclose(f);
    if (got != (size_t)N) {
        fprintf(stderr, "Error: '%s' holds %zu floats, expected %lld\n", datpath, got, N);
        sample_release(p, x);
        return -1;
    }
    const int rc = stage(p, x, N, ft->inf.dt, ft);
    sample_release(p, x);
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
