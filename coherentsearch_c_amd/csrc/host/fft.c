/*
 * Host complex FFT in double precision, any length: depth-first mixed-radix
 * decimation in time (radices 4, 2, 3, 5, 7) with one twiddle table for the
 * whole transform, Bluestein (chirp-z over a power of two) for lengths with
 * larger prime factors.  The top recursion level and its combine step run
 * on OpenMP threads.  Used for the .dat realfft staging and the candidate
 * duty-cycle profiles (in place of FFTW in the reference).
 */
#include "host_internal.h"

#include <omp.h>

#ifndef M_PI
#define M_PI 3.141592653589793238462643383279502884
#endif

typedef struct { double re, im; } cpx;

static int small_radix(long long n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    if (n % 3 == 0) return 3;
    if (n % 5 == 0) return 5;
    if (n % 7 == 0) return 7;
    return 0;
}

static int smooth(long long n)
{
    while (n > 1) {
        const int r = small_radix(n);
        if (!r) return 0;
        n /= r;
    }
    return 1;
}

/* out[0..n) = DFT of in[0], in[is], ...; tw[j] = e^{sign 2 pi i j / N},
 * the transform of length n uses tw[j * ts] with ts = N / n. */
static void rec(const cpx *in, cpx *out, long long n, long long is, long long ts,
                const cpx *tw, int depth)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    const int R = small_radix(n);
    const long long m = n / R;
    const int par = depth == 0 && n >= (1LL << 16);
    #pragma omp parallel for if (par) schedule(static)
    for (int r = 0; r < R; r++)
        rec(in + (long long)r * is, out + (long long)r * m, m, is * R, ts * R, tw, depth + 1);
    /* e^{sign 2 pi i s/R} for the butterfly: tw index s * (N/R) = s*m*ts */
    cpx wr[7];
    for (int s = 0; s < R; s++) wr[s] = tw[(long long)s * m * ts];
    #pragma omp parallel for if (par) schedule(static)
    for (long long k = 0; k < m; k++) {
        cpx u[7];
        for (int r = 0; r < R; r++) {
            const cpx y = out[(long long)r * m + k];
            if (r == 0 || k == 0) { u[r] = y; continue; }
            const cpx w = tw[(long long)r * k * ts];
            u[r].re = y.re * w.re - y.im * w.im;
            u[r].im = y.re * w.im + y.im * w.re;
        }
        if (R == 2) {
            out[k].re = u[0].re + u[1].re; out[k].im = u[0].im + u[1].im;
            out[k + m].re = u[0].re - u[1].re; out[k + m].im = u[0].im - u[1].im;
        } else if (R == 4) {
            /* wr[1] = e^{sign i pi/2} = (0, sign) */
            const double sg = wr[1].im;
            const cpx t0 = {u[0].re + u[2].re, u[0].im + u[2].im};
            const cpx t1 = {u[0].re - u[2].re, u[0].im - u[2].im};
            const cpx t2 = {u[1].re + u[3].re, u[1].im + u[3].im};
            const cpx t3 = {u[1].re - u[3].re, u[1].im - u[3].im};
            out[k].re = t0.re + t2.re; out[k].im = t0.im + t2.im;
            out[k + 2 * m].re = t0.re - t2.re; out[k + 2 * m].im = t0.im - t2.im;
            out[k + m].re = t1.re - sg * t3.im; out[k + m].im = t1.im + sg * t3.re;
            out[k + 3 * m].re = t1.re + sg * t3.im; out[k + 3 * m].im = t1.im - sg * t3.re;
        } else {
            for (int s = 0; s < R; s++) {
                double ar = 0.0, ai = 0.0;
                for (int r = 0; r < R; r++) {
                    const cpx w = wr[(r * s) % R];
                    ar += u[r].re * w.re - u[r].im * w.im;
                    ai += u[r].re * w.im + u[r].im * w.re;
                }
                out[k + (long long)s * m].re = ar;
                out[k + (long long)s * m].im = ai;
            }
        }
    }
}

static int fft_smooth(cpx *z, long long n, int sign)
{
    cpx *tw = (cpx *)malloc((size_t)n * sizeof(cpx));
    cpx *tmp = (cpx *)malloc((size_t)n * sizeof(cpx));
    if (!tw || !tmp) { free(tw); free(tmp); return -1; }
    #pragma omp parallel for if (n >= (1LL << 16)) schedule(static)
    for (long long j = 0; j < n; j++) {
        const double a = sign * 2.0 * M_PI * (double)j / (double)n;
        tw[j].re = cos(a);
        tw[j].im = sin(a);
    }
    rec(z, tmp, n, 1, 1, tw, 0);
    memcpy(z, tmp, (size_t)n * sizeof(cpx));
    free(tw);
    free(tmp);
    return 0;
}

/* Bluestein: DFT_n via a cyclic convolution of length 2^k >= 2n-1. */
static int fft_bluestein(cpx *z, long long n, int sign)
{
    long long L = 1;
    while (L < 2 * n - 1) L <<= 1;
    cpx *w = (cpx *)malloc((size_t)n * sizeof(cpx));
    cpx *a = (cpx *)calloc((size_t)L, sizeof(cpx));
    cpx *b = (cpx *)calloc((size_t)L, sizeof(cpx));
    if (!w || !a || !b) { free(w); free(a); free(b); return -1; }
    for (long long k = 0; k < n; k++) {
        const long long kk = (long long)(((__int128)k * k) % (2 * n));
        const double ang = sign * M_PI * (double)kk / (double)n;
        w[k].re = cos(ang);
        w[k].im = sin(ang);
    }
    for (long long k = 0; k < n; k++) {
        a[k].re = z[k].re * w[k].re - z[k].im * w[k].im;
        a[k].im = z[k].re * w[k].im + z[k].im * w[k].re;
    }
    b[0].re = w[0].re; b[0].im = -w[0].im;
    for (long long k = 1; k < n; k++) {
        b[k].re = b[L - k].re = w[k].re;
        b[k].im = b[L - k].im = -w[k].im;
    }
    if (fft_smooth(a, L, -1) || fft_smooth(b, L, -1)) { free(w); free(a); free(b); return -1; }
    for (long long k = 0; k < L; k++) {
        const double re = a[k].re * b[k].re - a[k].im * b[k].im;
        const double im = a[k].re * b[k].im + a[k].im * b[k].re;
        a[k].re = re;
        a[k].im = im;
    }
    if (fft_smooth(a, L, +1)) { free(w); free(a); free(b); return -1; }
    for (long long k = 0; k < n; k++) {
        const double re = a[k].re / (double)L, im = a[k].im / (double)L;
        z[k].re = re * w[k].re - im * w[k].im;
        z[k].im = re * w[k].im + im * w[k].re;
    }
    free(w); free(a); free(b);
    return 0;
}

int cs_fft_any(double *z, long long n, int sign)
{
    if (n < 1) return -1;
    if (n == 1) return 0;
    if (smooth(n)) return fft_smooth((cpx *)z, n, sign);
    return fft_bluestein((cpx *)z, n, sign);
}
