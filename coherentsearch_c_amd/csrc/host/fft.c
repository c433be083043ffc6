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

/* lengths from which the multi-level path is used */
#define LV_MIN 16384

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

/* ---- long transforms: multi-level, cache-blocked ----------------------
 *
 * n = L_0 * L_1 * ... * L_{m-1}, each level at most LV_LMAX points, split as
 * evenly as the prime factors allow.  Level l (B = L_0..L_{l-1} earlier,
 * S = L_{l+1}..L_{m-1} later) is a batch of L_l-point DFTs:
 *   - every level but the last transforms, in place, the columns
 *     z[b*L*S + a*S + i] (a < L) of each block b, and multiplies output row k
 *     of column i by e^{sign 2 pi i i*k/(L*S)};
 *   - the last level takes the contiguous runs z[b*L + a] of the blocks b
 *     whose digits reverse the output index r, and writes out[r + B*k].
 * Work items hold LV_C columns at once in split real/imaginary buffers, so
 * every butterfly is a vector loop over the columns.  The output is in
 * natural order. */
#define LV_LMAX 512
#define LV_C 16
#define LV_MAXLEV 6
#define LV_MAXRAD 12
#define LV_RESET 32

typedef struct {
    int L, nrad, rad[LV_MAXRAD];
    long long S, B;
    int nd, dL[LV_MAXLEV];
    long long dstride[LV_MAXLEV];
    double *twr, *twi;            /* e^{sign 2 pi i t/L}, t < L */
} Lev;

static int lv_plan(long long n, int *Ls)
{
    int pf[64], np = 0;
    long long m = n;
    for (int p = 7; p >= 2; p--) {
        if (p == 4 || p == 6) continue;
        while (m % p == 0 && np < 64) { pf[np++] = p; m /= p; }
    }
    if (m != 1) return 0;
    for (int nl = 1; nl <= LV_MAXLEV; nl++) {
        long long prod[LV_MAXLEV];
        int ok = 1;
        for (int i = 0; i < nl; i++) prod[i] = 1;
        for (int f = 0; f < np && ok; f++) {
            int best = -1;
            for (int i = 0; i < nl; i++)
                if (prod[i] * pf[f] <= LV_LMAX && (best < 0 || prod[i] < prod[best])) best = i;
            if (best < 0) ok = 0;
            else prod[best] *= pf[f];
        }
        if (ok) {
            for (int i = 0; i < nl; i++) Ls[i] = (int)prod[i];
            return nl;
        }
    }
    return 0;
}

static int lv_radices(int L, int *r)
{
    int n = 0;
    while (L % 8 == 0) { r[n++] = 8; L /= 8; }
    while (L % 4 == 0) { r[n++] = 4; L /= 4; }
    while (L % 2 == 0) { r[n++] = 2; L /= 2; }
    for (int p = 3; p <= 7; p += 2)
        while (L % p == 0) { r[n++] = p; L /= p; }
    return n;
}

/* one Stockham stage of radix R, Ns: rows j + r*L/R of x (twiddled by
 * e^{sign 2 pi i r (j mod Ns)/(Ns R)}) -> rows (j/Ns)*Ns*R + j mod Ns + r*Ns
 * of y, each row LV_C columns */
static void lv_stage(const Lev *a, int R, int Ns, int sign, const double *restrict xr,
                     const double *restrict xi, double *restrict yr, double *restrict yi)
{
    const int L = a->L, LR = L / R, tstride = L / (Ns * R);
    double vr[8][LV_C], vi[8][LV_C];
    double cr[7][7], ci[7][7];
    if (R == 3 || R == 5 || R == 7)
        for (int j = 0; j < R; j++)
            for (int k = 0; k < R; k++) {
                const double ang = 2.0 * M_PI * (double)((j * k) % R) / (double)R;
                cr[j][k] = cos(ang);
                ci[j][k] = sign * sin(ang);
            }
    const double h = 0.70710678118654752440, sg = (double)sign;
    for (int j = 0; j < LR; j++) {
        const int k = j % Ns, jq = j / Ns;
        for (int r = 0; r < R; r++) {
            const double *pr = xr + (size_t)(j + r * LR) * LV_C, *pi = xi + (size_t)(j + r * LR) * LV_C;
            if (r == 0 || k == 0) {
                for (int c = 0; c < LV_C; c++) { vr[r][c] = pr[c]; vi[r][c] = pi[c]; }
            } else {
                const int t = r * k * tstride;
                const double wr = a->twr[t], wi = a->twi[t];
                for (int c = 0; c < LV_C; c++) {
                    vr[r][c] = pr[c] * wr - pi[c] * wi;
                    vi[r][c] = pr[c] * wi + pi[c] * wr;
                }
            }
        }
        double *o_r[8], *o_i[8];
        for (int r = 0; r < R; r++) {
            const size_t row = (size_t)jq * Ns * R + k + (size_t)r * Ns;
            o_r[r] = yr + row * LV_C;
            o_i[r] = yi + row * LV_C;
        }
        if (R == 2) {
            for (int c = 0; c < LV_C; c++) {
                o_r[0][c] = vr[0][c] + vr[1][c]; o_i[0][c] = vi[0][c] + vi[1][c];
                o_r[1][c] = vr[0][c] - vr[1][c]; o_i[1][c] = vi[0][c] - vi[1][c];
            }
        } else if (R == 4) {
            for (int c = 0; c < LV_C; c++) {
                const double t0r = vr[0][c] + vr[2][c], t0i = vi[0][c] + vi[2][c];
                const double t1r = vr[0][c] - vr[2][c], t1i = vi[0][c] - vi[2][c];
                const double t2r = vr[1][c] + vr[3][c], t2i = vi[1][c] + vi[3][c];
                const double t3r = vr[1][c] - vr[3][c], t3i = vi[1][c] - vi[3][c];
                o_r[0][c] = t0r + t2r; o_i[0][c] = t0i + t2i;
                o_r[2][c] = t0r - t2r; o_i[2][c] = t0i - t2i;
                /* t1 + sign*i*t3, t1 - sign*i*t3 */
                o_r[1][c] = t1r - sg * t3i; o_i[1][c] = t1i + sg * t3r;
                o_r[3][c] = t1r + sg * t3i; o_i[3][c] = t1i - sg * t3r;
            }
        } else if (R == 8) {
            for (int c = 0; c < LV_C; c++) {
                /* 4-point DFTs of the even and odd inputs */
                const double a0r = vr[0][c] + vr[4][c], a0i = vi[0][c] + vi[4][c];
                const double a1r = vr[0][c] - vr[4][c], a1i = vi[0][c] - vi[4][c];
                const double a2r = vr[2][c] + vr[6][c], a2i = vi[2][c] + vi[6][c];
                const double a3r = vr[2][c] - vr[6][c], a3i = vi[2][c] - vi[6][c];
                const double e0r = a0r + a2r, e0i = a0i + a2i, e2r = a0r - a2r, e2i = a0i - a2i;
                const double e1r = a1r - sg * a3i, e1i = a1i + sg * a3r;
                const double e3r = a1r + sg * a3i, e3i = a1i - sg * a3r;
                const double b0r = vr[1][c] + vr[5][c], b0i = vi[1][c] + vi[5][c];
                const double b1r = vr[1][c] - vr[5][c], b1i = vi[1][c] - vi[5][c];
                const double b2r = vr[3][c] + vr[7][c], b2i = vi[3][c] + vi[7][c];
                const double b3r = vr[3][c] - vr[7][c], b3i = vi[3][c] - vi[7][c];
                const double f0r = b0r + b2r, f0i = b0i + b2i, f2r = b0r - b2r, f2i = b0i - b2i;
                const double f1r = b1r - sg * b3i, f1i = b1i + sg * b3r;
                const double f3r = b1r + sg * b3i, f3i = b1i - sg * b3r;
                /* w8^k f_k, w8 = e^{sign 2 pi i/8} = h + sign*h*i */
                const double g1r = h * (f1r - sg * f1i), g1i = h * (f1i + sg * f1r);
                const double g2r = -sg * f2i, g2i = sg * f2r;
                const double g3r = -h * (f3r + sg * f3i), g3i = h * (sg * f3r - f3i);
                o_r[0][c] = e0r + f0r; o_i[0][c] = e0i + f0i;
                o_r[4][c] = e0r - f0r; o_i[4][c] = e0i - f0i;
                o_r[1][c] = e1r + g1r; o_i[1][c] = e1i + g1i;
                o_r[5][c] = e1r - g1r; o_i[5][c] = e1i - g1i;
                o_r[2][c] = e2r + g2r; o_i[2][c] = e2i + g2i;
                o_r[6][c] = e2r - g2r; o_i[6][c] = e2i - g2i;
                o_r[3][c] = e3r + g3r; o_i[3][c] = e3i + g3i;
                o_r[7][c] = e3r - g3r; o_i[7][c] = e3i - g3i;
            }
        } else {
            for (int q = 0; q < R; q++) {
                double ar[LV_C], ai[LV_C];
                for (int c = 0; c < LV_C; c++) { ar[c] = vr[0][c]; ai[c] = vi[0][c]; }
                for (int r = 1; r < R; r++) {
                    const double wr = cr[r][q], wi = ci[r][q];
                    for (int c = 0; c < LV_C; c++) {
                        ar[c] += vr[r][c] * wr - vi[r][c] * wi;
                        ai[c] += vr[r][c] * wi + vi[r][c] * wr;
                    }
                }
                for (int c = 0; c < LV_C; c++) { o_r[q][c] = ar[c]; o_i[q][c] = ai[c]; }
            }
        }
    }
}

/* the level's stages on buffer x (L rows of LV_C columns), y as the other half
 * of the ping-pong; returns which one holds the result (0: x, 1: y) */
static int lv_stages(const Lev *a, int sign, double *xr, double *xi, double *yr, double *yi)
{
    int Ns = 1, in_y = 0;
    for (int s = 0; s < a->nrad; s++) {
        if (!in_y) lv_stage(a, a->rad[s], Ns, sign, xr, xi, yr, yi);
        else lv_stage(a, a->rad[s], Ns, sign, yr, yi, xr, xi);
        in_y ^= 1;
        Ns *= a->rad[s];
    }
    return in_y;
}

/* columns i0 .. i0+LV_C-1 of block b of a level other than the last, in place */
static void lv_columns(const Lev *a, cpx *z, long long b, long long i0, int sign, double *buf)
{
    const int L = a->L;
    const long long S = a->S, LS = (long long)L * S;
    const int ce = (int)(S - i0 < LV_C ? S - i0 : LV_C);
    double *xr = buf, *xi = xr + L * LV_C, *yr = xi + L * LV_C, *yi = yr + L * LV_C;
    cpx *base = z + b * LS + i0;
    for (int r = 0; r < L; r++) {
        const cpx *row = base + (long long)r * S;
        int c = 0;
        for (; c < ce; c++) { xr[r * LV_C + c] = row[c].re; xi[r * LV_C + c] = row[c].im; }
        for (; c < LV_C; c++) { xr[r * LV_C + c] = 0.0; xi[r * LV_C + c] = 0.0; }
    }
    const int in_y = lv_stages(a, sign, xr, xi, yr, yi);
    const double *rr = in_y ? yr : xr, *ri = in_y ? yi : xi;
    /* row k of column i times e^{sign 2 pi i i*k/LS}: a step recurrence per
     * column, restarted from the exact value every LV_RESET rows */
    double wr[LV_C], wi[LV_C], sr[LV_C], si[LV_C];
    for (int c = 0; c < ce; c++) {
        const double ang = 2.0 * M_PI * (double)((i0 + c) % LS) / (double)LS;
        sr[c] = cos(ang);
        si[c] = sign * sin(ang);
    }
    for (int k = 0; k < L; k++) {
        if (k % LV_RESET == 0)
            for (int c = 0; c < ce; c++) {
                const long long m = ((i0 + c) * (long long)k) % LS;
                const double ang = 2.0 * M_PI * (double)m / (double)LS;
                wr[c] = cos(ang);
                wi[c] = sign * sin(ang);
            }
        cpx *row = base + (long long)k * S;
        for (int c = 0; c < ce; c++) {
            const double xr_ = rr[k * LV_C + c], xi_ = ri[k * LV_C + c];
            row[c].re = xr_ * wr[c] - xi_ * wi[c];
            row[c].im = xr_ * wi[c] + xi_ * wr[c];
            const double t = wr[c] * sr[c] - wi[c] * si[c];
            wi[c] = wr[c] * si[c] + wi[c] * sr[c];
            wr[c] = t;
        }
    }
}

/* output columns r0 .. r0+LV_C-1 of the last level, z -> out */
static void lv_last(const Lev *a, const cpx *z, cpx *out, long long r0, int sign, double *buf)
{
    const int L = a->L;
    const long long B = a->B;
    const int ce = (int)(B - r0 < LV_C ? B - r0 : LV_C);
    double *xr = buf, *xi = xr + L * LV_C, *yr = xi + L * LV_C, *yi = yr + L * LV_C;
    for (int c = 0; c < LV_C; c++) {
        if (c >= ce) {
            for (int r = 0; r < L; r++) { xr[r * LV_C + c] = 0.0; xi[r * LV_C + c] = 0.0; }
            continue;
        }
        long long rr = r0 + c, b = 0;
        for (int d = 0; d < a->nd; d++) {
            b += (rr % a->dL[d]) * a->dstride[d];
            rr /= a->dL[d];
        }
        const cpx *src = z + b * L;
        for (int r = 0; r < L; r++) { xr[r * LV_C + c] = src[r].re; xi[r * LV_C + c] = src[r].im; }
    }
    const int in_y = lv_stages(a, sign, xr, xi, yr, yi);
    const double *rr = in_y ? yr : xr, *ri = in_y ? yi : xi;
    for (int k = 0; k < L; k++) {
        cpx *row = out + r0 + B * k;
        for (int c = 0; c < ce; c++) { row[c].re = rr[k * LV_C + c]; row[c].im = ri[k * LV_C + c]; }
    }
}

/* z (n points) -> z, with scratch of n points; -2 when n has a prime factor
 * above 7 or needs more than LV_MAXLEV levels */
static int fft_levels(cpx *z, cpx *scratch, long long n, int sign, int nthreads)
{
    int Ls[LV_MAXLEV];
    const int nl = lv_plan(n, Ls);
    if (nl == 0) return -2;
    Lev lev[LV_MAXLEV];
    memset(lev, 0, sizeof(lev));
    long long B = 1;
    int rc = 0;
    for (int l = 0; l < nl; l++) {
        Lev *a = &lev[l];
        a->L = Ls[l];
        a->nrad = lv_radices(a->L, a->rad);
        long long S = 1;
        for (int q = l + 1; q < nl; q++) S *= Ls[q];
        a->S = S;
        a->B = B;
        a->nd = nl - 1;
        for (int d = 0; d < nl - 1; d++) {
            a->dL[d] = Ls[d];
            long long st = 1;
            for (int q = d + 1; q < nl - 1; q++) st *= Ls[q];
            a->dstride[d] = st;
        }
        a->twr = (double *)malloc((size_t)a->L * sizeof(double));
        a->twi = (double *)malloc((size_t)a->L * sizeof(double));
        if (!a->twr || !a->twi) { rc = -1; goto done; }
        for (int t = 0; t < a->L; t++) {
            const double ang = 2.0 * M_PI * (double)t / (double)a->L;
            a->twr[t] = cos(ang);
            a->twi[t] = sign * sin(ang);
        }
        B *= a->L;
    }
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    for (int l = 0; l < nl; l++) {
        const Lev *a = &lev[l];
        const size_t bufn = (size_t)4 * a->L * LV_C;
        const int last = l == nl - 1;
        const long long nic = (a->S + LV_C - 1) / LV_C;
        const long long items = last ? (a->B + LV_C - 1) / LV_C : a->B * nic;
        int fail = 0;
        #pragma omp parallel num_threads(nt)
        {
            double *buf = (double *)malloc(bufn * sizeof(double));
            if (!buf) {
                #pragma omp atomic write
                fail = 1;
            }
            #pragma omp for schedule(static)
            for (long long it = 0; it < items; it++) {
                if (!buf) continue;
                if (last) lv_last(a, z, scratch, it * LV_C, sign, buf);
                else lv_columns(a, z, it / nic, (it % nic) * LV_C, sign, buf);
            }
            free(buf);
        }
        if (fail) { rc = -1; goto done; }
    }
    #pragma omp parallel for num_threads(nt) schedule(static)
    for (long long i = 0; i < n; i++) z[i] = scratch[i];
done:
    for (int l = 0; l < nl; l++) { free(lev[l].twr); free(lev[l].twi); }
    return rc;
}

int cs_fft_work(double *z, double *scratch, long long n, int sign, int nthreads)
{
    if (n < 1) return -1;
    if (n == 1) return 0;
    if (n >= LV_MIN && scratch) {
        const int rc = fft_levels((cpx *)z, (cpx *)scratch, n, sign, nthreads);
        if (rc != -2) return rc;
    }
    if (smooth(n)) return fft_smooth((cpx *)z, n, sign);
    return fft_bluestein((cpx *)z, n, sign);
}

int cs_fft_any(double *z, long long n, int sign)
{
    if (n < 1) return -1;
    if (n == 1) return 0;
    if (smooth(n)) return fft_smooth((cpx *)z, n, sign);
    return fft_bluestein((cpx *)z, n, sign);
}
