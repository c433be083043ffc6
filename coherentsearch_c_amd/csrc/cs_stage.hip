/*
 * Device staging of a raw series: PRESTO's realfft followed by rednoise
 * (src/datpipe.c:52-253), so a .dat goes from host samples to a resident
 * spectrum without the host FFT.
 *
 *   1. samples (N float) -> N/2 complex doubles;
 *   2. forward complex FFT of N/2 points in double precision: a multi-level
 *      four-step transform, each level a batch of <= 512-point sub-FFTs done
 *      in LDS (Stockham stages of radix 8/7/5/4/3/2), the inter-level
 *      twiddles exp(-2 pi i (i*k mod LS)/LS) from sincospi of an exact integer
 *      phase, the last level writing in transposed (natural) order.  When
 *      N/2 has a prime factor above 7 the transform is Bluestein's chirp-z:
 *      x_j w_j (w_j = exp(-i pi j^2/n), j^2 mod 2n exact) zero-padded to a
 *      7-smooth M >= 2n - 1, convolved with the conjugate chirp by two
 *      forward M-point transforms of the same machinery (the chirp's own
 *      transform computed once per plan; the inverse as a forward transform
 *      of the conjugate), then scaled by w_k / M;
 *   3. Numerical-Recipes unpack (datpipe.c:104-137): the FFT rounded to float,
 *      the half-sum/half-difference in float, the twiddle products in double,
 *      stored as float; DC and Nyquist packed into bin 0.  The twiddle of bin
 *      il is cos/sin(-2 pi il/N) evaluated directly, where the reference walks
 *      a trigonometric recurrence;
 *   4. rednoise (datpipe.c:143-253): the block layout depends only on N, dt
 *      and the widths, so the host tabulates it; one wave per block takes the
 *      exact lower median of the block's powers (element (n-1)/2 of the sorted
 *      block, by ranking) and mean = (float)(median / ln 2); one workgroup per
 *      output segment applies norm = invsqrtf(mean_old + dslope*(float)ii) with
 *      PRESTO's bit-trick inverse square root, in the reference's float
 *      operation order.  The reference builds datpipe.c with
 *      -ffp-contract=fast (src/Makefile:129-133); the operations gcc fuses
 *      there are fused here explicitly: the powers fmaf(r, r, im*im), the
 *      level fmaf(dslope, ii, mean_old), the Newton step
 *      y*fmaf(-(0.5x*y), y, 1.5), the unpack's outputs.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cs_gpu.h"
#include "cs_stage.h"

namespace {

constexpr int kThreads = 256;
constexpr int kLmax = 512;        /* longest sub-FFT of one level */
constexpr int kCL = 2048;         /* columns x length held in LDS (32 KB of double2) */
constexpr int kEPT = kCL / kThreads;  /* elements each thread moves to and from HBM */
constexpr int kMaxLevels = 4;
constexpr int kMaxStages = 10;

struct LevelArgs {
    int L;                        /* this level's length */
    int C;                        /* sub-FFTs per workgroup, a power of two */
    int lc;                       /* log2 C */
    int ldr;                      /* LDS row stride: C, or C + 1 on the last level, whose
                                   * loads run down the columns */
    int nrad;
    int rad[kMaxStages];
    unsigned nsmag[kMaxStages];   /* ceil(2^32 / Ns) before each stage (0 while Ns = 1) */
    unsigned Lmag;                /* ceil(2^32 / L) (0 when L = 1) */
    unsigned nic;                 /* column groups per block, (S + C - 1) / C */
    long long S;                  /* product of the later levels (1 on the last) */
    long long B;                  /* product of the earlier levels */
    int twoff;                    /* offset of exp(-2 pi i t/L), t < L, in the table */
    /* last level only: digits of the transposed index r = k1 + L1*k2 + ... */
    int nd;
    int dL[kMaxLevels];
    long long dstride[kMaxLevels];
};

/* x / d for x < 2^21 and 2 <= d <= 2^10 from m = ceil(2^32 / d): x*m/2^32 exceeds
 * x/d by less than 2^-11, while the fractional part of x/d is at most 1 - 1/d */
__device__ __forceinline__ unsigned fdiv(unsigned x, unsigned m, unsigned d)
{
    return d == 1 ? x : __umulhi(x, m);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

/* forward R-point DFT of v in place; w(t) = exp(-2 pi i t/R) from the table */
template <int R>
__device__ __forceinline__ void dft_r(double2 (&v)[R], const double2 *__restrict__ tw, int tstep)
{
    if constexpr (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        const double2 t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = make_double2(t1.x + t3.y, t1.y - t3.x);   /* t1 - i t3 */
        v[3] = make_double2(t1.x - t3.y, t1.y + t3.x);   /* t1 + i t3 */
    } else if constexpr (R == 8) {
        /* two 4-point DFTs of the even and odd inputs, combined with
         * w8^k = exp(-2 pi i k/8) */
        double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        dft_r<4>(e, tw, tstep);
        dft_r<4>(o, tw, tstep);
        const double h = 0.70710678118654752440;
        const double2 o1 = make_double2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
        const double2 o2 = make_double2(o[2].y, -o[2].x);
        const double2 o3 = make_double2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    } else {
        double2 y[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            double2 acc = v[0];
#pragma unroll
            for (int j = 1; j < R; j++) acc = cadd(acc, cmul(v[j], tw[((j * k) % R) * tstep]));
            y[k] = acc;
        }
#pragma unroll
        for (int k = 0; k < R; k++) v[k] = y[k];
    }
}

/* one Stockham stage of radix R over the C columns in LDS (row stride ldr):
 * butterfly j of column c reads a = j + r*L/R, twiddles by
 * exp(-2 pi i r (j mod Ns)/(Ns R)) -- table index r*(j mod Ns)*L/(Ns R) < L --
 * and writes a = (j/Ns)*Ns*R + j mod Ns + r*Ns */
template <int R>
__device__ void stage(double2 *lds, const LevelArgs &a, int Ns, unsigned nsmag,
                      const double2 *__restrict__ tw)
{
    constexpr int BPT = (kCL / R + kThreads - 1) / kThreads;
    const int L = a.L, lc = a.lc, ldr = a.ldr;
    const int nb = (L / R) << lc;
    const int LR = L / R;
    const int tstride = L / (Ns * R);
    double2 v[BPT][R];
    int dst[BPT];
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        const int u = threadIdx.x + q * kThreads;
        dst[q] = -1;
        if (u < nb) {
            const int c = u & ((1 << lc) - 1), j = u >> lc;
            const int jq = (int)fdiv((unsigned)j, nsmag, (unsigned)Ns);
            const int k = j - jq * Ns;
#pragma unroll
            for (int r = 0; r < R; r++) {
                double2 x = lds[(j + r * LR) * ldr + c];
                if (r > 0 && k > 0) x = cmul(x, tw[r * k * tstride]);
                v[q][r] = x;
            }
            dft_r<R>(v[q], tw, L / R);
            dst[q] = (jq * Ns * R + k) * ldr + c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BPT; q++) {
        if (dst[q] >= 0) {
#pragma unroll
            for (int r = 0; r < R; r++) lds[dst[q] + r * Ns * ldr] = v[q][r];
        }
    }
    __syncthreads();
}

__device__ __forceinline__ double2 phase(long long m, long long n)
{
    /* exp(-2 pi i m/n) for 0 <= m < n, reduced to |m/n| <= 1/2 */
    if (2 * m > n) m -= n;
    double s, c;
    sincospi(-2.0 * (double)m / (double)n, &s, &c);
    return make_double2(c, s);
}

/* stages of one level in LDS; POW2 levels instantiate the radix-8/4/2
 * butterflies only (fewer registers, higher occupancy) */
template <bool POW2>
__device__ __forceinline__ void level_stages(double2 *lds, const LevelArgs &a,
                                             const double2 *__restrict__ tw)
{
    int Ns = 1;
    for (int s = 0; s < a.nrad; s++) {
        const int R = a.rad[s];
        const unsigned m = a.nsmag[s];
        if (R == 8) stage<8>(lds, a, Ns, m, tw);
        else if (R == 4) stage<4>(lds, a, Ns, m, tw);
        else if (R == 2) stage<2>(lds, a, Ns, m, tw);
        else if constexpr (!POW2) {
            if (R == 3) stage<3>(lds, a, Ns, m, tw);
            else if (R == 5) stage<5>(lds, a, Ns, m, tw);
            else stage<7>(lds, a, Ns, m, tw);
        }
        Ns *= R;
    }
}

/* element idx of the level's input: the complex sequence z, or on the first
 * level the float series read as (x[2 idx], x[2 idx + 1]) */
template <bool FLOATIN>
__device__ __forceinline__ double2 load_in(const void *__restrict__ in, long long idx)
{
    if constexpr (FLOATIN) {
        const float2 v = ((const float2 *)in)[idx];
        return make_double2((double)v.x, (double)v.y);
    } else {
        return ((const double2 *)in)[idx];
    }
}

/* One level of the multi-level transform.  Every thread issues all of its
 * kEPT loads before the first LDS store, so a wave keeps kEPT 16-byte loads
 * in flight per lane. */
template <bool POW2, bool FLOATIN, bool LAST>
__global__ __launch_bounds__(kThreads) void fft_level_kernel(const void *__restrict__ in,
                                                             double2 *__restrict__ out,
                                                             const double2 *__restrict__ twall,
                                                             LevelArgs a)
{
    extern __shared__ double2 lds[];          /* L * ldr elements */
    const int L = a.L, lc = a.lc, C = 1 << lc, ldr = a.ldr, n = L << lc;
    const double2 *tw = twall + a.twoff;
    const int tid = threadIdx.x;
    double2 r[kEPT];
    if constexpr (!LAST) {
        /* block b, columns i0 .. i0+C-1 of the S inner indices; element
         * e = ai*C + c is row ai of column c, at LDS e */
        const unsigned b = blockIdx.x / a.nic;
        const long long i0 = (long long)(blockIdx.x - b * a.nic) << lc;
        const long long S = a.S;
        const long long base = (long long)b * L * S;
        const int c = tid & (C - 1);
        const long long i = i0 + c;
        const bool col = i < S;
#pragma unroll
        for (int q = 0; q < kEPT; q++) {
            const int e = tid + q * kThreads;
            r[q] = make_double2(0.0, 0.0);
            if (e < n && col) r[q] = load_in<FLOATIN>(in, base + (long long)(e >> lc) * S + i);
        }
#pragma unroll
        for (int q = 0; q < kEPT; q++) {
            const int e = tid + q * kThreads;
            if (e < n) lds[e] = r[q];
        }
        __syncthreads();
        level_stages<POW2>(lds, a, tw);
        /* output row k of column i, twiddled by exp(-2 pi i i*k/(L S)); this
         * thread's rows are k0 + q*dk, so the twiddle advances by a fixed step */
        const long long LS = (long long)L * S;
        const int k0 = tid >> lc, dk = kThreads >> lc;
        double2 w = phase(i * k0 % LS, LS);
        const double2 step = phase(i * dk % LS, LS);
#pragma unroll
        for (int q = 0; q < kEPT; q++) {
            const int e = tid + q * kThreads;
            if (e < n && col) {
                const int k = e >> lc;
                double2 x = lds[e];
                if (i * k) x = cmul(x, w);
                out[base + (long long)k * S + i] = x;
            }
            w = cmul(w, step);
        }
    } else {
        /* transposed indices r0 .. r0+C-1; r's digits give the block b */
        const long long r0 = (long long)blockIdx.x << lc;
        __shared__ long long sb[16];
        if (tid < C) {
            long long rr = r0 + tid, b = 0;
            for (int d = 0; d < a.nd; d++) {
                b += (rr % a.dL[d]) * a.dstride[d];
                rr /= a.dL[d];
            }
            sb[tid] = (r0 + tid < a.B) ? b : -1;
        }
        __syncthreads();
        int dst[kEPT];
#pragma unroll
        for (int q = 0; q < kEPT; q++) {
            const int e = tid + q * kThreads;
            r[q] = make_double2(0.0, 0.0);
            dst[q] = -1;
            if (e < n) {
                const int c = (int)fdiv((unsigned)e, a.Lmag, (unsigned)L), ai = e - c * L;
                const long long b = sb[c];
                if (b >= 0) r[q] = load_in<FLOATIN>(in, b * L + ai);
                dst[q] = ai * ldr + c;
            }
        }
#pragma unroll
        for (int q = 0; q < kEPT; q++)
            if (dst[q] >= 0) lds[dst[q]] = r[q];
        __syncthreads();
        level_stages<POW2>(lds, a, tw);
#pragma unroll
        for (int q = 0; q < kEPT; q++) {
            const int e = tid + q * kThreads;
            const int k = e >> lc, c = e & (C - 1);
            if (e < n && r0 + c < a.B) out[r0 + c + a.B * k] = lds[k * ldr + c];
        }
    }
}

/* NR unpack of the N/2-point transform W into the packed real spectrum d */
__global__ void unpack_kernel(const double2 *__restrict__ W, float2 *__restrict__ d, long long N)
{
    const long long n2 = N / 2;
    const long long il = 1 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (il == 1) {
        const float d0r = (float)W[0].x, d0i = (float)W[0].y;
        const double t = (double)d0r;
        d[0] = make_float2((float)(t + (double)d0i), (float)(t - (double)d0i));
    }
    if (il > (N >> 2)) return;
    const long long ih = n2 - il;
    const float lr = (float)W[il].x, li = (float)W[il].y;
    const float hr = (float)W[ih].x, hi = (float)W[ih].y;
    const double cc = -0.5;
    const double h1r = 0.5 * (double)(lr + hr);
    const double h1i = 0.5 * (double)(li - hi);
    const double h2r = -cc * (double)(li + hi);
    const double h2i = cc * (double)(lr - hr);
    double wi, wr;
    sincospi(-2.0 * (double)il / (double)N, &wi, &wr);
    if (il != ih)
        d[il] = make_float2((float)__builtin_fma(-h2i, wi, __builtin_fma(h2r, wr, h1r)),
                            (float)__builtin_fma(h2r, wi, __builtin_fma(h2i, wr, h1i)));
    d[ih] = make_float2((float)__builtin_fma(h2i, wi, __builtin_fma(-h2r, wr, h1r)),
                        (float)__builtin_fma(h2r, wi, __builtin_fma(h2i, wr, -h1i)));
}

/* lower median of each block's powers, mean = (float)(median / ln 2) */
/* The rednoise block layout (datpipe.c:143-226): block 0 starts at bin 1;
 * while the running bin count over T stays below endfreq each block is
 * startwidth*ln(binnum) long (the ramp, tabulated on the host), after that
 * every block is endwidth long up to the last bin (computed here). */
struct BlockLayout {
    const long long *bstart;      /* ramp blocks */
    const int *blen;
    int nramp;                    /* tabulated blocks */
    int nblocks;                  /* all blocks */
    int endwidth;
    long long tail;               /* first bin of block nramp */
    long long nc;                 /* bins */
};

__device__ __forceinline__ void block_of(const BlockLayout &L, int b, long long &s, int &n)
{
    if (b < L.nramp) {
        s = L.bstart[b];
        n = L.blen[b];
    } else {
        s = L.tail + (long long)(b - L.nramp) * L.endwidth;
        const long long avail = L.nc - s;
        n = avail < L.endwidth ? (int)avail : L.endwidth;
    }
}

/* ascending bitonic sort of 128 keys held as (k0, k1) = keys lane, lane + 64 */
__device__ __forceinline__ void bitonic128(unsigned &k0, unsigned &k1, int lane)
{
#pragma unroll
    for (int k = 2; k <= 128; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j == 64) {
                const unsigned lo = min(k0, k1), hi = max(k0, k1);
                k0 = lo;
                k1 = hi;
            } else {
                const bool lower = (lane & j) == 0;
                const unsigned y0 = (unsigned)__shfl_xor((int)k0, j, 64);
                const unsigned y1 = (unsigned)__shfl_xor((int)k1, j, 64);
                /* element i = lane + 64 r sorts ascending within its k-run
                 * when bit k of i is clear */
                const bool up0 = (lane & k) == 0;
                const bool up1 = ((lane + 64) & k) == 0;
                k0 = (lower == up0) ? min(k0, y0) : max(k0, y0);
                k1 = (lower == up1) ? min(k1, y1) : max(k1, y1);
            }
        }
    }
}

/* Lower median of each block's powers, mean = (float)(median / ln 2): element
 * (n-1)/2 of the block's powers in ascending order.  Powers are sums of
 * squares, >= +0 (or NaN), so their bit patterns order as unsigned integers
 * (a NaN above every number, as a comparison-based rank places it).  Blocks of
 * up to 128 bins sort in one wave's registers; longer ones rank in LDS. */
__global__ __launch_bounds__(64) void block_mean_kernel(const float2 *__restrict__ d,
                                                        BlockLayout L, float *__restrict__ mean)
{
    __shared__ float p[CS_STAGE_MAXBLK];
    const int b = blockIdx.x;
    if (b >= L.nblocks) return;
    long long s;
    int n;
    block_of(L, b, s, n);
    const int lane = threadIdx.x;
    const int t = (n - 1) / 2;
    if (n <= 128) {
        unsigned k0 = 0xffffffffu, k1 = 0xffffffffu;
        if (lane < n) {
            const float2 v = d[s + lane];
            k0 = __float_as_uint(__builtin_fmaf(v.x, v.x, v.y * v.y));
        }
        if (lane + 64 < n) {
            const float2 v = d[s + lane + 64];
            k1 = __float_as_uint(__builtin_fmaf(v.x, v.x, v.y * v.y));
        }
        bitonic128(k0, k1, lane);
        if (lane == (t & 63)) {
            const float x = __uint_as_float(t < 64 ? k0 : k1);
            mean[b] = (float)((double)x / 0.69314718055994530942);
        }
        return;
    }
    for (int e = lane; e < n; e += 64) {
        const float2 v = d[s + e];
        p[e] = __builtin_fmaf(v.x, v.x, v.y * v.y);
    }
    __syncthreads();
    for (int e = lane; e < n; e += 64) {
        const float x = p[e];
        int lt = 0, eqb = 0;
        for (int f = 0; f < n; f++) {
            const float y = p[f];
            lt += y < x ? 1 : 0;
            eqb += (y == x && f < e) ? 1 : 0;
        }
        if (lt + eqb == t) mean[b] = (float)((double)x / 0.69314718055994530942);
    }
}

__device__ __forceinline__ float presto_invsqrtf(float x)
{
    const int i = 0x5f3759df - (__float_as_int(x) >> 1);
    const float y = __int_as_float(i);
    const float hx = 0.5f * x;
    const float hxy = hx * y;
    return y * __builtin_fmaf(-hxy, y, 1.5f);
}

/* Segment sg of nblocks + 1: segment 0 is the first half of block 0,
 * normalised by the constant invsqrtf(mean[0]); segment k (1 <= k < nblocks)
 * runs from the middle of block k-1 to the middle of block k; the last one
 * from the middle of the last block to its end.  Bins of segment k are
 * normalised by invsqrtf(mean[k-1] + slope*(float)ii), ii = bin - first bin,
 * slope = f32((mean[k] - mean[k-1]) / (0.5*(len[k-1] + len[k]))) (the last
 * segment reuses the last pair, or slope 1 with a single block). */
__global__ __launch_bounds__(128) void rednoise_kernel(const float2 *__restrict__ d,
                                                       float2 *__restrict__ out, BlockLayout L,
                                                       const float *__restrict__ mean)
{
    const int sg = blockIdx.x;
    const int nbk = L.nblocks;
    if (sg > nbk) return;
    if (sg == 0 && threadIdx.x == 0) out[0] = make_float2(1.0f, 0.0f);
    long long s0;
    int n, mo;
    bool first = false;
    float slope = 1.0f;
    if (sg == 0) {
        long long bs;
        int bl;
        block_of(L, 0, bs, bl);
        s0 = 1;
        n = bl / 2;
        mo = 0;
        first = true;
    } else {
        const int k = sg < nbk ? sg : nbk - 1;      /* the block whose middle ends the segment */
        long long sa = 0, sb = 0;
        int la = 0, lb = 0;
        if (sg < nbk) {
            block_of(L, k - 1, sa, la);
            block_of(L, k, sb, lb);
            s0 = sa + la / 2;
            n = la - la / 2 + lb / 2;
            mo = k - 1;
        } else {
            block_of(L, k, sb, lb);
            s0 = sb + lb / 2;
            n = lb - lb / 2;
            mo = k;
            if (nbk >= 2) block_of(L, k - 1, sa, la);
        }
        if (sg < nbk || nbk >= 2) {
            const float diff = mean[k] - mean[k - 1];
            slope = (float)((double)diff / (0.5 * (double)(la + lb)));
        }
    }
    const float m0 = mean[mo];
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        float norm;
        if (first) {
            norm = presto_invsqrtf(m0);
        } else {
            norm = presto_invsqrtf(__builtin_fmaf(slope, (float)e, m0));
        }
        const float2 v = d[s0 + e];
        out[s0 + e] = make_float2(v.x * norm, v.y * norm);
    }
}

/* the chirp exp(-i pi j^2 / n) from the exact phase j^2 mod 2n */
__device__ __forceinline__ double2 chirp(long long j, long long n)
{
    long long m = (long long)(((unsigned long long)j * (unsigned long long)j) %
                              (unsigned long long)(2 * n));
    if (m > n) m -= 2 * n;
    double s, c;
    sincospi(-(double)m / (double)n, &s, &c);
    return make_double2(c, s);
}

/* Bluestein input: the float series as n complex points times the chirp,
 * zero-padded to M */
__global__ void chirp_in_kernel(const float2 *__restrict__ x, double2 *__restrict__ z,
                                long long n, long long M)
{
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    double2 v = make_double2(0.0, 0.0);
    if (j < n) {
        const float2 a = x[j];
        v = cmul(make_double2((double)a.x, (double)a.y), chirp(j, n));
    }
    z[j] = v;
}

/* the convolution kernel b: conj(chirp) at j and M - j (|j| < n), zero between */
__global__ void chirp_b_kernel(double2 *__restrict__ b, long long n, long long M)
{
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    double2 v = make_double2(0.0, 0.0);
    long long a = -1;
    if (j < n) a = j;
    else if (j > M - n) a = M - j;
    if (a >= 0) {
        const double2 w = chirp(a, n);
        v = make_double2(w.x, -w.y);
    }
    b[j] = v;
}

/* z <- conj(z * bf): the product of the transforms, conjugated so that the
 * next forward transform computes the inverse (conjugated) */
__global__ void chirp_mul_kernel(double2 *__restrict__ z, const double2 *__restrict__ bf,
                                 long long M)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    const double2 p = cmul(z[k], bf[k]);
    z[k] = make_double2(p.x, -p.y);
}

/* X_k = conj(y_k) / M * chirp(k), k < n */
__global__ void chirp_out_kernel(const double2 *__restrict__ y, double2 *__restrict__ X,
                                 long long n, double invM)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double2 v = y[k];
    X[k] = cmul(make_double2(v.x * invM, -v.y * invM), chirp(k, n));
}

bool factor7(long long n, std::vector<int> &pf)
{
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) { pf.push_back(p); n /= p; }
    return n == 1;
}

/* split the length into levels of at most kLmax, as balanced as possible,
 * in ascending order */
bool plan_levels(long long n, std::vector<int> &Ls)
{
    std::vector<int> pf;
    if (!factor7(n, pf)) return false;
    std::sort(pf.begin(), pf.end(), [](int a, int b) { return a > b; });
    for (int m = 1; m <= kMaxLevels; m++) {
        std::vector<long long> prod(m, 1);
        bool ok = true;
        for (int p : pf) {
            int best = -1;
            for (int i = 0; i < m; i++)
                if (prod[i] * p <= kLmax && (best < 0 || prod[i] < prod[best])) best = i;
            if (best < 0) { ok = false; break; }
            prod[best] *= p;
        }
        if (ok) {
            /* shortest level first: the first level reads the float series,
             * and a short level holds the most columns per workgroup, so its
             * row segments are the widest */
            std::sort(prod.begin(), prod.end());
            Ls.assign(prod.begin(), prod.end());
            return true;
        }
    }
    return false;
}

/* the smallest 2^a 3^b 5^c 7^d >= x that plan_levels accepts (Bluestein's
 * convolution length), 0 if none fits the levels */
long long smooth7_at_least(long long x)
{
    long long best = 0;
    const long long lim = (long long)1 << 40;
    for (long long a = 1; a < lim; a *= 2)
        for (long long b = a; b < lim; b *= 3)
            for (long long c = b; c < lim; c *= 5)
                for (long long d = c; d < lim; d *= 7) {
                    if (d < x) continue;
                    if (!best || d < best) {
                        std::vector<int> Ls;
                        if (plan_levels(d, Ls)) best = d;
                    }
                    break;
                }
    return best;
}

void radices(int L, std::vector<int> &r)
{
    while (L % 8 == 0) { r.push_back(8); L /= 8; }
    while (L % 4 == 0) { r.push_back(4); L /= 4; }
    while (L % 2 == 0) { r.push_back(2); L /= 2; }
    for (int p : {3, 5, 7})
        while (L % p == 0) { r.push_back(p); L /= p; }
}

}  // namespace

struct CsStage {
    long long N, n2;
    long long fl;                 /* complex FFT length: n2, or Bluestein's M */
    bool blue;                    /* N/2 has a prime factor above 7: chirp-z */
    double2 *d_bf;                /* Bluestein: transform of the conjugate chirp, fl points */
    double dt;
    int device;
    int nlev;
    LevelArgs lev[kMaxLevels];
    long long grid[kMaxLevels];
    /* every buffer is a part of one device block; the tables (twiddles and
     * rednoise layout) are uploaded with one copy */
    void *d_block;
    double2 *d_z0, *d_z1;         /* ping-pong, n2 each */
    double2 *d_tw;
    float *d_x;                   /* samples */
    float2 *d_packed;             /* unpacked realfft output */
    long long *d_bstart;          /* ramp blocks of the rednoise layout */
    int *d_blen;
    float *d_mean;                /* per block */
    int nblocks, nramp, endwidth;
    long long tail;               /* first bin after the ramp */
    hipStream_t st;
    int own_stream;
    bool async_alloc;             /* stream-ordered allocations on st */
    hipGraph_t graph;             /* staging kernels captured for output g_out */
    hipGraphExec_t gexec;
    float2 *g_out;
};

static hipError_t salloc(CsStage *g, void **p, size_t bytes)
{
    return g->async_alloc ? hipMallocAsync(p, bytes, g->st) : hipMalloc(p, bytes);
}

namespace {

/* per-thread pinned staging for table uploads */
struct UploadBuf {
    void *host = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
    ~UploadBuf()
    {
        if (pending && ev) hipEventSynchronize(ev);
        if (ev) hipEventDestroy(ev);
        if (host) hipHostFree(host);
    }
};
thread_local UploadBuf tl_upload;

}  // namespace

__attribute__((visibility("hidden"))) void *cs_upload_buffer(size_t bytes)
{
    UploadBuf &u = tl_upload;
    if (u.pending) {
        hipEventSynchronize(u.ev);
        u.pending = false;
    }
    if (bytes > u.cap) {
        if (u.host) hipHostFree(u.host);
        u.host = nullptr;
        u.cap = 0;
        const size_t c = std::max(std::max<size_t>(bytes + bytes / 2, 2 * u.cap),
                                  (size_t)1 << 20);
        if (hipHostMalloc(&u.host, c, hipHostMallocDefault) != hipSuccess) {
            u.host = nullptr;
            return nullptr;
        }
        u.cap = c;
    }
    return u.host;
}

__attribute__((visibility("hidden"))) hipError_t cs_upload(void *dst, size_t bytes,
                                                           hipStream_t st)
{
    UploadBuf &u = tl_upload;
    if (bytes == 0) return hipSuccess;
    if (!u.host || bytes > u.cap) return hipErrorInvalidValue;
    hipError_t e = hipMemcpyAsync(dst, u.host, bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    if (!u.ev) {
        e = hipEventCreateWithFlags(&u.ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            hipStreamSynchronize(st);
            return e;
        }
    }
    e = hipEventRecord(u.ev, st);
    if (e != hipSuccess) {
        hipStreamSynchronize(st);
        return e;
    }
    u.pending = true;
    return hipSuccess;
}

static void sfree(CsStage *g, void *p)
{
    if (!p) return;
    if (g->async_alloc) hipFreeAsync(p, g->st);
    else hipFree(p);
}

extern "C" __attribute__((visibility("hidden"))) long long cs_stage_N(const CsStage *g)
{
    return g ? g->N : -1;
}

extern "C" __attribute__((visibility("hidden"))) void cs_stage_bind_stream(CsStage *g,
                                                                            hipStream_t st)
{
    if (g && st && !g->own_stream) g->st = st;
}

extern "C" CS_API void cs_gpu_stage_free(CsStage *g)
{
    if (!g) return;
    hipSetDevice(g->device);
    if (g->gexec) hipGraphExecDestroy(g->gexec);
    if (g->graph) hipGraphDestroy(g->graph);
    sfree(g, g->d_block);
    if (g->own_stream && g->st) hipStreamDestroy(g->st);
    delete g;
}

#define SCHK(x)                                                                    \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            fprintf(stderr, "Error: HIP %s failed at %s:%d: %s\n", #x, __FILE__,   \
                    __LINE__, hipGetErrorString(e_));                              \
            goto fail;                                                             \
        }                                                                          \
    } while (0)

static double2 *run_fft(CsStage *g, const void *in, bool floatin, double2 *a_buf,
                        double2 *b_buf, hipStream_t st);

/* device bytes of a stage plan's transform buffers for N samples */
extern "C" __attribute__((visibility("hidden"))) long long cs_stage_fft_bytes(long long N)
{
    std::vector<int> Ls;
    const long long n2 = N / 2;
    if (plan_levels(n2, Ls)) return 32 * n2;
    return 48 * smooth7_at_least(2 * n2 - 1);
}

extern "C" __attribute__((visibility("hidden"))) CsStage *
cs_stage_create_on(long long N, double dt, int startwidth, int endwidth, double endfreq,
                   int device, hipStream_t ext)
{
    if (startwidth <= 0) startwidth = 6;
    if (endwidth <= 0) endwidth = 100;
    if (endfreq <= 0.0) endfreq = 6.0;
    if (N <= 0 || (N & 1)) {
        fprintf(stderr, "Error: device staging needs an even number of samples (%lld)\n", N);
        return nullptr;
    }
    if (endwidth > CS_STAGE_MAXBLK) {
        fprintf(stderr, "Error: rednoise endwidth %d exceeds the device limit %d\n", endwidth,
                CS_STAGE_MAXBLK);
        return nullptr;
    }
    std::vector<int> Ls;
    const long long n2 = N / 2;
    long long fl = n2;
    bool blue = false;
    if (!plan_levels(n2, Ls)) {
        blue = true;
        fl = smooth7_at_least(2 * n2 - 1);
        Ls.clear();
        if (fl <= 0 || !plan_levels(fl, Ls)) {
            fprintf(stderr, "Error: device realfft of N = %lld exceeds the transform's %d^%d "
                            "points\n", N, kLmax, kMaxLevels);
            return nullptr;
        }
    }
    /* host tables, packed into the pinned staging buffer at the end */
    std::vector<double2> h_tw;
    std::vector<long long> h_bstart;
    std::vector<int> h_blen;
    CsStage *g = new CsStage();
    g->N = N;
    g->n2 = n2;
    g->fl = fl;
    g->blue = blue;
    g->d_bf = nullptr;
    g->dt = dt;
    g->device = device;
    g->nlev = (int)Ls.size();
    if (hipSetDevice(device) != hipSuccess) goto fail;
    if (ext) {
        g->st = ext;
        g->async_alloc = true;
    } else {
        SCHK(hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking));
        g->own_stream = 1;
        g->async_alloc = true;        /* stream-ordered on its own stream too */
    }
    {
        std::vector<double2> &tw = h_tw;
        long long B = 1;
        for (int l = 0; l < g->nlev; l++) {
            LevelArgs &a = g->lev[l];
            a.L = Ls[l];
            a.lc = 0;
            while ((2 << a.lc) <= std::min(16, kCL / a.L)) a.lc++;
            a.C = 1 << a.lc;
            a.ldr = l == g->nlev - 1 ? a.C + 1 : a.C;
            std::vector<int> r;
            radices(a.L, r);
            a.nrad = (int)r.size();
            int Ns = 1;
            for (int i = 0; i < a.nrad; i++) {
                a.rad[i] = r[i];
                a.nsmag[i] = Ns == 1 ? 0u : (unsigned)((((unsigned long long)1 << 32) + Ns - 1) / Ns);
                Ns *= r[i];
            }
            a.Lmag = a.L == 1 ? 0u : (unsigned)((((unsigned long long)1 << 32) + a.L - 1) / a.L);
            long long S = 1;
            for (int q = l + 1; q < g->nlev; q++) S *= Ls[q];
            a.S = S;
            a.B = B;
            a.nic = (unsigned)((S + a.C - 1) / a.C);
            a.twoff = (int)tw.size();
            for (int t = 0; t < a.L; t++) {
                const double ang = -2.0 * M_PI * (double)t / (double)a.L;
                tw.push_back(make_double2(cos(ang), sin(ang)));
            }
            if (l == g->nlev - 1) {
                /* r = k1 + L1 k2 + ... + (L1..L_{m-2}) k_{m-1};
                 * b = k1 (L2..L_{m-1}) + k2 (L3..L_{m-1}) + ... + k_{m-1} */
                a.nd = g->nlev - 1;
                for (int d = 0; d < a.nd; d++) {
                    a.dL[d] = Ls[d];
                    long long st = 1;
                    for (int q = d + 1; q < g->nlev - 1; q++) st *= Ls[q];
                    a.dstride[d] = st;
                }
                g->grid[l] = (B + a.C - 1) / a.C;
            } else {
                g->grid[l] = B * ((S + a.C - 1) / a.C);
            }
            B *= a.L;
        }
    }
    {
        /* rednoise block layout (datpipe.c:143-226): depends on N, dt, widths
         * only; the ramp is tabulated, the endwidth blocks after it are not */
        const long long nc = n2;
        const float T = (float)((double)N * dt);
        int bufflen = startwidth;
        if (nc - 1 < bufflen) {
            fprintf(stderr, "Error: rednoise: too few Fourier bins (%lld)\n", nc);
            goto fail;
        }
        long long pos = 1;
        h_bstart.push_back(pos);
        h_blen.push_back(bufflen);
        pos += bufflen;
        long long binnum = 1 + bufflen;
        bufflen = (int)(startwidth * log((double)binnum));
        bool tail = false;
        g->endwidth = endwidth;
        g->tail = nc;
        long long ntail = 0;
        for (;;) {
            const long long avail = nc - pos;
            if (avail <= 0) break;
            if (tail) {
                g->tail = pos;
                ntail = (avail + endwidth - 1) / endwidth;
                break;
            }
            const int nb = avail < bufflen ? (int)avail : bufflen;
            h_bstart.push_back(pos);
            h_blen.push_back(nb);
            pos += nb;
            binnum += nb;
            if ((float)binnum / T < endfreq) {
                bufflen = (int)(startwidth * log((double)binnum));
            } else {
                bufflen = endwidth;
                tail = true;
            }
            if (bufflen > endwidth) {
                fprintf(stderr, "Error: rednoise block length %d exceeds buffer %d\n", bufflen,
                        endwidth);
                goto fail;
            }
        }
        if ((long long)h_bstart.size() + ntail > 0x7fffffffLL) {
            fprintf(stderr, "Error: rednoise layout too large\n");
            goto fail;
        }
        g->nramp = (int)h_bstart.size();
        g->nblocks = (int)(g->nramp + ntail);
    }
    {
        /* one device block: work buffers, then the tables in upload order */
        const size_t nbk = (size_t)g->nblocks, nramp = (size_t)g->nramp;
        size_t off = 0;
        const size_t o_z0 = cs_part(&off, (size_t)fl * sizeof(double2));
        const size_t o_z1 = cs_part(&off, (size_t)fl * sizeof(double2));
        const size_t o_bf = cs_part(&off, blue ? (size_t)fl * sizeof(double2) : 0);
        const size_t o_x = cs_part(&off, (size_t)N * sizeof(float));
        const size_t o_packed = cs_part(&off, (size_t)n2 * sizeof(float2));
        const size_t o_mean = cs_part(&off, nbk * sizeof(float));
        const size_t o_tab = off;
        size_t toff = 0;
        const size_t t_tw = cs_part(&toff, h_tw.size() * sizeof(double2));
        const size_t t_bstart = cs_part(&toff, nramp * sizeof(long long));
        const size_t t_blen = cs_part(&toff, nramp * sizeof(int));
        SCHK(salloc(g, &g->d_block, o_tab + toff));
        char *b = (char *)g->d_block;
        g->d_z0 = (double2 *)(b + o_z0);
        g->d_z1 = (double2 *)(b + o_z1);
        g->d_bf = blue ? (double2 *)(b + o_bf) : nullptr;
        g->d_x = (float *)(b + o_x);
        g->d_packed = (float2 *)(b + o_packed);
        g->d_mean = (float *)(b + o_mean);
        char *tb = b + o_tab;
        g->d_tw = (double2 *)(tb + t_tw);
        g->d_bstart = (long long *)(tb + t_bstart);
        g->d_blen = (int *)(tb + t_blen);
        char *h = (char *)cs_upload_buffer(toff);
        if (!h) {
            fprintf(stderr, "Error: pinned staging buffer of %zu bytes failed\n", toff);
            goto fail;
        }
        memcpy(h + t_tw, h_tw.data(), h_tw.size() * sizeof(double2));
        memcpy(h + t_bstart, h_bstart.data(), nramp * sizeof(long long));
        memcpy(h + t_blen, h_blen.data(), nramp * sizeof(int));
        SCHK(cs_upload(tb, toff, g->st));
    }
    if (blue) {
        /* the conjugate chirp's transform, once per plan */
        const unsigned nb = (unsigned)((fl + 255) / 256);
        chirp_b_kernel<<<nb, 256, 0, g->st>>>(g->d_z0, n2, fl);
        SCHK(hipGetLastError());
        double2 *R = run_fft(g, g->d_z0, false, g->d_z1, g->d_z0, g->st);
        if (!R) goto fail;
        SCHK(hipMemcpyAsync(g->d_bf, R, (size_t)fl * sizeof(double2), hipMemcpyDeviceToDevice,
                            g->st));
    }
    return g;
fail:
    cs_gpu_stage_free(g);
    return nullptr;
}

extern "C" CS_API CsStage *cs_gpu_stage_create(long long N, double dt, int startwidth,
                                               int endwidth, double endfreq, int device)
{
    return cs_stage_create_on(N, dt, startwidth, endwidth, endfreq, device, nullptr);
}

/* The complex FFT levels of length fl: level 0 reads `in` (the float series
 * when floatin), level l writes a then b then a ...; returns the buffer
 * holding the transform, NULL on a launch error. */
static double2 *run_fft(CsStage *g, const void *in, bool floatin, double2 *a_buf,
                        double2 *b_buf, hipStream_t st)
{
    const void *src = in;
    double2 *dst = a_buf;
    for (int l = 0; l < g->nlev; l++) {
        const LevelArgs &a = g->lev[l];
        const bool pow2 = (a.L & (a.L - 1)) == 0, first = l == 0 && floatin,
                   last = l == g->nlev - 1;
        const size_t lds = (size_t)a.L * a.ldr * sizeof(double2);
        const dim3 grid((unsigned)g->grid[l]);
#define CS_LEVEL(P, F, Z) \
    fft_level_kernel<P, F, Z><<<grid, kThreads, lds, st>>>(src, dst, g->d_tw, a)
        if (pow2) {
            if (first) { if (last) CS_LEVEL(true, true, true); else CS_LEVEL(true, true, false); }
            else { if (last) CS_LEVEL(true, false, true); else CS_LEVEL(true, false, false); }
        } else {
            if (first) { if (last) CS_LEVEL(false, true, true); else CS_LEVEL(false, true, false); }
            else { if (last) CS_LEVEL(false, false, true); else CS_LEVEL(false, false, false); }
        }
#undef CS_LEVEL
        if (hipGetLastError() != hipSuccess) return nullptr;
        src = dst;
        dst = dst == a_buf ? b_buf : a_buf;
    }
    return (double2 *)src;
}

/* the staging kernels (samples already in d_x) */
static int stage_kernels(CsStage *g, float2 *d_out, hipStream_t st)
{
    {
        double2 *src = nullptr;
        if (!g->blue) {
            src = run_fft(g, g->d_x, true, g->d_z0, g->d_z1, st);
            if (!src) goto fail;
        } else {
            const unsigned nb = (unsigned)((g->fl + 255) / 256);
            chirp_in_kernel<<<nb, 256, 0, st>>>((const float2 *)g->d_x, g->d_z0, g->n2, g->fl);
            SCHK(hipGetLastError());
            double2 *R1 = run_fft(g, g->d_z0, false, g->d_z1, g->d_z0, st);
            if (!R1) goto fail;
            chirp_mul_kernel<<<nb, 256, 0, st>>>(R1, g->d_bf, g->fl);
            SCHK(hipGetLastError());
            double2 *o1 = R1 == g->d_z0 ? g->d_z1 : g->d_z0;
            double2 *R2 = run_fft(g, R1, false, o1, R1, st);
            if (!R2) goto fail;
            src = R2 == g->d_z0 ? g->d_z1 : g->d_z0;
            chirp_out_kernel<<<(unsigned)((g->n2 + 255) / 256), 256, 0, st>>>(
                R2, src, g->n2, 1.0 / (double)g->fl);
            SCHK(hipGetLastError());
        }
        const long long nq = g->N >> 2;
        const long long nthr = nq > 0 ? nq : 1;
        unpack_kernel<<<(unsigned)((nthr + 255) / 256), 256, 0, st>>>(src, g->d_packed, g->N);
        SCHK(hipGetLastError());
        BlockLayout L;
        L.bstart = g->d_bstart;
        L.blen = g->d_blen;
        L.nramp = g->nramp;
        L.nblocks = g->nblocks;
        L.endwidth = g->endwidth;
        L.tail = g->tail;
        L.nc = g->n2;
        block_mean_kernel<<<(unsigned)g->nblocks, 64, 0, st>>>(g->d_packed, L, g->d_mean);
        SCHK(hipGetLastError());
        rednoise_kernel<<<(unsigned)g->nblocks + 1, 128, 0, st>>>(g->d_packed, d_out, L,
                                                                  g->d_mean);
        SCHK(hipGetLastError());
    }
    return 0;
fail:
    return -1;
}

extern "C" __attribute__((visibility("hidden"))) int cs_stage_enqueue(CsStage *g, const float *samples,
                                                                       float2 *d_out, hipStream_t st)
{
    if (!st) st = g->st;
    SCHK(hipMemcpyAsync(g->d_x, samples, g->N * sizeof(float), hipMemcpyDefault, st));
    return stage_kernels(g, d_out, st);
fail:
    return -1;
}

/* samples in host memory or in this device's memory (a de-dispersed series
 * resident in HBM): the copy's direction follows the pointer */
extern "C" __attribute__((visibility("hidden"))) int cs_stage_copy_in(CsStage *g, const float *samples,
                                                                       hipStream_t st)
{
    if (!st) st = g->st;
    SCHK(hipMemcpyAsync(g->d_x, samples, g->N * sizeof(float), hipMemcpyDefault, st));
    return 0;
fail:
    return -1;
}

extern "C" __attribute__((visibility("hidden"))) int cs_stage_run(CsStage *g, float2 *d_out,
                                                                   hipStream_t st)
{
    return stage_kernels(g, d_out, st ? st : g->st);
}

/* The samples' upload, then the staging kernels replayed from a graph
 * captured once per output buffer (kernels only).  0 on success. */
extern "C" __attribute__((visibility("hidden"))) int cs_stage_enqueue_graph(CsStage *g,
                                                                             const float *samples,
                                                                             float2 *d_out,
                                                                             hipStream_t st)
{
    if (!st) st = g->st;
    SCHK(hipMemcpyAsync(g->d_x, samples, g->N * sizeof(float), hipMemcpyDefault, st));
    if (!g->gexec || g->g_out != d_out) {
        if (g->gexec) hipGraphExecDestroy(g->gexec);
        if (g->graph) hipGraphDestroy(g->graph);
        g->gexec = nullptr;
        g->graph = nullptr;
        SCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = stage_kernels(g, d_out, st);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &gr);
        if (rc != 0 || e != hipSuccess) {
            fprintf(stderr, "Error: staging graph capture failed\n");
            if (e == hipSuccess && gr) hipGraphDestroy(gr);
            return -1;
        }
        g->graph = gr;
        SCHK(hipGraphInstantiate(&g->gexec, g->graph, nullptr, nullptr, 0));
        g->g_out = d_out;
    }
    SCHK(hipGraphLaunch(g->gexec, st));
    return 0;
fail:
    return -1;
}

extern "C" CS_API int cs_gpu_stage_run_host(CsStage *g, const float *samples, float *amps,
                                            float *packed)
{
    if (!g || !samples || !amps) return -1;
    float2 *d_out = nullptr;
    SCHK(hipSetDevice(g->device));
    /* stream-ordered: no device-wide synchronisation, so staging overlaps
     * work on other streams (a search in progress) */
    SCHK(hipMallocAsync((void **)&d_out, g->n2 * sizeof(float2), g->st));
    if (cs_stage_enqueue(g, samples, d_out, g->st) != 0) goto fail;
    SCHK(hipMemcpyAsync(amps, d_out, g->n2 * sizeof(float2), hipMemcpyDeviceToHost, g->st));
    if (packed)
        SCHK(hipMemcpyAsync(packed, g->d_packed, g->n2 * sizeof(float2), hipMemcpyDeviceToHost,
                            g->st));
    SCHK(hipFreeAsync(d_out, g->st));
    d_out = nullptr;
    SCHK(hipStreamSynchronize(g->st));
    return 0;
fail:
    if (d_out) hipFreeAsync(d_out, g->st);
    hipStreamSynchronize(g->st);
    return -1;
}
