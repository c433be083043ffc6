/*
 * Fused search kernel for gfx950, specialised on (nharms, m, max decimation).
 *
 * One workgroup = CS_TILE consecutive trials of one chunk, one trial per
 * lane (wave64).  Per workgroup:
 *   1. every filled harmonic's spectrum window for the tile is copied from
 *      HBM into LDS with coalesced loads (batches if the windows exceed LDS);
 *   2. each lane interpolates its trial at every harmonic h*r: m taps of
 *      amps from LDS times the phase's f32 weights (an L2-resident table
 *      indexed by the exact residue), explicit fmaf in tap order, then the
 *      unfused rotation -- the reference's arithmetic (src/interp.c:686-811,
 *      src/cuda.cu:121-135), so the stack is bit-identical to the CPU path;
 *   3. the harmonic stack stays in registers and is turned into profiles by
 *      a compile-time unrolled mixed-radix c2r (the project c2r, identical op
 *      sequence to the oracle's), for the base and every non-divisor
 *      decimation; divisor decimations fold the base profile
 *      (src/search.c:294-331);
 *   4. the boxcar matched filter runs on register-resident prefix sums over
 *      the enabled widths (src/boxcar.c:203-252) and above-threshold trials
 *      are appended with one atomic per wave.
 * Nothing but the spectrum, the tables and the records touches HBM.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "cs_device.h"
#include "cs_twiddles.h"

namespace csk {

template <typename T>
struct cx {
    T x, y;
};
using cf = cx<float>;

template <int... Is, class F>
__device__ __forceinline__ void sfor_(std::integer_sequence<int, Is...>, F &&f)
{
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void sfor(F &&f)
{
    sfor_(std::make_integer_sequence<int, N>{}, f);
}

/* smallest prime factor; the power of it dividing N (the prime-factor split's
 * first factor); the Cooley-Tukey radix of a prime power */
constexpr int spf(int N)
{
    if (N % 2 == 0) return 2;
    for (int p = 3; p * p <= N; p += 2)
        if (N % p == 0) return p;
    return N;
}
constexpr int pfa_n1(int N)
{
    int n1 = 1;
    while (N % (n1 * spf(N)) == 0) n1 *= spf(N);
    return n1;
}
constexpr int radix_of(int N) { return N % 4 == 0 ? 4 : (N % 2 == 0 ? 2 : spf(N)); }
constexpr int modinv(int a, int m)
{
    for (int x = 1; x < m; x++)
        if (a % m * x % m == 1) return x;
    return 0;
}
constexpr int twoff(int N) { return N * (N - 1) / 2; }

/* largest power of two <= n (n >= 1) */
constexpr int pow2_floor(int n)
{
    int p = 1;
    while (2 * p <= n) p *= 2;
    return p;
}
/* width candidates: the fsp = 1.5 bank 1,2,3,4,6,9,13,... (src/boxcar.c:78-92) */
constexpr int wseq(int i)
{
    int w = 1;
    for (int k = 0; k < i; k++) {
        int g = (int)(1.5 * (double)w);
        w = g > w + 1 ? g : w + 1;
    }
    return w;
}
constexpr int wcap(int nb)
{
    int c = (int)(0.3 * (double)nb);
    if (c < 1) c = 1;
    if (c > nb - 1) c = nb - 1;
    return c;
}
constexpr int ncand(int nb)
{
    int n = 0;
    while (n < CS_NSEQ && wseq(n) <= wcap(nb)) n++;
    return n;
}
constexpr int wcmax(int nb) { return ncand(nb) > 0 ? wseq(ncand(nb) - 1) : 0; }

/* Twiddles are compile-time constants (cs_twiddles.h, written at build time
 * from libm cos and sin exactly as the session code tabulates them, which
 * checks the two agree): every index below is a constant after unrolling, so
 * each twiddle becomes a scalar move or a literal operand -- no loads and no
 * waits on them.  The f32 path uses (float)cos/sin, the f64 path the double
 * values. */
struct TwF {};
struct TwD {};

__device__ __forceinline__ cx<float> ldtw(TwF, int i) { return {kTwF[2 * i], kTwF[2 * i + 1]}; }
__device__ __forceinline__ cx<double> ldtw(TwD, int i) { return {kTwD[2 * i], kTwD[2 * i + 1]}; }

__device__ __forceinline__ float fm(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fm(double a, double b, double c) { return __builtin_fma(a, b, c); }

/* y * w with the project c2r's fused form:
 * (fma(y.re, w.re, -(y.im*w.im)), fma(y.re, w.im, y.im*w.re)) */
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> w)
{
    const T p2 = a.y * w.y, p4 = a.y * w.x;
    return {fm(a.x, w.x, -p2), fm(a.x, w.y, p4)};
}

template <int R, typename T, typename TP>
__device__ __forceinline__ void bfly(const cx<T> (&u)[R], cx<T> (&v)[R], TP tw)
{
    if constexpr (R == 2) {
        v[0] = {u[0].x + u[1].x, u[0].y + u[1].y};
        v[1] = {u[0].x - u[1].x, u[0].y - u[1].y};
    } else if constexpr (R == 4) {
        const cx<T> t0 = {u[0].x + u[2].x, u[0].y + u[2].y};
        const cx<T> t1 = {u[0].x - u[2].x, u[0].y - u[2].y};
        const cx<T> t2 = {u[1].x + u[3].x, u[1].y + u[3].y};
        const cx<T> t3 = {u[1].x - u[3].x, u[1].y - u[3].y};
        v[0] = {t0.x + t2.x, t0.y + t2.y};
        v[2] = {t0.x - t2.x, t0.y - t2.y};
        v[1] = {t1.x - t3.y, t1.y + t3.x};
        v[3] = {t1.x + t3.y, t1.y - t3.x};
    } else {
        /* odd prime: sums and differences of mirrored inputs, then per
         * output pair A = u0 + sum sp_r c (fused, left to right) and
         * B = sum sm_r sn (first term a product, the rest fused) */
        constexpr int hh = (R - 1) / 2;
        cx<T> sp[hh + 1], sm[hh + 1];
        sfor<hh>([&](auto I) {
            constexpr int r = decltype(I)::value + 1;
            sp[r] = {u[r].x + u[R - r].x, u[r].y + u[R - r].y};
            sm[r] = {u[r].x - u[R - r].x, u[r].y - u[R - r].y};
        });
        cx<T> x0 = u[0];
        sfor<hh>([&](auto I) {
            constexpr int r = decltype(I)::value + 1;
            x0 = {x0.x + sp[r].x, x0.y + sp[r].y};
        });
        v[0] = x0;
        sfor<hh>([&](auto S) {
            constexpr int s = decltype(S)::value + 1;
            cx<T> A = u[0], B = {T(0), T(0)};
            sfor<hh>([&](auto I) {
                constexpr int r = decltype(I)::value + 1;
                constexpr int j = (r * s) % R;
                const cx<T> w = ldtw(tw, twoff(R) + j);
                A = {fm(sp[r].x, w.x, A.x), fm(sp[r].y, w.x, A.y)};
                if constexpr (r == 1) B = {sm[r].x * w.y, sm[r].y * w.y};
                else B = {fm(sm[r].x, w.y, B.x), fm(sm[r].y, w.y, B.y)};
            });
            v[s] = {A.x - B.y, A.y + B.x};
            v[R - s] = {A.x + B.y, A.y - B.x};
        });
    }
}

/* X = IDFT_N(x) (e^{+i}, unnormalised).  N with more than one prime: the
 * prime-factor split N = N1*N2 (N1 the full power of the smallest prime):
 * DFT_N1 over x[(N2*n1 + N1*n2) mod N] for every n2, then DFT_N2 over n2 for
 * every k1, written to the CRT index -- no twiddles, only renaming.  A prime
 * power: decimation in time with radix 4, 2 or the prime, twiddles cmul'd. */
template <int N, typename T, typename TP>
__device__ __forceinline__ void idft(const cx<T> (&x)[N], cx<T> (&X)[N], TP tw)
{
    if constexpr (N == 1) {
        X[0] = x[0];
    } else if constexpr (pfa_n1(N) < N) {
        constexpr int N1 = pfa_n1(N), N2 = N / N1;
        constexpr int E1 = N2 * modinv(N2, N1), E2 = N1 * modinv(N1, N2);
        cx<T> Tm[N2][N1];
        sfor<N2>([&](auto I2) {
            constexpr int n2 = decltype(I2)::value;
            cx<T> g[N1];
            sfor<N1>([&](auto I1) {
                constexpr int n1 = decltype(I1)::value;
                g[n1] = x[(N2 * n1 + N1 * n2) % N];
            });
            idft<N1>(g, Tm[n2], tw);
        });
        sfor<N1>([&](auto K1) {
            constexpr int k1 = decltype(K1)::value;
            cx<T> c[N2], o[N2];
            sfor<N2>([&](auto I2) {
                constexpr int n2 = decltype(I2)::value;
                c[n2] = Tm[n2][k1];
            });
            idft<N2>(c, o, tw);
            sfor<N2>([&](auto K2) {
                constexpr int k2 = decltype(K2)::value;
                X[(k1 * E1 + k2 * E2) % N] = o[k2];
            });
        });
    } else {
        constexpr int R = radix_of(N), M = N / R;
        cx<T> Y[R][M];
        sfor<R>([&](auto I) {
            constexpr int r = decltype(I)::value;
            cx<T> g[M];
            sfor<M>([&](auto J) {
                constexpr int j = decltype(J)::value;
                g[j] = x[r + j * R];
            });
            idft<M>(g, Y[r], tw);
        });
        sfor<M>([&](auto K) {
            constexpr int k = decltype(K)::value;
            cx<T> u[R], v[R];
            u[0] = Y[0][k];
            sfor<R - 1>([&](auto I) {
                constexpr int r = decltype(I)::value + 1;
                if constexpr (r * k == 0) u[r] = Y[r][k];
                else u[r] = cmul(Y[r][k], ldtw(tw, twoff(N) + r * k));
            });
            bfly<R>(u, v, tw);
            sfor<R>([&](auto I) {
                constexpr int s = decltype(I)::value;
                X[k + M * s] = v[s];
            });
        });
    }
}

/* Project c2r of X[0], X[XS], ..., X[H*XS] (the f32 stack, widened to T) into
 * x[0..2H-1] (unnormalised), in precision T. */
template <int H, int XS, int LEN, typename T, typename TP>
__device__ __forceinline__ void c2r(const cf (&X)[LEN], T (&x)[2 * H], TP tw)
{
    cx<T> Z[H];
    {
        const T x0 = (T)X[0].x, xh = (T)X[H * XS].x;
        Z[0] = {x0 + xh, x0 - xh};
    }
    sfor<H - 1>([&](auto I) {
        constexpr int k = decltype(I)::value + 1;
        const cx<T> a = {(T)X[k * XS].x, (T)X[k * XS].y};
        const cx<T> b = {(T)X[(H - k) * XS].x, (T)X[(H - k) * XS].y};
        const T er = a.x + b.x, ei = a.y - b.y;
        const cx<T> o = cmul(cx<T>{a.x - b.x, a.y + b.y}, ldtw(tw, twoff(2 * H) + k));
        Z[k] = {er - o.y, ei + o.x};
    });
    cx<T> z[H];
    idft<H>(Z, z, tw);
    sfor<H>([&](auto I) {
        constexpr int s = decltype(I)::value;
        x[2 * s] = z[s].x;
        x[2 * s + 1] = z[s].y;
    });
}

__device__ __forceinline__ float tmax(float m, float d) { return fmaxf(m, d); }
__device__ __forceinline__ double tmax(double m, double d) { return d > m ? d : m; }

/* Maximum window sum of width W over the NB wrapped windows of a profile:
 * max over p of P[p+W-1] - P[p-1], in the prefix order of src/boxcar.c:203-252.
 * f32: two interleaved maximum chains -- fmaxf ignores NaN and is exact, so
 * the maximum does not depend on the order, and the chains' independence
 * leaves the scheduler work to overlap. */
template <int NB, int W, int NP, typename T>
__device__ __forceinline__ T winmax(const T (&P)[NP])
{
    T m = P[W] - P[0];
    if constexpr (sizeof(T) == 4) {
        T m1 = -(T)INFINITY;
        sfor<NB - 1>([&](auto J) {
            constexpr int p = decltype(J)::value + 2;
            const T d = P[p + W - 1] - P[p - 1];
            if constexpr (p % 2 == 0) m1 = tmax(m1, d);
            else m = tmax(m, d);
        });
        m = tmax(m, m1);
    } else {
        sfor<NB - 1>([&](auto J) {
            constexpr int p = decltype(J)::value + 2;
            const T d = P[p + W - 1] - P[p - 1];
            m = tmax(m, d);
        });
    }
    return m;
}

/* Widths whose maximum window sum the f32 boxcar always computes; every
 * other width is first bounded by a sum of these (bx_parts). */
constexpr bool bx_anchor(int w) { return w == 1 || w == 4 || w == 13 || w == 42; }
/* Greedy split of width w into anchor widths: how many of anchor a. */
constexpr int bx_count(int w, int a)
{
    const int n42 = w / 42, r42 = w % 42;
    const int n13 = r42 / 13, r13 = r42 % 13;
    const int n4 = r13 / 4, n1 = r13 % 4;
    return a == 42 ? n42 : a == 13 ? n13 : a == 4 ? n4 : n1;
}
constexpr int bx_parts(int w)
{
    return bx_count(w, 42) + bx_count(w, 13) + bx_count(w, 4) + bx_count(w, 1);
}

/* Boxcar S/N of one profile over the enabled widths (mask over the fsp=1.5
 * candidates), src/boxcar.c:203-252 in precision T: wrapped prefix sums, max
 * window sum per width, (m - duty*total) * invsw_w with invsw_w =
 * invsigma*shape_w rounded to T.
 *
 * The caller only needs the exact value when it exceeds pthr (the candidate
 * threshold; -inf when every trial's metric is kept).  In f32 a width w that
 * is not an anchor (1, 4, 13, 42) is first bounded: a window of w splits into
 * k windows of anchor widths, so in real arithmetic its sum is at most the
 * sum of those widths' maxima.  Each computed window differs from the real
 * sum of its profile values by at most 2u(NB+WC+1)*max|P| (u = 2^-24: the
 * prefix sums' accumulated rounding plus the subtraction's), which gives
 *     m_w <= sum of anchor maxima + (k+1) * 2u(NB+WC+1) * max|P|;
 * the margin used, 4u(k+2)(NB+WC+2k+4) * max|P|, also covers the rounding of
 * the bound's own sum.  Rounding is monotone and invsw_w > 0, so the bound
 * put through the same (B - duty*total) * invsw_w is at least the width's
 * exact value: when it is at most pthr the width cannot raise the result
 * above pthr and its windows are skipped.  The value returned is then the
 * exact result whenever that exceeds pthr (and at most pthr otherwise).
 * NaN and infinite profiles make the bound NaN or +inf and are never
 * skipped. */
template <int NB, typename T>
__device__ __forceinline__ T boxcar(const T (&x)[NB], unsigned wmask,
                                    const T *__restrict__ invsw, int valid, double pthr)
{
    constexpr int NC = ncand(NB);
    constexpr int WC = wcmax(NB);
    constexpr int NP = NB + WC + 1;
    T P[NP];
    P[0] = T(0);
    sfor<NB + WC>([&](auto I) {
        constexpr int i = decltype(I)::value + 1;
        constexpr int idx = (i > NB ? i - NB : i) - 1;
        P[i] = P[i - 1] + x[idx];
    });
    const T stot = P[NB] - P[0];
    T best = -(T)INFINITY;
    if constexpr (sizeof(T) == 4) {
        /* the slot's per-width factors, read once through the constant
         * address space (uniform, so scalar loads issued ahead of use) */
        using cfloat = const __attribute__((address_space(4))) float;
        const cfloat *civ = (const cfloat *)invsw;
        float iv[NC];
        sfor<NC>([&](auto I) { iv[decltype(I)::value] = civ[decltype(I)::value]; });
        float pmax = 0.0f;
        sfor<NP>([&](auto I) { pmax = fmaxf(pmax, fabsf(P[decltype(I)::value])); });
        float ma1 = -INFINITY, ma4 = -INFINITY, ma13 = -INFINITY, ma42 = -INFINITY;
        sfor<NC>([&](auto I) {
            constexpr int wi = decltype(I)::value;
            constexpr int W = wseq(wi);
            constexpr T duty = (T)W / (T)NB;
            const T dt = duty * stot;
            if constexpr (bx_anchor(W)) {
                const T m = winmax<NB, W>(P);
                if constexpr (W == 1) ma1 = m;
                if constexpr (W == 4) ma4 = m;
                if constexpr (W == 13) ma13 = m;
                if constexpr (W == 42) ma42 = m;
                if (wmask & (1u << wi)) {
                    const T c = (m - dt) * iv[wi];
                    best = c > best ? c : best;
                }
            } else if (wmask & (1u << wi)) {
                constexpr int k = bx_parts(W);
                constexpr float cm = (float)((double)(4 * (k + 2) * (NB + WC + 2 * k + 4)) *
                                             0x1p-24);
                float sb = 0.0f;
                sfor<bx_count(W, 42)>([&](auto) { sb = sb + ma42; });
                sfor<bx_count(W, 13)>([&](auto) { sb = sb + ma13; });
                sfor<bx_count(W, 4)>([&](auto) { sb = sb + ma4; });
                sfor<bx_count(W, 1)>([&](auto) { sb = sb + ma1; });
                const float bnd = __builtin_fmaf(pmax, cm, sb);
                const float cb = (bnd - dt) * iv[wi];
                if (!(iv[wi] > 0.0f && (double)cb <= pthr)) {
                    const T m = winmax<NB, W>(P);
                    const T c = (m - dt) * iv[wi];
                    best = c > best ? c : best;
                }
            }
        });
    } else {
        sfor<NC>([&](auto I) {
            constexpr int wi = decltype(I)::value;
            constexpr int W = wseq(wi);
            if (wmask & (1u << wi)) {
                const T m = winmax<NB, W>(P);
                constexpr T duty = (T)W / (T)NB;
                const T dt = duty * stot;
                const T c = (m - dt) * invsw[wi];
                best = c > best ? c : best;
            }
        });
    }
    return valid ? best : T(0);
}

template <typename T>
__device__ __forceinline__ const T *invsw_of(const ChunkDesc *C, int slot);
template <>
__device__ __forceinline__ const float *invsw_of<float>(const ChunkDesc *C, int slot)
{
    return C->invsw[slot];
}
template <>
__device__ __forceinline__ const double *invsw_of<double>(const ChunkDesc *C, int slot)
{
    return C->invswd[slot];
}

struct FastArgs {
    const float2 *amps;
    long long namps;
    const float *wtab;
    const int2 *hinfo;        /* (rfloor0, carry residue threshold) per h */
    const float2 *tw;         /* (float)cos/sin table */
    const double2 *twd;       /* double cos/sin table (f64 path) */
    const int4 *tiles;
    int ntiles;
    int q;
    int pnum;
    int ws;                   /* weight row stride (floats) */
    int grp;                  /* trials of the tile per residue class, max(1, CS_TILE/q) */
    int m2e;                  /* first tap at base + 1 - m2e: M - m/2 (rows front-padded to M taps) */
    const ChunkDesc *chunks;
    CandRec *recs;
    unsigned int *counter;
    unsigned int cap;
    double threshold;
    int kslot[CS_MAXSLOT];    /* slot index of decimation k (k<=16), -1 off */
    unsigned int wmask[CS_MAXSLOT];   /* per slot */
    void *samp;               /* measured sigma sample pass: |profile| samples per
                                 (chunk, slot), CS_SIGCAP of T each; null otherwise */
    int nslot;
    int xcd_stripe;           /* 0: XCD x runs tiles [x*n/8, (x+1)*n/8); S > 0: stripes of
                                 S consecutive tiles dealt to the XCDs in turn */
    float2 *stack_out;        /* optional stack dump of trials [dump_t0, +dump_n) */
    long long dump_t0;
    long long dump_n;
    void *mout;               /* metric collection (--metricstats/--normalize): every valid
                                 trial's metric as T at mout[slot*mstride + i0 + j] instead
                                 of above-threshold records; null otherwise */
    long long mstride;
};

__device__ __forceinline__ void emit(bool pred, int chunk, int slot, int j, double metric,
                                     const FastArgs &a)
{
    const unsigned long long mask = __ballot(pred);
    if (mask == 0ull) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    unsigned int base = 0;
    if (lane == leader) base = atomicAdd(a.counter, (unsigned int)__popcll(mask));
    base = __shfl(base, leader);
    if (pred) {
        const unsigned int pos = base + (unsigned int)__popcll(mask & ((1ull << lane) - 1ull));
        if (pos < a.cap) {
            CandRec r;
            r.key = ((unsigned long long)chunk << 40) | ((unsigned long long)slot << 32) |
                    (unsigned int)j;
            r.metric = metric;
            a.recs[pos] = r;
        }
    }
}

/* The metric of trial j of the chunk starting at trial i0, in the dense
 * per-slot metric array of the collection pass. */
template <typename T>
__device__ __forceinline__ void metric_out(bool pred, int slot, long long i0, int j, T metric,
                                           const FastArgs &a)
{
    if (pred) ((T *)a.mout)[(long long)slot * a.mstride + i0 + j] = metric;
}

/* Measured sigma (src/boxcar.c:18-73, 138-167): the sample of (chunk, slot)
 * is the |value| of every profile bin of the valid trials when that is at
 * most CS_SIGCAP values, otherwise the whole profiles of the trials
 * j0 = c*nv/ncol, c < ncol = CS_SIGCAP/nb, in that order.  A lane writes its
 * profile at its row's place in the sample when its trial is one of them. */
template <int NB, typename T>
__device__ __forceinline__ void sample_out(const T (&x)[NB], bool active, int lj, int nv,
                                           int chunk, int slot, const FastArgs &a)
{
    if (!active || lj >= nv) return;
    long long pos;
    if ((long long)NB * nv <= CS_SIGCAP) {
        pos = (long long)lj * NB;
    } else {
        long long ncol = CS_SIGCAP / NB;
        if (ncol < 1) ncol = 1;
        if (ncol > nv) ncol = nv;
        const long long c = ((long long)lj * ncol + nv - 1) / nv;   /* ceil(lj*ncol/nv) */
        if (c >= ncol || (c * nv) / ncol != lj) return;
        pos = c * NB;
    }
    T *out = (T *)a.samp + ((long long)chunk * a.nslot + slot) * CS_SIGCAP + pos;
    sfor<NB>([&](auto I) {
        constexpr int i = decltype(I)::value;
        out[i] = x[i] < T(0) ? -x[i] : x[i];
    });
}

/* Occupancy target: f32 profiles 2 waves/SIMD for the largest stacks, 3
 * below, 1 above 64 harmonics; f64 profiles (twice the registers) 1 above 16
 * harmonics, 2 below. */
template <typename T>
constexpr int waves_for(int nh)
{
    return sizeof(T) == 4 ? (nh > 64 ? 1 : (nh > 32 ? 2 : 3)) : (nh > 16 ? 1 : 2);
}

/* Everything after the interpolated stack, for trial t (row lj of the chunk
 * at i0, the lane active when the trial is in its tile): the optional stack
 * dump, the c2r of every non-divisor decimation, the base profile and its
 * folds, then per decimation slot the boxcar metric with its record, metric
 * or measured-sigma sample output. */
template <int NH, int KMAX, typename T>
__device__ __forceinline__ void search_tail(const cf (&X)[NH + 1], const FastArgs &a,
                                            const ChunkDesc *__restrict__ C, int chunk,
                                            long long i0,
                                            int lj, long long t, bool active)
{
    if (a.stack_out != nullptr) {
        const long long d = t - a.dump_t0;
        if (active && d >= 0 && d < a.dump_n) {
            sfor<NH + 1>([&](auto I) {
                constexpr int h = decltype(I)::value;
                a.stack_out[d * (NH + 1) + h] = make_float2(X[h].x, X[h].y);
            });
        }
    }

    const double thr = a.threshold;
    /* the boxcar may skip widths that cannot exceed the threshold unless
     * every trial's metric is kept */
    const double pthr = a.mout ? -(double)INFINITY : thr;
    /* the c2r twiddles of this precision */
    using TwP = std::conditional_t<sizeof(T) == 4, TwF, TwD>;
    const TwP twp{};

    /* decimations k not dividing nharms: strided c2r of the stack */
    sfor<KMAX - 1>([&](auto I) {
        constexpr int k = decltype(I)::value + 2;
        if constexpr (NH / k >= 1 && NH % k != 0) {
            const int slot = a.kslot[k];
            if (slot >= 0) {
                constexpr int Hk = NH / k;
                T xk[2 * Hk];
                c2r<Hk, k, NH + 1>(X, xk, twp);
                if (a.samp) {
                    sample_out<2 * Hk>(xk, active, lj, C->nvalid[slot], chunk,
                                       slot, a);
                } else {
                    const T mv = boxcar<2 * Hk>(xk, a.wmask[slot], invsw_of<T>(C, slot),
                                                C->valid[slot], pthr);
                    if (a.mout)
                        metric_out<T>(active && lj < C->nvalid[slot], slot, i0, lj, mv, a);
                    else
                        emit(active && lj < C->nvalid[slot] && (double)mv > thr, chunk, slot,
                             lj, mv, a);
                }
            }
        }
    });

    /* base profile */
    T x[2 * NH];
    c2r<NH, 1, NH + 1>(X, x, twp);

    /* divisor decimations fold the base profile (src/search.c:298-331) */
    sfor<KMAX - 1>([&](auto I) {
        constexpr int k = decltype(I)::value + 2;
        if constexpr (NH / k >= 1 && NH % k == 0) {
            const int slot = a.kslot[k];
            if (slot >= 0) {
                constexpr int nk = 2 * (NH / k);
                constexpr T invk = T(1) / (T)k;
                T xk[nk];
                sfor<nk>([&](auto J) {
                    constexpr int j = decltype(J)::value;
                    T s = x[j];
                    sfor<k - 1>([&](auto Q) {
                        constexpr int qq = decltype(Q)::value + 1;
                        s = s + x[j + qq * nk];
                    });
                    xk[j] = s * invk;
                });
                if (a.samp) {
                    sample_out<nk>(xk, active, lj, C->nvalid[slot], chunk, slot, a);
                } else {
                    const T mv = boxcar<nk>(xk, a.wmask[slot], invsw_of<T>(C, slot),
                                            C->valid[slot], pthr);
                    if (a.mout)
                        metric_out<T>(active && lj < C->nvalid[slot], slot, i0, lj, mv, a);
                    else
                        emit(active && lj < C->nvalid[slot] && (double)mv > thr, chunk, slot,
                             lj, mv, a);
                }
            }
        }
    });

    if (a.samp) {
        sample_out<2 * NH>(x, active, lj, C->nvalid[0], chunk, 0, a);
        return;
    }
    const T mb = boxcar<2 * NH>(x, a.wmask[0], invsw_of<T>(C, 0), C->valid[0], pthr);
    if (a.mout) metric_out<T>(active, 0, i0, lj, mb, a);
    else emit(active && (double)mb > thr, chunk, 0, lj, mb, a);
}

template <int NH, int M, int KMAX, typename T>
__global__ __launch_bounds__(CS_TILE) __attribute__((amdgpu_waves_per_eu(waves_for<T>(NH))))
void fast_search_kernel(FastArgs a)
{
    static_assert(NH >= 1 && NH <= 128, "fast path holds up to 128 harmonics");
    static_assert(M % 4 == 0, "tap count multiple of 4");
    __shared__ float2 win[CS_WIN_CAP(NH, M)];
    __shared__ int s_off[NH + 1];
    __shared__ int s_lo[NH];
    __shared__ int s_C[NH];

    /* XCD-aware mapping: workgroup b runs on XCD b % 8; consecutive tiles
     * land on the same XCD's L2 (the grid is a multiple of 8, of 8*S with
     * stripes). */
    const int nblk = gridDim.x;
    int tile;
    if (a.xcd_stripe > 0) {
        const int S = a.xcd_stripe, i = blockIdx.x >> 3;
        tile = ((i / S) * 8 + (blockIdx.x & 7)) * S + i % S;
    } else {
        tile = (blockIdx.x & 7) * (nblk >> 3) + (blockIdx.x >> 3);
    }
    if (tile >= a.ntiles) return;

    const int4 Tl = a.tiles[tile];
    const ChunkDesc *C = a.chunks + Tl.x;
    const long long i0 = C->i0;
    const unsigned long long filled0 = C->filled[0], filled1 = C->filled[1];
    const int tid = threadIdx.x;
    /* Trials t and t+q have the same residue t*h*p mod q at every harmonic,
     * hence the same weight rows.  The first grp*q lanes take the tile's
     * trials in residue-class order (lane L -> (L % grp)*q + L / grp), so
     * neighbouring lanes gather the same rows; the LDS windows cover the
     * whole tile either way. */
    const int G = a.grp;
    const int slot_l = tid < G * a.q ? (tid % G) * a.q + tid / G : tid;
    const bool active = slot_l < Tl.z;
    const int lj = Tl.y + (active ? slot_l : Tl.z - 1);
    const long long t = i0 + lj;
    const int q = a.q;
    const int m2 = a.m2e;

    /* per-harmonic windows of this tile; harmonics that do not fit the chunk
     * get a window too (zero outside the spectrum) so that the interpolation
     * below is branch-free and their result is masked to zero */
    if (tid < NH) {
        const int h = tid + 1;
        const int2 hi = a.hinfo[tid];
        const long long tf = i0 + Tl.y, tl = i0 + Tl.y + Tl.z - 1;
        const long long af = tf * a.pnum, al = tl * a.pnum;
        const int rf = (int)(af % q), rl = (int)(al % q);
        const long long qf = af / q, ql = al / q;
        const int hrf = h * rf, hrl = h * rl;
        const int resf = hrf % q, resl = hrl % q;
        const long long bf = hi.x + (long long)h * qf + hrf / q + (resf >= hi.y ? 1 : 0);
        const long long bl = hi.x + (long long)h * ql + hrl / q + (resl >= hi.y ? 1 : 0);
        const int lo = (int)(bf + 1 - m2);
        const int len = (int)(bl - bf) + M;
        const int cc = hi.x + 1 - m2 - lo;
        s_lo[tid] = lo;
        s_C[tid] = cc;
        s_off[tid + 1] = len;
    }
    if (tid == 0) s_off[0] = 0;
    __syncthreads();
    if (tid < 64) {
        /* inclusive scan of lengths over the first wave; above 64 harmonics
         * each lane takes a pair (2*tid, 2*tid+1) */
        if constexpr (NH <= 64) {
            int v = (tid < NH) ? s_off[tid + 1] : 0;
            for (int d = 1; d < 64; d <<= 1) {
                const int u = __shfl_up(v, d);
                if (tid >= d) v += u;
            }
            if (tid < NH) s_off[tid + 1] = v;
        } else {
            const int i0 = 2 * tid, i1 = 2 * tid + 1;
            const int a0 = i0 < NH ? s_off[i0 + 1] : 0, a1 = i1 < NH ? s_off[i1 + 1] : 0;
            int v = a0 + a1;
            for (int d = 1; d < 64; d <<= 1) {
                const int u = __shfl_up(v, d);
                if (tid >= d) v += u;
            }
            if (i0 < NH) s_off[i0 + 1] = v - a1;
            if (i1 < NH) s_off[i1 + 1] = v;
        }
    }
    __syncthreads();
    if (tid < NH) s_C[tid] += s_off[tid];
    __syncthreads();

    /* per-lane exact residue walk: t*pnum = qb*q + rho */
    const long long tp = t * a.pnum;
    const int rho = (int)(tp % q);
    const int qb = (int)(tp / q);

    cf X[NH + 1];
    sfor<NH + 1>([&](auto I) { X[decltype(I)::value] = {0.0f, 0.0f}; });

    /* All windows of the tile at once: the host admits a configuration only
     * when sum_h (ceil(255*h*lodr) + m + 2) <= CS_WIN_CAP(NH, M), so one pass
     * suffices (a batched loop would let the compiler hoist every harmonic's
     * residue walk out of it and hold it in registers). */
    const int cnt = s_off[NH];
    /* every bin a lane stages in flight before its LDS stores: one memory
     * round trip for the tile's windows (SB covers CS_WIN_CAP, so the loop
     * runs once); each bin's harmonic is the last window starting at or
     * before it, found by a fixed-step binary search so that the searches
     * interleave */
    constexpr int SB = (CS_WIN_CAP(NH, M) + CS_TILE - 1) / CS_TILE;
    for (int base = tid; base < cnt; base += SB * CS_TILE) {
        float2 v[SB];
#pragma unroll
        for (int k = 0; k < SB; k++) {
            const int idx = base + k * CS_TILE;
            int lo = 0;
#pragma unroll
            for (int st = pow2_floor(NH); st >= 1; st >>= 1) {
                const int c = min(lo + st, NH - 1);
                lo = s_off[c] <= idx ? c : lo;
            }
            const long long ai = (long long)s_lo[lo] + (idx - s_off[lo]);
            v[k] = make_float2(0.0f, 0.0f);
            if (idx < cnt && ai >= 0 && ai < a.namps) v[k] = a.amps[ai];
        }
#pragma unroll
        for (int k = 0; k < SB; k++) {
            const int idx = base + k * CS_TILE;
            if (idx < cnt) win[idx] = v[k];
        }
    }
    __syncthreads();

    {
        /* Branch-free over harmonics so loads of later harmonics can be
         * scheduled under the FMA chains of earlier ones.  Weight rows are
         * CS_WSP(M) floats (an odd number of 16-byte units), addressed with a
         * 32-bit offset from the table base. */
        constexpr int WS4 = CS_WSP(M) / 4;
        const char *__restrict__ wbase = reinterpret_cast<const char *>(a.wtab);
        int res = 0, qi = 0;
        sfor<NH>([&](auto I) {
            constexpr int h = decltype(I)::value + 1;
            res += rho;
            qi += qb;
            if (res >= q) { res -= q; qi += 1; }
            const int carry = res >= a.hinfo[h - 1].y ? 1 : 0;
            const int wi = s_C[h - 1] + qi + carry;
            const unsigned wrow = (unsigned)((h - 1) * q + res) * (unsigned)(16 * WS4);
            const float4 *__restrict__ wr = reinterpret_cast<const float4 *>(wbase + wrow);
            float ar = 0.0f, ai = 0.0f;
            sfor<M / 4>([&](auto J) {
                constexpr int j = 4 * decltype(J)::value;
                const float4 w4 = wr[j / 4];
                const float2 a0 = win[wi + j], a1 = win[wi + j + 1];
                const float2 a2 = win[wi + j + 2], a3 = win[wi + j + 3];
                ar = __builtin_fmaf(w4.x, a0.x, ar); ai = __builtin_fmaf(w4.x, a0.y, ai);
                ar = __builtin_fmaf(w4.y, a1.x, ar); ai = __builtin_fmaf(w4.y, a1.y, ai);
                ar = __builtin_fmaf(w4.z, a2.x, ar); ai = __builtin_fmaf(w4.z, a2.y, ai);
                ar = __builtin_fmaf(w4.w, a3.x, ar); ai = __builtin_fmaf(w4.w, a3.y, ai);
            });
            const float2 A = reinterpret_cast<const float2 *>(wr + M / 4)[0];
            const float p1 = A.x * ar, p2 = A.y * ai, p3 = A.x * ai, p4 = A.y * ar;
            const bool fit = (((h - 1) < 64 ? filled0 : filled1) >> ((h - 1) & 63)) & 1ull;
            X[h] = {fit ? p1 - p2 : 0.0f, fit ? p3 + p4 : 0.0f};
        });
    }

    search_tail<NH, KMAX, T>(X, a, C, Tl.x, i0, lj, t, active);
}

}  // namespace csk

/* Instantiation + launch helper, one per translation unit. */
#define CS_FAST_INSTANTIATE(NH, M, KMAX)                                         \
    extern "C" __attribute__((visibility("hidden"))) hipError_t                   \
    cs_launch_fast_##NH##_##M##_##KMAX(const csk::FastArgs *a, int nblocks,      \
                                       hipStream_t st)                           \
    {                                                                            \
        hipLaunchKernelGGL((csk::fast_search_kernel<NH, M, KMAX, float>),        \
                           dim3(nblocks), dim3(CS_TILE), 0, st, *a);             \
        return hipGetLastError();                                                \
    }

#define CS_FAST64_INSTANTIATE(NH, M, KMAX)                                       \
    extern "C" __attribute__((visibility("hidden"))) hipError_t                   \
    cs_launch_fast64_##NH##_##M##_##KMAX(const csk::FastArgs *a, int nblocks,    \
                                         hipStream_t st)                         \
    {                                                                            \
        hipLaunchKernelGGL((csk::fast_search_kernel<NH, M, KMAX, double>),       \
                           dim3(nblocks), dim3(CS_TILE), 0, st, *a);             \
        return hipGetLastError();                                                \
    }
