/*
 * The general search kernel and its c2r plans (see cs_general.h).
 *
 * Per lane, one trial of a 256-trial tile, for every tile the persistent
 * workgroup is dealt:
 *   1. the interpolated harmonic stack: the m-tap f32 interpolation of every
 *      harmonic (fmaf in tap order, then the unfused rotation, the
 *      reference's arithmetic, src/interp.c:686-811), read straight from the
 *      resident spectrum (neighbouring lanes' windows share cache lines) and
 *      the weight-row arena, into the lane's stack in the work buffer;
 *   2. the c2r of the base and of every non-divisor decimation: the plan's
 *      ops in order -- the pre-step, then radix-2/4/p butterflies with their
 *      input twiddles -- on the lane's value slots (the plan is the same for
 *      every lane, so op words are scalar loads);
 *   3. per decimation slot the boxcar matched filter over the enabled widths
 *      (src/boxcar.c:203-252): prefix sums computed in order through a
 *      register block of B + WC + 1 values that slides along the profile, so
 *      every prefix value and window difference is the reference's own,
 *      read a complex slot (two profile values) per load; a divisor
 *      decimation's fold of the base profile (src/search.c:294-331) is
 *      written to its own value slots first, 32 bin pairs' loads in
 *      flight per term;
 *   4. above-threshold records (one atomic per wave), or the metric
 *      collection / measured-sigma sample outputs of the other passes.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "cs_general.h"

/* ---- host: the plan of a search's c2rs ---------------------------------- */
namespace {

int g_spf(int N)
{
    if (N % 2 == 0) return 2;
    for (int p = 3; p * p <= N; p += 2)
        if (N % p == 0) return p;
    return N;
}

int g_pfa_n1(int N)
{
    const int p = g_spf(N);
    int n1 = 1;
    while (N % (n1 * p) == 0) n1 *= p;
    return n1;
}

int g_radix(int N) { return N % 4 == 0 ? 4 : (N % 2 == 0 ? 2 : g_spf(N)); }

long g_modinv(long a, long m)
{
    a %= m;
    for (long x = 1; x < m; x++)
        if (a * x % m == 1) return x;
    return 0;
}

/* cos and sin as two separate libm calls (the twiddle definition), through
 * pointers the compiler cannot merge into one sincos call */
double (*volatile g_cos)(double) = cos;
double (*volatile g_sin)(double) = sin;

/* a symbolic op: operands are value ids */
struct SOp {
    int kind, R;
    std::vector<int> in, tw, out;
    int rtw = -1;             /* BFP: index of tw(0, R) (then tw(j, R) consecutive) */
    int H = 0, XS = 0, twb = -1;   /* PRE */
};

struct Build {
    CsGenPlan *P;
    std::vector<SOp> ops;
    int nval = 0;
    std::map<std::pair<int, int>, int> twmap;   /* (j, N) -> twiddle index */

    int twiddle(int j, int N)
    {
        auto it = twmap.find({j, N});
        if (it != twmap.end()) return it->second;
        const double a = 2.0 * M_PI * (double)j / (double)N;
        const int idx = (int)P->twc.size();
        P->twc.push_back(g_cos(a));
        P->tws.push_back(g_sin(a));
        twmap[{j, N}] = idx;
        return idx;
    }
    int fresh_twiddles(int first_j, int count, int N)
    {
        const int base = (int)P->twc.size();
        for (int j = first_j; j < first_j + count; j++) {
            const double a = 2.0 * M_PI * (double)j / (double)N;
            P->twc.push_back(g_cos(a));
            P->tws.push_back(g_sin(a));
        }
        return base;
    }

    /* IDFT_N of the values `in` (the project c2r's recursion); returns the
     * ids of the outputs in order */
    std::vector<int> dft(int N, const std::vector<int> &in)
    {
        if (N == 1) return in;
        const int N1 = g_pfa_n1(N);
        std::vector<int> out((size_t)N, -1);
        if (N1 < N) {
            const int N2 = N / N1;
            std::vector<std::vector<int>> Tm((size_t)N2);
            for (int n2 = 0; n2 < N2; n2++) {
                std::vector<int> g((size_t)N1);
                for (int n1 = 0; n1 < N1; n1++)
                    g[(size_t)n1] = in[(size_t)(((long)N2 * n1 + (long)N1 * n2) % N)];
                Tm[(size_t)n2] = dft(N1, g);
            }
            const long E1 = N2 * g_modinv(N2, N1), E2 = N1 * g_modinv(N1, N2);
            for (int k1 = 0; k1 < N1; k1++) {
                std::vector<int> c((size_t)N2);
                for (int n2 = 0; n2 < N2; n2++) c[(size_t)n2] = Tm[(size_t)n2][(size_t)k1];
                const std::vector<int> o = dft(N2, c);
                for (int n2 = 0; n2 < N2; n2++) out[(size_t)((k1 * E1 + n2 * E2) % N)] = o[(size_t)n2];
            }
            return out;
        }
        const int R = g_radix(N), M = N / R;
        std::vector<std::vector<int>> Y((size_t)R);
        for (int r = 0; r < R; r++) {
            std::vector<int> g((size_t)M);
            for (int j = 0; j < M; j++) g[(size_t)j] = in[(size_t)(r + j * R)];
            Y[(size_t)r] = dft(M, g);
        }
        int rtw = -1;
        if (R != 2 && R != 4) rtw = fresh_twiddles(0, R, R);
        for (int k = 0; k < M; k++) {
            SOp o;
            o.kind = R == 2 ? CS_OP_BF2 : (R == 4 ? CS_OP_BF4 : CS_OP_BFP);
            o.R = R;
            o.rtw = rtw;
            for (int r = 0; r < R; r++) {
                o.in.push_back(Y[(size_t)r][(size_t)k]);
                o.tw.push_back(r * k == 0 ? -1 : twiddle(r * k, N));
            }
            for (int s = 0; s < R; s++) {
                o.out.push_back(nval);
                out[(size_t)(k + M * s)] = nval++;
            }
            ops.push_back(o);
        }
        return out;
    }
};

}  // namespace

int cs_gen_build_plan(int nh, int nextra, const int *Hk, const int *ks, CsGenPlan *P)
{
    *P = CsGenPlan();
    if (nh < 1 || nh > CS_GEN_MAXH || nextra < 0 || nextra + 1 > CS_GEN_MAXC2R) return -1;
    Build b;
    b.P = P;
    struct Top {
        int H, XS, op0, nop;
        std::vector<int> z;
    };
    std::vector<Top> tops;
    for (int c = 0; c <= nextra; c++) {
        Top t;
        t.H = c == 0 ? nh : Hk[c - 1];
        t.XS = c == 0 ? 1 : ks[c - 1];
        if (t.H < 1 || (long)t.H * t.XS > nh) return -1;
        t.op0 = (int)b.ops.size();
        SOp pre;
        pre.kind = CS_OP_PRE;
        pre.R = t.H;
        pre.H = t.H;
        pre.XS = t.XS;
        pre.twb = t.H > 1 ? b.fresh_twiddles(1, t.H - 1, 2 * t.H) : 0;
        std::vector<int> Z;
        for (int k = 0; k < t.H; k++) {
            pre.out.push_back(b.nval);
            Z.push_back(b.nval++);
        }
        b.ops.push_back(pre);
        t.z = b.dft(t.H, Z);
        if (t.H == 1) {               /* IDFT_1 is the identity: copy into the output slot */
            SOp cp;
            cp.kind = CS_OP_COPY;
            cp.R = 1;
            cp.in.push_back(t.z[0]);
            cp.out.push_back(b.nval);
            t.z[0] = b.nval++;
            b.ops.push_back(cp);
        }
        t.nop = (int)b.ops.size() - t.op0;
        tops.push_back(t);
    }
    /* slots: every c2r's outputs in a reserved run, in output order; the
     * other values take free slots in op order, a slot freed after its
     * value's last reader (the op reads all operands before it writes) */
    std::vector<int> slot((size_t)b.nval, -1), last((size_t)b.nval, -1);
    std::vector<char> reserved((size_t)b.nval, 0);
    int nres = 0;
    for (Top &t : tops) {
        for (int s = 0; s < t.H; s++) {
            slot[(size_t)t.z[(size_t)s]] = nres + s;
            reserved[(size_t)t.z[(size_t)s]] = 1;
        }
        nres += t.H;
    }
    for (size_t i = 0; i < b.ops.size(); i++)
        for (int v : b.ops[i].in) last[(size_t)v] = (int)i;
    std::vector<int> freel;
    int next = nres;
    for (size_t i = 0; i < b.ops.size(); i++) {
        SOp &o = b.ops[i];
        for (int v : o.in)
            if (last[(size_t)v] == (int)i && !reserved[(size_t)v]) freel.push_back(slot[(size_t)v]);
        for (int v : o.out) {
            if (reserved[(size_t)v]) continue;
            if (!freel.empty()) {
                slot[(size_t)v] = freel.back();
                freel.pop_back();
            } else {
                slot[(size_t)v] = next++;
            }
        }
    }
    P->nslots = next;
    for (const SOp &o : b.ops)
        if (o.kind == CS_OP_BFP && o.R > 13) P->nscratch = std::max(P->nscratch, 2 * o.R);
    for (size_t i = 0; i < b.ops.size(); i++) {
        const SOp &o = b.ops[i];
        CsGenOp g;
        g.kind = o.kind;
        g.R = o.R;
        g.pool = (int)P->pool.size();
        g.pad = 0;
        if (o.kind == CS_OP_PRE) {
            P->pool.push_back(o.XS);
            P->pool.push_back(o.twb);
            for (int v : o.out) P->pool.push_back(slot[(size_t)v]);
        } else if (o.kind == CS_OP_COPY) {
            P->pool.push_back(slot[(size_t)o.in[0]]);
            P->pool.push_back(slot[(size_t)o.out[0]]);
        } else {
            for (int v : o.in) P->pool.push_back(slot[(size_t)v]);
            for (int t : o.tw) P->pool.push_back(t);
            for (int v : o.out) P->pool.push_back(slot[(size_t)v]);
            P->pool.push_back(o.rtw);
        }
        P->ops.push_back(g);
    }
    /* run lengths: op i leads the ops i .. i + pad - 1 when they are all
     * butterflies of its kind and radix and none reads a value another of
     * them writes (the kernel loads such a run's operands together) */
    for (size_t i = 0; i < b.ops.size(); i++) {
        const SOp &o = b.ops[i];
        if (o.kind != CS_OP_BF2 && o.kind != CS_OP_BF4 && o.kind != CS_OP_BFP) continue;
        std::vector<int> outs(o.out);
        size_t j = i + 1;
        for (; j < b.ops.size() && j - i < 64; j++) {
            const SOp &n = b.ops[j];
            if (n.kind != o.kind || n.R != o.R) break;
            bool dep = false;
            for (int v : n.in)
                if (std::find(outs.begin(), outs.end(), v) != outs.end()) dep = true;
            if (dep) break;
            outs.insert(outs.end(), n.out.begin(), n.out.end());
        }
        P->ops[i].pad = (int)(j - i);
    }
    nres = 0;
    for (const Top &t : tops) {
        CsGenC2R c;
        c.H = t.H;
        c.XS = t.XS;
        c.op0 = t.op0;
        c.nop = t.nop;
        c.zout = nres;
        c.pad = 0;
        nres += t.H;
        P->c2r.push_back(c);
    }
    return 0;
}

/* ---- device ------------------------------------------------------------- */
namespace gk {

template <typename T>
struct cx {
    T x, y;
};

__host__ __device__ __forceinline__ float fm(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__host__ __device__ __forceinline__ double fm(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__host__ __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> w)
{
    const T p2 = a.y * w.y, p4 = a.y * w.x;
    return {fm(a.x, w.x, -p2), fm(a.x, w.y, p4)};
}

__host__ __device__ __forceinline__ float tmax(float m, float d) { return fmaxf(m, d); }
__host__ __device__ __forceinline__ double tmax(double m, double d) { return d > m ? d : m; }

constexpr int wseq(int i)
{
    int w = 1;
    for (int k = 0; k < i; k++) {
        int g = (int)(1.5 * (double)w);
        w = g > w + 1 ? g : w + 1;
    }
    return w;
}
constexpr int ncand_upto(int wc)
{
    int n = 0;
    while (n < CS_NSEQ && wseq(n) <= wc) n++;
    return n;
}

template <int... Is, class F>
__host__ __device__ __forceinline__ void sfor_(std::integer_sequence<int, Is...>, F &&f)
{
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__host__ __device__ __forceinline__ void sfor(F &&f)
{
    sfor_(std::make_integer_sequence<int, N>{}, f);
}

/* the lane's views of the work buffer: value slot s and stack entry h */
template <typename T>
struct Lane {
    cx<T> *V;
    float2 *X;
    long long stride;
    __host__ __device__ __forceinline__ cx<T> &v(int s) const { return V[(long long)s * stride]; }
    __host__ __device__ __forceinline__ float2 &x(int h) const { return X[(long long)h * stride]; }
};

template <typename T>
__host__ __device__ __forceinline__ T invsw_at(const ChunkDesc *C, int slot, int wi);
template <>
__host__ __device__ __forceinline__ float invsw_at<float>(const ChunkDesc *C, int slot, int wi)
{
    return C->invsw[slot][wi];
}
template <>
__host__ __device__ __forceinline__ double invsw_at<double>(const ChunkDesc *C, int slot, int wi)
{
    return C->invswd[slot][wi];
}

__device__ __forceinline__ void emit(bool pred, int chunk, int slot, int j, double metric,
                                     const GenArgs &a)
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

/* one value of a slot's profile: the base or a non-divisor decimation reads
 * its c2r's outputs, a divisor decimation k folds the base (s = x[j];
 * s = s + x[j + q*nk]; s * (1/k), src/search.c:298-331) */
template <typename T>
struct Profile {
    const Lane<T> *L;
    int zout;           /* output slots of the profile's c2r (the base for folds) */
    int fold;           /* k for a fold, 0 otherwise */
    int nk;
    T invk;
    __host__ __device__ __forceinline__ T raw(int i) const
    {
        const cx<T> z = L->v(zout + (i >> 1));
        return (i & 1) ? z.y : z.x;
    }
    __host__ __device__ __forceinline__ T operator()(int i) const
    {
        if (!fold) return raw(i);
        T s = raw(i);
        for (int q = 1; q < fold; q++) s = s + raw(i + q * nk);
        return s * invk;
    }
    /* values i and i + 1 for an even i from one load of each complex slot
     * they share: nk is even, so i + q*nk and i + 1 + q*nk are the two parts
     * of one slot for every q, and each sum keeps the order of
     * operator() */
    __host__ __device__ __forceinline__ void pair(int i, T &e, T &o) const
    {
        const cx<T> z = L->v(zout + (i >> 1));
        e = z.x;
        o = z.y;
        if (!fold) return;
        for (int q = 1; q < fold; q++) {
            const cx<T> w = L->v(zout + ((i + q * nk) >> 1));
            e = e + w.x;
            o = o + w.y;
        }
        e = e * invk;
        o = o * invk;
    }
};

/* Boxcar S/N of an n-bin profile over the enabled width candidates:
 * P[i] = P[i-1] + x[(i-1) mod n], m_w = max_{p=1..n} (P[p+w-1] - P[p-1]),
 * best over widths of (m_w - (w/n) * (P[n] - P[0])) * invsw_w.  The window
 * Pw[j] = P[p0 - 1 + j] holds B + WC + 1 prefix values; every control
 * decision depends only on n, p0 and the mask, which are the same in every
 * lane of the wave. */
template <int WC, int B, typename T, class PX>
__host__ __device__ __forceinline__ T boxcar_stream(const PX &px, int n, unsigned wmask,
                                           const ChunkDesc *C, int slot, int valid)
{
    constexpr int NC = ncand_upto(WC);
    constexpr int LEN = B + WC + 1;
    static_assert(B % 2 == 0, "blocks of whole value pairs");
    T Pw[LEN];
    T m[NC];
    sfor<NC>([&](auto I) { m[decltype(I)::value] = -(T)INFINITY; });
    Pw[0] = T(0);
    T Pn = T(0);
    int idx = 0;              /* (I - 1) mod n of the next prefix index I */
    /* n is even and idx starts at 0, so the c-th value read (from 0) has an
     * idx of c's parity: an even idx loads the pair (idx, idx + 1) and keeps
     * the odd value for the next read */
    T odd = T(0);
    auto next = [&](auto C2) -> T {
        if constexpr (decltype(C2)::value == 0) {
            T e;
            px.pair(idx, e, odd);
            return e;
        } else {
            return odd;
        }
    };
    sfor<LEN - 1>([&](auto J) {
        constexpr int j = decltype(J)::value + 1;
        Pw[j] = Pw[j - 1] + next(std::integral_constant<int, (j - 1) & 1>{});
        if (j == n) Pn = Pw[j];
        idx = idx + 1 == n ? 0 : idx + 1;
    });
    /* the prefetch's B registers cost no occupancy for f32 at WC 63 (two
     * waves per SIMD either way) and 211 (one); at 141 they would halve it,
     * and in f64 they spill */
    constexpr bool PF = sizeof(T) == 4 && (WC < 100 || WC > 200);
    for (int p0 = 1; p0 <= n; p0 += B) {
        /* the next block's pair loads issued before this block's windows
         * (read jj of a block has idx + jj's parity, that of LEN - 1 + jj,
         * and the even reads are jj / 2 = 0 .. B/2 - 1 in order) */
        const bool more = p0 + B <= n;
        cx<T> nx[B / 2];
        if (PF && more) {
            sfor<B>([&](auto J) {
                constexpr int jj = decltype(J)::value;
                if constexpr (((LEN - 1 + jj) & 1) == 0) {
                    int ii = idx + jj;
                    while (ii >= n) ii -= n;
                    px.pair(ii, nx[jj / 2].x, nx[jj / 2].y);
                }
            });
        }
        sfor<NC>([&](auto I) {
            constexpr int wi = decltype(I)::value;
            constexpr int W = wseq(wi);
            if (wmask & (1u << wi)) {
                sfor<B>([&](auto Bb) {
                    constexpr int b = decltype(Bb)::value;
                    if (p0 + b <= n) m[wi] = tmax(m[wi], Pw[b + W] - Pw[b]);
                });
            }
        });
        if (more) {
            sfor<WC + 1>([&](auto J) {
                constexpr int j = decltype(J)::value;
                Pw[j] = Pw[j + B];
            });
            sfor<B>([&](auto J) {
                constexpr int jj = decltype(J)::value;
                constexpr int j = jj + WC + 1;
                T v;
                if constexpr (!PF) {
                    v = next(std::integral_constant<int, (LEN - 1 + jj) & 1>{});
                } else if constexpr (((LEN - 1 + jj) & 1) == 0) {
                    v = nx[jj / 2].x;
                    odd = nx[jj / 2].y;
                } else {
                    v = odd;
                }
                Pw[j] = Pw[j - 1] + v;
                if (p0 + B - 1 + j == n) Pn = Pw[j];
                idx = idx + 1 == n ? 0 : idx + 1;
            });
        }
    }
    const T stot = Pn - T(0);
    T best = -(T)INFINITY;
    sfor<NC>([&](auto I) {
        constexpr int wi = decltype(I)::value;
        constexpr int W = wseq(wi);
        if (wmask & (1u << wi)) {
            const T duty = (T)W / (T)n;
            const T dt = duty * stot;
            const T c = (m[wi] - dt) * invsw_at<T>(C, slot, wi);
            best = c > best ? c : best;
        }
    });
    return valid ? best : T(0);
}

/* The fold of the base profile by a divisor decimation k (src/search.c:298-331)
 * into value slots zout .. zout + Hk - 1: bin pair j of the fold is
 * (sum_q z[j + q*Hk]) * (1/k), q = 0 .. k-1 in order, both parts summed as
 * the profile reader's operator() sums them.  32 pairs at a time, so each
 * term's loads are in flight together instead of one dependent load per
 * term and bin. */
template <typename T>
__host__ __device__ __forceinline__ void fold_into(const Lane<T> &L, int zin, int k, int Hk,
                                                   int zout)
{
    constexpr int U = 32;
    const T invk = T(1) / (T)k;
    for (int j0 = 0; j0 < Hk; j0 += U) {
        cx<T> acc[U];
        sfor<U>([&](auto I) {
            constexpr int u = decltype(I)::value;
            acc[u] = L.v(zin + (j0 + u < Hk ? j0 + u : Hk - 1));
        });
        for (int q = 1; q < k; q++) {
            cx<T> w[U];
            sfor<U>([&](auto I) {
                constexpr int u = decltype(I)::value;
                w[u] = L.v(zin + (j0 + u < Hk ? j0 + u : Hk - 1) + q * Hk);
            });
            sfor<U>([&](auto I) {
                constexpr int u = decltype(I)::value;
                acc[u] = {acc[u].x + w[u].x, acc[u].y + w[u].y};
            });
        }
        sfor<U>([&](auto I) {
            constexpr int u = decltype(I)::value;
            if (j0 + u < Hk) L.v(zout + j0 + u) = {acc[u].x * invk, acc[u].y * invk};
        });
    }
}

/* A radix-R butterfly op of the project c2r with its input twiddles (the
 * oracle's dft_ step), R = 2, 4 or a small odd prime, in two halves: the
 * operand loads (each times its twiddle) and the outputs. */
template <int R, typename T>
__host__ __device__ __forceinline__ void bf_load(const Lane<T> &L, const int *__restrict__ pl,
                                                 const cx<T> *__restrict__ tw, cx<T> (&u)[R])
{
    sfor<R>([&](auto I) {
        constexpr int r = decltype(I)::value;
        const cx<T> y = L.v(pl[r]);
        const int t = pl[R + r];
        u[r] = t < 0 ? y : cmul(y, tw[t]);
    });
}

template <int R, typename T>
__host__ __device__ __forceinline__ void bf_store(const Lane<T> &L, const int *__restrict__ pl,
                                                  const cx<T> *__restrict__ tw,
                                                  const cx<T> (&u)[R])
{
    if constexpr (R == 2) {
        L.v(pl[4]) = {u[0].x + u[1].x, u[0].y + u[1].y};
        L.v(pl[5]) = {u[0].x - u[1].x, u[0].y - u[1].y};
    } else if constexpr (R == 4) {
        const cx<T> t0 = {u[0].x + u[2].x, u[0].y + u[2].y};
        const cx<T> t1 = {u[0].x - u[2].x, u[0].y - u[2].y};
        const cx<T> t2 = {u[1].x + u[3].x, u[1].y + u[3].y};
        const cx<T> t3 = {u[1].x - u[3].x, u[1].y - u[3].y};
        L.v(pl[8]) = {t0.x + t2.x, t0.y + t2.y};
        L.v(pl[10]) = {t0.x - t2.x, t0.y - t2.y};
        L.v(pl[9]) = {t1.x - t3.y, t1.y + t3.x};
        L.v(pl[11]) = {t1.x + t3.y, t1.y - t3.x};
    } else {
        /* odd prime: sums and differences of mirrored inputs, then per
         * output pair A = u0 + sum sp_r c (fused, left to right) and
         * B = sum sm_r sn (first term a product, the rest fused) */
        constexpr int hh = (R - 1) / 2;
        const int rtw = pl[3 * R];
        cx<T> sp[hh + 1], sm[hh + 1];
        sfor<hh>([&](auto I) {
            constexpr int r = decltype(I)::value + 1;
            sp[r] = {u[r].x + u[R - r].x, u[r].y + u[R - r].y};
            sm[r] = {u[r].x - u[R - r].x, u[r].y - u[R - r].y};
        });
        cx<T> o[R];
        sfor<hh>([&](auto S) {
            constexpr int s = decltype(S)::value + 1;
            cx<T> A = u[0], Bv = {T(0), T(0)};
            sfor<hh>([&](auto I) {
                constexpr int r = decltype(I)::value + 1;
                constexpr int j = (r * s) % R;
                const cx<T> w = tw[rtw + j];
                A = {fm(sp[r].x, w.x, A.x), fm(sp[r].y, w.x, A.y)};
                if constexpr (r == 1) Bv = {sm[r].x * w.y, sm[r].y * w.y};
                else Bv = {fm(sm[r].x, w.y, Bv.x), fm(sm[r].y, w.y, Bv.y)};
            });
            o[s] = {A.x - Bv.y, A.y + Bv.x};
            o[R - s] = {A.x + Bv.y, A.y - Bv.x};
        });
        cx<T> x0 = u[0];
        sfor<hh>([&](auto I) {
            constexpr int r = decltype(I)::value + 1;
            x0 = {x0.x + sp[r].x, x0.y + sp[r].y};
        });
        o[0] = x0;
        sfor<R>([&](auto I) {
            constexpr int s = decltype(I)::value;
            L.v(pl[2 * R + s]) = o[s];
        });
    }
}

/* G consecutive radix-R ops from op i, none reading another's outputs (the
 * plan marks such runs): every operand loaded before any output is stored,
 * so the G ops' loads are in flight together.  An op's reads of a slot a
 * later op of the run reuses for its outputs still come first. */
template <int R, int G, typename T>
__host__ __device__ __forceinline__ void bf_group(const Lane<T> &L, const GenArgs &a, int i,
                                                  const cx<T> *__restrict__ tw)
{
    cx<T> u[G][R];
    const int *pls[G];
    sfor<G>([&](auto Gi) {
        constexpr int g = decltype(Gi)::value;
        pls[g] = a.pool + a.ops[i + g].pool;
        bf_load<R>(L, pls[g], tw, u[g]);
    });
    sfor<G>([&](auto Gi) {
        constexpr int g = decltype(Gi)::value;
        bf_store<R>(L, pls[g], tw, u[g]);
    });
}

/* the same for any odd prime, its operands staged in two scratch runs of the
 * lane's value slots past the plan's own (nslots .. nslots + 2R) */
template <typename T>
__host__ __device__ void bfp_any(const Lane<T> &L, int R, const int *__restrict__ pl,
                        const cx<T> *__restrict__ tw, int scratch)
{
    const int hh = (R - 1) / 2;
    /* u_r at scratch + r; sp_r at scratch + R + r, sm_r at scratch + R + hh + r */
    for (int r = 0; r < R; r++) {
        const cx<T> y = L.v(pl[r]);
        const int t = pl[R + r];
        L.v(scratch + r) = t < 0 ? y : cmul(y, tw[t]);
    }
    const int rtw = pl[3 * R];
    for (int r = 1; r <= hh; r++) {
        const cx<T> a = L.v(scratch + r), b = L.v(scratch + R - r);
        L.v(scratch + R + r) = {a.x + b.x, a.y + b.y};
        L.v(scratch + R + hh + r) = {a.x - b.x, a.y - b.y};
    }
    const cx<T> u0 = L.v(scratch);
    for (int s = 1; s <= hh; s++) {
        cx<T> A = u0, Bv = {T(0), T(0)};
        for (int r = 1; r <= hh; r++) {
            const cx<T> w = tw[rtw + (r * s) % R];
            const cx<T> sp = L.v(scratch + R + r), sm = L.v(scratch + R + hh + r);
            A = {fm(sp.x, w.x, A.x), fm(sp.y, w.x, A.y)};
            if (r == 1) Bv = {sm.x * w.y, sm.y * w.y};
            else Bv = {fm(sm.x, w.y, Bv.x), fm(sm.y, w.y, Bv.y)};
        }
        L.v(pl[2 * R + s]) = {A.x - Bv.y, A.y + Bv.x};
        L.v(pl[2 * R + R - s]) = {A.x + Bv.y, A.y - Bv.x};
    }
    cx<T> x0 = u0;
    for (int r = 1; r <= hh; r++) {
        const cx<T> sp = L.v(scratch + R + r);
        x0 = {x0.x + sp.x, x0.y + sp.y};
    }
    L.v(pl[2 * R]) = x0;
}

/* run op i and the ops it leads: up to the plan's run length (op.pad) of
 * radix-R ops in groups of 4, 2 or 1; returns the ops done */
template <int R, typename T>
__host__ __device__ __forceinline__ int bf_run(const Lane<T> &L, const GenArgs &a, int i, int run,
                                               const cx<T> *__restrict__ tw)
{
    if (run >= 4) {
        bf_group<R, 4>(L, a, i, tw);
        return 4;
    }
    if (run >= 2) {
        bf_group<R, 2>(L, a, i, tw);
        return 2;
    }
    bf_group<R, 1>(L, a, i, tw);
    return 1;
}

template <typename T>
__host__ __device__ __forceinline__ void run_c2r(const Lane<T> &L, const GenArgs &a, const CsGenC2R &c)
{
    constexpr int PB = 16;    /* pre-step outputs per batch of loads */
    const cx<T> *__restrict__ tw = (const cx<T> *)a.tw;
    for (int i = c.op0; i < c.op0 + c.nop;) {
        const CsGenOp op = a.ops[i];
        const int *__restrict__ pl = a.pool + op.pool;
        const int run = op.pad;
        if (op.kind == CS_OP_PRE) {
            const int H = op.R, XS = pl[0], twb = pl[1];
            const int *outs = pl + 2;
            {
                const T x0 = (T)L.x(0).x, xh = (T)L.x(H * XS).x;
                L.v(outs[0]) = {x0 + xh, x0 - xh};
            }
            /* k = 1 .. H-1 in batches: every stack read of a batch before
             * its stores (the stack and the value slots do not overlap) */
            for (int k0 = 1; k0 < H; k0 += PB) {
                float2 fa[PB], fb[PB];
                cx<T> w[PB];
                sfor<PB>([&](auto U) {
                    constexpr int u = decltype(U)::value;
                    const int k = k0 + u < H ? k0 + u : H - 1;
                    fa[u] = L.x(k * XS);
                    fb[u] = L.x((H - k) * XS);
                    w[u] = tw[twb + k - 1];
                });
                sfor<PB>([&](auto U) {
                    constexpr int u = decltype(U)::value;
                    const int k = k0 + u;
                    if (k < H) {
                        const cx<T> A = {(T)fa[u].x, (T)fa[u].y}, Bv = {(T)fb[u].x, (T)fb[u].y};
                        const T er = A.x + Bv.x, ei = A.y - Bv.y;
                        const cx<T> o = cmul(cx<T>{A.x - Bv.x, A.y + Bv.y}, w[u]);
                        L.v(outs[k]) = {er - o.y, ei + o.x};
                    }
                });
            }
            i++;
        } else if (op.kind == CS_OP_BF2) {
            i += bf_run<2>(L, a, i, run, tw);
        } else if (op.kind == CS_OP_BF4) {
            i += bf_run<4>(L, a, i, run, tw);
        } else if (op.kind == CS_OP_BFP) {
            switch (op.R) {
            case 3: i += bf_run<3>(L, a, i, run, tw); break;
            case 5: i += bf_run<5>(L, a, i, run, tw); break;
            case 7: bf_group<7, 1>(L, a, i, tw); i++; break;
            case 11: bf_group<11, 1>(L, a, i, tw); i++; break;
            case 13: bf_group<13, 1>(L, a, i, tw); i++; break;
            default: bfp_any(L, op.R, pl, tw, a.nslots); i++; break;
            }
        } else {
            L.v(pl[1]) = L.v(pl[0]);
            i++;
        }
    }
}

template <typename T, int WC, int B>
__global__ __launch_bounds__(CS_TILE) void general_search_kernel(GenArgs a)
{
    const long long lane_g = (long long)blockIdx.x * CS_TILE + threadIdx.x;
    Lane<T> L;
    L.stride = a.nlanes;
    L.V = (cx<T> *)a.work + lane_g;
    L.X = (float2 *)((char *)a.work + (size_t)a.nvslots * a.nlanes * sizeof(cx<T>)) + lane_g;
    const int q = a.q;
    const int tbeg = a.given ? a.tile0 + (int)blockIdx.x : (int)blockIdx.x;
    const int tend = a.given ? (blockIdx.x < (unsigned)a.nslice ? tbeg + 1 : tbeg) : a.ntiles;
    for (int tile = tbeg; tile < tend; tile += gridDim.x) {
        const int4 Tl = a.tiles[tile];
        const ChunkDesc *C = a.chunks + Tl.x;
        const long long i0 = C->i0;
        const int tid = threadIdx.x;
        const int G = a.grp;
        const int slot_l = tid < G * q ? (tid % G) * q + tid / G : tid;
        const bool active = slot_l < Tl.z;
        const int lj = Tl.y + (active ? slot_l : Tl.z - 1);
        const long long t = i0 + lj;
        const long long tp = t * a.pnum;
        const int rho = (int)(tp % q);
        const int qb = (int)(tp / q);

        /* 1. the harmonic stack (unless the stack pass wrote it) */
        L.x(0) = make_float2(0.0f, 0.0f);
        if (!a.given) {
            int res = 0, qi = 0;
            for (int h = 1; h <= a.nh; h++) {
                res += rho;
                qi += qb;
                if (res >= q) {
                    res -= q;
                    qi += 1;
                }
                const int2 hi = a.hinfo[h - 1];
                const int carry = res >= hi.y ? 1 : 0;
                const float *__restrict__ row =
                    a.wtab + (size_t)((h - 1) * q + res) * (size_t)a.ws;
                const float *__restrict__ W = row + (a.mp - a.m);
                const long long b0 = (long long)hi.x + qi + carry + 1 - a.m / 2;
                float ar = 0.0f, ai = 0.0f;
                for (int j = 0; j < a.m; j++) {
                    const long long bj = b0 + j;
                    float2 v = make_float2(0.0f, 0.0f);
                    if (bj >= 0 && bj < a.namps) v = a.amps[bj];
                    const float w = W[j];
                    ar = __builtin_fmaf(w, v.x, ar);
                    ai = __builtin_fmaf(w, v.y, ai);
                }
                const float Are = row[a.mp], Aim = row[a.mp + 1];
                const float p1 = Are * ar, p2 = Aim * ai, p3 = Are * ai, p4 = Aim * ar;
                const bool fit = (C->filled[(h - 1) >> 6] >> ((h - 1) & 63)) & 1ull;
                L.x(h) = make_float2(fit ? p1 - p2 : 0.0f, fit ? p3 + p4 : 0.0f);
            }
        }
        if (a.stack_out != nullptr) {
            const long long d = t - a.dump_t0;
            if (active && d >= 0 && d < a.dump_n)
                for (int h = 0; h <= a.nh; h++) a.stack_out[d * (a.nh + 1) + h] = L.x(h);
        }

        /* 2. the c2rs */
        for (int c = 0; c < a.nc2r; c++) run_c2r<T>(L, a, a.c2r[c]);

        /* 3-4. boxcars, records */
        const double thr = a.threshold;
        const int chunk = Tl.x;
        for (int s = 0; s < a.nslot; s++) {
            const int k = a.slot_k[s], Hk = a.slot_Hk[s], nb = 2 * Hk;
            Profile<T> px;
            px.L = &L;
            px.fold = 0;
            px.nk = nb;
            px.invk = T(1);
            if (a.slot_c2r[s] >= 0) {
                px.zout = a.c2r[a.slot_c2r[s]].zout;
            } else {
                fold_into<T>(L, a.c2r[0].zout, k, Hk, a.fslot);
                px.zout = a.fslot;
            }
            const int nv = s == 0 ? C->n : C->nvalid[s];
            if (a.samp) {
                /* measured sigma: |profile| of the sampled trials
                 * (src/boxcar.c:18-73, 138-167) */
                if (active && lj < nv) {
                    long long pos = -1;
                    if ((long long)nb * nv <= CS_SIGCAP) {
                        pos = (long long)lj * nb;
                    } else {
                        long long ncol = CS_SIGCAP / nb;
                        if (ncol < 1) ncol = 1;
                        if (ncol > nv) ncol = nv;
                        const long long cc = ((long long)lj * ncol + nv - 1) / nv;
                        if (cc < ncol && (cc * nv) / ncol == lj) pos = cc * nb;
                    }
                    if (pos >= 0) {
                        T *out = (T *)a.samp + ((long long)chunk * a.nslot + s) * CS_SIGCAP + pos;
                        for (int i = 0; i < nb; i++) {
                            const T v = px(i);
                            out[i] = v < T(0) ? -v : v;
                        }
                    }
                }
                continue;
            }
            const T mv = boxcar_stream<WC, B, T>(px, nb, a.wmask[s], C, s, C->valid[s]);
            const bool ok = active && lj < nv;
            if (a.mout) {
                if (ok) ((T *)a.mout)[(long long)s * a.mstride + i0 + lj] = mv;
            } else {
                emit(ok && (double)mv > thr, chunk, s, lj, (double)mv, a);
            }
        }
    }
}

}  // namespace gk

extern "C" __attribute__((visibility("hidden"))) hipError_t cs_gen_launch(const GenArgs *a,
                                                                          bool f64,
                                                                          hipStream_t st)
{
    const int grid = (int)(a->nlanes / CS_TILE);
    int wmax = 0;
    for (int s = 0; s < a->nslot; s++)
        for (int wi = 0; wi < CS_NSEQ; wi++)
            if (a->wmask[s] & (1u << wi)) wmax = std::max(wmax, gk::wseq(wi));
    const dim3 g((unsigned)grid), b(CS_TILE);
    if (wmax <= 63) {
        if (f64) hipLaunchKernelGGL((gk::general_search_kernel<double, 63, 32>), g, b, 0, st, *a);
        else hipLaunchKernelGGL((gk::general_search_kernel<float, 63, 32>), g, b, 0, st, *a);
    } else if (wmax <= 141) {
        if (f64) hipLaunchKernelGGL((gk::general_search_kernel<double, 141, 32>), g, b, 0, st, *a);
        else hipLaunchKernelGGL((gk::general_search_kernel<float, 141, 32>), g, b, 0, st, *a);
    } else if (wmax <= 211) {
        if (f64) hipLaunchKernelGGL((gk::general_search_kernel<double, 211, 16>), g, b, 0, st, *a);
        else hipLaunchKernelGGL((gk::general_search_kernel<float, 211, 32>), g, b, 0, st, *a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
