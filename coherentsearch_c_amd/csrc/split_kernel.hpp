/*
 * Two-pass search for stacks too large for one lane's registers next to
 * their profiles (above 64 harmonics, or m above 32).
 *
 * The one-pass kernel (fast_kernel.hpp) keeps a trial's whole stack and
 * profiles in registers.  Above 64 harmonics that takes about 440 registers,
 * one wave per SIMD, and its interpolation loop -- LDS window reads and
 * weight-row gathers -- then has nothing to hide its latency behind.  Here
 * the work is split over a slice of tiles at a time:
 *   1. stack pass: per lane one trial, per workgroup one tile; the harmonics
 *      in groups whose spectrum windows fit a small LDS buffer, each group's
 *      windows staged coalesced, then the m-tap fmaf chain in tap order and
 *      the unfused rotation (src/interp.c:686-811) -- the one-pass kernel's
 *      arithmetic op for op, in few registers, so many waves per SIMD hide
 *      the loads.  Each harmonic's value is stored to the slice's stack
 *      buffer in HBM, coalesced across the wave;
 *   2. profile pass: per lane the same trial's stack read back into
 *      registers, then the one-pass kernel's c2r, folds, boxcar and output
 *      (search_tail), pure VALU work at one wave per SIMD.
 * The stack buffer holds one slice; slices run in turn on the stream.  The
 * stack pass also feeds the general kernel (cs_general.hip), writing into its
 * work buffer's stack entries.
 */
#pragma once

#include "fast_kernel.hpp"

namespace csk {

/* window bins a stack-pass lane stages per memory round trip */
constexpr int kSplitStage = 12;

struct SplitArgs {
    float2 *stk;          /* stack of the slice's tiles */
    long long tstride;    /* float2 elements from one tile's stack to the next */
    long long hstride;    /* float2 elements from harmonic h to h + 1 */
    int hbase;            /* harmonic h (1-based) at (h - 1 + hbase) * hstride */
    int tile0;            /* first tile of the slice */
    int nslice;           /* tiles in the slice */
    int nh;               /* harmonics */
    int hg;               /* harmonics per LDS window group, 1..64 */
    int m;                /* taps (the non-rational stack pass) */
    double lodr;          /* trial step (the non-rational stack pass) */
};

/* The slice-local tile of this workgroup: the one-pass kernel's XCD-aware
 * deal (workgroup b on XCD b % 8; the grid is a multiple of 8, of 8*S with
 * stripes of S tiles). */
__device__ __forceinline__ int slice_tile(const FastArgs &a)
{
    const int nblk = gridDim.x;
    if (a.xcd_stripe > 0) {
        const int S = a.xcd_stripe, i = blockIdx.x >> 3;
        return ((i / S) * 8 + (blockIdx.x & 7)) * S + i % S;
    }
    return (blockIdx.x & 7) * (nblk >> 3) + (blockIdx.x >> 3);
}

/* lane -> row of the tile: residue-class order over the first grp*q lanes */
__device__ __forceinline__ int lane_row(const FastArgs &a, int tid)
{
    const int G = a.grp;
    return tid < G * a.q ? (tid % G) * a.q + tid / G : tid;
}

/* Stack pass for M padded taps (rows of CS_WSP(M) floats, the first at
 * base + 1 - m2e); dynamic LDS holds one group's windows.  Each lane gathers
 * its residue's weight rows from the L2-resident table: the texture units
 * move 16 bytes per lane per load whether or not the lanes' rows coincide,
 * so staging a harmonic's rows through LDS moves as many bytes (q rows for
 * 256 lanes) and adds a barrier per harmonic -- measured 2.2x slower. */
template <int M>
__global__ __launch_bounds__(CS_TILE) void stack_pass_kernel(FastArgs a, SplitArgs s)
{
    static_assert(M % 4 == 0, "tap count multiple of 4");
    extern __shared__ float2 win[];
    __shared__ int s_off[65];
    __shared__ int s_lo[64];
    __shared__ int s_C[64];
    __shared__ int s_hy[64];                  /* first residue with a carry, per harmonic */
    __shared__ unsigned long long s_fit;      /* the group's filled-harmonic bits */

    const int u = slice_tile(a);
    if (u >= s.nslice) return;
    const int4 Tl = a.tiles[s.tile0 + u];
    const ChunkDesc *C = a.chunks + Tl.x;
    const long long i0 = C->i0;
    const int tid = threadIdx.x;
    const int slot_l = lane_row(a, tid);
    const bool active = slot_l < Tl.z;
    const int lj = Tl.y + (active ? slot_l : Tl.z - 1);
    const long long t = i0 + lj;
    const int q = a.q;
    const int m2 = a.m2e;

    /* first and last trial of the tile: t*pnum = q*qf + rf */
    const long long af = (i0 + Tl.y) * a.pnum, al = (i0 + Tl.y + Tl.z - 1) * a.pnum;
    const int rf = (int)(af % q), rl = (int)(al % q);
    const long long qf = af / q, ql = al / q;
    /* this lane: t*pnum = q*qb + rho, walked per harmonic with adds only */
    const long long tp = t * a.pnum;
    const int rho = (int)(tp % q);
    const int qb = (int)(tp / q);

    float2 *out = s.stk + (long long)u * s.tstride + tid;
    constexpr int WS4 = CS_WSP(M) / 4;
    const char *__restrict__ wbase = reinterpret_cast<const char *>(a.wtab);
    int res = 0, qi = 0;
    for (int g0 = 0; g0 < s.nh; g0 += s.hg) {
        const int gn = min(s.hg, s.nh - g0);
        /* the group's windows: harmonic h covers the bins of the tile's
         * first to last trial plus the taps; outside the spectrum zero */
        if (tid < gn) {
            const int h = g0 + tid + 1;
            const int2 hi = a.hinfo[h - 1];
            const int hrf = h * rf, hrl = h * rl;
            const int resf = hrf % q, resl = hrl % q;
            const long long bf = hi.x + (long long)h * qf + hrf / q + (resf >= hi.y ? 1 : 0);
            const long long bl = hi.x + (long long)h * ql + hrl / q + (resl >= hi.y ? 1 : 0);
            const int lo = (int)(bf + 1 - m2);
            s_lo[tid] = lo;
            s_C[tid] = hi.x + 1 - m2 - lo;
            s_hy[tid] = hi.y;
            s_off[tid + 1] = (int)(bl - bf) + M;
        }
        if (tid == 0) {
            s_off[0] = 0;
            /* bits g0 .. g0 + gn - 1 of the chunk's mask (a group never
             * straddles more than two 64-bit words) */
            const int w = g0 >> 6, b = g0 & 63;
            unsigned long long f = C->filled[w] >> b;
            if (b && w + 1 < 8) f |= C->filled[w + 1] << (64 - b);
            s_fit = gn < 64 ? f & ((1ull << gn) - 1ull) : f;
        }
        __syncthreads();
        if (tid < 64) {
            int v = tid < gn ? s_off[tid + 1] : 0;
            for (int d = 1; d < 64; d <<= 1) {
                const int w = __shfl_up(v, d);
                if (tid >= d) v += w;
            }
            if (tid < gn) s_off[tid + 1] = v;
        }
        __syncthreads();
        if (tid < gn) s_C[tid] += s_off[tid];
        const int cnt = s_off[gn];
        /* kSplitStage bins per thread in flight (a group's windows up to
         * 24 KB in one memory round trip): every load issued before the LDS
         * stores; each bin's harmonic by a fixed-step binary search (gn <=
         * 64), so that the searches interleave */
        for (int base = tid; base < cnt; base += kSplitStage * CS_TILE) {
            float2 v[kSplitStage];
            sfor<kSplitStage>([&](auto U) {
                constexpr int uu = decltype(U)::value;
                const int idx = base + uu * CS_TILE;
                v[uu] = make_float2(0.0f, 0.0f);
                int lo = 0;
#pragma unroll
                for (int st = 32; st >= 1; st >>= 1) {
                    const int c = min(lo + st, gn - 1);
                    lo = s_off[c] <= idx ? c : lo;
                }
                const long long ai = (long long)s_lo[lo] + (idx - s_off[lo]);
                if (idx < cnt && ai >= 0 && ai < a.namps) v[uu] = a.amps[ai];
            });
            sfor<kSplitStage>([&](auto U) {
                constexpr int uu = decltype(U)::value;
                const int idx = base + uu * CS_TILE;
                if (idx < cnt) win[idx] = v[uu];
            });
        }
        __syncthreads();
        const unsigned long long fitm = s_fit;
        /* two harmonics per step, both rows' loads issued before either
         * tap chain (the walk advances for the second only when it exists:
         * it is the group's last harmonic otherwise) */
        for (int i = 0; i < gn; i += 2) {
            const bool two = i + 1 < gn;
            int wi[2];
            unsigned wrow[2];
            sfor<2>([&](auto U) {
                constexpr int uu = decltype(U)::value;
                const int ii = i + uu;
                if (uu == 0 || two) {
                    res += rho;
                    qi += qb;
                    if (res >= q) { res -= q; qi += 1; }
                }
                const int ic = uu == 0 || two ? ii : i;
                const int carry = res >= s_hy[ic] ? 1 : 0;
                wi[uu] = s_C[ic] + qi + carry;
                wrow[uu] = (unsigned)((g0 + ic) * q + res) * (unsigned)(16 * WS4);
            });
            float4 w[2][M / 4];
            float2 Aw[2];
            sfor<2>([&](auto U) {
                constexpr int uu = decltype(U)::value;
                const float4 *__restrict__ wr =
                    reinterpret_cast<const float4 *>(wbase + wrow[uu]);
                sfor<M / 4>([&](auto J) { w[uu][decltype(J)::value] = wr[decltype(J)::value]; });
                Aw[uu] = reinterpret_cast<const float2 *>(wr + M / 4)[0];
            });
            float2 o[2];
            sfor<2>([&](auto U) {
                constexpr int uu = decltype(U)::value;
                float ar = 0.0f, ai = 0.0f;
                sfor<M / 4>([&](auto J) {
                    constexpr int j = 4 * decltype(J)::value;
                    const float4 w4 = w[uu][j / 4];
                    const float2 a0 = win[wi[uu] + j], a1 = win[wi[uu] + j + 1];
                    const float2 a2 = win[wi[uu] + j + 2], a3 = win[wi[uu] + j + 3];
                    ar = __builtin_fmaf(w4.x, a0.x, ar); ai = __builtin_fmaf(w4.x, a0.y, ai);
                    ar = __builtin_fmaf(w4.y, a1.x, ar); ai = __builtin_fmaf(w4.y, a1.y, ai);
                    ar = __builtin_fmaf(w4.z, a2.x, ar); ai = __builtin_fmaf(w4.z, a2.y, ai);
                    ar = __builtin_fmaf(w4.w, a3.x, ar); ai = __builtin_fmaf(w4.w, a3.y, ai);
                });
                const float2 A = Aw[uu];
                const float p1 = A.x * ar, p2 = A.y * ai, p3 = A.x * ai, p4 = A.y * ar;
                const bool fit = (fitm >> (i + uu)) & 1ull;
                o[uu] = make_float2(fit ? p1 - p2 : 0.0f, fit ? p3 + p4 : 0.0f);
            });
            out[(long long)(g0 + i + s.hbase) * s.hstride] = o[0];
            if (two) out[(long long)(g0 + i + 1 + s.hbase) * s.hstride] = o[1];
        }
        __syncthreads();
    }
}

/* Profile pass: the stack of the slice's tile back into registers, then
 * everything the one-pass kernel does after its interpolation. */
template <int NH, int KMAX, typename T>
__global__ __launch_bounds__(CS_TILE) __attribute__((amdgpu_waves_per_eu(1)))
void profile_pass_kernel(FastArgs a, SplitArgs s)
{
    const int u = slice_tile(a);
    if (u >= s.nslice) return;
    const int4 Tl = a.tiles[s.tile0 + u];
    const ChunkDesc *C = a.chunks + Tl.x;
    const long long i0 = C->i0;
    const int tid = threadIdx.x;
    const int slot_l = lane_row(a, tid);
    const bool active = slot_l < Tl.z;
    const int lj = Tl.y + (active ? slot_l : Tl.z - 1);
    const long long t = i0 + lj;

    const float2 *__restrict__ in = s.stk + (long long)u * s.tstride + tid;
    cf X[NH + 1];
    X[0] = {0.0f, 0.0f};
    sfor<NH>([&](auto I) {
        constexpr int h = decltype(I)::value + 1;
        const float2 v = in[(long long)(h - 1 + s.hbase) * s.hstride];
        X[h] = {v.x, v.y};
    });
    search_tail<NH, KMAX, T>(X, a, C, Tl.x, i0, lj, t, active);
}

}  // namespace csk

/* Profile-pass instantiation + launch helper (f32 profiles, decimations up
 * to KMAX = 16), one or more per translation unit. */
#define CS_SPLIT_INSTANTIATE(NH)                                                     \
    extern "C" __attribute__((visibility("hidden"))) hipError_t                       \
    cs_launch_profile_##NH(const csk::FastArgs *a, const csk::SplitArgs *s,          \
                           int nblocks, hipStream_t st)                              \
    {                                                                                \
        hipLaunchKernelGGL((csk::profile_pass_kernel<NH, 16, float>), dim3(nblocks), \
                           dim3(CS_TILE), 0, st, *a, *s);                            \
        return hipGetLastError();                                                    \
    }
