/*
 * De-dispersion of raw filterbank spectra into HBM-resident DM series
 * (include/cs_dedisp.h): the reference prep front end's hot loop, PRESTO
 * prepsubband's subband de-dispersion (src/prep_dedisp.c:1104-1427;
 * third_party/presto/src/dispersion.c:216-345, cuda.cu:505-768).
 *
 * The reference walks the data in work blocks of `worklen` spectra, each
 * de-dispersed against the next one (a two-block window), because its
 * readers return one block at a time.  The arithmetic is a continuous
 * formula:
 *   sub[s][t]  = sum over the channels c of subband s, ascending, of
 *                raw[t + idispdt[c]][c]                          (from 0.0f)
 *   subds[s][u] = (sum_{k<ds} sub[s][u*ds + k]) / ds             (ds > 1)
 *   dm[i][u]   = -0.0f + sum over s ascending of subds[s][u + off[i][s]]
 * for u below the length the reference writes, then padding with the running
 * mean of DM 0's written samples.  This file computes that formula in device
 * passes of `chunk` spectra (a multiple of the raw block and at least one
 * work block, sized to fill the GPU), so the pass size is the device's
 * choice, not the reader's: the sums, their order and the written length are
 * the reference's, the results bit-identical.
 *
 * Passes stream through two rings: raw spectra (three chunks: the pass reads
 * chunks j and j+1 while chunk j+2 is copied in on a second stream; skipped
 * when the spectra are already in HBM) and downsampled subbands (two chunks:
 * a DM pass reads its chunk and the next one).  The DM series go straight to
 * their resident [numdms][numout] array.
 *
 * Kernels: HBM-bound.  subband_kernel reads every raw sample once
 * (numchan * sizeof(T) bytes per spectrum) and writes nsub floats per
 * downsampled sample; dm_kernel reads nsub floats per output from the
 * subband ring (L2-resident) and writes one float per DM and sample.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "cs_dedisp.h"

#define DCHK(x)                                                                    \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            fprintf(stderr, "Error: HIP %s failed at %s:%d: %s\n", #x, __FILE__,   \
                    __LINE__, hipGetErrorString(e_));                              \
            goto fail;                                                             \
        }                                                                          \
    } while (0)

namespace {

constexpr int kWaves = 4;              /* waves per workgroup: one subband or DM each */
constexpr long long kChunkTarget = 1 << 20;   /* raw spectra per pass */
constexpr long long kRingBytes = 1ll << 30;    /* raw ring (host input) */
constexpr long long kSubBytes = 1ll << 29;     /* subband ring */

/* PRESTO's NEAREST_LONG: halves away from zero (src/prep_dedisp.c:37) */
long nearest_long(double x)
{
    return (long)(x < 0 ? std::ceil(x - 0.5) : std::floor(x + 0.5));
}

double delay_from_dm(double dm, double f)
{
    return f == 0.0 ? 0.0 : dm / (0.000241 * f * f);
}

/* delays of n channels of width w from lofreq up; the doppler factor
 * 1 + v/c is exactly 1 with v/c = 0 (dispersion.c:100-121) */
void chan_delays(int n, double dm, double lofreq, double w, double *out)
{
    for (int i = 0; i < n; i++) out[i] = delay_from_dm(dm, lofreq + i * w);
}

/* every subband's highest channel (dispersion.c:131-150) */
void subband_delays(int numchan, int nsub, double dm, double lofreq, double cw, double *out)
{
    const double sbw = cw * (numchan / nsub);
    chan_delays(nsub, dm, lofreq + sbw - cw, sbw, out);
}

bool smooth11(long long n)
{
    for (int f = 2; f <= 11; f++)
        while (n % f == 0) n /= f;
    return n == 1;
}

bool power_of_10(long long n)
{
    while (n > 9 && n % 10 == 0) n /= 10;
    return n == 1;
}

struct Tables {
    int bpr = 0, worklen = 0, maxoff = 0, maxd = 0;
    long long numout = 0, written = 0;
    double dsdt = 0.0;
    std::vector<double> dms;
    std::vector<int> idispdt, offsets;
};

int build_tables(const CsDedispConfig &c, long long nspec, Tables &t)
{
    const int nch = c.numchan, nsub = c.nsub, nd = c.numdms, ds = c.downsamp, P = c.spectra_per_subint;
    if (nch <= 0 || nsub <= 0 || nd <= 0 || ds <= 0 || P <= 0 || !(c.dt > 0.0)) {
        fprintf(stderr, "Error: de-dispersion needs positive channels, subbands, DMs, "
                        "downsampling, block length and dt\n");
        return -1;
    }
    if (nsub > nch || nch % nsub) {
        fprintf(stderr, "Error: %d subbands do not divide %d channels\n", nsub, nch);
        return -1;
    }
    if (P % ds) {
        fprintf(stderr, "Error:  The downsample factor (%d) must be a factor of the\n"
                        "        blocklength (%d).\n", ds, P);
        return -1;
    }
    if (nspec < P || nspec % P) {
        fprintf(stderr, "Error: %lld spectra are not a whole number of %d-spectrum blocks\n",
                nspec, P);
        return -1;
    }
    t.dms.resize(nd);
    double avgdm = 0.0;
    for (int i = 0; i < nd; i++) {
        t.dms[i] = c.lodm + i * c.dmstep;
        avgdm += t.dms[i];
    }
    avgdm /= nd;
    const double maxdm = t.dms[nd - 1];
    t.dsdt = ds * c.dt;
    const double bw = delay_from_dm(maxdm, c.lofreq) -
                      delay_from_dm(maxdm, c.lofreq + (nch - 1) * c.chan_wid);
    t.bpr = (int)(bw / c.dt) / P + 1;
    t.worklen = P * t.bpr;
    t.numout = c.numout > 0 ? c.numout : cs_choose_good_N(nspec / ds);
    std::vector<double> cd(nch), sd(nsub);
    const int cps = nch / nsub;
    subband_delays(nch, nsub, avgdm, c.lofreq, c.chan_wid, sd.data());
    chan_delays(nch, avgdm, c.lofreq, c.chan_wid, cd.data());
    for (int s = 0; s < nsub; s++)
        for (int j = 0; j < cps; j++) cd[s * cps + j] -= sd[s];
    t.idispdt.resize(nch);
    t.maxd = 0;
    for (int ch = 0; ch < nch; ch++) {
        const long d = nearest_long(cd[ch] / c.dt);
        if (d < 0 || d > P) {
            fprintf(stderr, "Error: there is more dispersion across a subband than time\n"
                            "in a block of data (channel %d: %ld samples, block %d).\n",
                    ch, d, P);
            return -1;
        }
        t.idispdt[ch] = (int)d;
        t.maxd = std::max(t.maxd, (int)d);
    }
    const int dsW = t.worklen / ds;
    t.offsets.resize((size_t)nd * nsub);
    t.maxoff = 0;
    for (int i = 0; i < nd; i++) {
        subband_delays(nch, nsub, t.dms[i], c.lofreq, c.chan_wid, sd.data());
        const double top = sd[nsub - 1];
        for (int s = 0; s < nsub; s++) {
            const long o = nearest_long((sd[s] - top) / t.dsdt);
            if (o < 0 || o > dsW) {
                fprintf(stderr, "Error: subband offset %ld of DM %g outside [0, %d]\n", o,
                        t.dms[i], dsW);
                return -1;
            }
            t.offsets[(size_t)i * nsub + s] = (int)o;
            t.maxoff = std::max(t.maxoff, (int)o);
        }
    }
    /* The reference writes work block m only once work block m + 1 was
     * read whole, and its first read takes two raw blocks: m runs while
     * (m + 2) * worklen + P <= nspec (src/prep_dedisp.c:485-526,
     * 1104-1290; backend_common.c:1706-1737). */
    long long nblocks = (nspec - P) / t.worklen - 1;
    if (nblocks < 0) nblocks = 0;
    t.written = std::min(nblocks * dsW, t.numout);
    return 0;
}

template <typename T>
__device__ __forceinline__ float sample(const T *p)
{
    return (float)*p;
}

/* Time-domain clipping as the subband kernels see it (clip_times replaces a
 * raw row whose zero-DM sum is more than trig from the running mean by the
 * block's channel levels): flag == nullptr means no clipping. */
struct ClipView {
    const unsigned char *flag; /* 1 for a clipped spectrum (absolute rows) */
    const float *levels;   /* [block][numchan]: the channel levels */
    long long row0;        /* absolute row of the pass's chunk start */
    int P;                 /* raw block length */
};

/* the clip test of every spectrum of raw blocks [row0 / P, (row0 + n) / P),
 * once: |zero-DM sum - running mean| > clip_sigma x running std */
__global__ void __launch_bounds__(256)
clip_flag_kernel(const float *__restrict__ zdm, const float *__restrict__ ravg,
                 const float *__restrict__ trig, long long row0, long long n, int P,
                 unsigned char *__restrict__ flag)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const long long row = row0 + i, b = row / P;
    flag[row] = fabsf(zdm[row] - ravg[b]) > trig[b] ? 1 : 0;
}

/* One downsampled subband sample per lane: wave w of the workgroup takes
 * subband blockIdx.y * kWaves + w, lanes take consecutive samples u.  ra/rb:
 * the pass's raw chunk and the next one (chunk spectra each, [t][numchan]). */
template <typename T>
__global__ void __launch_bounds__(64 * kWaves)
subband_kernel(const T *__restrict__ ra, const T *__restrict__ rb, long long chunk, int nchan,
               int cps, int nsub, int ds, const int *__restrict__ idisp, int nu,
               float *__restrict__ out, int dsc, ClipView cv)
{
    const int s = __builtin_amdgcn_readfirstlane((int)(blockIdx.y * kWaves + (threadIdx.x >> 6)));
    const int u = (int)(blockIdx.x * 64 + (threadIdx.x & 63));
    if (s >= nsub || u >= nu) return;
    const int c0 = s * cps;
    float acc = 0.0f;
    for (int k = 0; k < ds; k++) {
        const long long t0 = (long long)u * ds + k;
        float v = 0.0f;
        for (int j = 0; j < cps; j++) {
            const int ch = c0 + j;
            const long long t = t0 + idisp[ch];
            const T *row = t < chunk ? ra + t * nchan : rb + (t - chunk) * nchan;
            const long long at = cv.row0 + t;
            v += cv.flag && cv.flag[at] ? cv.levels[at / cv.P * nchan + ch] : sample(row + ch);
        }
        acc = ds > 1 ? acc + v : v;
    }
    out[(long long)s * dsc + u] = ds > 1 ? acc / (float)ds : acc;
}

/* One DM sample per lane: wave w takes DM blockIdx.y * kWaves + w.  sa/sb:
 * the pass's subband chunk and the next ([s][dsc] each). */
__global__ void __launch_bounds__(64 * kWaves)
dm_kernel(const float *__restrict__ sa, const float *__restrict__ sb, int dsc, int nsub,
          const int *__restrict__ offs, int numdms, int nu, float *__restrict__ out,
          long long numout, long long u0)
{
    const int i = __builtin_amdgcn_readfirstlane((int)(blockIdx.y * kWaves + (threadIdx.x >> 6)));
    const int u = (int)(blockIdx.x * 64 + (threadIdx.x & 63));
    if (i >= numdms || u >= nu) return;
    const int *o = offs + (long long)i * nsub;
    float acc = -0.0f;
    for (int s = 0; s < nsub; s++) {
        const int t = u + o[s];
        acc += t < dsc ? sa[(long long)s * dsc + t] : sb[(long long)s * dsc + (t - dsc)];
    }
    out[(long long)i * numout + u0 + u] = acc;
}

/* ---- LDS-tiled kernels -------------------------------------------------- */

constexpr int kRows = 256;            /* raw rows per subband tile: one per lane */
constexpr int kTileBytes = 32 * 1024; /* LDS budget of a subband tile */
constexpr int kLoadBatch = 8;          /* tile loads a lane keeps in flight */
constexpr int kSpanSub = 4;            /* DM tiles: subband spans loaded per step */
constexpr int kSpanLoads = 4;          /* DM tiles: samples per lane per span in flight */
constexpr int kTU = 512;              /* DM samples per DM tile: two per lane, 256 apart */
constexpr int kDmThreads = 256;
constexpr int kGMax = 32;             /* DMs per DM tile (register accumulators) */
constexpr int kSpanFloats = 8192;     /* LDS floats of one subband batch of a DM tile */
constexpr int kDmWorkgroups = 2048;   /* DM tiles per pass worth keeping the GPU busy with */

template <int VEC>
struct Vec;
template <>
struct Vec<16> {
    using type = uint4;
};
template <>
struct Vec<4> {
    using type = uint32_t;
};
template <>
struct Vec<1> {
    using type = uint8_t;
};

template <int VEC>
__device__ __forceinline__ void tile_store(unsigned char *dst, typename Vec<VEC>::type w)
{
    if constexpr (VEC == 16) {       /* rows are 4-byte aligned only: four dword stores */
        uint32_t *d = (uint32_t *)dst;
        d[0] = w.x;
        d[1] = w.y;
        d[2] = w.z;
        d[3] = w.w;
    } else {
        *(typename Vec<VEC>::type *)dst = w;
    }
}

/* Subbands from an LDS tile: the workgroup stages raw rows [t0, t0 + 256 +
 * maxd) of one group of gs subbands (row pitch an odd number of dwords, so
 * the lanes' rows fall in distinct banks), then lane r sums row t0 + r of
 * each subband over its channels in ascending order, and the ds lanes of one
 * downsampled sample add their sums in order (lane k = 0 writes).  rows:
 * raw rows the pass reads (from the chunk start; at most its samples' rows
 * plus the deepest delay, so at most two chunks).  VEC: bytes per
 * global load (16 when every row segment is 16-byte aligned). */
template <typename T, int VEC, bool CLIP>
__global__ void __launch_bounds__(kRows)
subband_tiled(const T *__restrict__ ra, const T *__restrict__ rb, long long chunk, int nchan,
              int cps, int gs, int nsub, int ds, const int *__restrict__ idisp, int maxd,
              int pitch, long long rows, int nu, float *__restrict__ out, int dsc, ClipView cv,
              int ngrp, int ntile)
{

    extern __shared__ __align__(16) unsigned char tile[];
    using W = typename Vec<VEC>::type;
    /* XCD-aware order: workgroup b runs on XCD b % 8, and the grid (a
     * multiple of 8) gives each XCD a contiguous run of (row tile, subband
     * group) pairs, groups fastest.  A row tile's groups and the next row
     * tile (whose first rows are this one's delay rows) then share one L2:
     * the row segments two groups share and the delay rows are read from
     * HBM once instead of once per XCD. */
    const int wg = (int)(blockIdx.x & 7) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
    if (wg >= ngrp * ntile) return;
    const int gi = wg % ngrp, ti = wg / ngrp;
    const int s0 = gi * gs, ns = min(gs, nsub - s0);
    const int c0 = s0 * cps;
    const long long t0 = (long long)ti * kRows;
    const int nr = kRows + maxd;
    /* with clipping: the rows' clip flags in LDS after the tile; a clipped
     * row's channel levels replace its samples where the sums read them */
    unsigned char *rflag = tile + (size_t)nr * pitch;
    int anyclip = 0;
    if constexpr (CLIP) {
        int mine = 0;
        for (int r = (int)threadIdx.x; r < nr; r += kRows) {
            const long long t = t0 + r;
            const unsigned char f = t < rows ? cv.flag[cv.row0 + t] : 0;
            rflag[r] = f;
            mine |= f;
        }
        /* the tile's rows all unclipped (the usual case): the plain sums */
        anyclip = __syncthreads_or(mine);
    }
    {
        /* loads per row nv; the lanes take rpi rows per step, the division by
         * nv done once (per load it would cost some 30 VALU instructions) */
        const int nv = ns * cps * (int)sizeof(T) / VEC;
        const int rpi = nv <= kRows ? kRows / nv : 0;
        if (rpi) {
            const int lr = (int)threadIdx.x / nv, v = (int)threadIdx.x - lr * nv;
            /* batches of kLoadBatch loads in flight before their LDS stores:
             * one memory round trip per batch instead of one per row */
            /* row t's segment sits at ra + t*rowb + col while t < chunk and
             * at rb + (t - chunk)*rowb + col after: per load one offset add
             * of a scalar step and a choice of base, instead of a 64-bit
             * multiply.  Rows are compared as 32-bit tile rows: r < rlim
             * exists, r < rch lies in chunk ra. */
            const long long rowb = (long long)nchan * (long long)sizeof(T);
            const long long col = (long long)c0 * (long long)sizeof(T) + (long long)v * VEC;
            const long long step = (long long)rpi * rowb, cb = chunk * rowb;
            const int rlim = (int)min((long long)nr, rows - t0);
            const int rch = (int)max(-1ll, min((long long)nr, chunk - t0));
            const unsigned char *rab = (const unsigned char *)ra, *rbb = (const unsigned char *)rb;
            if (lr < rpi)
                for (int rb0 = lr; rb0 < nr; rb0 += kLoadBatch * rpi) {
                    W w[kLoadBatch];
                    const long long o0 = (t0 + rb0) * rowb + col;
#pragma unroll
                    for (int i = 0; i < kLoadBatch; i++) {
                        const int r = rb0 + i * rpi;
                        w[i] = W{};
                        if (r < rlim) {
                            const long long o = o0 + (long long)i * step;
                            w[i] = *(const W *)(r < rch ? rab + o : rbb + (o - cb));
                        }
                    }
#pragma unroll
                    for (int i = 0; i < kLoadBatch; i++) {
                        const int r = rb0 + i * rpi;
                        if (r < nr) tile_store<VEC>(tile + (size_t)r * pitch + (size_t)v * VEC, w[i]);
                    }
                }
        } else {
            for (int idx = (int)threadIdx.x; idx < nr * nv; idx += kRows) {
                const int r = idx / nv, v = idx - r * nv;
                const long long t = t0 + r;
                W w{};
                if (t < rows) {
                    const T *row = t < chunk ? ra + t * nchan : rb + (t - chunk) * nchan;
                    w = *(const W *)((const unsigned char *)(row + c0) + (size_t)v * VEC);
                }
                tile_store<VEC>(tile + (size_t)r * pitch + (size_t)v * VEC, w);
            }
        }
    }
    __syncthreads();
    const int r = (int)threadIdx.x;
    const int k = r % ds, base = (int)(threadIdx.x & 63) - k;
    const long long u = (t0 + r) / ds;
    /* lane r's row in the tile; a channel's sample sits d*pitch + ch*size
     * further on, a uniform offset (scalar arithmetic), so each sample costs
     * one vector add for its address */
    const float rds = 1.0f / (float)ds;
    int lb = r * pitch;
    /* opaque to the optimiser, which would otherwise fold lb + d*pitch back
     * into (r + d)*pitch: a quarter-rate vector multiply per sample */
    asm volatile("" : "+v"(lb));
    /* the sums, with or without the clip test (the workgroup takes one
     * version: its loop stays as tight as the unclipped kernel's) */
    auto sums = [&](auto withclip) {
        for (int s = 0; s < ns; s++) {
            float v = 0.0f;
            if (sizeof(T) == 1 && !decltype(withclip)::value && cps <= 65535) {
                /* 8-bit samples: every partial sum of the float loop is an
                 * integer below 255 * 65536 < 2^24, so each float addition is
                 * exact and the sum equals the integer sum: accumulate in
                 * integers and convert once (the same bits) */
                unsigned vi = 0;
                for (int j = 0; j < cps; j++) {
                    const int ch = s * cps + j;
                    vi += (unsigned)*(const T *)(tile + lb + idisp[c0 + ch] * pitch + ch);
                }
                v = (float)vi;
            } else {
                for (int j = 0; j < cps; j++) {
                    const int ch = s * cps + j;
                    const int d = idisp[c0 + ch];
                    float x = (float)*(const T *)(tile + lb + (d * pitch + ch * (int)sizeof(T)));
                    if constexpr (decltype(withclip)::value) {
                        if (rflag[r + d]) {
                            const long long b = (cv.row0 + t0 + r + d) / cv.P;
                            x = cv.levels[b * nchan + c0 + ch];
                        }
                    }
                    v += x;
                }
            }
            float acc = v;
            if (ds > 1) {
                acc = 0.0f;
                for (int q = 0; q < ds; q++) acc += __shfl(v, base + q, 64);
                /* ds divides 64 here, a power of two: the quotient by ds
                 * and the product by its exact reciprocal round the same
                 * real number */
                acc *= rds;
            }
            if (k == 0 && u < nu) out[(long long)(s0 + s) * dsc + u] = acc;
        }
    };
    if (CLIP && __builtin_amdgcn_readfirstlane(anyclip))
        sums(std::true_type{});
    else
        sums(std::false_type{});
}

/* DM samples from LDS spans: the workgroup takes 512 consecutive samples of
 * a group of G consecutive DMs (lane l: samples l and l + 256, 2G
 * accumulators; the two reads of one offset share an address and issue as
 * one ds_read2st64).  Subbands go through LDS in batches of sb: subband s's span covers
 * the samples every DM of the group reads, [u0 + omin_s, u0 + 256 + omax_s);
 * each accumulator adds its subbands in ascending order across the batches.
 * gmin[g][s]: omin_s; gbase[g][s]: the span's LDS offset within its batch;
 * sofs[g][s][i] = gbase - omin + off[g * G + i][s] (the group's G offsets
 * of one subband contiguous: one vector scalar load).  tlim: subband samples of the pass
 * that exist (beyond: never read for a written sample). */
template <int G>
__global__ void __launch_bounds__(kDmThreads)
dm_tiled(const float *__restrict__ sa, const float *__restrict__ sb, int dsc, int nsub,
         int sbatch, const int *__restrict__ gmin, const int *__restrict__ glen,
         const int *__restrict__ gbase, const int *__restrict__ sofs, int numdms, int nu,
         int tlim, float *__restrict__ out, long long numout, long long v0, int ngrp,
         int ntile)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float *span = (float *)smem;
    /* XCD-aware order (the grid a multiple of 8, workgroup b on XCD b % 8):
     * each XCD takes a contiguous run of (sample tile, DM group) pairs,
     * groups fastest, so the groups of one sample tile -- which read the
     * same subband samples -- share an L2 */
    const int wg = (int)(blockIdx.x & 7) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
    if (wg >= ngrp * ntile) return;
    const int g = wg % ngrp, i0 = g * G, ng = min(G, numdms - i0);
    const int u0 = (wg / ngrp) * kTU, tid = (int)threadIdx.x;
    const int *mn = gmin + (size_t)g * nsub, *ln = glen + (size_t)g * nsub,
              *bs = gbase + (size_t)g * nsub;
    float acc[G], acc2[G];
#pragma unroll
    for (int i = 0; i < G; i++) acc[i] = acc2[i] = -0.0f;
    for (int sA = 0; sA < nsub; sA += sbatch) {
        const int sZ = min(nsub, sA + sbatch);
        if (sA) __syncthreads();
        /* kSpanSub subbands at a time, each lane's first kSpanLoads samples
         * of each loaded before any is stored: one memory round trip per
         * step instead of one per sample (longer spans finish in a loop) */
        for (int s0 = sA; s0 < sZ; s0 += kSpanSub) {
            float v[kSpanSub][kSpanLoads];
            int lj[kSpanSub], bj[kSpanSub], tj[kSpanSub];
#pragma unroll
            for (int j = 0; j < kSpanSub; j++) {
                const int s = s0 + j;
                lj[j] = s < sZ ? ln[s] : 0;
                bj[j] = s < sZ ? bs[s] : 0;
                tj[j] = s < sZ ? u0 + mn[s] : 0;
            }
#pragma unroll
            for (int j = 0; j < kSpanSub; j++) {
                const size_t so = (size_t)(s0 + j) * dsc;
#pragma unroll
                for (int k = 0; k < kSpanLoads; k++) {
                    const int q = tid + k * kDmThreads, t = tj[j] + q;
                    v[j][k] = 0.0f;
                    if (q < lj[j] && t < tlim) v[j][k] = t < dsc ? sa[so + t] : sb[so + (t - dsc)];
                }
            }
#pragma unroll
            for (int j = 0; j < kSpanSub; j++)
#pragma unroll
                for (int k = 0; k < kSpanLoads; k++) {
                    const int q = tid + k * kDmThreads;
                    if (q < lj[j]) span[bj[j] + q] = v[j][k];
                }
            for (int j = 0; j < kSpanSub; j++) {
                const size_t so = (size_t)(s0 + j) * dsc;
                for (int q = tid + kSpanLoads * kDmThreads; q < lj[j]; q += kDmThreads) {
                    const int t = tj[j] + q;
                    float x = 0.0f;
                    if (t < tlim) x = t < dsc ? sa[so + t] : sb[so + (t - dsc)];
                    span[bj[j] + q] = x;
                }
            }
        }
        __syncthreads();
        /* every accumulator, unguarded: a group's missing DMs have offset 0
         * (a valid span index) and are never stored, so the G offsets of a
         * subband come in one vector scalar load and the G reads issue
         * back to back */
        for (int s = sA; s < sZ; s++) {
            const int *so = sofs + ((size_t)g * nsub + s) * G;
            /* all 2G reads first (in flight together), then the adds */
            float x[G], y[G];
#pragma unroll
            for (int i = 0; i < G; i++) {
                const float *p = span + so[i] + tid;
                x[i] = p[0];
                y[i] = p[kDmThreads];
            }
#pragma unroll
            for (int i = 0; i < G; i++) {
                acc[i] += x[i];
                acc2[i] += y[i];
            }
        }
    }
    const int u = u0 + tid;
#pragma unroll
    for (int i = 0; i < G; i++)
        if (i < ng) {
            float *o = out + (size_t)(i0 + i) * numout + v0;
            if (u < nu) o[u] = acc[i];
            if (u + kDmThreads < nu) o[u + kDmThreads] = acc2[i];
        }
}

/* ---- time-domain clipping: clip_times of every raw block (clipping.c) ----
 * A block's statistics depend on its own samples only; the running state
 * (the 0.9/0.1 updates of the zero-DM mean, std and channel levels) is a
 * scan over blocks.  So: zero-DM sums and the exact lower median per block
 * (a workgroup each), the block's AS 52 statistics (one lane each, in
 * sample order), the channel means of its good rows (one lane per channel,
 * in row order), then the scan (one workgroup).  The subband kernels then
 * read a clipped row's channel levels in place of its samples. */

constexpr int kClipRcpMax = 4096;     /* block lengths whose reciprocals fit an LDS table */

struct BlockStat {
    double avg, std;       /* of the good rows' zero-DM sums */
    float lo, hi;          /* the good range, median -/+ 3 std of all rows */
    int ngood, pad_;
};

struct ClipState {
    float running_avg, running_std;
    int blocksread, pad_;
};

__device__ __forceinline__ unsigned order_key(float f)
{
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float key_float(unsigned k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

/* ascending bitonic sort of n2 (a power of two) keys in LDS by a workgroup
 * of 256 lanes; the keys are in place when every lane returns */
__device__ __forceinline__ void bitonic_sort_256(unsigned *keys, int n2)
{
    for (int kk = 2; kk <= n2; kk <<= 1)
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int i = (int)threadIdx.x; i < n2; i += 256) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const unsigned a = keys[i], c = keys[ixj];
                    if ((a > c) == ((i & kk) == 0)) {
                        keys[i] = c;
                        keys[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
}

/* LDS bytes of clip_zdm_kernel: the row tile, or the keys of the sort */
template <typename T>
constexpr int zdm_pitch()
{
    return (sizeof(T) == 1 ? 64 : 32) * (int)sizeof(T) + 4;      /* 17 or 33 dwords */
}
template <typename T>
static size_t zdm_smem(int P)
{
    int n2 = 1;
    while (n2 < P) n2 <<= 1;
    const size_t tile = 256 * (size_t)zdm_pitch<T>(), keys = P <= 8192 ? (size_t)n2 * 4 : 0;
    return tile > keys ? tile : keys;
}

/* zero-DM sums (float, channel order) and their element of rank (P-1)/2 by
 * an 8-bit radix select; one raw block per workgroup.  The sums go through
 * LDS: 256 rows x 64 (float: 32) channels at a time are loaded coalesced (rows
 * padded to an odd number of dwords), then lane r adds row r's channels in order,
 * its sum carried from one channel chunk to the next.  Where every row loads
 * as 16-byte vectors, the next tile's loads are in flight (in registers)
 * while the lanes add the current one.  LDS is sized to the block (dynamic),
 * so blocks of 2048 rows take 17 KB and eight share a CU. */
template <typename T>
__global__ void __launch_bounds__(256)
clip_zdm_kernel(const T *__restrict__ chunk, long long blk0, int P, int nch, int vec16,
                float *__restrict__ zdm, float *__restrict__ med)
{
    __shared__ unsigned hist[256];
    __shared__ unsigned s_prefix, s_k;
    constexpr int CW = sizeof(T) == 1 ? 64 : 32;             /* channels per chunk */
    constexpr int PITCH = zdm_pitch<T>();
    constexpr int NSORT = 8192;                              /* blocks sorted in LDS */
    /* the row tile of the sums, then the keys of the sort */
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned char *rows_ = smem;
    const int b = (int)blockIdx.x;
    const T *blk = chunk + (long long)b * P * nch;
    float *z = zdm + (blk0 + b) * P;
    if (vec16 && nch % CW == 0) {
        /* 16 bytes a lane: the address arithmetic of one load serves 16
         * (float: 4) samples; four dword stores, as the padded rows are
         * 4-byte aligned.  Tiles in order (row chunk, then channel chunk). */
        constexpr int NV = CW * (int)sizeof(T) / 16, RPI = 256 / NV, NL = 256 / RPI;
        const int lr = (int)threadIdx.x / NV, lv = (int)threadIdx.x % NV;
        const int nc0 = nch / CW, nt = (P + 255) / 256 * nc0;
        uint4 w[NL];
        auto load = [&](int t) {
            const int i0 = t / nc0 * 256, c0 = t % nc0 * CW, nrow = min(256, P - i0);
            /* rows past the block's end load its last row (unconditional
             * loads, which the waits can count) and are not stored */
#pragma unroll
            for (int k = 0; k < NL; k++) {
                const int r = min(lr + k * RPI, nrow - 1);
                w[k] = *(const uint4 *)((const unsigned char *)(blk + (long long)(i0 + r) * nch +
                                                                c0) + lv * 16);
            }
        };
        load(0);
        float sum = 0.0f;
        for (int t = 0; t < nt; t++) {
            const int i0 = t / nc0 * 256, nrow = min(256, P - i0);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < NL; k++) {
                const int r = lr + k * RPI;
                if (r < nrow) {
                    uint32_t *dst = (uint32_t *)(rows_ + r * PITCH + lv * 16);
                    dst[0] = w[k].x;
                    dst[1] = w[k].y;
                    dst[2] = w[k].z;
                    dst[3] = w[k].w;
                }
            }
            __syncthreads();
            if (t + 1 < nt) load(t + 1);
            if (t % nc0 == 0) sum = 0.0f;
            if ((int)threadIdx.x < nrow) {
                const unsigned char *row = rows_ + threadIdx.x * PITCH;
#pragma unroll 16
                for (int e = 0; e < CW; e++) sum += (float)*(const T *)(row + e * (int)sizeof(T));
                if (t % nc0 == nc0 - 1) z[i0 + threadIdx.x] = sum;
            }
        }
    } else {
        for (int i0 = 0; i0 < P; i0 += 256) {
            const int nrow = min(256, P - i0);
            float sum = 0.0f;
            for (int c0 = 0; c0 < nch; c0 += CW) {
                const int cw = min(CW, nch - c0);
                __syncthreads();
                for (int idx = (int)threadIdx.x; idx < nrow * CW; idx += 256) {
                    const int r = idx / CW, e = idx % CW;
                    if (e < cw)
                        *(T *)(rows_ + r * PITCH + e * (int)sizeof(T)) =
                            blk[(long long)(i0 + r) * nch + c0 + e];
                }
                __syncthreads();
                if ((int)threadIdx.x < nrow) {
                    const unsigned char *row = rows_ + threadIdx.x * PITCH;
                    for (int e = 0; e < cw; e++) sum += (float)*(const T *)(row + e * (int)sizeof(T));
                }
            }
            if ((int)threadIdx.x < nrow) z[i0 + threadIdx.x] = sum;
        }
    }
    /* the median: blocks of up to 8192 spectra by a bitonic sort of their
     * order-preserving keys in LDS (sums of nearby values would pile the
     * radix select's histogram atomics onto one bin) */
    if (P <= NSORT) {
        unsigned *keys = (unsigned *)smem;
        int n2 = 1;
        while (n2 < P) n2 <<= 1;
        __syncthreads();
        for (int i = (int)threadIdx.x; i < n2; i += 256)
            keys[i] = i < P ? order_key(z[i]) : 0xFFFFFFFFu;
        __syncthreads();
        bitonic_sort_256(keys, n2);
        if (threadIdx.x == 0) med[blk0 + b] = key_float(keys[(P - 1) / 2]);
        return;
    }
    __syncthreads();
    unsigned prefix = 0, mask = 0, k = (unsigned)(P - 1) / 2;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[threadIdx.x] = 0;
        __syncthreads();
        for (int i = (int)threadIdx.x; i < P; i += 256) {
            const unsigned key = order_key(z[i]);
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned acc = 0;
            int d = 0;
            for (; d < 255; d++) {
                if (acc + hist[d] > k) break;
                acc += hist[d];
            }
            s_prefix = prefix | ((unsigned)d << shift);
            s_k = k - acc;
        }
        __syncthreads();
        prefix = s_prefix;
        k = s_k;
        mask |= 255u << shift;
        __syncthreads();
    }
    if (threadIdx.x == 0) med[blk0 + b] = key_float(prefix);
}

/* AS 52 mean and variance in double, in sample order (misc_utils.c:641) */
__device__ __forceinline__ void as52_step(double x, int &n, double &m, double &v, double &an1)
{
    if (n == 0) {
        m = x;
    } else {
        const double an = (double)(n + 1);
        an1 = (double)n;
        const double dx = (x - m) / an;
        v += an * an1 * dx * dx;
        m += dx;
    }
    n++;
}

/* d / a correctly rounded, given y = RN(1/a): q = RN(d y) is within an ulp
 * of d / a, the residual d - q a is exact, and RN(q + r y) is the correctly
 * rounded quotient (Markstein, IBM J. Res. Dev. 34 (1990) 111-119; the
 * theorem needs y within half an ulp of 1/a and no underflow, which the
 * zero-DM sums and divisors up to a block length give).  Three dependent
 * operations where a division takes a dozen. */
__device__ __forceinline__ double div_rcp(double d, double a, double y)
{
    const double q = d * y;
    const double r = __builtin_fma(-q, a, d);
    return __builtin_fma(r, y, q);
}

/* as52_step after the first sample (n >= 1); y = RN(1 / (n + 1)) */
__device__ __forceinline__ void as52_next(double x, int &n, double &m, double &v, double &an1,
                                          double y)
{
    const double an = (double)(n + 1);
    an1 = (double)n;
    const double dx = div_rcp(x - m, an, y);
    v += an * an1 * dx * dx;
    m += dx;
    n++;
}

/* as52_step of x where take holds, else nothing; by selects, so a run of
 * steps has no branch (the arithmetic of a taken step is as52_step's);
 * y = RN(1 / (n + 1)) */
__device__ __forceinline__ void as52_if(bool take, double x, int &n, double &m, double &v,
                                        double &an1, double y)
{
    const double an = (double)(n + 1), a1 = (double)n;
    const double dx = div_rcp(x - m, an, y);
    const double vn = v + an * a1 * dx * dx, mn = m + dx;
    const bool upd = take && n > 0;
    m = take ? (n > 0 ? mn : x) : m;
    v = upd ? vn : v;
    an1 = upd ? a1 : an1;
    n += take ? 1 : 0;
}

/* one lane per block: the statistics of all rows, the good range, and the
 * statistics of the good rows, in sample order.  Every division of AS 52
 * is by a step count, so the workgroup first tabulates RN(1/k) for k up to
 * the block length in LDS; the first pass reads its reciprocal as one
 * broadcast, the second (whose count advances on good rows only) counts its
 * good rows 16 at a time and gathers their reciprocals ahead of the
 * recurrence.  The sums are loaded 16 at a time ahead of their steps, and
 * the first step of the first pass (the mean := the first sum) is peeled
 * off. */
__global__ void __launch_bounds__(64) clip_seq_kernel(const float *__restrict__ zdm,
                                                      const float *__restrict__ med,
                                                      long long blk0, int nblk, int P,
                                                      BlockStat *__restrict__ bs)
{
    __shared__ double rcp[kClipRcpMax + 1];
    for (int k = (int)threadIdx.x + 1; k <= P; k += 64) rcp[k] = 1.0 / (double)k;
    __syncthreads();
    const int b = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= nblk) return;
    const float *z = zdm + (blk0 + b) * P;
    constexpr int G = 16;
    int n = 1;
    double m = (double)z[0], v = 0.0, an1 = 0.0;
    int i = 1;
    for (; i + G <= P; i += G) {
        float xs[G];
#pragma unroll
        for (int q = 0; q < G; q++) xs[q] = z[i + q];
#pragma unroll
        for (int q = 0; q < G; q++) as52_next((double)xs[q], n, m, v, an1, rcp[i + q + 1]);
    }
    for (; i < P; i++) as52_next((double)z[i], n, m, v, an1, rcp[i + 1]);
    if (n > 1) v /= an1;
    const double sd = sqrt(v);
    const float md = med[blk0 + b];
    const float lo = md - 3.0 * sd, hi = md + 3.0 * sd;
    int g = 0;
    double gm = 0.0, gv = 0.0, gan1 = 0.0;
    for (i = 0; i + G <= P; i += G) {
        float xs[G];
#pragma unroll
        for (int q = 0; q < G; q++) xs[q] = z[i + q];
        bool tk[G];
        double ys[G];
        int gg = g;
#pragma unroll
        for (int q = 0; q < G; q++) {
            tk[q] = xs[q] > lo && xs[q] < hi;
            ys[q] = rcp[gg + 1];
            gg += tk[q] ? 1 : 0;
        }
#pragma unroll
        for (int q = 0; q < G; q++) as52_if(tk[q], (double)xs[q], g, gm, gv, gan1, ys[q]);
    }
    for (; i < P; i++) {
        const float x = z[i];
        as52_if(x > lo && x < hi, (double)x, g, gm, gv, gan1, rcp[g + 1]);
    }
    if (g > 1) gv /= gan1;
    BlockStat r;
    r.avg = gm;
    r.std = sqrt(gv);
    r.lo = lo;
    r.hi = hi;
    r.ngood = g;
    r.pad_ = 0;
    bs[blk0 + b] = r;
}

/* one lane per block, dividing as written, for blocks longer than the table */
__global__ void clip_seq_lane_kernel(const float *__restrict__ zdm, const float *__restrict__ med,
                                     long long blk0, int nblk, int P, BlockStat *__restrict__ bs)
{
    const int b = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= nblk) return;
    const float *z = zdm + (blk0 + b) * P;
    int n = 0;
    double m = 0.0, v = 0.0, an1 = 0.0;
    for (int i = 0; i < P; i++) as52_step((double)z[i], n, m, v, an1);
    if (n > 1) v /= an1;
    const double sd = sqrt(v);
    const float md = med[blk0 + b];
    const float lo = md - 3.0 * sd, hi = md + 3.0 * sd;
    int g = 0;
    double gm = 0.0, gv = 0.0, gan1 = 0.0;
    for (int i = 0; i < P; i++) {
        const float x = z[i];
        if (x > lo && x < hi) as52_step((double)x, g, gm, gv, gan1);
    }
    if (g > 1) gv /= gan1;
    BlockStat r;
    r.avg = gm;
    r.std = sqrt(gv);
    r.lo = lo;
    r.hi = hi;
    r.ngood = g;
    r.pad_ = 0;
    bs[blk0 + b] = r;
}

/* one lane per (channel, block): the channel's mean over the block's good
 * rows, summed in double in row order; the rows' good flags go through LDS
 * 4096 at a time */
template <typename T>
__global__ void __launch_bounds__(256)
clip_chavg_kernel(const T *__restrict__ chunk, long long blk0, int P, int nch,
                  const float *__restrict__ zdm, const BlockStat *__restrict__ bs,
                  double *__restrict__ chavg)
{
    __shared__ unsigned char good[4096];
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x), b = (int)blockIdx.y;
    const BlockStat st = bs[blk0 + b];
    const T *blk = chunk + (long long)b * P * nch;
    const float *z = zdm + (blk0 + b) * P;
    double acc = 0.0;
    for (int i0 = 0; i0 < P; i0 += 4096) {
        const int n = min(4096, P - i0);
        __syncthreads();
        for (int i = (int)threadIdx.x; i < n; i += 256) {
            const float x = z[i0 + i];
            good[i] = x > st.lo && x < st.hi;
        }
        __syncthreads();
        /* loads unconditional (in flight together), adds of 0.0 for bad
         * rows: acc starts at +0.0 and never becomes -0.0, so adding +0.0
         * leaves it unchanged, as skipping the row does */
        if (c < nch) {
            const T *col = blk + (long long)i0 * nch + c;
#pragma unroll 8
            for (int i = 0; i < n; i++) {
                const double x = (double)(float)col[(long long)i * nch];
                acc += good[i] ? x : 0.0;
            }
        }
    }
    if (c < nch) chavg[(blk0 + b) * nch + c] = st.ngood > 0 ? acc / st.ngood : 0.0;
}

/* clip_chavg_kernel with four channels a lane: one 4-byte (float: 16-byte)
 * load per row, masked to zero on bad rows (the same +0.0 addends), each
 * channel's mean still summed in double in row order.  The rows' masks go
 * through LDS 2048 at a time. */
template <typename T>
__global__ void __launch_bounds__(256)
clip_chavg4_kernel(const T *__restrict__ chunk, long long blk0, int P, int nch,
                   const float *__restrict__ zdm, const BlockStat *__restrict__ bs,
                   double *__restrict__ chavg)
{
    using V = typename std::conditional<sizeof(T) == 1, uint32_t, uint4>::type;
    __shared__ uint32_t gmask[2048];
    const int c = 4 * (int)(blockIdx.x * blockDim.x + threadIdx.x), b = (int)blockIdx.y;
    const BlockStat st = bs[blk0 + b];
    const T *blk = chunk + (long long)b * P * nch;
    const float *z = zdm + (blk0 + b) * P;
    const int cc = min(c, nch - 4);             /* lanes past the channels load the last four */
    const V *col = (const V *)(blk + cc);
    const long long vstride = nch / 4;          /* V per row */
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i0 = 0; i0 < P; i0 += 2048) {
        const int n = min(2048, P - i0);
        __syncthreads();
        for (int i = (int)threadIdx.x; i < n; i += 256) {
            const float x = z[i0 + i];
            gmask[i] = (x > st.lo && x < st.hi) ? 0xFFFFFFFFu : 0u;
        }
        __syncthreads();
        const V *row = (const V *)((const T *)col + (long long)i0 * nch);
#pragma unroll 8
        for (int i = 0; i < n; i++) {
            const uint32_t g = gmask[i];
            if constexpr (sizeof(T) == 1) {
                const uint32_t w = row[(long long)i * vstride] & g;
#pragma unroll
                for (int e = 0; e < 4; e++) acc[e] += (double)((w >> (8 * e)) & 0xFFu);
            } else {
                const uint4 w = row[(long long)i * vstride];
                acc[0] += (double)__uint_as_float(w.x & g);
                acc[1] += (double)__uint_as_float(w.y & g);
                acc[2] += (double)__uint_as_float(w.z & g);
                acc[3] += (double)__uint_as_float(w.w & g);
            }
        }
    }
    if (c < nch) {
        double *o = chavg + (blk0 + b) * nch + c;
#pragma unroll
        for (int e = 0; e < 4; e++) o[e] = st.ngood > 0 ? acc[e] / st.ngood : 0.0;
    }
}

/* uint8 spectra: every sum clip_times takes of samples is an integer -- a
 * spectrum's zero-DM sum below 2^24 (numchan <= 4096), a channel's block sum
 * below 2^32 -- so its float and double additions are exact in any order.
 * One read of a block then gives its zero-DM sums, their lower median (an
 * integer sort in LDS: the order of non-negative floats), and every
 * channel's sum over all the block's rows; clip_badrows_u8_kernel takes the
 * bad rows' samples off for the means of the good rows.  NV = numchan / 16
 * lanes per spectrum (16 bytes each), a power of two up to 256; a row's
 * lanes add their byte sums by shuffles (and, past one wave, LDS atomics). */
template <int NV>
__global__ void __launch_bounds__(256)
clip_zdm_u8_kernel(const uint8_t *__restrict__ chunk, long long blk0, int P, int nch,
                   float *__restrict__ zdm, float *__restrict__ med, uint32_t *__restrict__ chsum)
{
    extern __shared__ uint32_t zsm[];    /* n2 row sums (the sort's keys), then nch channel sums */
    constexpr int RPI = NV >= 256 ? 1 : 256 / NV, U = 4, GL = NV < 64 ? NV : 64;
    const int b = (int)blockIdx.x, tid = (int)threadIdx.x;
    const uint8_t *blk = chunk + (long long)b * P * nch;
    int n2 = 1;
    while (n2 < P) n2 <<= 1;
    uint32_t *cs = zsm + n2;
    for (int i = tid; i < nch; i += 256) cs[i] = 0;
    if constexpr (NV > 64)
        for (int i = tid; i < P; i += 256) zsm[i] = 0;
    __syncthreads();
    const int lr = tid / NV, lv = tid % NV;
    uint32_t acc[16];
#pragma unroll
    for (int j = 0; j < 16; j++) acc[j] = 0;
    for (int r0 = 0; r0 < P; r0 += RPI * U) {
        /* U rows a lane in flight; rows past the block load its last row,
         * which is not counted */
        uint4 w[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const int r = min(r0 + q * RPI + lr, P - 1);
            w[q] = *(const uint4 *)(blk + (long long)r * nch + lv * 16);
        }
#pragma unroll
        for (int q = 0; q < U; q++) {
            const int r = r0 + q * RPI + lr;
            const bool ok = r < P;
            const uint32_t keep = ok ? 0xFFFFFFFFu : 0u;
            const uint32_t wv[4] = {w[q].x, w[q].y, w[q].z, w[q].w};
            uint32_t sum = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                sum = __builtin_amdgcn_sad_u8(wv[k], 0u, sum);
                const uint32_t x = wv[k] & keep;
#pragma unroll
                for (int e = 0; e < 4; e++) acc[4 * k + e] += (x >> (8 * e)) & 0xFFu;
            }
#pragma unroll
            for (int o = 1; o < GL; o <<= 1) sum += __shfl_xor(sum, o, 64);
            if constexpr (NV <= 64) {
                if (lv == 0 && ok) zsm[r] = sum;
            } else {
                if ((lv & 63) == 0 && ok) atomicAdd(&zsm[r], sum);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 16; j++) atomicAdd(&cs[lv * 16 + j], acc[j]);
    __syncthreads();
    for (int i = tid; i < nch; i += 256) chsum[(blk0 + b) * nch + i] = cs[i];
    float *z = zdm + (blk0 + b) * P;
    for (int i = tid; i < n2; i += 256) {
        if (i < P)
            z[i] = (float)zsm[i];
        else
            zsm[i] = 0xFFFFFFFFu;
    }
    __syncthreads();
    bitonic_sort_256(zsm, n2);
    if (tid == 0) med[blk0 + b] = (float)zsm[(P - 1) / 2];
}

constexpr int kBadWords = 8192 / 64;    /* bad-row masks of a block of up to 8192 spectra */

/* the channel means over each block's good rows from the all-row sums: the
 * bad rows (zero-DM sum outside (lo, hi)) are marked 64 at a time by ballot,
 * then every lane walks the marks (the same for all lanes) and takes each bad
 * row's samples off its four channels, four rows' loads at a time */
__global__ void __launch_bounds__(256)
clip_badrows_u8_kernel(const uint8_t *__restrict__ chunk, long long blk0, int P, int nch,
                       const float *__restrict__ zdm, const BlockStat *__restrict__ bs,
                       const uint32_t *__restrict__ chsum, double *__restrict__ chavg)
{
    __shared__ unsigned long long bm[kBadWords];
    const int b = (int)blockIdx.x, tid = (int)threadIdx.x;
    const BlockStat st = bs[blk0 + b];
    const float *z = zdm + (blk0 + b) * P;
    const int nw = (P + 63) / 64;
    for (int w0 = 0; w0 < nw; w0 += 4) {
        const int wi = w0 + tid / 64, i = wi * 64 + (tid & 63);
        bool bad = false;
        if (i < P) {
            const float x = z[i];
            bad = !(x > st.lo && x < st.hi);
        }
        const unsigned long long m = __ballot(bad);
        if (wi < nw && (tid & 63) == 0) bm[wi] = m;
    }
    __syncthreads();
    const uint8_t *blk = chunk + (long long)b * P * nch;
    for (int c = 4 * tid; c < nch; c += 1024) {
        uint32_t sb[4] = {0, 0, 0, 0};
        if (st.ngood > 0) {
            int wi = 0;
            unsigned long long m = bm[0];
            for (;;) {
                int idx[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    while (!m && wi + 1 < nw) m = bm[++wi];
                    idx[q] = m ? wi * 64 + __builtin_ctzll(m) : -1;
                    m &= m - 1;
                }
                if (idx[0] < 0) break;
                uint32_t w[4];
#pragma unroll
                for (int q = 0; q < 4; q++)
                    w[q] = idx[q] >= 0 ? *(const uint32_t *)(blk + (long long)idx[q] * nch + c) : 0u;
#pragma unroll
                for (int q = 0; q < 4; q++)
#pragma unroll
                    for (int e = 0; e < 4; e++) sb[e] += (w[q] >> (8 * e)) & 0xFFu;
            }
        }
        const uint32_t *csum = chsum + (blk0 + b) * nch + c;
        double *o = chavg + (blk0 + b) * nch + c;
#pragma unroll
        for (int e = 0; e < 4; e++)
            o[e] = st.ngood > 0 ? (double)(csum[e] - sb[e]) / st.ngood : 0.0;
    }
}

/* one block of clip_times' running state (clipping.c): the zero-DM mean
 * and std, and one channel's level, each 0.9 old + 0.1 new from the second
 * block on; a block without good rows carries the old values */
__device__ __forceinline__ void scan_step(const BlockStat &s, double chmean, int &nread,
                                          float &ra, float &rs, float &cr)
{
    const bool none = s.ngood < 1;
    const double ca = none ? (double)ra : s.avg, cs = none ? (double)rs : s.std;
    const double cv = none ? (double)cr : chmean;
    if (nread) {
        ra = 0.9 * ra + 0.1 * ca;
        rs = 0.9 * rs + 0.1 * cs;
        cr = (float)(0.9 * cr + 0.1 * cv);
    } else {
        ra = ca;
        rs = cs;
        cr = (float)cv;
    }
    nread++;
}

/* the running state over the blocks [blk0, blk0 + nblk), continuing from
 * `in` into `out` (two buffers: every workgroup reads the state the previous
 * launch left).  One lane per channel walks the blocks: the channel levels'
 * recurrences are independent of each other, and the zero-DM mean and std
 * (uniform) are recomputed by every lane from the block statistics, which go
 * through LDS 512 blocks at a time. */
__global__ void __launch_bounds__(256)
clip_scan_kernel(long long blk0, int nblk, int nch, float clip_sigma,
                 const BlockStat *__restrict__ bs, const double *__restrict__ chavg,
                 const ClipState *__restrict__ in, ClipState *__restrict__ out,
                 float *__restrict__ crun, float *__restrict__ ravg, float *__restrict__ trig,
                 float *__restrict__ levels)
{
    __shared__ BlockStat sb[512];
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    float ra = in->running_avg, rs = in->running_std;
    int nread = in->blocksread;
    float cr = c < nch ? crun[c] : 0.0f;
    for (int b0 = 0; b0 < nblk; b0 += 512) {
        const int n = min(512, nblk - b0);
        __syncthreads();
        for (int i = (int)threadIdx.x; i < n; i += 256) sb[i] = bs[blk0 + b0 + i];
        __syncthreads();
        /* 16 blocks a group: their channel means are loaded a group ahead,
         * and the group's levels and running values are stored after its
         * steps.  A wave's loads and stores complete in order, so a store
         * between a load and its use would make the use wait for the store;
         * here the wait for a group's means passes over the previous group's
         * stores.  The loads are unconditional (block and channel indices
         * clamped into the arrays), so the waits count them exactly; blocks
         * past the last full group take the plain loop. */
        constexpr int G = 16;
        const int nfull = n / G * G;
        const int cc = min(c, nch - 1);
        const double *cva = chavg + (blk0 + b0) * nch + cc;
        /* one group's steps on the means in cur, loading the next group's
         * into nxt; two arrays in turn, so no copy waits for the loads */
        auto group = [&](const double (&cur)[G], double (&nxt)[G], int i0) {
#pragma unroll
            for (int q = 0; q < G; q++) nxt[q] = cva[(long long)min(i0 + G + q, n - 1) * nch];
            float lv[G], av[G], tv[G];
#pragma unroll
            for (int q = 0; q < G; q++) {
                scan_step(sb[i0 + q], cur[q], nread, ra, rs, cr);
                lv[q] = cr;
                av[q] = ra;
                tv[q] = clip_sigma * rs;
            }
            const long long B0 = blk0 + b0 + i0;
            if (c < nch) {
#pragma unroll
                for (int q = 0; q < G; q++) levels[(B0 + q) * nch + c] = lv[q];
            }
            if (lead) {
#pragma unroll
                for (int q = 0; q < G; q++) {
                    ravg[B0 + q] = av[q];
                    trig[B0 + q] = tv[q];
                }
            }
        };
        double ca_[G], cb_[G];
#pragma unroll
        for (int q = 0; q < G; q++) ca_[q] = cva[(long long)min(q, n - 1) * nch];
        int i0 = 0;
        for (; i0 + 2 * G <= nfull; i0 += 2 * G) {
            group(ca_, cb_, i0);
            group(cb_, ca_, i0 + G);
        }
        if (i0 < nfull) group(ca_, cb_, i0);
        for (int i = nfull; i < n; i++) {
            const long long B = blk0 + b0 + i;
            scan_step(sb[i], cva[(long long)i * nch], nread, ra, rs, cr);
            if (c < nch) levels[B * nch + c] = cr;
            if (lead) {
                ravg[B] = ra;
                trig[B] = clip_sigma * rs;
            }
        }
    }
    if (c < nch) crun[c] = cr;
    if (lead) {
        out->running_avg = ra;
        out->running_std = rs;
        out->blocksread = nread;
    }
}

__global__ void pad_kernel(float *__restrict__ out, long long numout, long long written,
                           int numdms, float v)
{
    const long long n = numout - written;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n * numdms;
         k += (long long)gridDim.x * blockDim.x) {
        const long long i = k / n, u = k - i * n;
        out[i * numout + written + u] = v;
    }
}

}  // namespace

struct CsDedisp {
    CsDedispConfig c;
    Tables t;
    int device = 0;
    long long nspec = 0, chunk = 0;
    int dsc = 0;                   /* downsampled samples per chunk */
    hipStream_t st = nullptr, cst = nullptr;
    hipStream_t ost = nullptr;     /* DM 0's samples back to the host */
    int *d_idisp = nullptr, *d_offs = nullptr;
    float *d_sub = nullptr;        /* 2 x nsub x dsc */
    float *d_out = nullptr;        /* numdms x numout */
    void *d_ring = nullptr;        /* 3 chunks of raw spectra (host input) */
    size_t ring_bytes = 0;
    float *h_dm0 = nullptr;        /* pinned: DM 0's written samples */
    bool simple = false;           /* the untiled kernels only */
    /* DM tiles: G DMs per group, subband batches of sbatch, tables
     * [ngroups][nsub] (gmin, glen, gbase) and [numdms][nsub] (sofs) */
    int dm_g = 0, dm_ngroups = 0, dm_sbatch = 0, dm_lds = 0;
    int *d_gtab = nullptr;
    /* clipping (clip_sigma > 0): per spectrum, per raw block, state */
    float *d_zdm = nullptr, *d_med = nullptr, *d_ravg = nullptr, *d_trig = nullptr;
    float *d_levels = nullptr, *d_crun = nullptr;
    unsigned char *d_cflag = nullptr;
    uint32_t *d_chsum = nullptr;
    double *d_chavg = nullptr;
    BlockStat *d_bs = nullptr;
    ClipState *d_cst = nullptr;
    float pad = 0.0f;
    double last_ms = 0.0, last_sub_ms = 0.0, last_dm_ms = 0.0;
};

extern "C" CS_API long long cs_choose_good_N(long long n)
{
    if (n <= 0) return 0;
    /* the first four digits of n (src: misc_utils.c:334-373); the good
     * factors are 1000 and the multiples of 16 in (1000, 10000] without a
     * prime factor above 11, ascending */
    char digits[32];
    snprintf(digits, sizeof(digits), "%lld", n);
    digits[4] = '\0';
    const long long first4 = atoll(digits);
    long long g = 0;
    for (long long v = 1000; v <= 10000; v++) {
        if (v != 1000 && (v % 16 || !smooth11(v))) continue;
        g = v;
        if (v == first4 && n % v == 0 && power_of_10(n / v)) break;
        if (v > first4) break;
    }
    while (g < n) g *= 10;
    long long two = 2;
    while (two < n) two *= 2;
    return two < g ? two : g;
}

extern "C" CS_API double cs_delay_from_dm(double dm, double freq_mhz)
{
    return delay_from_dm(dm, freq_mhz);
}

extern "C" CS_API void cs_gpu_dedisp_free(CsDedisp *d)
{
    if (!d) return;
    if (d->device < 0) {
        delete d;
        return;
    }
    hipSetDevice(d->device);
    if (d->st) hipStreamSynchronize(d->st);
    if (d->cst) hipStreamSynchronize(d->cst);
    if (d->ost) hipStreamSynchronize(d->ost);
    hipFree(d->d_idisp);
    hipFree(d->d_offs);
    hipFree(d->d_sub);
    hipFree(d->d_out);
    hipFree(d->d_ring);
    hipFree(d->d_gtab);
    for (void *q : {(void *)d->d_zdm, (void *)d->d_med, (void *)d->d_ravg, (void *)d->d_trig,
                    (void *)d->d_levels, (void *)d->d_crun, (void *)d->d_chavg, (void *)d->d_bs,
                    (void *)d->d_cst, (void *)d->d_cflag, (void *)d->d_chsum})
        hipFree(q);
    if (d->h_dm0) hipHostFree(d->h_dm0);
    if (d->st) hipStreamDestroy(d->st);
    if (d->cst) hipStreamDestroy(d->cst);
    if (d->ost) hipStreamDestroy(d->ost);
    delete d;
}

extern "C" CS_API CsDedisp *cs_gpu_dedisp_create(const CsDedispConfig *c, long long nspec,
                                                 int device)
{
    if (!c) return nullptr;
    CsDedisp *d = new CsDedisp();
    d->c = *c;
    d->device = device;
    d->nspec = nspec;
    if (build_tables(*c, nspec, d->t) != 0) {
        delete d;
        return nullptr;
    }
    {
        const Tables &t = d->t;
        const int P = c->spectra_per_subint, nsub = c->nsub;
        /* a pass over spectra in HBM: whole raw blocks, at least one work
         * block, about 2^20 spectra, the subband ring within 512 MiB (passes
         * over host spectra may be shorter: run_passes) */
        /* CS_DEDISP_CHUNK: spectra per pass (tests of several passes) */
        const char *ec = getenv("CS_DEDISP_CHUNK");
        const long long target = ec && atoll(ec) > 0 ? atoll(ec) : kChunkTarget;
        long long want = std::max<long long>(t.worklen, target);
        const long long cap = kSubBytes * c->downsamp / (8ll * nsub);
        if (want > cap) want = std::max<long long>(t.worklen, cap);
        d->chunk = (want + P - 1) / P * P;
        d->dsc = (int)(d->chunk / c->downsamp);
        if (device < 0) return d;     /* tables only */
        const char *env = getenv("CS_DEDISP_SIMPLE");
        d->simple = env && *env && *env != '0';
        DCHK(hipSetDevice(device));
        DCHK(hipStreamCreateWithFlags(&d->st, hipStreamNonBlocking));
        DCHK(hipStreamCreateWithFlags(&d->cst, hipStreamNonBlocking));
        DCHK(hipStreamCreateWithFlags(&d->ost, hipStreamNonBlocking));
        DCHK(hipMalloc(&d->d_idisp, sizeof(int) * t.idispdt.size()));
        DCHK(hipMalloc(&d->d_offs, sizeof(int) * t.offsets.size()));
        DCHK(hipMemcpy(d->d_idisp, t.idispdt.data(), sizeof(int) * t.idispdt.size(),
                       hipMemcpyHostToDevice));
        DCHK(hipMemcpy(d->d_offs, t.offsets.data(), sizeof(int) * t.offsets.size(),
                       hipMemcpyHostToDevice));
        DCHK(hipMalloc(&d->d_sub, sizeof(float) * 2 * (size_t)nsub * d->dsc));
        DCHK(hipMalloc(&d->d_out, sizeof(float) * (size_t)c->numdms * t.numout));
        if (t.written > 0)
            DCHK(hipHostMalloc((void **)&d->h_dm0, sizeof(float) * (size_t)t.written,
                               hipHostMallocDefault));
        if (c->clip_sigma > 0.0f) {
            const size_t nb = (size_t)(nspec / P), nc = (size_t)c->numchan;
            DCHK(hipMalloc(&d->d_zdm, sizeof(float) * (size_t)nspec));
            DCHK(hipMalloc(&d->d_cflag, (size_t)nspec));
            DCHK(hipMalloc(&d->d_med, sizeof(float) * nb));
            DCHK(hipMalloc(&d->d_ravg, sizeof(float) * nb));
            DCHK(hipMalloc(&d->d_trig, sizeof(float) * nb));
            DCHK(hipMalloc(&d->d_levels, sizeof(float) * nb * nc));
            DCHK(hipMalloc(&d->d_crun, sizeof(float) * nc));
            DCHK(hipMalloc(&d->d_chavg, sizeof(double) * nb * nc));
            DCHK(hipMalloc(&d->d_chsum, sizeof(uint32_t) * nb * nc));
            DCHK(hipMalloc(&d->d_bs, sizeof(BlockStat) * nb));
            DCHK(hipMalloc(&d->d_cst, 2 * sizeof(ClipState)));
        }
        /* DM tiles: the largest group (up to 32 DMs, a power of two) that
         * still gives a pass about kDmWorkgroups tiles and whose widest span
         * fits one batch */
        const int nd = c->numdms;
        const long long tiles = (d->dsc + kTU - 1) / kTU;
        int G = 1;
        while (G < kGMax && G < nd) G *= 2;
        while (G > 4 && tiles * ((nd + G - 1) / G) < kDmWorkgroups) G /= 2;
        std::vector<int> gmin, glen, gbase, sofs;
        for (;; G /= 2) {
            const int ng = (nd + G - 1) / G;
            gmin.assign((size_t)ng * nsub, 0);
            glen.assign((size_t)ng * nsub, 0);
            int maxlen = 0;
            for (int g = 0; g < ng; g++)
                for (int s = 0; s < nsub; s++) {
                    int lo = 1 << 30, hi = -1;
                    for (int i = g * G; i < std::min(nd, (g + 1) * G); i++) {
                        lo = std::min(lo, t.offsets[(size_t)i * nsub + s]);
                        hi = std::max(hi, t.offsets[(size_t)i * nsub + s]);
                    }
                    gmin[(size_t)g * nsub + s] = lo;
                    glen[(size_t)g * nsub + s] = kTU + hi - lo;
                    maxlen = std::max(maxlen, kTU + hi - lo);
                }
            if (maxlen <= kSpanFloats || G == 1) {
                d->dm_g = G;
                d->dm_ngroups = ng;
                d->dm_sbatch = std::max(1, std::min(nsub, kSpanFloats / maxlen));
                d->dm_lds = d->dm_sbatch * maxlen * (int)sizeof(float);
                break;
            }
        }
        if (d->dm_lds > kSpanFloats * (int)sizeof(float)) d->dm_g = 0;   /* untiled */
        if (d->dm_g) {
            const int ng = d->dm_ngroups;
            gbase.assign((size_t)ng * nsub, 0);
            for (int g = 0; g < ng; g++)
                for (int sA = 0; sA < nsub; sA += d->dm_sbatch) {
                    int b = 0;
                    for (int s = sA; s < std::min(nsub, sA + d->dm_sbatch); s++) {
                        gbase[(size_t)g * nsub + s] = b;
                        b += glen[(size_t)g * nsub + s];
                    }
                }
            sofs.assign((size_t)ng * nsub * d->dm_g, 0);
            for (int i = 0; i < nd; i++) {
                const int g = i / d->dm_g, k = i % d->dm_g;
                for (int s = 0; s < nsub; s++)
                    sofs[((size_t)g * nsub + s) * d->dm_g + k] =
                        gbase[(size_t)g * nsub + s] - gmin[(size_t)g * nsub + s] +
                        t.offsets[(size_t)i * nsub + s];
            }
            std::vector<int> all;
            all.insert(all.end(), gmin.begin(), gmin.end());
            all.insert(all.end(), glen.begin(), glen.end());
            all.insert(all.end(), gbase.begin(), gbase.end());
            all.insert(all.end(), sofs.begin(), sofs.end());
            DCHK(hipMalloc(&d->d_gtab, sizeof(int) * all.size()));
            DCHK(hipMemcpy(d->d_gtab, all.data(), sizeof(int) * all.size(),
                           hipMemcpyHostToDevice));
        }
    }
    return d;
fail:
    cs_gpu_dedisp_free(d);
    return nullptr;
}

extern "C" CS_API int cs_gpu_dedisp_info(const CsDedisp *d, CsDedispInfo *o)
{
    if (!d || !o) return -1;
    memset(o, 0, sizeof(*o));
    o->blocksperread = d->t.bpr;
    o->worklen = d->t.worklen;
    o->nsub = d->c.nsub;
    o->max_offset = d->t.maxoff;
    o->numout = d->t.numout;
    o->written = d->t.written;
    o->chunk = d->chunk;
    o->dsdt = d->t.dsdt;
    o->pad = d->pad;
    return 0;
}

extern "C" CS_API int cs_gpu_dedisp_tables(const CsDedisp *d, double *dms, int *idispdt,
                                           int *offsets)
{
    if (!d) return -1;
    if (dms) memcpy(dms, d->t.dms.data(), sizeof(double) * d->t.dms.size());
    if (idispdt) memcpy(idispdt, d->t.idispdt.data(), sizeof(int) * d->t.idispdt.size());
    if (offsets) memcpy(offsets, d->t.offsets.data(), sizeof(int) * d->t.offsets.size());
    return 0;
}

extern "C" CS_API const float *cs_gpu_dedisp_series(const CsDedisp *d, int i)
{
    if (!d || !d->d_out || i < 0 || i >= d->c.numdms) return nullptr;
    return d->d_out + (size_t)i * d->t.numout;
}

extern "C" CS_API int cs_gpu_dedisp_copy_out(const CsDedisp *d, int i0, int n, float *host)
{
    if (!d || !d->d_out || !host || i0 < 0 || n < 0 || i0 + n > d->c.numdms) return -1;
    DCHK(hipSetDevice(d->device));
    DCHK(hipMemcpy(host, d->d_out + (size_t)i0 * d->t.numout,
                   sizeof(float) * (size_t)n * d->t.numout, hipMemcpyDeviceToHost));
    return 0;
fail:
    return -1;
}

extern "C" CS_API double cs_gpu_dedisp_last_ms(const CsDedisp *d, double *subband_ms,
                                               double *dm_ms)
{
    if (!d) return -1.0;
    if (subband_ms) *subband_ms = d->last_sub_ms;
    if (dm_ms) *dm_ms = d->last_dm_ms;
    return d->last_ms;
}

template <typename T>
static int run_passes(CsDedisp *d, const T *spectra, bool on_device)
{
    const Tables &t = d->t;
    const CsDedispConfig &c = d->c;
    const int nch = c.numchan, nsub = c.nsub, cps = nch / nsub, ds = c.downsamp;
    const int P = c.spectra_per_subint;
    const bool clip = c.clip_sigma > 0.0f;
    /* passes over host spectra: a raw ring of three chunks within 1 GiB */
    long long C = d->chunk;
    if (!on_device) {
        long long want = std::max<long long>(t.worklen, kRingBytes / (3ll * nch * (long long)sizeof(T)));
        want = (want + P - 1) / P * P;
        C = std::min(C, want);
    }
    const long long dsc = C / ds;
    const long long tsub = t.written > 0 ? t.written + t.maxoff : 0;
    /* raw spectra the passes read (whole raw blocks with clipping: a block's
     * statistics take all of its rows) */
    long long rows = tsub * ds + t.maxd;
    if (clip && tsub > 0) rows = std::min(d->nspec, (rows + P - 1) / P * P);
    const long long nk1 = (tsub + dsc - 1) / dsc, nk2 = (t.written + dsc - 1) / dsc;
    const size_t row_bytes = sizeof(T) * (size_t)nch;
    std::vector<hipEvent_t> ev;
    hipEvent_t raw_ready[3] = {}, k1_done[3] = {}, t_start = nullptr;
    int rc = -1;
    auto new_event = [&](hipEvent_t *e, unsigned flags) {
        if (hipEventCreateWithFlags(e, flags) != hipSuccess) return false;
        ev.push_back(*e);
        return true;
    };
    std::vector<hipEvent_t> k1s(nk1), k1e(nk1), k2s(nk2), k2e(nk2), cp(nk2);
    const T *ring = nullptr;
    auto chunk_ptr = [&](long long j) -> const T * {
        return on_device ? spectra + j * C * nch : ring + (j % 3) * C * nch;
    };
    /* host spectra: chunk j of the input into ring slot j % 3 once the pass
     * that last read that slot (j - 3's pass reads j - 3 and j - 2) is done */
    auto copy_chunk = [&](long long j) -> bool {
        const long long r0 = j * C, nr = std::min(C, rows - r0);
        if (on_device || nr <= 0) return true;
        if (j >= 3 && hipStreamWaitEvent(d->cst, k1_done[j % 3], 0) != hipSuccess) return false;
        if (hipMemcpyAsync((void *)(ring + (j % 3) * C * nch), spectra + r0 * nch,
                           row_bytes * (size_t)nr, hipMemcpyDefault, d->cst) != hipSuccess)
            return false;
        return hipEventRecord(raw_ready[j % 3], d->cst) == hipSuccess;
    };
    int sb_gs = 0, sb_pitch = 0, sb_vec = 0;
    auto clip_view = [&](long long j) {
        ClipView v{};
        if (clip) {
            v.flag = d->d_cflag;
            v.levels = d->d_levels;
            v.row0 = j * C;
            v.P = P;
        }
        return v;
    };
    int nscan = 0;              /* scans so far: the state's buffer */
    /* clip_times' statistics and state for the raw blocks of chunk j */
    auto clip_chunk = [&](long long j) -> hipError_t {
        /* j = -1: every block the passes read (spectra in HBM) */
        const long long r0 = j < 0 ? 0 : j * C, r1 = j < 0 ? rows : std::min((j + 1) * C, rows);
        if (!clip || r1 <= r0) return hipSuccess;
        const long long blk0 = r0 / P;
        const int nb = (int)((r1 - r0) / P);
        const T *src = j < 0 ? spectra : chunk_ptr(j);
        const int vec16 = (nch * (int)sizeof(T)) % 16 == 0 && (uintptr_t)src % 16 == 0;
        /* uint8 rows of 16 x 2^k channels (up to 4096) in blocks of up to
         * 8192: one read for the zero-DM sums and the channel sums */
        const int nv = nch / 16;
        const bool fused = sizeof(T) == 1 && nch % 16 == 0 && nv <= 256 && !(nv & (nv - 1)) &&
                           P <= 8192 && (uintptr_t)src % 16 == 0;
        if (fused) {
            if constexpr (sizeof(T) == 1) {
                int n2 = 1;
                while (n2 < P) n2 <<= 1;
                const size_t lds = 4 * ((size_t)n2 + (size_t)nch);
                const uint8_t *s8 = (const uint8_t *)src;
#define CS_ZU8(V)                                                                          \
    clip_zdm_u8_kernel<V><<<nb, 256, lds, d->st>>>(s8, blk0, P, nch, d->d_zdm, d->d_med, \
                                                   d->d_chsum)
                switch (nv) {
                case 1: CS_ZU8(1); break;
                case 2: CS_ZU8(2); break;
                case 4: CS_ZU8(4); break;
                case 8: CS_ZU8(8); break;
                case 16: CS_ZU8(16); break;
                case 32: CS_ZU8(32); break;
                case 64: CS_ZU8(64); break;
                case 128: CS_ZU8(128); break;
                default: CS_ZU8(256); break;
                }
#undef CS_ZU8
            }
        } else {
            clip_zdm_kernel<T><<<nb, 256, zdm_smem<T>(P), d->st>>>(src, blk0, P, nch, vec16,
                                                                   d->d_zdm, d->d_med);
        }
        if (P <= kClipRcpMax)
            clip_seq_kernel<<<(nb + 63) / 64, 64, 0, d->st>>>(d->d_zdm, d->d_med, blk0, nb, P,
                                                              d->d_bs);
        else
            clip_seq_lane_kernel<<<(nb + 63) / 64, 64, 0, d->st>>>(d->d_zdm, d->d_med, blk0, nb, P,
                                                                   d->d_bs);
        /* four channels a lane where rows are whole vectors */
        const bool four = nch % 4 == 0 && (uintptr_t)src % (sizeof(T) == 1 ? 4 : 16) == 0;
        if (fused) {
            clip_badrows_u8_kernel<<<nb, 256, 0, d->st>>>((const uint8_t *)src, blk0, P, nch,
                                                          d->d_zdm, d->d_bs, d->d_chsum,
                                                          d->d_chavg);
        }
        for (int b0 = 0; !fused && b0 < nb; b0 += 65535) {  /* grid y: at most 65535 */
            const int n = std::min(65535, nb - b0);
            const T *s0 = src + (long long)b0 * P * nch;
            if (four)
                clip_chavg4_kernel<T><<<dim3((unsigned)((nch / 4 + 255) / 256), (unsigned)n), 256,
                                        0, d->st>>>(s0, blk0 + b0, P, nch, d->d_zdm, d->d_bs,
                                                    d->d_chavg);
            else
                clip_chavg_kernel<T><<<dim3((unsigned)((nch + 255) / 256), (unsigned)n), 256, 0,
                                       d->st>>>(s0, blk0 + b0, P, nch, d->d_zdm, d->d_bs,
                                                d->d_chavg);
        }
        const ClipState *sin = d->d_cst + (nscan & 1);
        ClipState *sout = d->d_cst + ((nscan + 1) & 1);
        nscan++;
        clip_scan_kernel<<<(nch + 255) / 256, 256, 0, d->st>>>(
            blk0, nb, nch, c.clip_sigma, d->d_bs, d->d_chavg, sin, sout, d->d_crun, d->d_ravg,
            d->d_trig, d->d_levels);
        const long long nrow = (long long)nb * P;
        clip_flag_kernel<<<(unsigned)((nrow + 255) / 256), 256, 0, d->st>>>(
            d->d_zdm, d->d_ravg, d->d_trig, blk0 * P, nrow, P, d->d_cflag);
        return hipGetLastError();
    };
    /* the subband pass of chunk j: nu downsampled samples into `sub` */
    auto launch_sub = [&](long long j, long long nu, float *sub) -> hipError_t {
        const T *ra = chunk_ptr(j), *rb = chunk_ptr(j + 1);
        /* grid y holds the row tiles: at most 65535 */
        if (sb_gs && (nu * ds + kRows - 1) / kRows <= 65535) {
            /* rows the tiles may load: those the pass reads (its samples'
             * rows plus the deepest delay), never past the data */
            const long long nrow = nu * ds, here = std::min(rows - j * C, nrow + t.maxd);
            const int ngrp = (nsub + sb_gs - 1) / sb_gs;
            const int ntile = (int)((nrow + kRows - 1) / kRows);
            const dim3 grid((unsigned)(((long long)ngrp * ntile + 7) / 8 * 8));
            const size_t lds = (size_t)(kRows + t.maxd) * sb_pitch +
                               (clip ? (size_t)(kRows + t.maxd + 3) / 4 * 4 : 0);
#define CS_SUB(V, K)                                                                       \
    subband_tiled<T, V, K><<<grid, kRows, lds, d->st>>>(ra, rb, C, nch, cps, sb_gs, nsub, ds, \
                                                        d->d_idisp, t.maxd, sb_pitch, here, \
                                                        (int)nu, sub, (int)dsc, clip_view(j), \
                                                        ngrp, ntile)
            if (clip) {
                if (sb_vec == 16) CS_SUB(16, true);
                else if (sb_vec == 4 || sizeof(T) != 1) CS_SUB(4, true);
                else if constexpr (sizeof(T) == 1) CS_SUB(1, true);
            } else if (sb_vec == 16) CS_SUB(16, false);
            else if (sb_vec == 4 || sizeof(T) != 1) CS_SUB(4, false);
            else if constexpr (sizeof(T) == 1) CS_SUB(1, false);
#undef CS_SUB
        } else {
            const dim3 grid((unsigned)((nu + 63) / 64), (unsigned)((nsub + kWaves - 1) / kWaves));
            subband_kernel<T><<<grid, 64 * kWaves, 0, d->st>>>(ra, rb, C, nch, cps, nsub, ds,
                                                               d->d_idisp, (int)nu, sub, (int)dsc,
                                                               clip_view(j));
        }
        return hipGetLastError();
    };
    /* DM 0's samples of DM pass m back to the pinned host buffer on their
     * own stream, so the host's running mean follows the passes */
    auto copy_dm0 = [&](long long m, long long nv, long long v0) -> hipError_t {
        hipError_t e = hipStreamWaitEvent(d->ost, k2e[m], 0);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d->h_dm0 + v0, d->d_out + v0, sizeof(float) * (size_t)nv,
                               hipMemcpyDeviceToHost, d->ost);
        if (e == hipSuccess) e = hipEventRecord(cp[m], d->ost);
        return e;
    };
    /* the DM pass of subband chunk m: nv samples from v0 on */
    auto launch_dm = [&](long long m, long long nv, long long v0) -> hipError_t {
        const float *sa = d->d_sub + (m & 1) * (size_t)nsub * dsc;
        const float *sb = d->d_sub + ((m + 1) & 1) * (size_t)nsub * dsc;
        if (d->dm_g && (nv + kTU - 1) / kTU <= 65535) {
            const int ng = d->dm_ngroups;
            const int *gmin = d->d_gtab, *glen = gmin + (size_t)ng * nsub,
                      *gbase = glen + (size_t)ng * nsub, *sofs = gbase + (size_t)ng * nsub;
            const int tlim = (int)(nv + t.maxoff);
            const int ntile = (int)((nv + kTU - 1) / kTU);
            const dim3 grid((unsigned)(((long long)ng * ntile + 7) / 8 * 8));
#define CS_DM(G)                                                                           \
    dm_tiled<G><<<grid, kDmThreads, (size_t)d->dm_lds, d->st>>>(sa, sb, (int)dsc, nsub, d->dm_sbatch, \
                                                         gmin, glen, gbase, sofs, c.numdms,  \
                                                         (int)nv, tlim, d->d_out, t.numout, v0, \
                                                         ng, ntile)
            switch (d->dm_g) {
            case 32: CS_DM(32); break;
            case 16: CS_DM(16); break;
            case 8: CS_DM(8); break;
            case 4: CS_DM(4); break;
            case 2: CS_DM(2); break;
            default: CS_DM(1); break;
            }
#undef CS_DM
        } else {
            const dim3 grid((unsigned)((nv + 63) / 64), (unsigned)((c.numdms + kWaves - 1) / kWaves));
            dm_kernel<<<grid, 64 * kWaves, 0, d->st>>>(sa, sb, (int)dsc, nsub, d->d_offs, c.numdms,
                                                       (int)nv, d->d_out, t.numout, v0);
        }
        return hipGetLastError();
    };
    DCHK(hipSetDevice(d->device));
    for (int i = 0; i < 3; i++) {
        if (!new_event(&raw_ready[i], hipEventDisableTiming) ||
            !new_event(&k1_done[i], hipEventDisableTiming))
            goto fail;
    }
    for (auto *v : {&k1s, &k1e, &k2s, &k2e})
        for (auto &e : *v)
            if (!new_event(&e, hipEventDefault)) goto fail;
    for (auto &e : cp)
        if (!new_event(&e, hipEventDisableTiming)) goto fail;
    if (!new_event(&t_start, hipEventDefault)) goto fail;
    if (!on_device && nk1 > 0) {
        const size_t need = 3 * (size_t)C * row_bytes;
        if (d->ring_bytes < need) {
            hipFree(d->d_ring);
            d->d_ring = nullptr;
            d->ring_bytes = 0;
            DCHK(hipMalloc(&d->d_ring, need));
            d->ring_bytes = need;
        }
        ring = (const T *)d->d_ring;
    }
    /* subband tiles for this sample type: the widest group of subbands
     * whose tile fits the LDS budget, loads as wide as the rows' alignment
     * allows (64 % ds == 0: a downsampled sample's lanes share a wave) */
    if (!d->simple && 64 % ds == 0) {
        /* loads: VEC bytes of samples (alignment of the raw rows); tile
         * elements: floats with clipping, samples otherwise */
        const int rs = (int)sizeof(T), es = rs;
        const bool al16 = !on_device || (uintptr_t)spectra % 16 == 0;
        sb_vec = (cps * rs) % 16 == 0 && (nch * rs) % 16 == 0 && al16 ? 16
                 : (cps * rs) % 4 == 0 && (nch * rs) % 4 == 0 &&
                           (!on_device || (uintptr_t)spectra % 4 == 0)
                     ? 4
                     : 1;
        for (int gs = nsub; gs >= 1; gs--) {
            const int pitch = (gs * cps * es + 15) / 16 * 16 + 4;    /* odd dwords */
            if ((long long)(kRows + t.maxd) * pitch <= kTileBytes) {
                sb_gs = gs;
                sb_pitch = pitch;
                break;
            }
        }
    }
    DCHK(hipEventRecord(t_start, d->st));
    if (clip) {
        DCHK(hipMemsetAsync(d->d_cst, 0, 2 * sizeof(ClipState), d->st));
        DCHK(hipMemsetAsync(d->d_crun, 0, sizeof(float) * (size_t)nch, d->st));
    }
    if (!copy_chunk(0) || !copy_chunk(1)) goto fail;
    for (long long j = 0; j < nk1; j++) {
        const long long u0 = j * dsc, nu = std::min(dsc, tsub - u0);
        float *sub = d->d_sub + (j & 1) * (size_t)nsub * dsc;
        if (!on_device) {
            DCHK(hipStreamWaitEvent(d->st, raw_ready[j % 3], 0));
            if ((j + 1) * C < rows) DCHK(hipStreamWaitEvent(d->st, raw_ready[(j + 1) % 3], 0));
        }
        /* the pass reads the first rows of chunk j + 1: its blocks' clip
         * levels first (chunk 0's before the first pass).  Spectra in HBM:
         * every block's at once, so the per-block statistics run on all
         * blocks in parallel */
        if (on_device) {
            if (j == 0) DCHK(clip_chunk(-1));
        } else {
            if (j == 0) DCHK(clip_chunk(0));
            DCHK(clip_chunk(j + 1));
        }
        DCHK(hipEventRecord(k1s[j], d->st));
        DCHK(launch_sub(j, nu, sub));
        DCHK(hipEventRecord(k1e[j], d->st));
        DCHK(hipEventRecord(k1_done[j % 3], d->st));
        if (j >= 1 && j - 1 < nk2) {
            const long long m = j - 1, v0 = m * dsc, nv = std::min(dsc, t.written - v0);
            DCHK(hipEventRecord(k2s[m], d->st));
            DCHK(launch_dm(m, nv, v0));
            DCHK(hipEventRecord(k2e[m], d->st));
            DCHK(copy_dm0(m, nv, v0));
        }
        if (!copy_chunk(j + 2)) goto fail;
    }
    /* the last DM pass: its subbands end inside the last subband chunk */
    for (long long m = std::max<long long>(nk1 - 1, 0); m < nk2; m++) {
        const long long v0 = m * dsc, nv = std::min(dsc, t.written - v0);
        DCHK(hipEventRecord(k2s[m], d->st));
        DCHK(launch_dm(m, nv, v0));
        DCHK(hipEventRecord(k2e[m], d->st));
        DCHK(copy_dm0(m, nv, v0));
    }
    /* the padding value: the running mean of DM 0's written samples, in
     * order, as update_stats keeps it (misc_utils.h:114-127) */
    {
        double avg = 0.0;
        for (long long m = 0; m < nk2; m++) {        /* while later passes run */
            DCHK(hipEventSynchronize(cp[m]));
            const long long v1 = std::min(t.written, (m + 1) * dsc);
            for (long long u = m * dsc; u < v1; u++) {
                const double dev = (double)d->h_dm0[u] - avg;
                avg += dev / ((double)u + 1.0);
            }
        }
        d->pad = (float)avg;
        if (t.numout > t.written) {
            pad_kernel<<<1024, 256, 0, d->st>>>(d->d_out, t.numout, t.written, c.numdms, d->pad);
            DCHK(hipGetLastError());
        }
        DCHK(hipStreamSynchronize(d->st));
        DCHK(hipStreamSynchronize(d->cst));
        DCHK(hipStreamSynchronize(d->ost));
    }
    {
        double s1 = 0.0, s2 = 0.0;
        float ms = 0.0f;
        for (long long j = 0; j < nk1; j++)
            if (hipEventElapsedTime(&ms, k1s[j], k1e[j]) == hipSuccess) s1 += ms;
        for (long long m = 0; m < nk2; m++)
            if (hipEventElapsedTime(&ms, k2s[m], k2e[m]) == hipSuccess) s2 += ms;
        d->last_sub_ms = s1;
        d->last_dm_ms = s2;
        d->last_ms = 0.0;
        /* from the start of the run (clipping included) to the last DM pass */
        if (nk1 > 0 && hipEventElapsedTime(&ms, t_start, nk2 > 0 ? k2e[nk2 - 1] : k1e[nk1 - 1]) ==
                           hipSuccess)
            d->last_ms = ms;
    }
    rc = 0;
fail:
    if (rc != 0) {
        hipStreamSynchronize(d->st);
        hipStreamSynchronize(d->cst);
        hipStreamSynchronize(d->ost);
    }
    for (hipEvent_t e : ev) hipEventDestroy(e);
    return rc;
}

extern "C" CS_API int cs_gpu_dedisp_run(CsDedisp *d, const void *spectra, int is_uint8)
{
    if (!d || !spectra) return -1;
    if (d->device < 0) {
        fprintf(stderr, "Error: a tables-only de-dispersion plan (device -1) cannot run\n");
        return -1;
    }
    hipPointerAttribute_t attr;
    bool on_device = false;
    if (hipPointerGetAttributes(&attr, spectra) == hipSuccess)
        on_device = attr.type == hipMemoryTypeDevice && attr.device == d->device;
    (void)hipGetLastError();
    return is_uint8 ? run_passes(d, (const uint8_t *)spectra, on_device)
                    : run_passes(d, (const float *)spectra, on_device);
}
