/*
 * libcs_hip.so session layer: the reference's cs_gpu_* C-ABI (src/cuda.cu
 * :488-1266) re-designed for one MI355X.
 *
 * create  : plan arena (per harmonic x residue weight rows), twiddles, the
 *           chunk table with per-chunk filled masks, Nyquist counts and
 *           1/(sigma*sd_w) for every slot and width, and the tile table --
 *           everything a search needs stays resident in HBM.
 * run     : one upload of the spectrum, one fused kernel launch over every
 *           tile of every chunk, one readback of the above-threshold records,
 *           host sort into the reference's candidate order.
 * The noise-scale shape follows the reference CPU path's count-keyed cache
 * (src/search.c:36-46) walked in chunk order, i.e. one CPU thread.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "cands.h"
#include "core.h"
#include "cs_device.h"
#include "cs_general.h"
#include "cs_gpu.h"
#include "cs_stage.h"
#include "fast_kernel.hpp"
#include "split_kernel.hpp"

#define HIDDEN __attribute__((visibility("hidden")))

typedef hipError_t (*fast_launch_fn)(const csk::FastArgs *, int, hipStream_t);

#define CS_DECL_FAST(NH, M, K)                                                    \
    extern "C" HIDDEN hipError_t cs_launch_fast_##NH##_##M##_##K(const csk::FastArgs *, \
                                                                int, hipStream_t);
#define CS_DECL_FAST64(NH, M, K)                                                  \
    extern "C" HIDDEN hipError_t cs_launch_fast64_##NH##_##M##_##K(const csk::FastArgs *, \
                                                                  int, hipStream_t);
#define CS_DECL_NH(NH) CS_DECL_FAST(NH, 16, 16) CS_DECL_FAST(NH, 32, 16) \
    CS_DECL_FAST64(NH, 16, 16) CS_DECL_FAST64(NH, 32, 16)
typedef hipError_t (*profile_launch_fn)(const csk::FastArgs *, const csk::SplitArgs *, int,
                                        hipStream_t);
#define CS_DECL_NH_BIG(NH)                                                        \
    extern "C" HIDDEN hipError_t cs_launch_profile_##NH(const csk::FastArgs *,    \
                                                        const csk::SplitArgs *, int, \
                                                        hipStream_t);
CS_NH_LIST(CS_DECL_NH)
CS_NH_BIG_LIST(CS_DECL_NH_BIG)

namespace {

struct FastVariant {
    int nh, m, kmax;
    fast_launch_fn fn;        /* f32 profiles */
    fast_launch_fn fn64;      /* f64 profiles */
};

#define CS_ENTRY_FAST(NH, M, K)                                                    \
    {NH, M, K, cs_launch_fast_##NH##_##M##_##K, cs_launch_fast64_##NH##_##M##_##K},
#define CS_ENTRY_NH(NH) CS_ENTRY_FAST(NH, 16, 16) CS_ENTRY_FAST(NH, 32, 16)
const FastVariant kFast[] = {CS_NH_LIST(CS_ENTRY_NH)};

/* profile passes of the two-pass search, f32, decimations up to 16 */
struct ProfileVariant {
    int nh;
    profile_launch_fn fn;
};
#define CS_ENTRY_NH_BIG(NH) {NH, cs_launch_profile_##NH},
const ProfileVariant kProfile[] = {CS_NH_BIG_LIST(CS_ENTRY_NH_BIG)};

#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            fprintf(stderr, "Error: HIP %s failed at %s:%d: %s\n", #x, __FILE__,   \
                    __LINE__, hipGetErrorString(e_));                              \
            goto fail;                                                             \
        }                                                                          \
    } while (0)

int wseq_index(int w)
{
    int v = 1;
    for (int i = 0; i < CS_NSEQ; i++) {
        if (v == w) return i;
        int g = (int)(1.5 * (double)v);
        v = g > v + 1 ? g : v + 1;
    }
    return -1;
}

int wcap_of(int nb)
{
    int c = (int)(0.3 * (double)nb);
    if (c < 1) c = 1;
    if (c > nb - 1) c = nb - 1;
    return c;
}

}  // namespace

/* Measured sigma of one (chunk, slot) per workgroup (src/boxcar.c:18-73,
 * 138-167; src/search.c:1005-1050): the sample (|profile| values of the
 * sampled rows, written by the search kernel's sample pass) is sorted in LDS
 * by a bitonic network, sigma = 1.4826 * median (the mean of the two middle
 * values for an even count), and the chunk's per-width factors
 * invsigma*shape_w and valid flag are written for the search pass. */
struct SigmaArgs {
    int nslot;
    int nb[CS_MAXSLOT];
    unsigned int wmask[CS_MAXSLOT];
};

template <typename T>
__global__ __launch_bounds__(256) void sigma_kernel(const T *__restrict__ samp,
                                                    ChunkDesc *__restrict__ chunks,
                                                    const double *__restrict__ shp, SigmaArgs sa)
{
    __shared__ T v[CS_SIGCAP];
    const int c = blockIdx.x / sa.nslot, s = blockIdx.x % sa.nslot;
    ChunkDesc *C = chunks + c;
    const int nb = sa.nb[s], nv = C->nvalid[s];
    int ns;
    if ((long long)nb * nv <= CS_SIGCAP) {
        ns = nb * nv;
    } else {
        int ncol = CS_SIGCAP / nb;
        if (ncol < 1) ncol = 1;
        if (ncol > nv) ncol = nv;
        ns = ncol * nb;
    }
    int P = 1;
    while (P < ns) P <<= 1;
    const T *src = samp + ((long long)c * sa.nslot + s) * CS_SIGCAP;
    for (int i = threadIdx.x; i < P; i += blockDim.x) v[i] = i < ns ? src[i] : (T)INFINITY;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const T x = v[i], y = v[ixj];
                    if ((x > y) == ((i & k) == 0)) {
                        v[i] = y;
                        v[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    if (threadIdx.x != 0) return;
    T sig = 0;
    if (ns > 0) {
        const int half = ns >> 1;
        const T med = (ns & 1) ? v[half] : (T)0.5 * (v[half - 1] + v[half]);
        sig = (T)1.4826 * med;
    }
    const T invsigma = sig > (T)0 ? (T)1 / sig : (T)0;
    C->valid[s] = invsigma > (T)0;
    const double *sh = shp + ((long long)c * sa.nslot + s) * CS_NSEQ;
    for (int wi = 0; wi < CS_NSEQ; wi++) {
        if (!(sa.wmask[s] & (1u << wi))) continue;
        if (sizeof(T) == 4) C->invsw[s][wi] = (float)((double)invsigma * sh[wi]);
        else C->invswd[s][wi] = (double)invsigma * sh[wi];
    }
}

/* HBM the general kernel's work buffer may take per session: with 4 GiB a
 * slice of the two-pass search at nharms 240 is 2048 workgroups, four whole
 * rounds of the chip's wave slots at two waves per SIMD (at 1.5 GiB it was
 * 790, and each launch's second round ran half full; 8 GiB and 4096
 * workgroups measured no faster) */
#define CS_GEN_WORK_BUDGET (4096ll << 20)
/* HBM the two-pass search's stack buffer may take per session (one slice) */
#define CS_SPLIT_BUDGET (1024ll << 20)
/* dynamic LDS of the stack pass: a group's windows up to the first bound
 * when a group of several harmonics fits it, else up to the second */
#define CS_SPLIT_LDS_SOFT (24 << 10)
#define CS_SPLIT_LDS_MAX (48 << 10)

struct GpuSession {
    const SearchParams *p;
    int device;
    int Nprof;
    long long total, nchunks, N, namps;
    double T, r_lo, lodr, nyq;
    int nh, m, q, pnum;
    int nslot;
    int slot_k[CS_MAXSLOT], slot_Hk[CS_MAXSLOT];
    std::vector<long long> cstarts;
    std::vector<ChunkDesc> chunks;
    std::vector<int> tiles;   /* 4 ints per tile */
    long long ht;             /* filled harmonic-trials */
    double alg_bytes;
    const FastVariant *fast;
    hipStream_t st;
    hipEvent_t ev0, ev1;
    void *d_block;                /* spectrum, records, samples and tables in one block */
    bool recs_own;                /* d_recs regrown outside the block */
    float2 *d_amps;
    float *d_wtab;
    int2 *d_hinfo;
    float2 *d_tw;
    int4 *d_tiles;
    ChunkDesc *d_chunks;
    CandRec *d_recs;
    unsigned int *d_counter;
    float2 *d_stack;
    unsigned int cap;
    csk::FastArgs args;
    double last_ms;
    bool f64;                     /* --precision f64: profiles and metrics in double */
    /* the search's launch sequence captured once and replayed while the
     * threshold and the record buffer stay the same */
    bool use_graph;
    hipGraph_t graph;
    hipGraphExec_t gexec;
    double g_thr;
    CandRec *g_recs;
    unsigned int g_cap;
    bool measured;                /* --sigma measured: sample pass + per-chunk sigma */
    std::vector<double> h_shp;    /* shape_w per (chunk, slot, width candidate) */
    double *d_shp;
    void *d_samp;
    bool async_alloc;             /* stream-ordered allocations on st (engine sessions) */
    bool own_stream;
    std::vector<float> h_wtab;    /* host copies kept alive for asynchronous uploads */
    std::vector<int2> h_hinfo;
    void *d_mout;                 /* every trial's metric per slot (metric statistics) */
    bool mout_ready;
    /* the general kernel (configurations no register kernel holds) */
    bool gen;
    CsGenPlan plan;
    void *d_plan;                 /* ops, pool, twiddles */
    void *d_work;                 /* per-lane value slots and stacks */
    size_t work_bytes;
    std::vector<char> h_planbuf;  /* host copy kept alive for the asynchronous upload */
    GenArgs gargs;
    /* the two-pass search (split_kernel.hpp): 0 off; 1 stack pass + register
     * profile pass; 2 stack pass + the general kernel's c2r and boxcar */
    int split;
    int split_m;                  /* padded taps of the stack pass */
    size_t split_lds;             /* its dynamic LDS bytes */
    profile_launch_fn prof;
    csk::SplitArgs sargs;         /* per launch: tile0, nslice */
    int slice;                    /* tiles per slice */
    float2 *d_stk;                /* stack of one slice (split 1) */
    bool slow;                    /* a trial step that is no simple rational (P = 0):
                                     stack_slow_kernel, hr_lo per harmonic in d_wtab */
};

/* Harmonics per window group of the stack pass for nh harmonics, trial step
 * lodr and M padded taps (0 when not even one harmonic's window fits), and
 * the LDS bytes of the largest group.  A group's windows take at most
 * sum_h (ceil(255*h*lodr) + M + 2) bins. */
static int split_group(int nh, double lodr, int M, size_t *lds)
{
    for (int pass = 0; pass < 2; pass++) {
        const double cap = (pass == 0 ? CS_SPLIT_LDS_SOFT : CS_SPLIT_LDS_MAX) / 8.0;
        for (int hg = 64; hg >= (pass == 0 ? 2 : 1); hg >>= 1) {
            double worst = 0.0;
            for (int g0 = 0; g0 < nh; g0 += hg) {
                double need = 0.0;
                for (int h = g0 + 1; h <= std::min(nh, g0 + hg); h++)
                    need += ceil((CS_TILE - 1) * h * lodr) + M + 2;
                worst = std::max(worst, need);
            }
            if (worst <= cap) {
                *lds = ((size_t)worst * 8 + 255) / 256 * 256;
                return hg;
            }
        }
    }
    return 0;
}

/* stack pass launch for the session's padded tap count */
static hipError_t launch_stack_pass(const csk::FastArgs *a, const csk::SplitArgs *sa, int M,
                                    int nb, size_t lds, hipStream_t st)
{
    if (M == 16)
        hipLaunchKernelGGL(csk::stack_pass_kernel<16>, dim3(nb), dim3(CS_TILE), lds, st, *a, *sa);
    else if (M == 32)
        hipLaunchKernelGGL(csk::stack_pass_kernel<32>, dim3(nb), dim3(CS_TILE), lds, st, *a, *sa);
    else if (M == 64)
        hipLaunchKernelGGL(csk::stack_pass_kernel<64>, dim3(nb), dim3(CS_TILE), lds, st, *a, *sa);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

/* cos and sin as two separate libm calls, through pointers the compiler cannot
 * merge into one sincos call (which differs from them in the last bit for some
 * angles): the c2r twiddle definition is libm cos and sin. */
static double (*volatile cs_cos_fn)(double) = cos;
static double (*volatile cs_sin_fn)(double) = sin;

/* The project c2r's twiddles tw(j, N) = (cos, sin)(2 pi j / N), N < CS_TW_NMAX:
 * one table per device, shared by every session and held until the process
 * ends. */
static float2 *shared_twiddles(int device)
{
    static std::mutex mu;
    static float2 *tab[64] = {};
    if (device < 0 || device >= 64) return NULL;
    std::lock_guard<std::mutex> lk(mu);
    if (!tab[device]) {
        std::vector<float2> tw(CS_TW_SIZE);
        for (int Nn = 1; Nn < CS_TW_NMAX; Nn++)
            for (int j = 0; j < Nn; j++) {
                const double ang = 2.0 * M_PI * (double)j / (double)Nn;
                tw[(size_t)cs_tw_off(Nn) + j] =
                    make_float2((float)cs_cos_fn(ang), (float)cs_sin_fn(ang));
            }
        /* the kernels' compile-time twiddles must be this table */
        static_assert(sizeof(kTwF) == sizeof(float) * 2 * CS_TW_SIZE, "twiddle table size");
        if (memcmp(tw.data(), kTwF, sizeof(kTwF)) != 0) {
            fprintf(stderr, "Error: this host's libm cos/sin differ from the twiddles the "
                            "kernels were built with; rebuild the library on this host\n");
            return NULL;
        }
        float2 *d = NULL;
        if (hipMalloc(&d, tw.size() * sizeof(float2)) != hipSuccess) return NULL;
        if (hipMemcpy(d, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice) !=
            hipSuccess) {
            hipFree(d);
            return NULL;
        }
        tab[device] = d;
    }
    return tab[device];
}

/* the same table in double precision for the f64 path */
static double2 *shared_twiddles64(int device)
{
    static std::mutex mu;
    static double2 *tab[64] = {};
    if (device < 0 || device >= 64) return NULL;
    std::lock_guard<std::mutex> lk(mu);
    if (!tab[device]) {
        std::vector<double2> tw(CS_TW_SIZE);
        for (int Nn = 1; Nn < CS_TW_NMAX; Nn++)
            for (int j = 0; j < Nn; j++) {
                const double ang = 2.0 * M_PI * (double)j / (double)Nn;
                tw[(size_t)cs_tw_off(Nn) + j] = make_double2(cs_cos_fn(ang), cs_sin_fn(ang));
            }
        static_assert(sizeof(kTwD) == sizeof(double) * 2 * CS_TW_SIZE, "twiddle table size");
        if (memcmp(tw.data(), kTwD, sizeof(kTwD)) != 0) {
            fprintf(stderr, "Error: this host's libm cos/sin differ from the twiddles the "
                            "kernels were built with; rebuild the library on this host\n");
            return NULL;
        }
        double2 *d = NULL;
        if (hipMalloc(&d, tw.size() * sizeof(double2)) != hipSuccess) return NULL;
        if (hipMemcpy(d, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice) !=
            hipSuccess) {
            hipFree(d);
            return NULL;
        }
        tab[device] = d;
    }
    return tab[device];
}

static unsigned char *scratch_on(int device);

static hipError_t dalloc(GpuSession *S, void **p, size_t bytes)
{
    return S->async_alloc ? hipMallocAsync(p, bytes, S->st) : hipMalloc(p, bytes);
}

static void dfree(GpuSession *S, void *p)
{
    if (!p) return;
    if (S->async_alloc) hipFreeAsync(p, S->st);
    else hipFree(p);
}

static void session_release(GpuSession *S)
{
    if (!S) return;
    if (S->recs_own) dfree(S, S->d_recs);
    dfree(S, S->d_block);
    dfree(S, S->d_stack);
    dfree(S, S->d_mout);
    dfree(S, S->d_plan);
    dfree(S, S->d_work);
    dfree(S, S->d_stk);
    if (S->gexec) hipGraphExecDestroy(S->gexec);
    if (S->graph) hipGraphDestroy(S->graph);
    if (S->ev0) hipEventDestroy(S->ev0);
    if (S->ev1) hipEventDestroy(S->ev1);
    if (S->own_stream && S->st) hipStreamDestroy(S->st);
    delete S;
}
extern "C" CS_API int cs_gpu_available(int device, char *errbuf, size_t errlen)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        if (errbuf) snprintf(errbuf, errlen, "no HIP device available (%s)",
                             e != hipSuccess ? hipGetErrorString(e) : "0 devices");
        return -1;
    }
    if (device < 0 || device >= count) {
        if (errbuf) snprintf(errbuf, errlen,
                             "HIP device %d requested but %d device(s) present",
                             device, count);
        return -1;
    }
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        if (errbuf) snprintf(errbuf, errlen, "cannot select HIP device %d: %s", device,
                             hipGetErrorString(e));
        return -1;
    }
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        if (errbuf) snprintf(errbuf, errlen, "hipGetDeviceProperties failed: %s",
                             hipGetErrorString(e));
        return -1;
    }
    if (errbuf) snprintf(errbuf, errlen, "%s (device %d, %s)", prop.name, device,
                         prop.gcnArchName);
    return 1;
}

/* Per-chunk slot data: filled mask, Nyquist counts, analytic 1/(sigma sd_w). */
static int build_chunks(GpuSession *S, const DirectPlanSet *dp)
{
    const SearchParams *p = S->p;
    const int nh = S->nh;
    std::vector<uint8_t> filled(nh);
    int widths[CS_MAXSLOT][256], nw[CS_MAXSLOT];
    int wsel[CS_MAXSLOT][256];
    for (int s = 0; s < S->nslot; s++) {
        const int nb = 2 * S->slot_Hk[s];
        nw[s] = csx_ladder_widths(nb, S->slot_k[s], p, widths[s], 256);
        unsigned mask = 0;
        for (int i = 0; i < nw[s]; i++) {
            const int wi = wseq_index(widths[s][i]);
            if (wi < 0 || widths[s][i] > wcap_of(nb)) return -2;
            wsel[s][i] = wi;
            mask |= 1u << wi;
        }
        S->args.wmask[s] = mask;
    }
    /* count-keyed shape cache, one per slot (src/search.c:36-46) */
    int shapeH[CS_MAXSLOT];
    static thread_local double shape[CS_MAXSLOT][256];
    for (int s = 0; s < CS_MAXSLOT; s++) shapeH[s] = -1;
    S->chunks.resize((size_t)S->nchunks);
    if (S->measured) S->h_shp.assign((size_t)S->nchunks * S->nslot * CS_NSEQ, 0.0);
    S->tiles.clear();
    S->tiles.reserve((size_t)4 * (size_t)(S->total / CS_TILE + S->nchunks + 1));
    S->ht = 0;
    for (long long c = 0; c < S->nchunks; c++) {
        ChunkDesc &C = S->chunks[(size_t)c];
        memset(&C, 0, sizeof(C));
        const long long i0 = S->cstarts[(size_t)c];
        const int n = (int)((c + 1 < S->nchunks ? S->cstarts[(size_t)c + 1] : S->total) - i0);
        C.i0 = i0;
        C.n = n;
        const double rstart = S->r_lo + (double)i0 * S->lodr;
        int nf_all = 0;
        for (int h = 1; h <= nh; h++) {
            const DirectPlan *pl = &dp->plans[h - 1];
            const int f = csx_harmonic_fits(pl, S->lodr, S->namps, S->N, i0) &&
                          csx_harmonic_fits(pl, S->lodr, S->namps, S->N, i0 + n - 1);
            filled[h - 1] = (uint8_t)f;
            if (f) C.filled[(h - 1) >> 6] |= 1ull << ((h - 1) & 63);
            nf_all += f;
        }
        S->ht += (long long)nf_all * n;
        for (int s = 0; s < S->nslot; s++) {
            const int k = S->slot_k[s], Hk = S->slot_Hk[s], nb = 2 * Hk;
            int nvalid = n;
            if (s > 0) {
                int lo = 0, hi = n;
                while (lo < hi) {
                    const int mid = lo + ((hi - lo) >> 1);
                    if ((double)k * (rstart + (double)mid * S->lodr) < S->nyq) lo = mid + 1;
                    else hi = mid;
                }
                nvalid = lo;
            }
            C.nvalid[s] = nvalid;
            int nf = 0;
            for (int j = 1; j <= Hk; j++) nf += filled[(size_t)j * k - 1];
            if (shapeH[s] != nf) {
                csx_boxcar_shape(shape[s], nb, widths[s], nw[s], filled.data(), k, Hk);
                shapeH[s] = nf;
            }
            /* f32: sigma and 1/sigma in float, invsw = f32(invsigma*shape_w);
             * f64: everything in double (src/search.c:1001-1002, boxcar.c) */
            const double sigd = csx_analytic_sigma(filled.data(), k, Hk);
            const float sig = (float)sigd;
            const float invsigma = sig > 0.0f ? 1.0f / sig : 0.0f;
            const double invsd = sigd > 0.0 ? 1.0 / sigd : 0.0;
            C.valid[s] = S->f64 ? invsd > 0.0 : invsigma > 0.0f;
            for (int i = 0; i < nw[s]; i++) {
                C.invsw[s][wsel[s][i]] = (float)((double)invsigma * shape[s][i]);
                C.invswd[s][wsel[s][i]] = invsd * shape[s][i];
            }
            if (S->measured)
                for (int i = 0; i < nw[s]; i++)
                    S->h_shp[((size_t)c * S->nslot + s) * CS_NSEQ + wsel[s][i]] = shape[s][i];
        }
        for (int j0 = 0; j0 < n; j0 += CS_TILE) {
            S->tiles.push_back((int)c);
            S->tiles.push_back(j0);
            S->tiles.push_back(std::min(CS_TILE, n - j0));
            S->tiles.push_back(0);
        }
    }
    return 0;
}

/* Launch shape of the search kernel: lane order (0 residue classes, 1
 * natural) and the XCD tile mapping (0 contiguous ranges, S > 0 stripes of S
 * tiles).  Neither changes any result; a captured graph is dropped. */
__attribute__((visibility("hidden"))) void cs_session_set_launch(GpuSession *S, int lane_order,
                                                                 int xcd_stripe)
{
    if (!S) return;
    S->args.grp = lane_order == 0 && S->q < CS_TILE ? CS_TILE / S->q : 1;
    S->args.xcd_stripe = xcd_stripe > 0 ? xcd_stripe : 0;
    if (S->gexec) hipGraphExecDestroy(S->gexec);
    if (S->graph) hipGraphDestroy(S->graph);
    S->gexec = NULL;
    S->graph = NULL;
}

/* the persistent grid's workgroups for a plan of nvslots value slots */
static int gen_grid(int nvslots, int nh, bool f64, int ntiles)
{
    const double lane_bytes = (double)cs_gen_work_bytes(nvslots, nh, 1, f64);
    long long g = (long long)((double)CS_GEN_WORK_BUDGET / (lane_bytes * CS_TILE));
    g = std::min<long long>(g, CS_GEN_GRID);
    g = std::min<long long>(g, std::max(ntiles, 1));
    return (int)std::max<long long>(g, 1);
}

/* The general kernel's plan, tables and work buffer, and its fixed
 * arguments (the per-launch ones are set in enqueue_search_into). */
static int gen_setup(GpuSession *S)
{
    int Hk[CS_MAXSLOT], ks[CS_MAXSLOT], nextra = 0;
    GenArgs &g = S->gargs;
    memset((void *)&g, 0, sizeof(g));
    for (int s = 0; s < S->nslot; s++) {
        const int k = S->slot_k[s];
        g.slot_k[s] = k;
        g.slot_Hk[s] = S->slot_Hk[s];
        g.wmask[s] = S->args.wmask[s];
        if (s == 0) {
            g.slot_c2r[s] = 0;
        } else if (S->nh % k == 0) {
            g.slot_c2r[s] = -1;                 /* a fold of the base profile */
        } else {
            Hk[nextra] = S->slot_Hk[s];
            ks[nextra] = k;
            g.slot_c2r[s] = ++nextra;
        }
        if (g.wmask[s] >> 14) {
            fprintf(stderr, "Error: boxcar widths above 211 are not supported by the GPU "
                            "search\n");
            return -1;
        }
    }
    if (cs_gen_build_plan(S->nh, nextra, Hk, ks, &S->plan) != 0) {
        fprintf(stderr, "Error: c2r plan for %d harmonics failed\n", S->nh);
        return -1;
    }
    const CsGenPlan &P = S->plan;
    const size_t tsz = S->f64 ? 16 : 8;
    size_t off = 0;
    const size_t o_ops = cs_part(&off, P.ops.size() * sizeof(CsGenOp));
    const size_t o_pool = cs_part(&off, P.pool.size() * sizeof(int));
    const size_t o_tw = cs_part(&off, std::max<size_t>(P.twc.size(), 1) * tsz);
    const size_t bytes = off;
    S->h_planbuf.assign(bytes, 0);
    char *h = S->h_planbuf.data();
    memcpy(h + o_ops, P.ops.data(), P.ops.size() * sizeof(CsGenOp));
    memcpy(h + o_pool, P.pool.data(), P.pool.size() * sizeof(int));
    for (size_t i = 0; i < P.twc.size(); i++) {
        if (S->f64) {
            const double v[2] = {P.twc[i], P.tws[i]};
            memcpy(h + o_tw + i * tsz, v, tsz);
        } else {
            const float v[2] = {(float)P.twc[i], (float)P.tws[i]};
            memcpy(h + o_tw + i * tsz, v, tsz);
        }
    }
    if (dalloc(S, &S->d_plan, bytes) != hipSuccess ||
        hipMemcpyAsync(S->d_plan, h, bytes, hipMemcpyHostToDevice, S->st) != hipSuccess) {
        fprintf(stderr, "Error: c2r plan upload failed\n");
        return -1;
    }
    /* value slots, butterfly scratch, then the fold of the base profile
     * (nh / 2 complex slots: a fold by k has nh / k bin pairs) */
    const int nv = P.nslots + P.nscratch + S->nh / 2;
    const int grid = gen_grid(nv, S->nh, S->f64, S->args.ntiles);
    g.nlanes = (long long)grid * CS_TILE;
    const size_t wbytes = cs_gen_work_bytes(nv, S->nh, g.nlanes, S->f64);
    if (dalloc(S, &S->d_work, wbytes) != hipSuccess) {
        fprintf(stderr, "Error: general-kernel work buffer of %zu bytes failed\n", wbytes);
        return -1;
    }
    S->work_bytes = wbytes;
    g.amps = S->d_amps;
    g.namps = S->namps;
    g.wtab = S->d_wtab;
    g.hinfo = S->d_hinfo;
    g.tiles = S->d_tiles;
    g.ntiles = S->args.ntiles;
    g.q = S->q;
    g.pnum = S->pnum;
    g.ws = S->args.ws;
    g.m = S->m;
    g.mp = S->args.m2e + S->m / 2;      /* the table's padded tap count */
    g.nh = S->nh;
    g.chunks = S->d_chunks;
    g.nslot = S->nslot;
    g.ops = (const CsGenOp *)((char *)S->d_plan + o_ops);
    g.pool = (const int *)((char *)S->d_plan + o_pool);
    g.tw = (const void *)((char *)S->d_plan + o_tw);
    g.nc2r = (int)P.c2r.size();
    for (int c = 0; c < g.nc2r; c++) g.c2r[c] = P.c2r[(size_t)c];
    g.nslots = P.nslots;
    g.nvslots = nv;
    g.fslot = P.nslots + P.nscratch;
    g.work = S->d_work;
    return 0;
}

/* Device bytes beyond the tables that a session of p at trial step lodr
 * takes for the two-pass search's stack buffer or the general kernel's work
 * buffer (at most), 0 when a one-pass register kernel serves it -- the
 * engine's admission estimate. */
__attribute__((visibility("hidden"))) long long cs_session_general_bytes(const SearchParams *p,
                                                                        double lodr)
{
    const int nh = p->nharms, m = p->m, mp = m <= 16 ? 16 : (m <= 32 ? 32 : 0);
    const bool f64 = p->precision == CS_PREC_F64;
    int pn = 0, qd = 0;
    const bool rational = csx_trial_grid_rational(lodr, &pn, &qd) != 0;
    for (const FastVariant &v : kFast) {
        if (!rational) break;
        if (v.nh != nh || v.m != mp || (f64 && !v.fn64)) continue;
        double need = 0.0;
        for (int h = 1; h <= nh; h++) need += ceil((CS_TILE - 1) * h * lodr) + mp + 2;
        if (need <= (double)CS_WIN_CAP(nh, mp)) return 0;
    }
    size_t lds = 0;
    const int ms = m <= 16 ? 16 : (m <= 32 ? 32 : 64);
    if (!f64 && nh >= 65 && nh <= 128 && (!rational || split_group(nh, lodr, ms, &lds) > 0))
        return CS_SPLIT_BUDGET + (16ll << 20);
    return CS_GEN_WORK_BUDGET + (16ll << 20);
}

/* ext: run on the caller's stream with stream-ordered allocations (the
 * engine's concurrent sessions); null: own stream, plain allocations. */
__attribute__((visibility("hidden"))) GpuSession *cs_session_create_on(const FFTFile *ft, const SearchParams *p,
                                 const DirectPlanSet *dplans, const long long *cstarts,
                                 long long total, int Nprof, long long nchunks, double r_lo,
                                 double lodr, int device, hipStream_t ext)
{
    if (!ft || !p || !dplans || (!cstarts && nchunks > 0)) return NULL;
    GpuSession *S = new GpuSession();
    memset((void *)&S->args, 0, sizeof(S->args));
    S->p = p;
    S->device = device;
    S->Nprof = Nprof;
    S->total = total;
    S->nchunks = nchunks;
    S->N = ft->N;
    S->namps = ft->namps;
    S->T = ft->T;
    S->r_lo = r_lo;
    S->lodr = lodr;
    S->nyq = (double)ft->N / 2.0;
    S->nh = p->nharms;
    S->m = p->m;
    S->cstarts.assign(cstarts, cstarts + nchunks);
    S->d_block = NULL; S->recs_own = false;
    S->d_amps = NULL; S->d_wtab = NULL; S->d_hinfo = NULL; S->d_tw = NULL;
    S->d_tiles = NULL; S->d_chunks = NULL; S->d_recs = NULL; S->d_counter = NULL;
    S->d_stack = NULL; S->st = NULL; S->ev0 = NULL; S->ev1 = NULL;
    S->async_alloc = false;
    S->own_stream = false;
    S->f64 = p->precision == CS_PREC_F64;
    S->use_graph = ext == NULL;
    S->graph = NULL;
    S->gexec = NULL;
    S->g_recs = NULL;
    S->g_cap = 0;
    S->g_thr = 0.0;
    S->measured = p->sigma == CS_SIGMA_MEASURED;
    S->d_shp = NULL;
    S->d_samp = NULL;
    S->last_ms = 0.0;
    S->fast = NULL;
    S->gen = false;
    S->d_plan = NULL;
    S->d_work = NULL;
    S->work_bytes = 0;
    memset((void *)&S->gargs, 0, sizeof(S->gargs));
    S->split = 0;
    S->split_m = 0;
    S->split_lds = 0;
    S->prof = NULL;
    memset((void *)&S->sargs, 0, sizeof(S->sargs));
    S->slice = 0;
    S->d_stk = NULL;
    S->slow = false;

    const int nh = p->nharms, m = p->m;
    std::vector<float> wtab;
    std::vector<int2> hinfo;
    int ws;
    int rc;
    int mp = 0;

    if (dplans->n != nh || nh < 1) {
        fprintf(stderr, "Error: plan set has %d harmonics, expected %d\n", dplans->n, nh);
        goto fail;
    }
    if ((p->precision != CS_PREC_F32 && p->precision != CS_PREC_F64) ||
        (p->sigma != CS_SIGMA_ANALYTIC && p->sigma != CS_SIGMA_MEASURED)) {
        fprintf(stderr, "Error: unknown --precision or --sigma\n");
        goto fail;
    }
    /* slots: base, then decimations k > 1 in SearchParams order */
    S->nslot = 1;
    S->slot_k[0] = 1;
    S->slot_Hk[0] = nh;
    for (int i = 0; i < CS_MAXSLOT; i++) S->args.kslot[i] = -1;
    for (int i = 0; i < p->ndecim; i++) {
        const int k = p->decimations[i];
        if (k <= 1) continue;
        if (S->nslot >= CS_MAXSLOT || k >= CS_MAXSLOT || S->args.kslot[k] >= 0 || nh / k < 1) {
            fprintf(stderr, "Error: decimation %d is not supported by the GPU search\n", k);
            goto fail;
        }
        S->args.kslot[k] = S->nslot;
        S->slot_k[S->nslot] = k;
        S->slot_Hk[S->nslot] = nh / k;
        S->nslot++;
    }
    /* a trial step that is no simple rational: per-trial FP64 positions and
     * coefficients (the stack pass's non-rational form), in f32 only -- the
     * reference keeps a double stack for it in f64 (src/interp.c:464-467) */
    S->slow = dplans->plans[0].P == 0;
    if (S->slow && S->f64) {
        fprintf(stderr, "Error: --precision f64 needs a simple rational --hidr on the GPU "
                        "(the non-rational form keeps a double stack)\n");
        goto fail;
    }
    /* the register kernel of the smallest padded tap count M >= m, when one
     * holds this configuration and the tile's windows fit its LDS buffer;
     * otherwise the general kernel */
    mp = m <= 16 ? 16 : (m <= 32 ? 32 : 0);
    for (const FastVariant &v : kFast) {
        if (S->slow) break;
        if (v.nh != nh || v.m != mp || (S->f64 && !v.fn64)) continue;
        bool ok = true;
        for (int s = 1; s < S->nslot; s++) ok = ok && S->slot_k[s] <= v.kmax;
        if (ok) { S->fast = &v; break; }
    }
    if (S->fast) {
        double need = 0.0;
        for (int h = 1; h <= nh; h++) need += ceil((CS_TILE - 1) * h * lodr) + mp + 2;
        if (need > (double)CS_WIN_CAP(nh, mp)) S->fast = NULL;
    }
    if (!S->fast) {
        if (nh > CS_GEN_MAXH || m < 2 || m > 64 || (m & 1)) {
            fprintf(stderr, "Error: no GPU kernel for nharms=%d m=%d (the GPU search holds up to "
                            "%d harmonics and an even m up to 64)\n", nh, m, CS_GEN_MAXH);
            goto fail;
        }
        for (int s = 1; s < S->nslot; s++)
            if (S->slot_k[s] > CS_KMAX) {
                fprintf(stderr, "Error: decimation %d is not supported by the GPU search\n",
                        S->slot_k[s]);
                goto fail;
            }
        /* the two-pass search whenever a window group fits the stack pass's
         * LDS: its register profile pass for 65-128 harmonics in f32, the
         * general kernel's c2r and boxcar otherwise */
        const int ms = m <= 16 ? 16 : (m <= 32 ? 32 : 64);
        const int hg = S->slow ? 64 : split_group(nh, lodr, ms, &S->split_lds);
        if (S->slow) S->split_lds = 0;
        if (hg > 0) {
            S->split_m = ms;
            S->sargs.hg = hg;
            S->sargs.nh = nh;
            if (!S->f64 && nh >= 65 && nh <= 128) {
                S->split = 1;
                S->prof = kProfile[nh - 65].fn;
            } else {
                S->split = 2;
            }
        }
        S->gen = S->split != 1;
        mp = S->split ? ms : (m + 3) / 4 * 4;
    }
    S->q = S->slow ? 1 : dplans->plans[0].q;
    S->pnum = S->slow ? 0 : dplans->plans[0].pnum;
    for (int h = 0; h < nh; h++)
        if ((dplans->plans[h].P == 0) != S->slow || dplans->plans[h].m != m ||
            (!S->slow && (dplans->plans[h].P < 0 || dplans->plans[h].q != S->q ||
                          dplans->plans[h].pnum != S->pnum))) {
            fprintf(stderr, "Error: inconsistent plan set for the GPU search\n");
            goto fail;
        }
    /* 32-bit residue walk bounds: h*rho < 2^31, h*qbase < 2^31 */
    if ((double)nh * (double)S->q >= 2147483647.0 ||
        (double)nh * ((double)total * S->pnum / S->q + 2.0) >= 2147483000.0) {
        fprintf(stderr, "Error: trial grid too large for the GPU residue walk\n");
        goto fail;
    }

    /* plan arena: row (h, res) = M - m zeros, W[0..m-1], A.re, A.im, zero
     * padding to CS_WSP(M) floats; row index (h-1)*q + res must fit 32-bit
     * offsets */
    ws = CS_WSP(mp);
    if ((double)nh * S->q * ws >= 4294967295.0 / 4) {
        fprintf(stderr, "Error: weight table too large\n");
        goto fail;
    }
    wtab.assign((size_t)nh * S->q * ws, 0.0f);
    hinfo.resize((size_t)nh);
    if (S->slow) {
        /* hr_lo = rfloor0 + dr0 per harmonic, as double, in the table's place */
        wtab.assign((size_t)2 * nh, 0.0f);
        for (int h = 1; h <= nh; h++) {
            const DirectPlan *pl = &dplans->plans[h - 1];
            const double v = (double)(pl->rfloor0 + pl->dr0);
            memcpy(&wtab[(size_t)2 * (h - 1)], &v, sizeof(double));
            hinfo[(size_t)h - 1] = make_int2(pl->rfloor0, 1);
        }
    }
    for (int h = 1; h <= nh && !S->slow; h++) {
        const DirectPlan *pl = &dplans->plans[h - 1];
        int cthr = S->q;
        for (int res = 0; res < S->q; res++) {
            if (pl->row[res] == 0) continue;
            const int ph = pl->row[res] - 1;
            float *row = &wtab[((size_t)(h - 1) * S->q + res) * ws];
            memcpy(row + (mp - m), pl->W + (size_t)ph * m, (size_t)m * sizeof(float));
            row[mp] = pl->A_re[ph];
            row[mp + 1] = pl->A_im[ph];
            if (pl->carry[ph] && res < cthr) cthr = res;
        }
        for (int res = 0; res < S->q; res++) {
            if (pl->row[res] == 0) continue;
            if ((pl->carry[pl->row[res] - 1] != 0) != (res >= cthr)) {
                fprintf(stderr, "Error: harmonic %d carry table is not monotone\n", h);
                goto fail;
            }
        }
        hinfo[(size_t)h - 1] = make_int2(pl->rfloor0, cthr);
    }
    rc = build_chunks(S, dplans);
    if (rc == -2) {
        fprintf(stderr, "Error: boxcar width bank outside the GPU kernel's candidates "
                        "(fsp 1.5, maxfrac <= 0.3)\n");
        goto fail;
    }
    {
        /* algorithmic bytes of the reference GPU dataflow (per trial) */
        bool anyfold = false;
        double b = 16.0 * (nh + 1);
        for (int s = 1; s < S->nslot; s++) {
            const int k = S->slot_k[s], Hk = S->slot_Hk[s];
            if (nh % k == 0) anyfold = true;
            else b += 8.0 * (Hk + 1) + 16.0 * Hk;
        }
        b += 8.0 * nh * (2 + (anyfold ? 1 : 0));
        S->alg_bytes = b * (double)total + 8.0 * (double)S->namps;
    }

    HIPCHK(hipSetDevice(device));
    (void)scratch_on(device);     /* set-up of the no-allocation search's scratch */
    if (ext) {
        S->st = ext;
        S->async_alloc = true;
    } else {
        HIPCHK(hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking));
        S->own_stream = true;
    }
    HIPCHK(hipEventCreate(&S->ev0));
    HIPCHK(hipEventCreate(&S->ev1));
    S->h_wtab.swap(wtab);
    S->h_hinfo.swap(hinfo);
    S->d_tw = shared_twiddles(device);
    S->args.twd = S->f64 ? shared_twiddles64(device) : NULL;
    if (!S->d_tw || (S->f64 && !S->args.twd)) {
        fprintf(stderr, "Error: twiddle table upload failed\n");
        goto fail;
    }
    S->cap = 1u << 16;
    {
        /* one device block: spectrum, records, counter, measured-sigma
         * samples, then the tables in upload order (one pinned copy) */
        const bool samp = S->measured && S->nchunks > 0;
        const size_t nsamp = samp ? (size_t)S->nchunks * S->nslot * CS_SIGCAP : 0;
        size_t off = 0;
        const size_t o_amps =
            cs_part(&off, (size_t)std::max<long long>(S->namps, 1) * sizeof(float2));
        const size_t o_recs = cs_part(&off, (size_t)S->cap * sizeof(CandRec));
        const size_t o_cnt = cs_part(&off, sizeof(unsigned int));
        const size_t o_samp = cs_part(&off, nsamp * (S->f64 ? sizeof(double) : sizeof(float)));
        const size_t o_tab = off;
        size_t toff = 0;
        const size_t t_wtab = cs_part(&toff, S->h_wtab.size() * sizeof(float));
        const size_t t_hinfo = cs_part(&toff, S->h_hinfo.size() * sizeof(int2));
        const size_t t_chunks = cs_part(&toff, std::max<size_t>(S->chunks.size(), 1) * sizeof(ChunkDesc));
        const size_t t_tiles = cs_part(&toff, std::max<size_t>(S->tiles.size(), 4) * sizeof(int));
        const size_t t_shp = cs_part(&toff, (samp ? S->h_shp.size() : 0) * sizeof(double));
        HIPCHK(dalloc(S, &S->d_block, o_tab + toff));
        char *b = (char *)S->d_block;
        S->d_amps = (float2 *)(b + o_amps);
        S->d_recs = (CandRec *)(b + o_recs);
        S->d_counter = (unsigned int *)(b + o_cnt);
        S->d_samp = samp ? (void *)(b + o_samp) : NULL;
        char *tb = b + o_tab;
        S->d_wtab = (float *)(tb + t_wtab);
        S->d_hinfo = (int2 *)(tb + t_hinfo);
        S->d_chunks = (ChunkDesc *)(tb + t_chunks);
        S->d_tiles = (int4 *)(tb + t_tiles);
        S->d_shp = samp ? (double *)(tb + t_shp) : NULL;
        char *h = (char *)cs_upload_buffer(toff);
        if (!h) {
            fprintf(stderr, "Error: pinned staging buffer of %zu bytes failed\n", toff);
            goto fail;
        }
        memset(h, 0, toff);
        memcpy(h + t_wtab, S->h_wtab.data(), S->h_wtab.size() * sizeof(float));
        memcpy(h + t_hinfo, S->h_hinfo.data(), S->h_hinfo.size() * sizeof(int2));
        if (!S->chunks.empty())
            memcpy(h + t_chunks, S->chunks.data(), S->chunks.size() * sizeof(ChunkDesc));
        if (!S->tiles.empty())
            memcpy(h + t_tiles, S->tiles.data(), S->tiles.size() * sizeof(int));
        if (samp) memcpy(h + t_shp, S->h_shp.data(), S->h_shp.size() * sizeof(double));
        HIPCHK(cs_upload(tb, toff, S->st));
    }

    S->args.amps = S->d_amps;
    S->args.namps = S->namps;
    S->args.wtab = S->d_wtab;
    S->args.hinfo = S->d_hinfo;
    S->args.tw = S->d_tw;
    S->args.tiles = S->d_tiles;
    S->args.ntiles = (int)(S->tiles.size() / 4);
    S->args.q = S->q;
    S->args.pnum = S->pnum;
    S->args.ws = ws;
    S->args.grp = S->q < CS_TILE ? CS_TILE / S->q : 1;
    S->args.m2e = mp - m / 2;
    S->args.nslot = S->nslot;
    S->args.samp = NULL;
    S->args.chunks = S->d_chunks;
    S->args.recs = S->d_recs;
    S->args.counter = S->d_counter;
    S->args.cap = S->cap;
    S->args.xcd_stripe = 0;
    if (S->gen) {
        if (gen_setup(S) != 0) goto fail;
    }
    if (S->split == 1) {
        /* the stack of a slice of tiles, up to the budget */
        const long long per = (long long)nh * CS_TILE * (long long)sizeof(float2);
        S->slice = (int)std::max<long long>(
            1, std::min<long long>(std::max(S->args.ntiles, 1), CS_SPLIT_BUDGET / per));
        HIPCHK(dalloc(S, (void **)&S->d_stk, (size_t)S->slice * (size_t)per));
        S->sargs.stk = S->d_stk;
        S->sargs.tstride = (long long)nh * CS_TILE;
        S->sargs.hstride = CS_TILE;
        S->sargs.hbase = 0;
    } else if (S->split == 2) {
        /* straight into the general kernel's stack entries: slice tile u is
         * its workgroup u, harmonic h at X + h*nlanes + u*256 + lane */
        S->slice = (int)(S->gargs.nlanes / CS_TILE);
        S->sargs.stk = (float2 *)((char *)S->d_work + (size_t)S->gargs.nvslots *
                                                          (size_t)S->gargs.nlanes *
                                                          (S->f64 ? 16 : 8));
        S->sargs.tstride = CS_TILE;
        S->sargs.hstride = S->gargs.nlanes;
        S->sargs.hbase = 1;
    }
    if (S->split) {
        S->sargs.m = m;
        S->sargs.lodr = lodr;
        /* CS_SPLIT_SLICE=n: at most n tiles per slice (tests of the slice loop) */
        if (const char *e = getenv("CS_SPLIT_SLICE")) {
            const int n = atoi(e);
            if (n > 0 && n < S->slice) S->slice = n;
        }
        /* stripes of 16 tiles per XCD: neighbouring tiles' windows share L2 */
        S->args.xcd_stripe = 16;
    } else if (!S->gen) {
        /* a tuned launch for this kernel class, when there is one */
        CsTuneResult tr;
        if (cs_tune_lookup(device, nh, mp, S->f64, S->measured, &tr))
            cs_session_set_launch(S, tr.lane_order, tr.xcd_stripe);
    }
    return S;
fail:
    session_release(S);
    /* the failed call's error must not be reported by a later launch check */
    (void)hipGetLastError();
    return NULL;
}

extern "C" CS_API GpuSession *cs_gpu_session_create(const FFTFile *ft, const SearchParams *p,
                                                    const DirectPlanSet *dplans,
                                                    const long long *cstarts, long long total,
                                                    int Nprof, long long nchunks,
                                                    double r_lo, double lodr, int device)
{
    return cs_session_create_on(ft, p, dplans, cstarts, total, Nprof, nchunks, r_lo, lodr,
                                device, NULL);
}

extern "C" CS_API int cs_gpu_session_matches(const GpuSession *S, const FFTFile *ft,
                                             const SearchParams *p, int Nprof,
                                             long long total, double r_lo, double lodr,
                                             int device)
{
    if (!S) return 0;
    return S->p == p && S->device == device && S->Nprof == Nprof && S->total == total &&
           S->T == ft->T && S->N == ft->N && S->namps == ft->namps && S->r_lo == r_lo &&
           S->lodr == lodr;
}

__global__ void cs_zero_u32(unsigned int *p) { p[0] = 0u; }

/* The launch sequence of one search into a record buffer of `cap` records
 * and its counter: the counter cleared by a kernel (also inside a captured
 * graph), the measured-sigma passes, the search kernel. */
namespace csk {

/* sinpi / cospi of the reference (src/interp.c:17-43): exact at the
 * multiples of 1/2, sin and cos of pi*r otherwise (r = x mod 2) */
__device__ __forceinline__ double ref_sinpi(double x)
{
    if (!isfinite(x)) return x - x;
    double r = fmod(x, 2.0);
    if (r < 0) r += 2.0;
    if (r == 0.0 || r == 1.0) return 0.0 * r;
    if (r == 0.5) return 1.0;
    if (r == 1.5) return -1.0;
    return sin(M_PI * r);
}
__device__ __forceinline__ double ref_cospi(double x)
{
    if (!isfinite(x)) return x - x;
    double r = fmod(x, 2.0);
    if (r < 0) r += 2.0;
    if (r == 0.0) return 1.0;
    if (r == 0.5 || r == 1.5) return 0.0;
    if (r == 1.0) return -1.0;
    return cos(M_PI * r);
}

/* Stack pass for a trial step that is no simple rational (plans with P = 0,
 * src/interp.c:427-475): per trial and harmonic the FP64 position
 * r = hr_lo[h] + t*(h*lodr), the taps around floor(r + 1e-15), weights
 * 1/(dr - (i - m/2)) and the rotation (sin cos, -sin^2)/pi of the fractional
 * offset, every product rounded and then added in double (no contraction),
 * rounded to the float stack.  No table: the spectrum is read straight from
 * HBM, neighbouring lanes' taps sharing cache lines.  hr_lo[h-1] =
 * rfloor0 + dr0 of harmonic h, as double. */
__global__ __launch_bounds__(CS_TILE) void stack_slow_kernel(FastArgs a, SplitArgs s,
                                                             const double *__restrict__ hr_lo)
{
    const int u = slice_tile(a);
    if (u >= s.nslice) return;
    const int4 Tl = a.tiles[s.tile0 + u];
    const ChunkDesc *C = a.chunks + Tl.x;
    const int tid = threadIdx.x;
    const bool active = tid < Tl.z;
    const int lj = Tl.y + (active ? tid : Tl.z - 1);
    const long long t = C->i0 + lj;
    const int m = s.m, m2 = m / 2;
    const long long namps = a.namps;
    float2 *out = s.stk + (long long)u * s.tstride + tid;
    for (int h = 1; h <= s.nh; h++) {
        float2 v = make_float2(0.0f, 0.0f);
        if ((C->filled[(h - 1) >> 6] >> ((h - 1) & 63)) & 1ull) {
            const double dh = (double)h * s.lodr;
            const double r = hr_lo[h - 1] + (double)t * dh;
            const long long rint0 = (long long)floor(r + 1e-15) + 1;
            double dr = fmod(r, 1.0);
            if (dr < 0) dr += 1.0;
            const long long base = rint0 - m2;        /* 1-based */
            double xr, xi;
            if (dr == 0.0) {
                const long long idx = base + m2 - 1;
                const float2 b = idx >= 0 && idx < namps ? a.amps[idx] : make_float2(0.0f, 0.0f);
                xr = (double)b.x;
                xi = (double)b.y;
            } else {
                double sre = 0.0, sim = 0.0;
                for (int i = 1; i <= m; i++) {
                    const double w = 1.0 / (dr - (double)(i - m2));
                    const long long idx = base + i - 1;
                    const float2 b =
                        idx >= 0 && idx < namps ? a.amps[idx] : make_float2(0.0f, 0.0f);
                    sre += w * (double)b.x;
                    sim += w * (double)b.y;
                }
                const double sn = ref_sinpi(dr), cs = ref_cospi(dr);
                const double pr = sn * cs / M_PI;
                const double pi_ = -sn * sn / M_PI;
                xr = pr * sre - pi_ * sim;
                xi = pr * sim + pi_ * sre;
            }
            v = make_float2((float)xr, (float)xi);
        }
        out[(long long)(h - 1 + s.hbase) * s.hstride] = v;
    }
}

}  // namespace csk

/* the session's stack pass over one slice */
static hipError_t launch_stack(const GpuSession *S, const csk::FastArgs *a,
                               const csk::SplitArgs *sa, int nb, hipStream_t st)
{
    if (!S->slow) return launch_stack_pass(a, sa, S->split_m, nb, S->split_lds, st);
    hipLaunchKernelGGL(csk::stack_slow_kernel, dim3(nb), dim3(CS_TILE), 0, st, *a, *sa,
                       (const double *)S->d_wtab);
    return hipGetLastError();
}

static int enqueue_search_into(GpuSession *S, double threshold, float2 *stack_out,
                               CandRec *recs, unsigned int *counter, unsigned int cap)
{
    S->args.stack_out = stack_out;
    S->args.threshold = threshold;
    S->args.recs = recs;
    S->args.counter = counter;
    S->args.cap = cap;
    const int ntiles = S->args.ntiles;
    const int unit = 8 * std::max(S->args.xcd_stripe, 1);
    const int nblk = ((ntiles + unit - 1) / unit) * unit;
    auto fn = [S, unit](const csk::FastArgs *a, int nb, hipStream_t st) -> hipError_t {
        if (S->split == 1) {
            /* per slice: the stack pass, then the profile pass */
            for (int t0 = 0; t0 < a->ntiles; t0 += S->slice) {
                csk::SplitArgs sa = S->sargs;
                sa.tile0 = t0;
                sa.nslice = std::min(S->slice, a->ntiles - t0);
                const int sb = (sa.nslice + unit - 1) / unit * unit;
                hipError_t e = launch_stack(S, a, &sa, sb, st);
                if (e != hipSuccess) return e;
                e = S->prof(a, &sa, sb, st);
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        }
        if (!S->gen) return (S->f64 ? S->fast->fn64 : S->fast->fn)(a, nb, st);
        GenArgs &g = S->gargs;
        g.recs = a->recs;
        g.counter = a->counter;
        g.cap = a->cap;
        g.threshold = a->threshold;
        g.samp = a->samp;
        g.mout = a->mout;
        g.mstride = a->mstride;
        g.stack_out = a->stack_out;
        g.dump_t0 = a->dump_t0;
        g.dump_n = a->dump_n;
        g.grp = a->grp;
        if (S->split == 2) {
            /* per slice: the stack pass into the work buffer, then the
             * general kernel's c2r and boxcar, workgroup u on slice tile u */
            for (int t0 = 0; t0 < a->ntiles; t0 += S->slice) {
                csk::SplitArgs sa = S->sargs;
                sa.tile0 = t0;
                sa.nslice = std::min(S->slice, a->ntiles - t0);
                const int sb = (sa.nslice + unit - 1) / unit * unit;
                hipError_t e = launch_stack(S, a, &sa, sb, st);
                if (e != hipSuccess) return e;
                g.tile0 = t0;
                g.nslice = sa.nslice;
                g.given = 1;
                e = cs_gen_launch(&g, S->f64, st);
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        }
        g.given = 0;
        return cs_gen_launch(&g, S->f64, st);
    };
    hipLaunchKernelGGL(cs_zero_u32, dim3(1), dim3(1), 0, S->st, counter);
    if (hipGetLastError() != hipSuccess) return -1;
    if (S->measured && ntiles > 0) {
        /* sample pass, then sigma and the per-width factors of every chunk */
        S->args.samp = S->d_samp;
        float2 *keep = S->args.stack_out;
        S->args.stack_out = NULL;
        const hipError_t e = fn(&S->args, nblk, S->st);
        S->args.samp = NULL;
        S->args.stack_out = keep;
        if (e != hipSuccess) return -1;
        SigmaArgs sa;
        memset(&sa, 0, sizeof(sa));
        sa.nslot = S->nslot;
        for (int s = 0; s < S->nslot; s++) {
            sa.nb[s] = 2 * S->slot_Hk[s];
            sa.wmask[s] = S->args.wmask[s];
        }
        const unsigned grid = (unsigned)(S->nchunks * S->nslot);
        if (S->f64)
            hipLaunchKernelGGL(sigma_kernel<double>, dim3(grid), dim3(256), 0, S->st,
                               (const double *)S->d_samp, S->d_chunks, S->d_shp, sa);
        else
            hipLaunchKernelGGL(sigma_kernel<float>, dim3(grid), dim3(256), 0, S->st,
                               (const float *)S->d_samp, S->d_chunks, S->d_shp, sa);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    if (ntiles > 0 && fn(&S->args, nblk, S->st) != hipSuccess) return -1;
    return 0;
}

static int enqueue_search(GpuSession *S, double threshold, float2 *stack_out)
{
    return enqueue_search_into(S, threshold, stack_out, S->d_recs, S->d_counter, S->cap);
}

/* One search: the counter reset, the measured-sigma passes and the search
 * kernel, between the timing events.  Without a stack dump, sessions on their
 * own stream replay a captured graph of the sequence (kernels only). */
static int launch_search(GpuSession *S, double threshold, float2 *stack_out)
{
    hipEventRecord(S->ev0, S->st);
    if (!S->use_graph || stack_out) {
        if (enqueue_search(S, threshold, stack_out) != 0) return -1;
        hipEventRecord(S->ev1, S->st);
        return 0;
    }
    /* the graph holds the record buffer and its capacity by value: a regrown
     * buffer (possibly at the same address) needs a new capture */
    if (!S->gexec || S->g_thr != threshold || S->g_recs != S->d_recs || S->g_cap != S->cap) {
        if (S->gexec) hipGraphExecDestroy(S->gexec);
        if (S->graph) hipGraphDestroy(S->graph);
        S->gexec = NULL;
        S->graph = NULL;
        hipError_t e = hipStreamBeginCapture(S->st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            fprintf(stderr, "Error: search graph capture: %s\n", hipGetErrorString(e));
            return -1;
        }
        const int rc = enqueue_search(S, threshold, NULL);
        e = hipStreamEndCapture(S->st, &S->graph);
        if (e != hipSuccess || rc != 0) {
            fprintf(stderr, "Error: search graph capture (%d): %s\n", rc, hipGetErrorString(e));
            if (e == hipSuccess && S->graph) hipGraphDestroy(S->graph);
            S->graph = NULL;
            return -1;
        }
        e = hipGraphInstantiate(&S->gexec, S->graph, NULL, NULL, 0);
        if (e != hipSuccess) {
            fprintf(stderr, "Error: search graph instantiate: %s\n", hipGetErrorString(e));
            S->gexec = NULL;
            return -1;
        }
        S->g_thr = threshold;
        S->g_recs = S->d_recs;
        S->g_cap = S->cap;
    }
    const hipError_t el = hipGraphLaunch(S->gexec, S->st);
    if (el != hipSuccess) {
        fprintf(stderr, "Error: search graph launch: %s\n", hipGetErrorString(el));
        return -1;
    }
    hipEventRecord(S->ev1, S->st);
    return 0;
}

extern "C" CS_API int cs_gpu_session_upload(GpuSession *S, const FFTFile *ft)
{
    if (!S || !ft) return -1;
    if (ft->namps != S->namps || ft->N != S->N) {
        fprintf(stderr, "Error: GPU session built for N=%lld/%lld amps but the input has "
                        "N=%lld/%lld\n", S->N, S->namps, ft->N, ft->namps);
        return -1;
    }
    if (hipSetDevice(S->device) != hipSuccess) return -1;
    if (S->namps > 0 &&
        hipMemcpy(S->d_amps, ft->amps, (size_t)S->namps * sizeof(float2),
                  hipMemcpyHostToDevice) != hipSuccess) {
        fprintf(stderr, "Error: spectrum upload failed\n");
        return -1;
    }
    S->T = ft->T;
    return 0;
}

extern "C" CS_API int cs_gpu_session_stage(GpuSession *S, CsStage *g, const float *samples,
                                           float *amps_host)
{
    if (!S || !g || !samples) return -1;
    if (cs_stage_N(g) != S->N || S->N / 2 != S->namps) {
        fprintf(stderr, "Error: stage plan for N=%lld does not match the session's N=%lld/%lld "
                        "amps\n", cs_stage_N(g), S->N, S->namps);
        return -1;
    }
    if (hipSetDevice(S->device) != hipSuccess) return -1;
    if (cs_stage_enqueue(g, samples, S->d_amps, S->st) != 0) return -1;
    if (amps_host &&
        hipMemcpyAsync(amps_host, S->d_amps, (size_t)S->namps * sizeof(float2),
                       hipMemcpyDeviceToHost, S->st) != hipSuccess)
        return -1;
    if (hipStreamSynchronize(S->st) != hipSuccess) {
        fprintf(stderr, "Error: device staging failed: %s\n", hipGetErrorString(hipGetLastError()));
        return -1;
    }
    return 0;
}

extern "C" CS_API CandidateList cs_gpu_session_run(GpuSession *S, const FFTFile *ft,
                                                   double threshold)
{
    CandidateList out;
    out.v = NULL; out.n = 0; out.cap = 0;
    if (!S || cs_gpu_session_upload(S, ft) != 0) return out;
    return cs_gpu_session_search_resident(S, threshold);
}

/* Candidates of `count` records sorted by key (chunk, slot, trial: the
 * reference's harvest order, src/cuda.cu:833-881), above threshold and, for
 * decimations, below the Nyquist frequency, written to out[0..].  Record i is
 * read before candidate i is written, so out may overlap the records when
 * they sit at the end of out's count candidate slots.  Returns the number
 * written, -1 for a record no search could have produced. */
static long long records_to_candidates(const GpuSession *S, const CandRec *recs,
                                       unsigned int count, double threshold, Candidate *out)
{
    long long n = 0;
    for (unsigned int i = 0; i < count; i++) {
        CandRec r;
        memcpy(&r, recs + i, sizeof(r));
        const long long c = (long long)(r.key >> 40);
        const int slot = (int)((r.key >> 32) & 0xff);
        const long long j = (long long)(r.key & 0xffffffffull);
        if (c >= S->nchunks || slot >= S->nslot || j >= (long long)S->chunks[(size_t)c].n) {
            fprintf(stderr, "Error: GPU search returned an invalid record (chunk %lld, "
                            "slot %d, trial %lld)\n", c, slot, j);
            return -1;
        }
        const double metric = (double)r.metric;
        if (!(metric > threshold)) continue;
        const double rstart = S->r_lo + (double)S->cstarts[(size_t)c] * S->lodr;
        double rr;
        int nharm;
        if (slot == 0) {
            rr = rstart + (double)j * S->lodr;
            nharm = S->nh;
        } else {
            rr = (double)S->slot_k[slot] * (rstart + (double)j * S->lodr);
            nharm = S->slot_Hk[slot];
            if (!(rr < S->nyq)) continue;
        }
        Candidate cd;
        memset(&cd, 0, sizeof(cd));
        cd.freq = rr / S->T;
        cd.metric = metric;
        cd.r = rr;
        cd.nharm = nharm;
        cd.ducy = NAN;
        memcpy(out + n, &cd, sizeof(cd));
        n++;
    }
    return n;
}

static bool key_less(const CandRec &a, const CandRec &b) { return a.key < b.key; }

/* The records of a pass, ordered by key, as candidates above threshold. */
static CandidateList harvest(GpuSession *S, unsigned int count, double threshold)
{
    CandidateList out;
    out.v = NULL; out.n = 0; out.cap = 0;
    if (count == 0) return out;
    std::vector<CandRec> recs(count);
    if (hipMemcpyAsync(recs.data(), S->d_recs, (size_t)count * sizeof(CandRec),
                       hipMemcpyDeviceToHost, S->st) != hipSuccess ||
        hipStreamSynchronize(S->st) != hipSuccess)
        return out;
    std::sort(recs.begin(), recs.end(), key_less);
    out.v = (Candidate *)malloc((size_t)count * sizeof(Candidate));
    if (!out.v) return out;
    out.cap = count;
    const long long n = records_to_candidates(S, recs.data(), count, threshold, out.v);
    if (n < 0) {
        free(out.v);
        out.v = NULL;
        out.cap = 0;
        return out;
    }
    out.n = (size_t)n;
    return out;
}

extern "C" CS_API CandidateList cs_gpu_session_search_resident(GpuSession *S,
                                                               double threshold)
{
    CandidateList out;
    out.v = NULL; out.n = 0; out.cap = 0;
    if (!S || hipSetDevice(S->device) != hipSuccess) return out;
    unsigned int count = 0;
    for (int attempt = 0; attempt < 2; attempt++) {
        if (launch_search(S, threshold, NULL) != 0) {
            fprintf(stderr, "Error: GPU search launch failed\n");
            return out;
        }
        if (hipMemcpyAsync(&count, S->d_counter, sizeof(count), hipMemcpyDeviceToHost,
                           S->st) != hipSuccess ||
            hipStreamSynchronize(S->st) != hipSuccess) {
            fprintf(stderr, "Error: GPU search failed: %s\n",
                    hipGetErrorString(hipGetLastError()));
            return out;
        }
        float ms = 0.0f;
        hipEventElapsedTime(&ms, S->ev0, S->ev1);
        S->last_ms = ms;
        if (count <= S->cap) break;
        if (attempt == 1) {
            /* the grown buffer was searched and still overflowed: the count
             * changed between two searches of the same spectrum */
            fprintf(stderr, "Error: GPU search found %u records after a search that found "
                            "fewer; not harvesting a partly written buffer\n", count);
            return out;
        }
        /* never truncate: grow to the exact count and search again */
        if (S->recs_own) dfree(S, S->d_recs);
        S->d_recs = NULL;
        S->recs_own = true;
        S->cap = count;
        if (dalloc(S, (void **)&S->d_recs, (size_t)S->cap * sizeof(CandRec)) != hipSuccess) {
            fprintf(stderr, "Error: cannot hold %u candidate records\n", count);
            S->cap = 0;
            return out;
        }
    }
    return harvest(S, count, threshold);
}

/* ---- a search that allocates nothing -------------------------------------
 * The record buffer and its counter of cs_gpu_session_search_noalloc live in
 * one file-scope device array per GPU, held until the process ends:
 * CS_NOALLOC_MAXREC records of 16 bytes after a 256-byte counter slot.  In
 * host code `scratch` names a host-side stand-in; its device address on each
 * GPU comes from hipGetSymbolAddress, called once per device with that device
 * current, and only that address is handed to kernels and copies. */
#define CS_NOALLOC_MAXREC (1u << 18)
static __device__ __align__(256) unsigned char scratch[256 + (size_t)CS_NOALLOC_MAXREC * sizeof(CandRec)];

/* the device address of `scratch` on the current device `device`, NULL if the
 * runtime cannot give it (message on stderr) */
static unsigned char *scratch_on(int device)
{
    static std::mutex mu;
    static unsigned char *addr[64] = {};
    if (device < 0 || device >= 64) return NULL;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        void *ptr = NULL;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            fprintf(stderr, "Error: device address of the search scratch on device %d: %s\n",
                    device, hipGetErrorString(e));
            return NULL;
        }
        addr[device] = (unsigned char *)ptr;
    }
    return addr[device];
}

extern "C" CS_API long long cs_gpu_session_search_noalloc(GpuSession *S, double threshold,
                                                         Candidate *out, long long cap)
{
    if (!S || cap < 0 || (!out && cap > 0)) {
        fprintf(stderr, "Error: cs_gpu_session_search_noalloc: bad arguments\n");
        return -1;
    }
    if (hipSetDevice(S->device) != hipSuccess) return -1;
    unsigned char *base = scratch_on(S->device);
    if (!base) return -1;
    unsigned int *counter = (unsigned int *)base;
    CandRec *recs = (CandRec *)(base + 256);
    hipEventRecord(S->ev0, S->st);
    if (enqueue_search_into(S, threshold, NULL, recs, counter, CS_NOALLOC_MAXREC) != 0) {
        fprintf(stderr, "Error: GPU search launch failed\n");
        return -1;
    }
    hipEventRecord(S->ev1, S->st);
    unsigned int count = 0;
    if (hipMemcpyAsync(&count, counter, sizeof(count), hipMemcpyDeviceToHost, S->st) !=
            hipSuccess ||
        hipStreamSynchronize(S->st) != hipSuccess) {
        fprintf(stderr, "Error: GPU search failed: %s\n", hipGetErrorString(hipGetLastError()));
        return -1;
    }
    float ms = 0.0f;
    hipEventElapsedTime(&ms, S->ev0, S->ev1);
    S->last_ms = ms;
    if (count > CS_NOALLOC_MAXREC) {
        fprintf(stderr, "Error: the search found %u records; the search that allocates "
                        "nothing holds at most %u (raise the threshold or use "
                        "cs_gpu_session_search_resident)\n", count, CS_NOALLOC_MAXREC);
        return -1;
    }
    if ((long long)count > cap) {
        fprintf(stderr, "Error: the search found %u records but the output holds %lld "
                        "candidates\n", count, cap);
        return -1;
    }
    if (count == 0) return 0;
    /* the records go to the end of the caller's array, are sorted there and
     * turned into candidates from its start (a candidate never overtakes the
     * record it is made from, since count <= cap) */
    CandRec *tail = (CandRec *)((char *)out + (size_t)cap * sizeof(Candidate) -
                                (size_t)count * sizeof(CandRec));
    if (hipMemcpyAsync(tail, recs, (size_t)count * sizeof(CandRec), hipMemcpyDeviceToHost,
                       S->st) != hipSuccess ||
        hipStreamSynchronize(S->st) != hipSuccess) {
        fprintf(stderr, "Error: record readback failed\n");
        return -1;
    }
    std::sort(tail, tail + count, key_less);
    return records_to_candidates(S, tail, count, threshold, out);
}

extern "C" CS_API void cs_gpu_session_free(GpuSession *S) { session_release(S); }

/* Engine sessions: run on another worker's stream, replay launch graphs.
 * The stream-ordered allocations stay valid on any stream of the device. */
__attribute__((visibility("hidden"))) void cs_session_bind_stream(GpuSession *S, hipStream_t st,
                                                                  bool graphs)
{
    if (!S || S->own_stream || !st) return;
    S->st = st;
    S->use_graph = graphs;
}

/* stage samples into the resident spectrum with the staging graph (engine
 * plans reused across requests); no host copy, no synchronisation */
__attribute__((visibility("hidden"))) int cs_session_stage_graph(GpuSession *S, CsStage *g,
                                                                 const float *samples)
{
    if (!S || !g || !samples || cs_stage_N(g) != S->N || S->N / 2 != S->namps) return -1;
    return cs_stage_enqueue_graph(g, samples, S->d_amps, S->st);
}

/* stage samples already copied into g (cs_stage_copy_in on the session's
 * stream) into the resident spectrum; the spectrum is copied to amps_host when
 * given; synchronises */
__attribute__((visibility("hidden"))) int cs_session_stage_loaded(GpuSession *S, CsStage *g,
                                                                  float *amps_host)
{
    if (!S || !g || cs_stage_N(g) != S->N || S->N / 2 != S->namps) return -1;
    if (cs_stage_run(g, S->d_amps, S->st) != 0) return -1;
    if (amps_host &&
        hipMemcpyAsync(amps_host, S->d_amps, (size_t)S->namps * sizeof(float2),
                       hipMemcpyDeviceToHost, S->st) != hipSuccess)
        return -1;
    if (hipStreamSynchronize(S->st) != hipSuccess) {
        fprintf(stderr, "Error: device staging failed: %s\n", hipGetErrorString(hipGetLastError()));
        return -1;
    }
    return 0;
}

extern "C" CS_API int cs_gpu_session_set_graphs(GpuSession *S, int enable)
{
    if (!S) return -1;
    if (!S->own_stream) return enable ? -1 : 0;
    S->use_graph = enable != 0;
    return 0;
}

extern "C" CS_API CandidateList cs_gpu_search_region(const FFTFile *ft, const SearchParams *p,
                                                     const DirectPlanSet *dplans,
                                                     const long long *cstarts,
                                                     long long total, int Nprof,
                                                     long long nchunks, double r_lo,
                                                     double r_hi, double lodr,
                                                     double threshold, int device)
{
    (void)r_hi;
    GpuSession *S = cs_gpu_session_create(ft, p, dplans, cstarts, total, Nprof, nchunks,
                                          r_lo, lodr, device);
    if (!S) {
        CandidateList e;
        e.v = NULL; e.n = 0; e.cap = 0;
        return e;
    }
    CandidateList c = cs_gpu_session_run(S, ft, threshold);
    cs_gpu_session_free(S);
    return c;
}

extern "C" CS_API int cs_gpu_session_dump_stack(GpuSession *S, const FFTFile *ft,
                                                long long t0, long long n, float *out)
{
    if (!S || !ft || !out || n <= 0 || t0 < 0 || t0 + n > S->total) return -1;
    if (ft->namps != S->namps || ft->N != S->N) return -1;
    if (hipSetDevice(S->device) != hipSuccess) return -1;
    const size_t bytes = (size_t)n * (S->nh + 1) * sizeof(float2);
    float2 *d = NULL;
    int rc = -1;
    if (hipMemcpy(S->d_amps, ft->amps, (size_t)S->namps * sizeof(float2),
                  hipMemcpyHostToDevice) != hipSuccess)
        return -1;
    if (hipMalloc(&d, bytes) != hipSuccess) return -1;
    if (hipMemsetAsync(d, 0, bytes, S->st) == hipSuccess) {
        S->args.dump_t0 = t0;
        S->args.dump_n = n;
        if (launch_search(S, INFINITY, d) == 0 &&
            hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, S->st) == hipSuccess &&
            hipStreamSynchronize(S->st) == hipSuccess)
            rc = 0;
    }
    S->args.stack_out = NULL;
    hipFree(d);
    return rc;
}

extern "C" CS_API double cs_gpu_session_last_kernel_ms(const GpuSession *S)
{
    return S ? S->last_ms : 0.0;
}

extern "C" CS_API int cs_gpu_session_work(const GpuSession *S, long long *ht,
                                          long long *trials, double *alg_bytes)
{
    if (!S) return -1;
    if (ht) *ht = S->ht;
    if (trials) *trials = S->total;
    if (alg_bytes) *alg_bytes = S->alg_bytes;
    return 0;
}

/* ---- duty cycles of the reported candidates, on the device ----------------
 * src/cli.c:1059-1061 (sort by S/N, truncate) and src/candidate.c:305-329
 * (measure_ducy) with candidate_profile (src/candidate.c:208-271) and
 * boxcar_best_width (src/boxcar.c:736-785), read from the resident spectrum so
 * it never has to come back to the host. */
namespace {

constexpr int kDucyM = 64;        /* exact interpolation taps */
constexpr int kDucyHmax = 1024;   /* harmonics per candidate */
constexpr int kDucyWcap = 256;    /* boxcar widths per profile */
constexpr int kDucyThreads = 256;
/* prefix sums over n + wmax bins, wmax <= 0.3 n (boxcar_widths' maxfrac cap) */
constexpr int kDucyPs = 2 * kDucyHmax + (2 * kDucyHmax * 3) / 10 + 2;

struct DucyCand {
    double r;
    int H;                        /* leading harmonics whose 64-bin window fits */
    int n;                        /* profile bins, 2H */
    int nw;                       /* boxcar widths of an n-bin profile */
    int woff;                     /* their offset in the width/shape tables */
};

/* One workgroup per candidate.  Threads stride over the harmonics, each
 * interpolating harmonic h exactly (m = 64 taps of sinc(off) e^{-i pi off},
 * double); over the bins, each forming the profile by a direct real inverse
 * transform of the H harmonics (harmonic H taken as the real Nyquist term)
 * scaled by 1/2H; thread 0 takes the prefix sums of the mean-removed profile
 * over n + wmax wrapped bins; threads over widths take each width's best
 * window, (max - duty*total) * shape_w; thread 0 keeps the first width with
 * the largest value. */
__global__ __launch_bounds__(kDucyThreads) void ducy_kernel(const float2 *__restrict__ amps,
                                                  long long namps,
                                                  const DucyCand *__restrict__ cands,
                                                  const int *__restrict__ widths,
                                                  const double *__restrict__ shapes,
                                                  int *__restrict__ bestw)
{
    __shared__ double Xr[kDucyHmax + 1], Xi[kDucyHmax + 1];
    __shared__ double prof[2 * kDucyHmax];
    __shared__ double ps[kDucyPs];
    __shared__ double cval[kDucyWcap];
    const DucyCand c = cands[blockIdx.x];
    const int lane = threadIdx.x;
    const int H = c.H, n = c.n;
    for (int h = lane + 1; h <= H; h += kDucyThreads) {
        const double rh = c.r * (double)h;
        double dr = fmod(rh, 1.0);
        if (dr < 0) dr += 1.0;
        const long long first = (long long)floor(rh + 1e-15) + 1 - kDucyM / 2;
        double sre = 0.0, sim = 0.0;
        for (int j = 0; j < kDucyM; j++) {
            const double off = dr - (double)(-kDucyM / 2 + 1 + j);
            const double sinc = off == 0.0 ? 1.0 : sin(M_PI * off) / (M_PI * off);
            const double cr = sinc * cos(M_PI * off), ci = -sinc * sin(M_PI * off);
            const long long idx = first + j;
            double bre = 0.0, bim = 0.0;
            if (idx >= 0 && idx < namps) {
                const float2 a = amps[idx];
                bre = (double)a.x;
                bim = (double)a.y;
            }
            sre += cr * bre - ci * bim;
            sim += cr * bim + ci * bre;
        }
        Xr[h] = sre;
        Xi[h] = sim;
    }
    __syncthreads();
    const double inv = 1.0 / (double)n;
    for (int i = lane; i < n; i += kDucyThreads) {
        double acc = 0.0;
        for (int h = 1; h < H; h++) {
            double sn, cs;
            sincospi(2.0 * (double)(((long long)h * i) % n) / (double)n, &sn, &cs);
            acc += 2.0 * (Xr[h] * cs - Xi[h] * sn);
        }
        acc += (i & 1) ? -Xr[H] : Xr[H];
        prof[i] = acc * inv;
    }
    __syncthreads();
    const int wmax = widths[c.woff + c.nw - 1];
    if (lane == 0) {
        double sum = 0.0;
        for (int i = 0; i < n; i++) sum += prof[i];
        const double mean = sum / (double)n;
        ps[0] = 0.0;
        for (int i = 1; i <= n + wmax; i++) {
            const int idx = i > n ? i - n : i;
            ps[i] = ps[i - 1] + (prof[idx - 1] - mean);
        }
    }
    __syncthreads();
    const double stot = ps[n] - ps[0];
    for (int k = lane; k < c.nw; k += kDucyThreads) {
        const int w = widths[c.woff + k];
        const double duty = (double)w / (double)n;
        double mx = ps[w] - ps[0];
        for (int q = 2; q <= n; q++) {
            const double d = ps[q + w - 1] - ps[q - 1];
            if (d > mx) mx = d;
        }
        cval[k] = (mx - duty * stot) * shapes[c.woff + k];
    }
    __syncthreads();
    if (lane == 0) {
        double best = -INFINITY;
        int bw = widths[c.woff];
        for (int k = 0; k < c.nw; k++)
            if (cval[k] > best) {
                best = cval[k];
                bw = widths[c.woff + k];
            }
        bestw[blockIdx.x] = bw;
    }
}

}  // namespace

__attribute__((visibility("hidden"))) CandidateList
cs_session_finalize(GpuSession *S, const CandidateList *in, int ncands)
{
    CandidateList out;
    csx_candlist_init(&out);
    if (!S || !in) return out;
    for (size_t i = 0; i < in->n; i++) csx_candlist_push(&out, in->v[i]);
    if (out.n != in->n) {
        fprintf(stderr, "Error: out of host memory for %zu candidates\n", in->n);
        csx_candlist_free(&out);
        return out;
    }
    csx_candlist_sort_metric_desc(&out);
    if (ncands >= 0 && out.n > (size_t)ncands) out.n = (size_t)ncands;
    const double fsp = S->p->boxcar_fsp, maxfrac = S->p->boxcar_maxfrac;
    const long long Nhalf = S->N / 2;
    std::vector<DucyCand> dc;
    std::vector<size_t> dc_of;           /* list index of each device candidate */
    std::vector<int> wt;
    std::vector<double> st;
    std::vector<int> woff_of_n(2 * kDucyHmax + 1, -1), nw_of_n(2 * kDucyHmax + 1, 0);
    for (size_t i = 0; i < out.n; i++) {
        Candidate *c = &out.v[i];
        c->ducy = NAN;
        if (!(c->nharm >= 1 && c->r > 0.0)) continue;
        if (c->nharm > kDucyHmax) {
            fprintf(stderr, "Error: duty cycle of a %d-harmonic candidate (limit %d)\n",
                    c->nharm, kDucyHmax);
            csx_candlist_free(&out);
            return out;
        }
        int H = 0;
        for (int h = 1; h <= c->nharm; h++) {
            const double rh = c->r * (double)h;
            const long long rint0 = (long long)floor(rh + 1e-15) + 1;
            if (!(rint0 - kDucyM / 2 >= 0 && rint0 + kDucyM / 2 <= S->namps &&
                  rh < (double)Nhalf))
                break;
            H++;
        }
        const int n = H < 1 ? 2 : 2 * H;
        if (woff_of_n[n] < 0) {
            int w[kDucyWcap];
            const int nw = csx_boxcar_widths(n, fsp, maxfrac, w, kDucyWcap);
            woff_of_n[n] = (int)wt.size();
            nw_of_n[n] = nw;
            if (nw > 0) {
                std::vector<uint8_t> filled((size_t)n / 2, 1);
                std::vector<double> sh((size_t)nw);
                csx_boxcar_shape(sh.data(), n, w, nw, filled.data(), 1, n / 2);
                wt.insert(wt.end(), w, w + nw);
                st.insert(st.end(), sh.begin(), sh.end());
            }
        }
        if (nw_of_n[n] <= 0) {
            c->ducy = 1.0 / (double)n;   /* no widths: one bin */
            continue;
        }
        if (H < 1) {
            /* an all-zero profile: every width scores 0, the first one wins */
            c->ducy = (double)wt[(size_t)woff_of_n[n]] / (double)n;
            continue;
        }
        if (n + wt[(size_t)(woff_of_n[n] + nw_of_n[n] - 1)] + 1 > kDucyPs) {
            fprintf(stderr, "Error: duty cycle of a %d-bin profile with widths up to %d "
                            "exceeds the device pass (%d prefix sums)\n", n,
                    wt[(size_t)(woff_of_n[n] + nw_of_n[n] - 1)], kDucyPs);
            csx_candlist_free(&out);
            return out;
        }
        DucyCand d;
        d.r = c->r;
        d.H = H;
        d.n = n;
        d.nw = nw_of_n[n];
        d.woff = woff_of_n[n];
        dc.push_back(d);
        dc_of.push_back(i);
    }
    if (dc.empty()) return out;
    if (hipSetDevice(S->device) != hipSuccess) {
        csx_candlist_free(&out);
        return out;
    }
    const size_t bc = dc.size() * sizeof(DucyCand), bw = wt.size() * sizeof(int),
                 bs = st.size() * sizeof(double), bo = dc.size() * sizeof(int);
    const size_t o_w = (bc + 255) / 256 * 256, o_s = o_w + (bw + 255) / 256 * 256,
                 o_o = o_s + (bs + 255) / 256 * 256, total = o_o + bo;
    std::vector<unsigned char> h_in(o_o);
    memcpy(h_in.data(), dc.data(), bc);
    memcpy(h_in.data() + o_w, wt.data(), bw);
    memcpy(h_in.data() + o_s, st.data(), bs);
    std::vector<int> best(dc.size(), 0);
    unsigned char *d = NULL;
    int ok = 0;
    if (dalloc(S, (void **)&d, total) == hipSuccess) {
        if (hipMemcpyAsync(d, h_in.data(), o_o, hipMemcpyHostToDevice, S->st) == hipSuccess) {
            hipLaunchKernelGGL(ducy_kernel, dim3((unsigned)dc.size()), dim3(kDucyThreads), 0, S->st,
                               (const float2 *)S->d_amps, S->namps, (const DucyCand *)d,
                               (const int *)(d + o_w), (const double *)(d + o_s),
                               (int *)(d + o_o));
            if (hipGetLastError() == hipSuccess &&
                hipMemcpyAsync(best.data(), d + o_o, bo, hipMemcpyDeviceToHost, S->st) ==
                    hipSuccess &&
                hipStreamSynchronize(S->st) == hipSuccess)
                ok = 1;
        }
        dfree(S, d);
    }
    if (!ok) {
        fprintf(stderr, "Error: device duty-cycle pass failed\n");
        csx_candlist_free(&out);
        return out;
    }
    for (size_t k = 0; k < dc.size(); k++)
        out.v[dc_of[k]].ducy = (double)best[k] / (double)dc[k].n;
    return out;
}

extern "C" CS_API CandidateList cs_gpu_session_finalize(GpuSession *S, const CandidateList *cands,
                                                        int ncands)
{
    return cs_session_finalize(S, cands, ncands);
}

/* ---- zaplist on the resident spectrum (src/zap.c:94-187) ------------------
 * The intervals are applied in order by one workgroup: the median windows of
 * an interval are read after every earlier interval's bins are written. */
namespace {

__global__ __launch_bounds__(256) void zap_kernel(float2 *amps, const CsZapOp *__restrict__ ops,
                                                  int nops, double ln2)
{
    __shared__ float pw[2][CS_ZAP_MEDIANBINS];
    __shared__ float med[2];
    const int tid = threadIdx.x;
    for (int k = 0; k < nops; k++) {
        const CsZapOp o = ops[k];
        if (tid < 2) med[tid] = 0.0f;
        for (int i = tid; i < CS_ZAP_MEDIANBINS; i += blockDim.x) {
            if (i < o.n_lo) {
                const float2 a = amps[o.s_lo + i];
                pw[0][i] = __builtin_fmaf(a.x, a.x, a.y * a.y);
            }
            if (i < o.n_hi) {
                const float2 a = amps[o.s_hi + i];
                pw[1][i] = __builtin_fmaf(a.x, a.x, a.y * a.y);
            }
        }
        __syncthreads();
        /* lower median by rank: the value whose stable rank is (n-1)/2 */
        for (int idx = tid; idx < 2 * CS_ZAP_MEDIANBINS; idx += blockDim.x) {
            const int w = idx >= CS_ZAP_MEDIANBINS;
            const int i = idx - w * CS_ZAP_MEDIANBINS;
            const int n = w ? o.n_hi : o.n_lo;
            if (i < n) {
                const float x = pw[w][i];
                int r = 0;
                for (int j = 0; j < n; j++) {
                    const float y = pw[w][j];
                    r += (y < x) || (y == x && j < i);
                }
                if (r == (n - 1) / 2) med[w] = x;
            }
        }
        __syncthreads();
        const float avg = (float)sqrt(0.5 * (double)(med[0] + med[1]) / ln2);
        for (long long i = o.ilo + tid; i < o.ihi; i += blockDim.x) amps[i] = make_float2(avg, 0.0f);
        /* the next interval's windows read what this one wrote */
        __threadfence();
        __syncthreads();
        __threadfence();
    }
}

}  // namespace

extern "C" CS_API int cs_gpu_session_zap(GpuSession *S, const ZapList *z, double baryv,
                                         float *amps_host)
{
    if (!S || hipSetDevice(S->device) != hipSuccess) return -1;
    CsZapOp *ops = NULL;
    const int n = csx_zap_ops(z, S->T, S->N, S->namps, baryv, &ops);
    if (n < 0) return -1;
    int rc = 0;
    if (n > 0) {
        CsZapOp *d = NULL;
        const size_t bytes = (size_t)n * sizeof(CsZapOp);
        rc = -1;
        if (dalloc(S, (void **)&d, bytes) == hipSuccess) {
            if (hipMemcpyAsync(d, ops, bytes, hipMemcpyHostToDevice, S->st) == hipSuccess) {
                hipLaunchKernelGGL(zap_kernel, dim3(1), dim3(256), 0, S->st, S->d_amps,
                                   (const CsZapOp *)d, n, -log(0.5));
                if (hipGetLastError() == hipSuccess) rc = 0;
            }
            /* ops must outlive the copy */
            if (hipStreamSynchronize(S->st) != hipSuccess) rc = -1;
            dfree(S, d);
        }
    }
    free(ops);
    if (rc == 0 && amps_host &&
        hipMemcpyAsync(amps_host, S->d_amps, (size_t)S->namps * sizeof(float2),
                       hipMemcpyDeviceToHost, S->st) != hipSuccess)
        rc = -1;
    if (hipStreamSynchronize(S->st) != hipSuccess) rc = -1;
    if (rc != 0) {
        fprintf(stderr, "Error: zaplist on the device failed\n");
        return -1;
    }
    return n;
}

/* ---- metric statistics and normalisation -------------------------------
 *
 * The collection pass is the search kernel with mout set: every valid
 * trial's metric lands at mout[slot*total + trial] (T = float or double, the
 * profile precision, so the value is the metric exactly).  Then one
 * workgroup per (chunk, slot) summarises its metrics (metric_stats_kernel),
 * and a later pass scores them against per-window normalisations
 * (metric_norm_kernel). */

namespace {

constexpr int kStatThreads = 256;
constexpr int kStatMaxBins = 4096;
constexpr int kDigitBits = 11;

struct StatArgs {
    const ChunkDesc *chunks;
    const void *mout;
    long long mstride;
    int nslot;
    const int *win;
    int nwin;
    double lo, hi, invw;
    int nb;
    CsMetricBlock *blocks;
    unsigned long long *counts;
};

/* order-preserving unsigned keys of float / double values */
__device__ __forceinline__ unsigned int okey(float x)
{
    const unsigned int u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ unsigned long long okey(double x)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ float ofloat(unsigned int k)
{
    return __uint_as_float((k & 0x80000000u) ? (k ^ 0x80000000u) : ~k);
}
__device__ __forceinline__ double ofloat(unsigned long long k)
{
    return __longlong_as_double((long long)((k & 0x8000000000000000ull)
                                                ? (k ^ 0x8000000000000000ull) : ~k));
}

/* sum over the workgroup of one value per thread, in a fixed order */
template <typename V>
__device__ V wg_sum(V v, V *sh)
{
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    V t = sh[0];
    for (int i = 1; i < kStatThreads / 64; i++) t += sh[i];
    return t;
}
template <typename V>
__device__ V wg_min(V v, V *sh)
{
    for (int d = 32; d > 0; d >>= 1) {
        const V o = __shfl_down(v, d, 64);
        v = o < v ? o : v;
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    V t = sh[0];
    for (int i = 1; i < kStatThreads / 64; i++) t = sh[i] < t ? sh[i] : t;
    return t;
}

/* The value of rank r (0-based, ascending) among v[0..n): a most-significant
 * digit radix selection over the values' order-preserving keys, kDigitBits
 * per pass, counting in LDS. */
template <typename T, typename K>
__device__ T select_rank(const T *__restrict__ v, int n, long long r, unsigned int *cnt,
                         unsigned long long *sh)
{
    constexpr int BITS = 8 * (int)sizeof(K);
    K prefix = 0;
    int done = 0;
    while (done < BITS) {
        const int width = BITS - done < kDigitBits ? BITS - done : kDigitBits;
        const int shift = BITS - done - width;
        const int nd = 1 << width;
        for (int b = threadIdx.x; b < nd; b += kStatThreads) cnt[b] = 0u;
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += kStatThreads) {
            const K k = okey(v[i]);
            if (done == 0 || ((k ^ prefix) >> (BITS - done)) == 0)
                atomicAdd(&cnt[(int)((k >> shift) & (K)(nd - 1))], 1u);
        }
        __syncthreads();
        /* each thread sums its run of digits; thread 0 finds the run, then the
         * digit, holding rank r */
        const int per = (nd + kStatThreads - 1) / kStatThreads;
        unsigned long long part = 0;
        for (int b = threadIdx.x * per; b < (threadIdx.x + 1) * per && b < nd; b++) part += cnt[b];
        __syncthreads();
        sh[threadIdx.x] = part;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long rem = r;
            int t = 0;
            while (t < kStatThreads - 1 && rem >= (long long)sh[t]) rem -= (long long)sh[t++];
            int d = t * per;
            while (d < nd - 1 && rem >= (long long)cnt[d]) rem -= (long long)cnt[d++];
            sh[kStatThreads] = (unsigned long long)d;
            sh[kStatThreads + 1] = (unsigned long long)rem;
        }
        __syncthreads();
        prefix |= (K)sh[kStatThreads] << shift;
        r = (long long)sh[kStatThreads + 1];
        done += width;
        __syncthreads();
    }
    return (T)ofloat(prefix);
}

template <typename T, typename K>
__global__ __launch_bounds__(kStatThreads) void metric_stats_kernel(StatArgs a)
{
    __shared__ unsigned int cnt[kStatMaxBins];
    __shared__ unsigned long long sh[kStatThreads + 2];
    __shared__ double shd[kStatThreads / 64];
    __shared__ long long shl[kStatThreads / 64];
    const int c = blockIdx.x / a.nslot, s = blockIdx.x % a.nslot;
    const ChunkDesc *C = a.chunks + c;
    int n = C->n;
    if (s > 0 && C->nvalid[s] < n) n = C->nvalid[s];
    if (n < 0) n = 0;
    const T *v = (const T *)a.mout + (long long)s * a.mstride + C->i0;
    for (int b = threadIdx.x; b < a.nb; b += kStatThreads) cnt[b] = 0u;
    __syncthreads();
    double mn = INFINITY, mx = -INFINITY, sum = 0.0, sq = 0.0;
    long long under = 0, over = 0;
    for (int i = threadIdx.x; i < n; i += kStatThreads) {
        const double x = (double)v[i];
        mn = x < mn ? x : mn;
        mx = x > mx ? x : mx;
        sum += x;
        sq += x * x;
        if (x < a.lo) {
            under++;
        } else if (x >= a.hi) {
            over++;
        } else {
            int idx = (int)floor((x - a.lo) * a.invw);
            if (idx < 0) idx = 0;
            if (idx >= a.nb) idx = a.nb - 1;
            atomicAdd(&cnt[idx], 1u);
        }
    }
    mn = wg_min(mn, shd);
    mx = -wg_min(-mx, shd);
    sum = wg_sum(sum, shd);
    sq = wg_sum(sq, shd);
    under = wg_sum(under, shl);
    over = wg_sum(over, shl);
    const double mean = n > 0 ? sum / (double)n : 0.0;
    double ss = 0.0;
    for (int i = threadIdx.x; i < n; i += kStatThreads) {
        const double d = (double)v[i] - mean;
        ss += d * d;
    }
    ss = wg_sum(ss, shd);
    const int w = a.win[(long long)c * a.nslot + s];
    __syncthreads();
    if (w >= 1 && w <= a.nwin) {
        unsigned long long *dst = a.counts + ((long long)s * a.nwin + (w - 1)) * a.nb;
        for (int b = threadIdx.x; b < a.nb; b += kStatThreads)
            if (cnt[b]) atomicAdd(&dst[b], (unsigned long long)cnt[b]);
    }
    __syncthreads();
    double med = NAN;
    if (n > 0) {
        const int half = n >> 1;
        const T upper = select_rank<T, K>(v, n, half, cnt, sh);
        if (n & 1) {
            med = (double)upper;
        } else {
            const T lower = select_rank<T, K>(v, n, half - 1, cnt, sh);
            med = 0.5 * ((double)lower + (double)upper);
        }
    }
    if (threadIdx.x == 0) {
        CsMetricBlock b;
        b.n = n;
        b.min = n > 0 ? mn : NAN;
        b.max = n > 0 ? mx : NAN;
        b.sum = sum;
        b.sumsq = sq;
        b.ss = ss;
        b.median = med;
        b.under = under;
        b.over = over;
        a.blocks[blockIdx.x] = b;
    }
}

struct NormArgs {
    const ChunkDesc *chunks;
    const void *mout;
    long long mstride;
    int nslot;
    int slot_k[CS_MAXSLOT];
    int nwin;
    double r_lo, lodr, T, threshold;
    const int *has_norm;
    const double *edges;          /* nwin + 1 per slot */
    const double *loc, *scale;    /* nwin per slot */
    CandRec *recs;
    unsigned int *counter;
    unsigned int cap;
};

/* The metrics of (chunk, slot) scored (M - loc)/scale in their frequency
 * window (src/search.c:814-823, window_index src/search.c:479-492) and
 * appended above threshold, one atomic per wave. */
template <typename T>
__global__ __launch_bounds__(kStatThreads) void metric_norm_kernel(NormArgs a)
{
    __shared__ double se[CS_MAXWIN + 1];
    const int c = blockIdx.x / a.nslot, s = blockIdx.x % a.nslot;
    const ChunkDesc *C = a.chunks + c;
    int n = C->n;
    if (s > 0 && C->nvalid[s] < n) n = C->nvalid[s];
    const int norm = a.has_norm[s];
    for (int i = threadIdx.x; i <= a.nwin; i += kStatThreads)
        se[i] = a.edges[(long long)s * (a.nwin + 1) + i];
    __syncthreads();
    const T *v = (const T *)a.mout + (long long)s * a.mstride + C->i0;
    const double rstart = a.r_lo + (double)C->i0 * a.lodr;
    const int k = a.slot_k[s];
    for (int j0 = 0; j0 < n; j0 += kStatThreads) {
        const int j = j0 + threadIdx.x;
        const bool in = j < n;
        double score = in ? (double)v[j] : 0.0;
        if (in && norm) {
            const double rr = s == 0 ? rstart + (double)j * a.lodr
                                     : (double)k * (rstart + (double)j * a.lodr);
            const double f = rr / a.T;
            int lo = 0, hi = a.nwin + 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (se[mid] <= f) lo = mid + 1;
                else hi = mid;
            }
            int w = lo;
            if (w < 1) w = 1;
            if (w > a.nwin) w = a.nwin;
            const long long o = (long long)s * a.nwin + (w - 1);
            score = (score - a.loc[o]) / a.scale[o];
        }
        const bool pred = in && score > a.threshold;
        const unsigned long long mask = __ballot(pred);
        if (mask == 0ull) continue;
        const int lane = __lane_id();
        const int leader = __ffsll((long long)mask) - 1;
        unsigned int base = 0;
        if (lane == leader) base = atomicAdd(a.counter, (unsigned int)__popcll(mask));
        base = __shfl(base, leader);
        if (pred) {
            const unsigned int pos = base + (unsigned int)__popcll(mask & ((1ull << lane) - 1ull));
            if (pos < a.cap) {
                CandRec rec;
                rec.key = ((unsigned long long)c << 40) | ((unsigned long long)s << 32) |
                          (unsigned int)j;
                rec.metric = score;
                a.recs[pos] = rec;
            }
        }
    }
}

}  // namespace

extern "C" CS_API int cs_gpu_session_metric_stats(GpuSession *S, const FFTFile *ft,
                                                  const int *win, int nwin, double hist_lo,
                                                  double hist_hi, int hist_nb,
                                                  CsMetricBlock *blocks, long long *counts)
{
    if (!S || !win || !blocks || !counts) return -1;
    if (nwin < 1 || nwin > CS_MAXWIN || hist_nb < 1 || hist_nb > kStatMaxBins ||
        !(hist_hi > hist_lo)) {
        fprintf(stderr, "Error: metric statistics need 1..%d windows, 1..%d histogram bins "
                        "and hist_hi > hist_lo\n", CS_MAXWIN, kStatMaxBins);
        return -1;
    }
    if (ft && cs_gpu_session_upload(S, ft) != 0) return -1;
    if (hipSetDevice(S->device) != hipSuccess) return -1;
    const size_t esz = S->f64 ? sizeof(double) : sizeof(float);
    const long long nb2 = S->nchunks * S->nslot;
    const size_t ncount = (size_t)S->nslot * nwin * hist_nb;
    int *d_win = NULL;
    CsMetricBlock *d_blocks = NULL;
    unsigned long long *d_counts = NULL;
    int rc = -1;
    if (!S->d_mout &&
        dalloc(S, &S->d_mout, (size_t)S->nslot * (size_t)S->total * esz) != hipSuccess) {
        fprintf(stderr, "Error: cannot hold %lld x %d trial metrics on the device\n", S->total,
                S->nslot);
        S->d_mout = NULL;
        return -1;
    }
    S->mout_ready = false;
    if (hipMalloc(&d_win, (size_t)nb2 * sizeof(int)) != hipSuccess ||
        hipMalloc(&d_blocks, (size_t)nb2 * sizeof(CsMetricBlock)) != hipSuccess ||
        hipMalloc(&d_counts, ncount * sizeof(unsigned long long)) != hipSuccess)
        goto done;
    if (hipMemcpyAsync(d_win, win, (size_t)nb2 * sizeof(int), hipMemcpyHostToDevice, S->st) !=
            hipSuccess ||
        hipMemcpyAsync(d_counts, counts, ncount * sizeof(unsigned long long),
                       hipMemcpyHostToDevice, S->st) != hipSuccess)
        goto done;
    {
        S->args.mout = S->d_mout;
        S->args.mstride = S->total;
        hipEventRecord(S->ev0, S->st);
        const int e = enqueue_search(S, INFINITY, NULL);
        S->args.mout = NULL;
        if (e != 0) goto done;
        hipEventRecord(S->ev1, S->st);
        if (nb2 > 0) {
            StatArgs a;
            a.chunks = S->d_chunks;
            a.mout = S->d_mout;
            a.mstride = S->total;
            a.nslot = S->nslot;
            a.win = d_win;
            a.nwin = nwin;
            a.lo = hist_lo;
            a.hi = hist_hi;
            a.invw = (double)hist_nb / (hist_hi - hist_lo);
            a.nb = hist_nb;
            a.blocks = d_blocks;
            a.counts = d_counts;
            if (S->f64)
                hipLaunchKernelGGL((metric_stats_kernel<double, unsigned long long>),
                                   dim3((unsigned)nb2), dim3(kStatThreads), 0, S->st, a);
            else
                hipLaunchKernelGGL((metric_stats_kernel<float, unsigned int>),
                                   dim3((unsigned)nb2), dim3(kStatThreads), 0, S->st, a);
            if (hipGetLastError() != hipSuccess) goto done;
        }
    }
    if (hipMemcpyAsync(blocks, d_blocks, (size_t)nb2 * sizeof(CsMetricBlock),
                       hipMemcpyDeviceToHost, S->st) != hipSuccess ||
        hipMemcpyAsync(counts, d_counts, ncount * sizeof(unsigned long long),
                       hipMemcpyDeviceToHost, S->st) != hipSuccess ||
        hipStreamSynchronize(S->st) != hipSuccess) {
        fprintf(stderr, "Error: GPU metric statistics failed: %s\n",
                hipGetErrorString(hipGetLastError()));
        goto done;
    }
    {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, S->ev0, S->ev1);
        S->last_ms = ms;
    }
    S->mout_ready = true;
    rc = 0;
done:
    hipStreamSynchronize(S->st);
    if (d_win) hipFree(d_win);
    if (d_blocks) hipFree(d_blocks);
    if (d_counts) hipFree(d_counts);
    return rc;
}

extern "C" CS_API CandidateList cs_gpu_session_metric_candidates(GpuSession *S, double threshold,
                                                                 int nwin, const int *has_norm,
                                                                 const double *edges,
                                                                 const double *loc,
                                                                 const double *scale)
{
    CandidateList out;
    out.v = NULL; out.n = 0; out.cap = 0;
    if (!S || !has_norm || !edges || !loc || !scale) return out;
    if (!S->mout_ready) {
        fprintf(stderr, "Error: no kept metrics (run cs_gpu_session_metric_stats first)\n");
        return out;
    }
    if (nwin < 1 || nwin > CS_MAXWIN) {
        fprintf(stderr, "Error: 1..%d normalisation windows\n", CS_MAXWIN);
        return out;
    }
    if (hipSetDevice(S->device) != hipSuccess) return out;
    const int ns = S->nslot;
    const size_t o_e = 0, o_l = o_e + (size_t)ns * (nwin + 1) * sizeof(double);
    const size_t o_s = o_l + (size_t)ns * nwin * sizeof(double);
    const size_t o_h = o_s + (size_t)ns * nwin * sizeof(double);
    const size_t bytes = o_h + (size_t)ns * sizeof(int);
    char *d = NULL;
    unsigned int count = 0;
    if (hipMalloc(&d, bytes) != hipSuccess) return out;
    NormArgs a;
    memset(&a, 0, sizeof(a));
    a.chunks = S->d_chunks;
    a.mout = S->d_mout;
    a.mstride = S->total;
    a.nslot = ns;
    for (int s = 0; s < ns; s++) a.slot_k[s] = S->slot_k[s];
    a.nwin = nwin;
    a.r_lo = S->r_lo;
    a.lodr = S->lodr;
    a.T = S->T;
    a.threshold = threshold;
    a.edges = (const double *)(d + o_e);
    a.loc = (const double *)(d + o_l);
    a.scale = (const double *)(d + o_s);
    a.has_norm = (const int *)(d + o_h);
    if (hipMemcpyAsync(d + o_e, edges, (size_t)ns * (nwin + 1) * sizeof(double),
                       hipMemcpyHostToDevice, S->st) != hipSuccess ||
        hipMemcpyAsync(d + o_l, loc, (size_t)ns * nwin * sizeof(double), hipMemcpyHostToDevice,
                       S->st) != hipSuccess ||
        hipMemcpyAsync(d + o_s, scale, (size_t)ns * nwin * sizeof(double),
                       hipMemcpyHostToDevice, S->st) != hipSuccess ||
        hipMemcpyAsync(d + o_h, has_norm, (size_t)ns * sizeof(int), hipMemcpyHostToDevice,
                       S->st) != hipSuccess)
        goto done;
    for (int attempt = 0; attempt < 2; attempt++) {
        a.recs = S->d_recs;
        a.counter = S->d_counter;
        a.cap = S->cap;
        hipLaunchKernelGGL(cs_zero_u32, dim3(1), dim3(1), 0, S->st, S->d_counter);
        if (S->nchunks * ns > 0) {
            if (S->f64)
                hipLaunchKernelGGL(metric_norm_kernel<double>, dim3((unsigned)(S->nchunks * ns)),
                                   dim3(kStatThreads), 0, S->st, a);
            else
                hipLaunchKernelGGL(metric_norm_kernel<float>, dim3((unsigned)(S->nchunks * ns)),
                                   dim3(kStatThreads), 0, S->st, a);
        }
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&count, S->d_counter, sizeof(count), hipMemcpyDeviceToHost,
                           S->st) != hipSuccess ||
            hipStreamSynchronize(S->st) != hipSuccess) {
            fprintf(stderr, "Error: GPU metric normalisation failed: %s\n",
                    hipGetErrorString(hipGetLastError()));
            goto done;
        }
        if (count <= S->cap) break;
        if (attempt == 1) {
            fprintf(stderr, "Error: GPU metric normalisation found %u records after a pass "
                            "that found fewer; not harvesting a partly written buffer\n", count);
            goto done;
        }
        if (S->recs_own) dfree(S, S->d_recs);
        S->d_recs = NULL;
        S->recs_own = true;
        S->cap = count;
        if (dalloc(S, (void **)&S->d_recs, (size_t)S->cap * sizeof(CandRec)) != hipSuccess) {
            fprintf(stderr, "Error: cannot hold %u candidate records\n", count);
            S->cap = 0;
            goto done;
        }
    }
    out = harvest(S, count, threshold);
done:
    hipStreamSynchronize(S->st);
    hipFree(d);
    return out;
}
