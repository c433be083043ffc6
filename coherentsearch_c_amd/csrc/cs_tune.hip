/*
 * Autotuner of the search launch: for one kernel class it times the search
 * of a synthetic noise spectrum under every candidate launch shape and chunk
 * length and keeps the fastest.  The candidates change only where trials run
 * (lane order in a tile, tile-to-XCD mapping) and, for analytic sigma, where
 * chunks split (the analytic noise scale depends on the filled-harmonic set,
 * which is constant between the fit change points that are chunk starts at
 * every blocksize), so every choice gives the same results.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cands.h"
#include "core.h"
#include "cs_gpu.h"
#include "cs_stage.h"

namespace {

struct TuneKey {
    std::string arch;
    int device, nh, mp, f64, measured;
    bool operator==(const TuneKey &o) const
    {
        return arch == o.arch && device == o.device && nh == o.nh && mp == o.mp &&
               f64 == o.f64 && measured == o.measured;
    }
};

std::mutex g_mu;                                    /* the table */
std::mutex g_tuning;                                /* one tuning run at a time */
std::vector<std::pair<TuneKey, CsTuneResult>> g_tab;

std::string arch_of(int device)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "?";
    return prop.gcnArchName;
}

bool find(const TuneKey &k, CsTuneResult *out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto &e : g_tab)
        if (e.first == k) {
            if (out) *out = e.second;
            return true;
        }
    return false;
}

void store(const TuneKey &k, const CsTuneResult &r)
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &e : g_tab)
        if (e.first == k) {
            e.second = r;
            return;
        }
    g_tab.push_back({k, r});
}

int padded_taps(int m) { return m <= 16 ? 16 : (m <= 32 ? 32 : 0); }

/* unit-power complex Gaussian noise (E|X|^2 = 1), xorshift64* + Box-Muller */
void noise_spectrum(std::vector<float> &amps, unsigned long long seed)
{
    unsigned long long x = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&]() {
        x ^= x >> 12;
        x ^= x << 25;
        x ^= x >> 27;
        return (double)((x * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
    };
    for (size_t i = 0; i + 1 < amps.size(); i += 2) {
        const double u = std::max(next(), 1e-300), v = next();
        const double r = std::sqrt(-std::log(u));   /* per component variance 1/2 */
        amps[i] = (float)(r * std::cos(2.0 * M_PI * v));
        amps[i + 1] = (float)(r * std::sin(2.0 * M_PI * v));
    }
    amps[0] = 1.0f;
    amps[1] = 0.0f;
}

/* median kernel time of `reps` searches after one untimed */
double time_search(GpuSession *S, int reps)
{
    std::vector<double> t;
    for (int i = 0; i <= reps; i++) {
        CandidateList c = cs_gpu_session_search_resident(S, 8.0);
        free(c.v);
        if (i > 0) t.push_back(cs_gpu_session_last_kernel_ms(S));
    }
    std::sort(t.begin(), t.end());
    return t.empty() ? INFINITY : t[t.size() / 2];
}

}  // namespace

__attribute__((visibility("hidden"))) bool cs_tune_lookup(int device, int nh, int mp, bool f64,
                                                          bool measured, CsTuneResult *out)
{
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (g_tab.empty()) return false;
    }
    return find({arch_of(device), device, nh, mp, f64 ? 1 : 0, measured ? 1 : 0}, out);
}

extern "C" CS_API int cs_gpu_autotune(int device, const SearchParams *p, CsTuneResult *out)
{
    if (!p) return -1;
    const int mp = padded_taps(p->m);
    if (mp == 0) return -1;
    const bool f64 = p->precision == CS_PREC_F64, measured = p->sigma == CS_SIGMA_MEASURED;
    const TuneKey key{arch_of(device), device, p->nharms, mp, f64 ? 1 : 0, measured ? 1 : 0};
    if (find(key, out)) return 0;
    std::lock_guard<std::mutex> tl(g_tuning);
    if (find(key, out)) return 0;

    /* the tuning spectrum: 2^22 samples at 64 us, band 0.1 .. 125 Hz */
    const long long N = 1ll << 22;
    const double dt = 64e-6, T = (double)N * dt;
    std::vector<float> amps((size_t)N);
    noise_spectrum(amps, 12345u + (unsigned)p->nharms);
    FFTFile ft;
    memset(&ft, 0, sizeof(ft));
    ft.N = N;
    ft.T = T;
    ft.df = 1.0 / T;
    ft.namps = N / 2;
    ft.amps = amps.data();
    ft.DC = 1.0;
    SearchParams pc = *p;
    const double lodr = pc.hidr / (double)pc.nharms, r_lo = 0.1 * T, r_hi = 125.0 * T;
    const long long total = (long long)floor((r_hi - r_lo) / lodr) + 1;
    std::vector<int> blocks;
    if (measured) blocks = {CS_GPU_DEFAULT_BLOCKSIZE};
    else blocks = {16384, CS_GPU_DEFAULT_BLOCKSIZE, 262144};
    const int lanes[2] = {0, 1}, stripes[3] = {0, 16, 64};

    CsTuneResult r;
    memset(&r, 0, sizeof(r));
    r.nharms = pc.nharms;
    r.m_padded = mp;
    r.precision = pc.precision;
    r.sigma = pc.sigma;
    r.default_ms = INFINITY;
    r.best_ms = INFINITY;
    DirectPlanSet set;
    memset(&set, 0, sizeof(set));
    if (csx_plans_build(&set, &pc, r_lo) != 0) return -1;
    int rc = -1;
    for (int bs : blocks) {
        long long *cstarts = NULL, nchunks = 0;
        if (csx_chunk_starts(&set, lodr, ft.namps, N, total, bs, &cstarts, &nchunks) != 0)
            break;
        GpuSession *S = cs_gpu_session_create(&ft, &pc, &set, cstarts, total, bs, nchunks,
                                              r_lo, lodr, device);
        if (!S) {
            free(cstarts);
            break;
        }
        if (cs_gpu_session_upload(S, &ft) == 0) {
            rc = 0;
            for (int lo : lanes)
                for (int st : stripes) {
                    cs_session_set_launch(S, lo, st);
                    const double ms = time_search(S, 3);
                    r.ntried++;
                    if (bs == CS_GPU_DEFAULT_BLOCKSIZE && lo == 0 && st == 0) r.default_ms = ms;
                    if (ms < r.best_ms) {
                        r.best_ms = ms;
                        r.lane_order = lo;
                        r.xcd_stripe = st;
                        r.blocksize = bs;
                    }
                }
        }
        cs_gpu_session_free(S);
        free(cstarts);
    }
    csx_plans_free(&set);
    if (rc != 0 || !(r.best_ms < INFINITY)) return -1;
    store(key, r);
    if (out) *out = r;
    return 0;
}

extern "C" CS_API int cs_gpu_tune_save(const char *path)
{
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    fprintf(f, "# arch device nharms m_padded f64 measured lane_order xcd_stripe blocksize "
               "default_ms best_ms\n");
    for (const auto &e : g_tab)
        fprintf(f, "%s %d %d %d %d %d %d %d %d %.6f %.6f\n", e.first.arch.c_str(),
                e.first.device, e.first.nh, e.first.mp, e.first.f64, e.first.measured,
                e.second.lane_order, e.second.xcd_stripe, e.second.blocksize,
                e.second.default_ms, e.second.best_ms);
    return fclose(f) == 0 ? 0 : -1;
}

extern "C" CS_API int cs_gpu_tune_load(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char line[512];
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '#') continue;
        char arch[128];
        TuneKey k;
        CsTuneResult r;
        memset(&r, 0, sizeof(r));
        if (sscanf(line, "%127s %d %d %d %d %d %d %d %d %lf %lf", arch, &k.device, &k.nh, &k.mp,
                   &k.f64, &k.measured, &r.lane_order, &r.xcd_stripe, &r.blocksize,
                   &r.default_ms, &r.best_ms) != 11)
            continue;
        if (k.nh < 1 || k.nh > 128 || (k.mp != 16 && k.mp != 32) || r.xcd_stripe < 0 ||
            r.xcd_stripe > 4096 || r.lane_order < 0 || r.lane_order > 1 || r.blocksize < 1)
            continue;
        k.arch = arch;
        r.nharms = k.nh;
        r.m_padded = k.mp;
        r.precision = k.f64 ? CS_PREC_F64 : CS_PREC_F32;
        r.sigma = k.measured ? CS_SIGMA_MEASURED : CS_SIGMA_ANALYTIC;
        store(k, r);
    }
    fclose(f);
    return 0;
}

static_assert(sizeof(CsTuneResult) == 48, "CsTuneResult layout");
