/*
 * Search engine: continuous batching of whole searches on one GPU.
 *
 * A request is a raw series in host memory plus the search options of
 * search_file (src/search.c:1247-1406 -- grid, plans, chunk table) and the
 * datpipe staging (src/datpipe.c:52-253).  The engine keeps a FIFO of
 * submitted requests and `nslots` worker threads, each owning one HIP
 * stream.  A worker admits the oldest request as soon as its HBM footprint
 * fits the budget next to the requests already running, then, on its own
 * stream:
 *   host   plans (per-harmonic residue tables) and chunk starts;
 *   device stream-ordered allocation of the session and staging buffers,
 *          realfft + rednoise of the uploaded series, the fused search
 *          kernel, the record readback, stream-ordered frees.
 * Requests that arrive while others run are admitted as slots and memory free
 * up, and their transfers and kernels overlap on the device.  Each request's
 * result is the raw candidate list of cs_gpu_session_run -- the same
 * arithmetic whether it ran alone or beside others.
 *
 * Plans stay resident: the session (spectrum, weight tables, chunk and tile
 * tables, records) and the stage plan of a finished request are kept, keyed
 * by the geometry (length, sample time, search options, band, chunk length),
 * and a later request of the same geometry -- the next DM trial of a sweep --
 * runs on them with no host planning, allocation or table upload, replaying
 * its staging and search launches from captured hipGraphs.  Idle plans count
 * against the HBM budget and are evicted, least recently used first, when a
 * request needs the room or the cache is full.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cands.h"
#include "core.h"
#include "cs_device.h"
#include "cs_gpu.h"
#include "cs_stage.h"

static_assert(sizeof(CsSearchRequest) == 120, "CsSearchRequest layout");
static_assert(sizeof(CsSearchInfo) == 96, "CsSearchInfo layout");
static_assert(sizeof(CsEngineStats) == 72, "CsEngineStats layout");

namespace {

using clk = std::chrono::steady_clock;

double ms_between(clk::time_point a, clk::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct EngReq {
    long long id = 0;
    CsSearchRequest in{};
    SearchParams p{};
    long long footprint = 0;
    std::vector<ZapBird> zap;     /* the request's zaplist, copied at submit */
    ZapList zl{};
    int state = 0;                /* 0 queued, 1 running, 2 done */
    long long reserved = 0;       /* HBM the request holds in the budget */
    CandidateList out{};
    CsSearchInfo info{};
    clk::time_point t_submit, t_start, t_end;
};

/* A resident plan: the session and stage plan of one geometry. */
struct Plan {
    long long N = 0;
    double dt = 0.0, r_lo = 0.0, r_hi = 0.0;
    int Nprof = 0;
    SearchParams p{};             /* the session points here */
    DirectPlanSet set{};
    long long *cstarts = NULL;
    GpuSession *S = NULL;
    CsStage *g = NULL;
    long long footprint = 0;
    bool busy = false;
    unsigned long long last_use = 0;
};

}  // namespace

struct CsEngine {
    int device = 0;
    long long budget = 0, in_use = 0, peak_in_use = 0;
    int nslots = 0, running = 0, peak_running = 0;
    long long completed = 0;
    int max_plans = 8;            /* resident plans kept; 0: none */
    std::vector<Plan *> plans;
    unsigned long long use_clock = 0;
    long long plan_hits = 0;
    hipStream_t idle_st = NULL;   /* idle plans are bound here (their frees run on it) */
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<EngReq *> queue;
    std::unordered_map<long long, EngReq *> reqs;
    std::deque<long long> done_order;
    std::vector<std::thread> workers;
    bool stop = false;
    long long next_id = 0;
};

namespace {

/* HBM bytes one request holds while it runs: series, FFT ping-pong, realfft
 * output, spectrum, rednoise tables, plan arena, chunk/tile tables, records */
long long footprint_of(const CsSearchRequest *r, const SearchParams *p)
{
    const long long N = r->N;
    long long b = 4 * N + cs_stage_fft_bytes(N) + 4 * N + 4 * N + N / 2;
    int pnum = 0, q = 0;
    if (!csx_trial_grid_rational(p->hidr / (double)p->nharms, &pnum, &q)) q = 1;
    b += (long long)p->nharms * q * 4 * (p->m + 12);
    const double T = (double)N * r->dt;
    const double lodr = p->hidr / (double)p->nharms;
    double r_lo = r->lofreq * T;
    if (r->lobin != 100) r_lo = (double)r->lobin;
    const double trials = std::max(0.0, (r->hifreq * T - r_lo) / lodr + 1.0);
    const int nprof = r->blocksize > 0 ? r->blocksize : CS_GPU_DEFAULT_BLOCKSIZE;
    const double nchunks = trials / nprof + 2.0 * p->nharms + 2.0;
    b += (long long)(trials / 256.0 * 16.0 + nchunks * (16.0 + (double)sizeof(ChunkDesc)));
    b += (1 << 16) * 16 + (1 << 20);
    b += cs_session_general_bytes(p, lodr);
    return b;
}

void plan_free(Plan *P)
{
    if (!P) return;
    if (P->g) cs_gpu_stage_free(P->g);
    if (P->S) cs_gpu_session_free(P->S);
    csx_plans_free(&P->set);
    free(P->cstarts);
    delete P;
}

bool plan_matches(const Plan *P, long long N, double dt, const SearchParams *p, double r_lo,
                  double r_hi, int Nprof)
{
    return P->N == N && P->dt == dt && P->r_lo == r_lo && P->r_hi == r_hi &&
           P->Nprof == Nprof && memcmp(&P->p, p, sizeof(SearchParams)) == 0;
}

/* under E->mu: drop idle plans, least recently used first, until `need`
 * more bytes fit the budget (need < 0: until at most max_plans remain).  The
 * dropped plans go to `out`, freed by the caller once it has released
 * E->mu: freeing device memory, streams' graphs and events can wait on the
 * device, and under the lock it would hold up every other worker. */
void evict_idle(CsEngine *E, long long need, std::vector<Plan *> &out)
{
    for (;;) {
        const bool over = need >= 0 ? E->in_use + need > E->budget
                                    : (int)E->plans.size() > E->max_plans;
        if (!over) return;
        int victim = -1;
        for (int i = 0; i < (int)E->plans.size(); i++)
            if (!E->plans[i]->busy &&
                (victim < 0 || E->plans[i]->last_use < E->plans[victim]->last_use))
                victim = i;
        if (victim < 0) return;
        Plan *P = E->plans[victim];
        E->plans.erase(E->plans.begin() + victim);
        E->in_use -= P->footprint;
        out.push_back(P);
    }
}

void free_plans(CsEngine *E, std::vector<Plan *> &ps)
{
    if (ps.empty()) return;
    hipSetDevice(E->device);
    for (Plan *P : ps) plan_free(P);
    ps.clear();
}

int run_request(CsEngine *E, EngReq *r, hipStream_t st)
{
    const SearchParams *p = &r->p;
    const long long N = r->in.N;
    const double dt = r->in.dt, T = (double)N * dt;
    const double lodr = p->hidr / (double)p->nharms;
    double r_lo = r->in.lofreq * T;
    if (r->in.lobin != 100) r_lo = (double)r->in.lobin;          /* src/search.c:1259-1260 */
    const double r_hi = r->in.hifreq * T;
    long long total = (long long)floor((r_hi - r_lo) / lodr) + 1;
    if (total < 0) total = 0;
    int Nprof = r->in.blocksize > 0 ? r->in.blocksize : CS_GPU_DEFAULT_BLOCKSIZE;
    if (r->in.blocksize < 0 && p->sigma == CS_SIGMA_ANALYTIC) {
        /* the class's tuned chunk length, when it has one: the analytic
         * metrics do not depend on it, the raw list's chunk order does */
        CsTuneResult tr;
        const int mp = p->m <= 16 ? 16 : 32;
        if (cs_tune_lookup(E->device, p->nharms, mp, p->precision == CS_PREC_F64, false, &tr))
            Nprof = tr.blocksize;
    }

    const clk::time_point t0 = clk::now();
    clk::time_point t1, t2, t3;
    Plan *P = NULL;
    bool reused = false;
    int rc = -1;
    {
        std::lock_guard<std::mutex> lk(E->mu);
        for (Plan *c : E->plans)
            if (!c->busy && plan_matches(c, N, dt, p, r_lo, r_hi, Nprof)) {
                P = c;
                break;
            }
        if (P) {
            /* the plan's memory is already in the budget */
            P->busy = true;
            reused = true;
            E->plan_hits++;
            E->in_use -= r->reserved;
            r->reserved = 0;
        }
    }
    if (!P) {
        P = new Plan();
        P->N = N;
        P->dt = dt;
        P->r_lo = r_lo;
        P->r_hi = r_hi;
        P->Nprof = Nprof;
        P->p = *p;
        P->busy = true;
        FFTFile ft;
        memset(&ft, 0, sizeof(ft));
        ft.N = N;
        ft.T = T;
        ft.df = 1.0 / T;
        ft.namps = N / 2;
        ft.DC = 1.0;
        long long nchunks = 0;
        /* the stage plan first and the samples' copy started on the stream,
         * so that the DMA runs under the host's planning of the search */
        if (!(P->g = cs_stage_create_on(N, dt, 6, 100, 6.0, E->device, st)) ||
            cs_stage_copy_in(P->g, r->in.samples, st) != 0 ||
            csx_plans_build(&P->set, &P->p, r_lo) != 0 ||
            csx_chunk_starts(&P->set, lodr, ft.namps, N, total, Nprof, &P->cstarts,
                             &nchunks) != 0 ||
            !(P->S = cs_session_create_on(&ft, &P->p, &P->set, P->cstarts, total, Nprof,
                                          nchunks, r_lo, lodr, E->device, st))) {
            plan_free(P);
            hipStreamSynchronize(st);      /* the copy may still read the samples */
            return -1;
        }
    }
    r->info.reused = reused ? 1 : 0;
    GpuSession *S = P->S;
    CsStage *g = P->g;
    /* the plan runs on this worker's stream; graphs once it is reused */
    cs_session_bind_stream(S, st, reused);
    cs_stage_bind_stream(g, st);
    t1 = clk::now();
    r->info.plan_ms = ms_between(t0, t1);
    if (reused) {
        if (cs_session_stage_graph(S, g, r->in.samples) != 0) goto done;
        if ((r->zl.n > 0 || r->in.amps_out) &&
            cs_gpu_session_zap(S, &r->zl, r->in.baryv, r->in.amps_out) < 0)
            goto done;
    } else if (r->zl.n > 0) {
        if (cs_session_stage_loaded(S, g, NULL) != 0 ||
            cs_gpu_session_zap(S, &r->zl, r->in.baryv, r->in.amps_out) < 0)
            goto done;
    } else if (cs_session_stage_loaded(S, g, r->in.amps_out) != 0) {
        goto done;
    }
    t2 = clk::now();
    r->info.stage_ms = ms_between(t1, t2);
    r->out = cs_gpu_session_search_resident(S, r->in.threshold);
    t3 = clk::now();
    r->info.search_ms = ms_between(t2, t3);
    r->info.kernel_ms = cs_gpu_session_last_kernel_ms(S);
    if (r->in.ncands > 0) {
        /* src/search.c:1467-1478, then src/cli.c:1059-1061 on the device */
        CandidateList c = r->out;
        if (r->in.remove) {
            CandidateList kept = csx_remove_duplicates(&c, r->in.dr_tol);
            csx_candlist_free(&c);
            c = kept;
        } else {
            csx_candlist_sort_freq(&c);
        }
        if (r->in.harm_remove) {
            CandidateList kept = csx_remove_harmonics(&c, r->in.numharm, r->in.harm_tol);
            csx_candlist_free(&c);
            c = kept;
        }
        const size_t nin = c.n;
        r->out = cs_session_finalize(S, &c, r->in.ncands);
        csx_candlist_free(&c);
        if (nin > 0 && r->out.n == 0) goto done;
        r->info.post_ms = ms_between(t3, clk::now());
    }
    {
        long long ht = 0, tr = 0;
        double ab = 0.0;
        cs_gpu_session_work(S, &ht, &tr, &ab);
        r->info.harmonic_trials = ht;
        r->info.trials = tr;
    }
    rc = 0;
done:
    hipStreamSynchronize(st);
    std::vector<Plan *> dropped;
    {
        std::lock_guard<std::mutex> lk(E->mu);
        if (rc == 0 && E->max_plans > 0) {
            /* keep the plan resident, idle, bound to the engine's idle stream */
            cs_session_bind_stream(S, E->idle_st, false);
            cs_stage_bind_stream(g, E->idle_st);
            P->busy = false;
            P->last_use = ++E->use_clock;
            if (!reused) {
                P->footprint = r->reserved;       /* the request's share moves to the plan */
                r->reserved = 0;
                E->plans.push_back(P);
            }
            evict_idle(E, -1, dropped);
        } else {
            if (reused) {
                for (size_t i = 0; i < E->plans.size(); i++)
                    if (E->plans[i] == P) {
                        E->plans.erase(E->plans.begin() + (long)i);
                        break;
                    }
                E->in_use -= P->footprint;
            }
            dropped.push_back(P);
        }
    }
    free_plans(E, dropped);
    return rc;
}

void worker(CsEngine *E)
{
    hipSetDevice(E->device);
    hipStream_t st = NULL;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = NULL;
    /* this thread's pinned table-upload buffer at its working size up front:
     * growing it later (hipHostFree + hipHostMalloc) stalls every stream */
    (void)cs_upload_buffer((size_t)32 << 20);
    for (;;) {
        EngReq *r = NULL;
        std::vector<Plan *> dropped;
        {
            std::unique_lock<std::mutex> lk(E->mu);
            E->cv_work.wait(lk, [&] {
                if (E->queue.empty()) return E->stop;
                const EngReq *h = E->queue.front();
                if (E->in_use + h->footprint > E->budget) evict_idle(E, h->footprint, dropped);
                return E->in_use == 0 || E->in_use + h->footprint <= E->budget;
            });
            if (E->queue.empty()) {
                lk.unlock();
                free_plans(E, dropped);
                break;
            }
            r = E->queue.front();
            E->queue.pop_front();
            E->in_use += r->footprint;
            r->reserved = r->footprint;
            E->peak_in_use = std::max(E->peak_in_use, E->in_use);
            E->running++;
            E->peak_running = std::max(E->peak_running, E->running);
            r->state = 1;
            r->t_start = clk::now();
        }
        /* memory of plans evicted for this request: freed now, the lock
         * released (the budget already counts it as returned) */
        free_plans(E, dropped);
        const int rc = st ? run_request(E, r, st) : -1;
        {
            std::lock_guard<std::mutex> lk(E->mu);
            r->t_end = clk::now();
            r->info.status = rc;
            r->info.footprint = r->footprint;
            r->info.queue_ms = ms_between(r->t_submit, r->t_start);
            r->info.run_ms = ms_between(r->t_start, r->t_end);
            r->info.latency_ms = ms_between(r->t_submit, r->t_end);
            r->state = 2;
            E->in_use -= r->reserved;
            r->reserved = 0;
            E->running--;
            E->completed++;
            E->done_order.push_back(r->id);
        }
        E->cv_work.notify_all();
        E->cv_done.notify_all();
    }
    if (st) hipStreamDestroy(st);
}

}  // namespace

extern "C" CS_API CsEngine *cs_engine_create(int device, long long hbm_budget, int nslots)
{
    if (hipSetDevice(device) != hipSuccess) {
        fprintf(stderr, "Error: cannot select HIP device %d\n", device);
        return NULL;
    }
    if (hbm_budget <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return NULL;
        hbm_budget = (long long)(0.8 * (double)fr);
    }
    {
        /* keep freed request buffers in the device's pool for the next request
         * instead of returning them to the driver at every synchronisation */
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
            uint64_t thr = (uint64_t)hbm_budget;
            hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
    }
    CsEngine *E = new CsEngine();
    E->device = device;
    E->budget = hbm_budget;
    E->nslots = nslots > 0 ? nslots : 4;
    if (hipStreamCreateWithFlags(&E->idle_st, hipStreamNonBlocking) != hipSuccess) {
        fprintf(stderr, "Error: cannot create the engine's stream\n");
        delete E;
        return NULL;
    }
    for (int i = 0; i < E->nslots; i++) E->workers.emplace_back(worker, E);
    return E;
}

extern "C" CS_API long long cs_engine_submit(CsEngine *E, const CsSearchRequest *r)
{
    if (!E || !r || !r->samples || !r->p || r->N <= 0 || (r->N & 1) || !(r->dt > 0.0))
        return -1;
    EngReq *q = new EngReq();
    q->in = *r;
    q->p = *r->p;
    q->in.p = &q->p;
    if (r->zaplist && r->zaplist->n > 0 && r->zaplist->v) {
        q->zap.assign(r->zaplist->v, r->zaplist->v + r->zaplist->n);
        q->zl.v = q->zap.data();
        q->zl.n = (int)q->zap.size();
    }
    q->in.zaplist = NULL;
    q->footprint = footprint_of(r, &q->p);
    q->t_submit = clk::now();
    {
        std::lock_guard<std::mutex> lk(E->mu);
        if (E->stop) {
            delete q;
            return -1;
        }
        q->id = E->next_id++;
        E->reqs[q->id] = q;
        E->queue.push_back(q);
    }
    E->cv_work.notify_all();
    return q->id;
}

extern "C" CS_API int cs_engine_poll(CsEngine *E, long long ticket)
{
    if (!E) return -1;
    std::lock_guard<std::mutex> lk(E->mu);
    auto it = E->reqs.find(ticket);
    if (it == E->reqs.end()) return -1;
    return it->second->state == 2 ? 1 : 0;
}

extern "C" CS_API long long cs_engine_next_done(CsEngine *E, double timeout_ms)
{
    if (!E) return -1;
    std::unique_lock<std::mutex> lk(E->mu);
    auto ready = [&] { return !E->done_order.empty(); };
    if (timeout_ms < 0) E->cv_done.wait(lk, ready);
    else E->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready);
    if (E->done_order.empty()) return -1;
    const long long id = E->done_order.front();
    E->done_order.pop_front();
    return id;
}

extern "C" CS_API int cs_engine_wait(CsEngine *E, long long ticket, CandidateList *out,
                                     CsSearchInfo *info)
{
    if (!E) return -1;
    EngReq *r = NULL;
    {
        std::unique_lock<std::mutex> lk(E->mu);
        auto it = E->reqs.find(ticket);
        if (it == E->reqs.end()) return -1;
        r = it->second;
        E->cv_done.wait(lk, [&] { return r->state == 2; });
        E->reqs.erase(it);
        for (auto d = E->done_order.begin(); d != E->done_order.end(); ++d)
            if (*d == ticket) {
                E->done_order.erase(d);
                break;
            }
    }
    if (out) *out = r->out;
    else free(r->out.v);
    if (info) *info = r->info;
    const int st = r->info.status;
    delete r;
    return st;
}

extern "C" CS_API void cs_engine_stats(CsEngine *E, CsEngineStats *s)
{
    if (!E || !s) return;
    std::lock_guard<std::mutex> lk(E->mu);
    s->budget = E->budget;
    s->in_use = E->in_use;
    s->peak_in_use = E->peak_in_use;
    s->nslots = E->nslots;
    s->running = E->running;
    s->peak_running = E->peak_running;
    s->queued = (long long)E->queue.size();
    s->completed = E->completed;
    s->plan_hits = E->plan_hits;
    s->plans_resident = (long long)E->plans.size();
}

extern "C" CS_API void cs_engine_free(CsEngine *E)
{
    if (!E) return;
    {
        std::lock_guard<std::mutex> lk(E->mu);
        E->stop = true;
    }
    E->cv_work.notify_all();
    for (std::thread &t : E->workers) t.join();
    for (auto &kv : E->reqs) {
        free(kv.second->out.v);
        delete kv.second;
    }
    hipSetDevice(E->device);
    for (Plan *P : E->plans) plan_free(P);
    E->plans.clear();
    if (E->idle_st) {
        hipStreamSynchronize(E->idle_st);
        hipStreamDestroy(E->idle_st);
    }
    delete E;
}

extern "C" CS_API void cs_engine_set_plan_cache(CsEngine *E, int max_plans)
{
    if (!E) return;
    std::vector<Plan *> dropped;
    {
        std::lock_guard<std::mutex> lk(E->mu);
        E->max_plans = max_plans > 0 ? max_plans : 0;
        evict_idle(E, -1, dropped);
    }
    free_plans(E, dropped);
}

extern "C" CS_API void *cs_host_alloc(long long bytes)
{
    void *p = NULL;
    if (bytes <= 0 || hipHostMalloc(&p, (size_t)bytes, hipHostMallocDefault) != hipSuccess)
        return NULL;
    return p;
}

extern "C" CS_API void cs_host_free(void *p)
{
    if (p) hipHostFree(p);
}
