"""Continuous-batching search engine (the cs_engine_* C-ABI) and the request
loops built on it.

`Engine` wraps one engine: `nslots` worker streams on one GPU, admission of
queued requests by HBM footprint, device staging of each raw series
(realfft + rednoise) and the fused search.  Three loops share it:

  run_plain    the one-file-at-a-time GPU form the reference runs
               (src/cli.c:590-700): host realfft + rednoise, search_file
               (session per file, search, collapse), duty cycles, in turn;
  run_reference_flow  the same with the reference list mode's prefetch:
               producer threads stage upcoming series on the host while the
               current one is searched;
  run_batched  every series submitted at once; each request collapses its
               list and measures duty cycles on the device from the resident
               spectrum (or on host threads from a copied-back spectrum);
  serve        requests submitted at their arrival times (a served stream):
               whatever arrives while others run is admitted as slots and HBM
               free up; per-request latency is arrival -> final list ready.
"""
import concurrent.futures as cf
import ctypes as C
import os
import threading
import time

import numpy as np

from . import abi
from .search import (Spectrum, Zaplist, collapse, finalize, make_params, redfft, search_file,
                     take_list)


class PinnedSeries:
    """A float32 buffer in pinned host memory (cs_host_alloc), for series
    and staged spectra that the engine copies with DMA."""

    def __init__(self, n):
        self.L = abi.host()
        self.n = int(n)
        self.ptr = self.L.cs_host_alloc(4 * self.n)
        if not self.ptr:
            raise MemoryError("pinned allocation of %d floats failed" % self.n)
        self.a = np.ctypeslib.as_array((C.c_float * self.n).from_address(self.ptr))

    def close(self):
        if self.ptr:
            self.a = None
            self.L.cs_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    def __init__(self, device=0, hbm_budget=0, nslots=4, plan_cache=8):
        """plan_cache: resident plans kept between requests (0: none); a
        request with the geometry of an idle plan reuses it, replaying its
        staging and search from graphs."""
        self.L = abi.host()
        self.h = self.L.cs_engine_create(device, int(hbm_budget), int(nslots))
        if not self.h:
            raise RuntimeError("cs_engine_create failed (see stderr)")
        self.L.cs_engine_set_plan_cache(self.h, int(plan_cache))
        self._keep = {}
        self._lock = threading.Lock()

    def submit(self, samples, dt, params, lofreq=0.1, hifreq=125.0, lobin=100, blocksize=0,
               threshold=8.0, spectrum=None, out=None, ncands=0, numharm=16, dr_tol=1.0,
               harm_tol=1.0, remove=True, harm_remove=True, zaplist=None, baryv=0.0):
        """Queue one search of a raw float32 series (host memory, a
        PinnedSeries, or a series resident on the engine's device such as a
        de-dispersed DM series, dedisp.DeviceSeries); returns its ticket.

        ncands > 0: the request returns the final list (collapsed, sorted by
        S/N, cut to ncands, duty cycles measured on the device) and the
        spectrum stays on the device.  ncands == 0: the raw list.
        With `spectrum` (default: only for raw lists) the staged spectrum
        comes back too, into `out` (N float32, e.g. pinned) when given.
        `zaplist` (Zaplist or (freq, width, bary) tuples) is applied to the
        staged spectrum before the search."""
        if spectrum is None:
            spectrum = ncands <= 0
        if hasattr(samples, "ptr") and hasattr(samples, "n"):
            # a series already in this device's memory (dedisp.DeviceSeries)
            s, ptr, n = samples, int(samples.ptr), int(samples.n)
        else:
            s = samples.a if isinstance(samples, PinnedSeries) else samples
            s = np.ascontiguousarray(s, dtype=np.float32)
            ptr, n = s.ctypes.data, s.size
        if n % 2:
            raise ValueError("the series needs an even number of samples")
        amps = None
        if spectrum:
            amps = out.a if isinstance(out, PinnedSeries) else out
            if amps is None:
                amps = np.zeros(n, dtype=np.float32)
            if amps.dtype != np.float32 or amps.size != n or not amps.flags.c_contiguous:
                raise ValueError("spectrum buffer must be N contiguous float32")
        req = abi.CsSearchRequest()
        req.samples = ptr
        req.N = n
        req.dt = float(dt)
        req.p = C.pointer(params)
        req.lofreq, req.hifreq, req.lobin = float(lofreq), float(hifreq), int(lobin)
        req.blocksize, req.threshold = int(blocksize), float(threshold)
        req.amps_out = amps.ctypes.data if spectrum else None
        req.ncands, req.numharm = int(ncands), int(numharm)
        req.remove, req.harm_remove = int(bool(remove)), int(bool(harm_remove))
        req.dr_tol, req.harm_tol = float(dr_tol), float(harm_tol)
        if zaplist is not None and len(zaplist):
            z = zaplist if isinstance(zaplist, Zaplist) else Zaplist(zaplist)
            req.zaplist = C.pointer(z.zl)     # copied by the engine at submit
        req.baryv = float(baryv)
        # the ticket is recorded before any waiter can see it finish
        with self._lock:
            t = self.L.cs_engine_submit(self.h, C.byref(req))
            if t < 0:
                raise ValueError("request refused")
            self._keep[t] = (n, amps, params, float(dt), samples)
        return t

    def next_done(self, timeout_ms=-1.0):
        t = self.L.cs_engine_next_done(self.h, float(timeout_ms))
        return None if t < 0 else t

    def wait(self, ticket):
        """(raw candidates, info dict, staged Spectrum or None)"""
        cl = abi.CandidateList()
        info = abi.CsSearchInfo()
        self.L.cs_engine_wait(self.h, ticket, C.byref(cl), C.byref(info))
        with self._lock:
            n, amps, params, dt, _ = self._keep.pop(ticket)
        raw = take_list(cl)
        d = {k: getattr(info, k) for k, _ in abi.CsSearchInfo._fields_ if k != "pad"}
        spec = None
        if amps is not None and info.status == 0:
            spec = Spectrum(amps, n, dt)
        return raw, d, spec

    def stats(self):
        st = abi.CsEngineStats()
        self.L.cs_engine_stats(self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in abi.CsEngineStats._fields_ if k != "pad"}

    def close(self):
        if self.h:
            self.L.cs_engine_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _submit_job(engine, j, ncands=0):
    """ncands > 0: the request finishes on the device (final list with duty
    cycles); 0: the raw list and the staged spectrum come back."""
    return engine.submit(j["samples"], j["dt"], j["params"], lofreq=j.get("lofreq", 0.1),
                         hifreq=j.get("hifreq", 125.0), lobin=j.get("lobin", 100),
                         blocksize=j.get("blocksize", 0), threshold=j.get("threshold", 8.0),
                         out=j.get("out") if ncands <= 0 else None, ncands=ncands)


def sweep_jobs(nfiles=64, seed=4, pinned=True, threshold=8.0, lo=22, hi=25, same_length=None):
    """The batched DM sweep: `nfiles` series, lengths log-uniform in
    [2^lo, 2^hi] rounded up to even 7-smooth sizes, per-file seeded N(0,1)
    noise, a 3.7 ms pulsar in 8 files around the middle DM (S/N peaking at
    18), nharms cycling 16/32/60, m 16, maxdecim 6, blocksize 65536 (all of
    length `same_length` when given).  With `pinned`, series and spectrum
    buffers are pinned host memory."""
    from . import synth
    jobs = []
    for m in synth.dm_sweep(nfiles, seed, lo=lo, hi=hi, same_length=same_length):
        x = synth.series_with_pulsar(m["N"], m["period"], m["duty"], m["snr"], m["seed"])
        s, o = x, None
        if pinned:
            s = PinnedSeries(m["N"])
            s.a[:] = x
            o = PinnedSeries(m["N"])
        jobs.append(dict(samples=s, out=o, dt=synth.DT, threshold=threshold, blocksize=65536,
                         params=make_params(nharms=m["nharms"], m=16, maxdecim=6),
                         snr=m["snr"], period=m["period"], nharms=m["nharms"]))
    return jobs


def serving_jobs(n=500, rate=20.0, seed=5, lo=20, hi=25, pool=8, threshold=8.0):
    """The served stream: Poisson arrivals at `rate` per second, lengths
    log-uniform in [2^lo, 2^hi] (even 7-smooth), nharms uniform over
    16/32/60, 10 % with a pulsar (period log-uniform 1-500 ms, S/N 8-20).
    Noise-only requests read windows of `pool` pinned noise buffers; pulsar
    requests get their own series.  Returns (jobs, arrival seconds)."""
    from . import synth
    meta = synth.serving_requests(n, rate, seed, lo, hi)
    nmax = max(m["N"] for m in meta)
    bufs = []
    for b in range(pool):
        ps = PinnedSeries(nmax)
        ps.a[:] = np.random.default_rng(seed * 7919 + b).standard_normal(nmax, dtype=np.float32)
        bufs.append(ps)
    jobs = []
    for i, m in enumerate(meta):
        if m["snr"] > 0:
            x = synth.series_with_pulsar(m["N"], m["period"], m["duty"], m["snr"], m["seed"])
        else:
            b = bufs[i % pool]
            off = (i * 7919) % (nmax - m["N"] + 1)
            x = b.a[off:off + m["N"]]
        jobs.append(dict(samples=x, dt=synth.DT, threshold=threshold, blocksize=65536,
                         params=make_params(nharms=m["nharms"], m=16, maxdecim=6),
                         snr=m["snr"], period=m["period"], nharms=m["nharms"],
                         _pool=bufs))
    return jobs, np.array([m["arrival"] for m in meta])


def run_plain_device(engine, jobs, ncands=100, device_post=True):
    """One file at a time through a one-slot engine: device staging, search,
    collapse and duty cycles of each file before the next is submitted
    (duty cycles on the device, or on the host from the copied-back spectrum
    with device_post=False).  Returns (final lists, infos, wall seconds)."""
    finals, infos = [], []
    t0 = time.perf_counter()
    for j in jobs:
        lst, info, spec = engine.wait(_submit_job(engine, j, ncands if device_post else 0))
        if info["status"] != 0:
            raise RuntimeError("request failed")
        finals.append(lst if device_post else _post(spec, j["params"], lst, ncands))
        infos.append(info)
    return finals, infos, time.perf_counter() - t0


def _post(spec, params, raw, ncands):
    return finalize(spec, params, collapse(raw), ncands)


def run_plain(jobs, device=0, ncands=100):
    """One file at a time: host staging, search_file, duty cycles.  jobs:
    dicts with samples, dt, params and optional lofreq/hifreq/lobin/blocksize/
    threshold.  Returns (final lists, wall seconds)."""
    out = []
    t0 = time.perf_counter()
    for j in jobs:
        s = j["samples"].a if isinstance(j["samples"], PinnedSeries) else j["samples"]
        spec = redfft(s, j["dt"])
        c = search_file(spec, j["params"], lofreq=j.get("lofreq", 0.1),
                        hifreq=j.get("hifreq", 125.0), lobin=j.get("lobin", 100),
                        blocksize=j.get("blocksize", 0), threshold=j.get("threshold", 8.0),
                        device=device)
        out.append(finalize(spec, j["params"], c, ncands))
    return out, time.perf_counter() - t0


class _HostSpectrum:
    """A staged spectrum held in a DatPipe's buffer, passed to search_file and
    measure_ducy as it is (no copy)."""

    def __init__(self, ft):
        self.ft = ft


def run_reference_flow(jobs, device=0, ncands=100, producers=2, slots=4, threads=None,
                       device_staging=False):
    """The reference's list mode on in-memory series (src/cli.c:556-686,
    996-1082): `producers` threads realfft + rednoise upcoming series on the
    host, input i into DatPipe slot i % `slots` once input i - `slots` is
    released, while this thread takes the inputs in order: search_file on the
    GPU with one SearchCache shared across inputs (a session is rebuilt
    whenever the geometry changes), sort by S/N, keep `ncands`, duty cycles
    on the host.  Each producer's transforms run on `threads` OpenMP threads
    (default: an equal share of this process's cores); with device_staging
    the producers stage on the GPU instead (datpipe_set_device), as this
    library's CLI does.  Returns (final lists, wall seconds, per-input
    staging seconds)."""
    L = abi.host()
    n = len(jobs)
    nslots, nprod = min(slots, n), min(producers, n)
    if threads is None:
        threads = max(1, len(os.sched_getaffinity(0)) // nprod)
    pipes = [L.datpipe_create(6, 100, 6.0) for _ in range(nslots)]
    for pp in pipes:
        L.datpipe_set_threads(pp, threads)
        if device_staging:
            L.datpipe_set_device(pp, device)
    fts = [abi.FFTFile() for _ in range(nslots)]
    held, ok = [-1] * nslots, [False] * nslots
    released, stop = [0], [False]
    cv = threading.Condition()
    stage_s = [0.0] * n

    def producer(first):
        for i in range(first, n, nprod):
            slot = i % nslots
            with cv:
                while not stop[0] and i - released[0] >= nslots:
                    cv.wait()
                if stop[0]:
                    return
            s = jobs[i]["samples"]
            a = s.a if isinstance(s, PinnedSeries) else np.ascontiguousarray(s, np.float32)
            t = time.perf_counter()
            rc = L.datpipe_redfft_mem(pipes[slot], a.ctypes.data_as(C.c_void_p), a.size,
                                      jobs[i]["dt"], b"memory", C.byref(fts[slot]))
            stage_s[i] = time.perf_counter() - t
            with cv:
                ok[slot] = rc == 0
                held[slot] = i
                cv.notify_all()

    cache = abi.SearchCache()
    out = []
    t0 = time.perf_counter()
    th = [threading.Thread(target=producer, args=(t,), daemon=True) for t in range(nprod)]
    for t in th:
        t.start()
    try:
        for i, j in enumerate(jobs):
            slot = i % nslots
            with cv:
                while held[slot] != i:
                    cv.wait()
            if not ok[slot]:
                raise RuntimeError("host staging of input %d failed" % i)
            spec = _HostSpectrum(fts[slot])
            c = search_file(spec, j["params"], lofreq=j.get("lofreq", 0.1),
                            hifreq=j.get("hifreq", 125.0), lobin=j.get("lobin", 100),
                            blocksize=j.get("blocksize", 0), threshold=j.get("threshold", 8.0),
                            device=device, cache=cache)
            out.append(finalize(spec, j["params"], c, ncands))
            L.fftfile_close(C.byref(fts[slot]))
            with cv:
                released[0] += 1
                cv.notify_all()
        wall = time.perf_counter() - t0
    finally:
        with cv:
            stop[0] = True
            cv.notify_all()
        for t in th:
            t.join()
        L.search_cache_free(C.byref(cache))
        for pp in pipes:
            L.datpipe_free(pp)
    return out, wall, stage_s


def run_batched(engine, jobs, post_workers=8, ncands=100, device_post=True):
    """Every job submitted at once.  With device_post each request finishes
    on the device (collapse, then duty cycles from the resident spectrum);
    otherwise each raw result is collapsed and finished on a host thread as
    soon as its search completes.  Returns (final lists in job order, infos,
    wall seconds)."""
    t0 = time.perf_counter()
    tickets = {}
    for i, j in enumerate(jobs):
        tickets[_submit_job(engine, j, ncands if device_post else 0)] = i
    finals = [None] * len(jobs)
    infos = [None] * len(jobs)
    futs = {}
    with cf.ThreadPoolExecutor(post_workers) as pool:
        for _ in range(len(jobs)):
            t = engine.next_done()
            lst, info, spec = engine.wait(t)
            i = tickets[t]
            infos[i] = info
            if info["status"] != 0:
                raise RuntimeError("request %d failed" % i)
            if device_post:
                finals[i] = lst
            else:
                futs[pool.submit(_post, spec, jobs[i]["params"], lst, ncands)] = i
        for f in cf.as_completed(futs):
            finals[futs[f]] = f.result()
    return finals, infos, time.perf_counter() - t0


def serve(engine, jobs, arrivals, post_workers=8, ncands=100, device_post=True):
    """Submit job i at arrivals[i] seconds after the start; returns (final
    lists, per-request dicts with latency_s = arrival -> final list ready,
    wall seconds).  With device_post requests finish on the device (final
    list with duty cycles); otherwise on host threads from the copied-back
    spectrum."""
    n = len(jobs)
    order = np.argsort(arrivals, kind="stable")
    start = time.perf_counter()
    tickets, t_arr = {}, {}
    lock = threading.Lock()

    def feeder():
        for i in order:
            delay = start + float(arrivals[i]) - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            j = jobs[i]
            with lock:
                t_arr[i] = start + float(arrivals[i])
                tickets[_submit_job(engine, j, ncands if device_post else 0)] = i

    th = threading.Thread(target=feeder, daemon=True)
    th.start()
    finals = [None] * n
    recs = [None] * n

    def post(i, spec, raw, info):
        f = _post(spec, jobs[i]["params"], raw, ncands)
        done = time.perf_counter()
        return i, f, done, info

    def record(i, fin, done, info):
        finals[i] = fin
        info = dict(info)
        info["latency_s"] = done - t_arr[i]
        recs[i] = info

    futs = []
    with cf.ThreadPoolExecutor(post_workers) as pool:
        for _ in range(n):
            t = engine.next_done()
            lst, info, spec = engine.wait(t)
            done = time.perf_counter()
            with lock:
                i = tickets[t]
            if info["status"] != 0:
                raise RuntimeError("request %d failed" % i)
            if device_post:
                record(i, lst, done, info)
            else:
                futs.append(pool.submit(post, i, spec, lst, info))
        for f in cf.as_completed(futs):
            record(*f.result())
    th.join()
    return finals, recs, time.perf_counter() - start


__all__ = ["Engine", "PinnedSeries", "run_plain", "run_plain_device", "run_batched", "serve",
           "sweep_jobs", "serving_jobs", "make_params"]
