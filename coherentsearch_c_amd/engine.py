"""Continuous-batching search engine (the cs_engine_* C-ABI) and the request
loops built on it.

`Engine` wraps one engine: `nslots` worker streams on one GPU, admission of
queued requests by HBM footprint, device staging of each raw series
(realfft + rednoise) and the fused search.  Three loops share it:

  run_plain    the one-file-at-a-time GPU form the reference runs
               (src/cli.c:590-700): host realfft + rednoise, search_file
               (session per file, search, collapse), duty cycles, in turn;
  run_batched  every series submitted at once; results collapsed and finished
               on host threads as they complete, overlapping the GPU work;
  serve        requests submitted at their arrival times (a served stream):
               whatever arrives while others run is admitted as slots and HBM
               free up; per-request latency is arrival -> final list ready.
"""
import concurrent.futures as cf
import ctypes as C
import threading
import time

import numpy as np

from . import abi
from .search import (Spectrum, collapse, finalize, make_params, redfft, search_file,
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
    def __init__(self, device=0, hbm_budget=0, nslots=4):
        self.L = abi.host()
        self.h = self.L.cs_engine_create(device, int(hbm_budget), int(nslots))
        if not self.h:
            raise RuntimeError("cs_engine_create failed (see stderr)")
        self._keep = {}
        self._lock = threading.Lock()

    def submit(self, samples, dt, params, lofreq=0.1, hifreq=125.0, lobin=100, blocksize=0,
               threshold=8.0, spectrum=True):
        """Queue one search of a raw float32 series; returns its ticket.
        With `spectrum`, the staged spectrum comes back for host
        post-processing (duty cycles)."""
        s = samples.a if isinstance(samples, PinnedSeries) else samples
        s = np.ascontiguousarray(s, dtype=np.float32)
        if s.size % 2:
            raise ValueError("the series needs an even number of samples")
        amps = np.zeros(s.size, dtype=np.float32) if spectrum else None
        req = abi.CsSearchRequest()
        req.samples = s.ctypes.data
        req.N = s.size
        req.dt = float(dt)
        req.p = C.pointer(params)
        req.lofreq, req.hifreq, req.lobin = float(lofreq), float(hifreq), int(lobin)
        req.blocksize, req.threshold = int(blocksize), float(threshold)
        req.amps_out = amps.ctypes.data if spectrum else None
        # the ticket is recorded before any waiter can see it finish
        with self._lock:
            t = self.L.cs_engine_submit(self.h, C.byref(req))
            if t < 0:
                raise ValueError("request refused")
            self._keep[t] = (s, amps, params, float(dt), samples)
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
            s, amps, params, dt, _ = self._keep.pop(ticket)
        raw = take_list(cl)
        d = {k: getattr(info, k) for k, _ in abi.CsSearchInfo._fields_ if k != "pad"}
        spec = None
        if amps is not None and info.status == 0:
            spec = Spectrum(amps, s.size, dt)
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


def run_batched(engine, jobs, post_workers=8, ncands=100):
    """Every job submitted at once; each result collapsed and finished on a
    host thread as soon as its search completes.  Returns (final lists in job
    order, infos, wall seconds)."""
    t0 = time.perf_counter()
    tickets = {}
    for i, j in enumerate(jobs):
        t = engine.submit(j["samples"], j["dt"], j["params"], lofreq=j.get("lofreq", 0.1),
                          hifreq=j.get("hifreq", 125.0), lobin=j.get("lobin", 100),
                          blocksize=j.get("blocksize", 0), threshold=j.get("threshold", 8.0))
        tickets[t] = i
    finals = [None] * len(jobs)
    infos = [None] * len(jobs)
    futs = {}
    with cf.ThreadPoolExecutor(post_workers) as pool:
        for _ in range(len(jobs)):
            t = engine.next_done()
            raw, info, spec = engine.wait(t)
            i = tickets[t]
            infos[i] = info
            if info["status"] != 0:
                raise RuntimeError("request %d failed" % i)
            futs[pool.submit(_post, spec, jobs[i]["params"], raw, ncands)] = i
        for f in cf.as_completed(futs):
            finals[futs[f]] = f.result()
    return finals, infos, time.perf_counter() - t0


def serve(engine, jobs, arrivals, post_workers=8, ncands=100):
    """Submit job i at arrivals[i] seconds after the start; returns (final
    lists, per-request dicts with latency_s = arrival -> final list ready,
    wall seconds)."""
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
                t = engine.submit(j["samples"], j["dt"], j["params"],
                                  lofreq=j.get("lofreq", 0.1), hifreq=j.get("hifreq", 125.0),
                                  lobin=j.get("lobin", 100), blocksize=j.get("blocksize", 0),
                                  threshold=j.get("threshold", 8.0))
                tickets[t] = i

    th = threading.Thread(target=feeder, daemon=True)
    th.start()
    finals = [None] * n
    recs = [None] * n

    def post(i, spec, raw, info):
        f = _post(spec, jobs[i]["params"], raw, ncands)
        done = time.perf_counter()
        return i, f, done, info

    futs = []
    with cf.ThreadPoolExecutor(post_workers) as pool:
        for _ in range(n):
            t = engine.next_done()
            raw, info, spec = engine.wait(t)
            with lock:
                i = tickets[t]
            if info["status"] != 0:
                raise RuntimeError("request %d failed" % i)
            futs.append(pool.submit(post, i, spec, raw, info))
        for f in cf.as_completed(futs):
            i, fin, done, info = f.result()
            finals[i] = fin
            info = dict(info)
            info["latency_s"] = done - t_arr[i]
            recs[i] = info
    th.join()
    return finals, recs, time.perf_counter() - start


__all__ = ["Engine", "PinnedSeries", "run_plain", "run_batched", "serve", "make_params"]
