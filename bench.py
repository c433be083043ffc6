#!/usr/bin/env python3
"""Benchmark: harmonic-trials/s of the coherent harmonic search on one MI355X.

Workload (one "step"): one full search of a 2^24-sample series at dt = 64 us
(5.3 ms von Mises pulsar, duty 0.03, S/N 12, realfft + rednoise in memory),
nharms 60, m 16, maxdecim 6, blocksize 65536, hifreq 125 Hz, threshold 8 --
interpolation of every filled harmonic of 16.1 M trial fundamentals, c2r
profiles, boxcar S/N over all widths and decimations, candidate list.  The
spectrum is resident in HBM before the timed region (uploaded once); each step
is the fused kernel plus the candidate readback and host ordering.

Prints one JSON line.  Extra keys: the kernel's HBM roofline from HIP events
(algorithmic bytes of the reference GPU dataflow per launch / average kernel
time), the CPU baseline (the oracle's restatement of the reference CPU path on
a bounded sample of the same workload), the upload-inclusive rate and the
parity of the step's candidates.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(spec, params_kw, budget_s=12.0, threads=None):
    """Oracle (reference CPU path restated, f32, blocksize 2048, OpenMP
    round-robin chunks, `threads` host threads, default every one) on the same
    spectrum: passes over the chunks from a quarter of the band to its end
    (all harmonics filled, like most of the search) until the time budget is
    spent; harmonic-trials/s over the timed searches only."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    threads = threads or len(os.sched_getaffinity(0))
    cfg = po.config(blocksize=2048, **params_kw)
    o = po.Oracle(spec.amps, spec.N, spec.T, cfg)
    nchunks = o.nchunks
    c_start = nchunks // 4
    # harmonic-trials per chunk from the fit predicate at the chunk ends
    # (fits are fixed within a chunk by construction of the chunk starts)
    ht = []
    for c in range(c_start, nchunks):
        i0, n, _ = o.chunk(c)
        nf = sum(po.lib().orc_harmonic_fits(o.h, h, i0) and
                 po.lib().orc_harmonic_fits(o.h, h, i0 + n - 1)
                 for h in range(1, cfg.nharms + 1))
        ht.append(nf * n)
    step = max(threads * 4, 8)
    done_ht, elapsed, passes, c0 = 0, 0.0, 0, c_start
    while elapsed < budget_s:
        c1 = min(c0 + step, nchunks)
        t = time.perf_counter()
        o.search(8.0, threads, c0, c1)
        elapsed += time.perf_counter() - t
        done_ht += sum(ht[c0 - c_start:c1 - c_start])
        c0 = c1
        if c0 >= nchunks:
            c0, passes = c_start, passes + 1
    o.close()
    return dict(value=done_ht / elapsed, unit="harmonic-trials/s", cores=threads, kind="port",
                sample="chunks %d..%d of the same search (blocksize 2048) searched %.2f times "
                       "over, %.1f s of CPU search, %.3g harmonic-trials"
                       % (c_start, nchunks, done_ht / max(sum(ht), 1), elapsed, done_ht))


def _found(final, period, tol=2e-3):
    f0 = 1.0 / period
    return bool(final.size and any(abs(fr / f0 - round(fr / f0)) < tol and round(fr / f0) >= 1
                                   for fr in final["freq"][:3]))


def _same_lists(a, b):
    return a.size == b.size and all(np.array_equal(a[f], b[f])
                                    for f in ("r", "nharm", "metric", "ducy"))


def dm_sweep(nfiles, slots, warm_all=True):
    """The batched DM sweep: the batched/served form (engine, `slots`
    streams, every file submitted at once, each request finishing on the
    device: collapse, duty cycles from the resident spectrum) against the
    one-file-at-a-time forms on the same series: the reference's list mode
    (two producer threads realfft + rednoise the next series on the host while
    search_file runs the current one on the GPU, then duty cycles on the
    host), and the same one-at-a-time order with device staging (a one-slot
    engine, each file finished before the next starts)."""
    from coherentsearch_c_amd import engine as eng
    t = time.perf_counter()
    jobs = eng.sweep_jobs(nfiles)
    log("dm sweep: %d series, %.3g samples, generated in %.1f s"
        % (nfiles, sum(j["samples"].n for j in jobs), time.perf_counter() - t))
    with eng.Engine(nslots=1) as E:          # code objects and memory pool, untimed
        eng.run_plain_device(E, jobs[:2])
    if warm_all:
        # one untimed batched pass: the device's stream-ordered memory pool
        # grows to the sweep's working set once, as it has in a long-running
        # engine (growing it costs a request up to 12 ms of setup); every
        # form below runs on the warm pool
        with eng.Engine(nslots=slots) as E:
            eng.run_batched(E, jobs, post_workers=16)
    with eng.Engine(nslots=slots) as E:
        fb, ib, wb = eng.run_batched(E, jobs, post_workers=16)
        st = E.stats()
    ht = sum(i["harmonic_trials"] for i in ib)
    log("  batched: %.3f s, %.4g ht/s" % (wb, ht / wb))
    with eng.Engine(nslots=1) as E:
        fp, _, wp = eng.run_plain_device(E, jobs)
    log("  plain, device staging: %.3f s, %.4g ht/s" % (wp, ht / wp))
    threads = max(1, len(os.sched_getaffinity(0)) // 2)
    fh, wh, stage_s = eng.run_reference_flow(jobs, producers=2, slots=4, threads=threads)
    log("  plain, reference list mode: %.3f s, %.4g ht/s (host staging %.3f s per file on %d "
        "threads per producer)" % (wh, ht / wh, float(np.mean(stage_s)), threads))
    fl, wl, stage_l = eng.run_reference_flow(jobs, producers=2, slots=4, device_staging=True)
    log("  plain, this library's list mode (producers stage on the GPU): %.3f s, %.4g ht/s"
        % (wl, ht / wl))
    # host staging differs from device staging in the last float bits of a few
    # bins (both FFTs are double precision): lists agree within the tolerance
    host_ok = True
    for a, b in zip(fb, fh):
        if a.size != b.size or not np.array_equal(a["nharm"], b["nharm"]) or \
                np.any(np.abs(a["metric"] - b["metric"]) > 1e-4 * np.maximum(np.abs(b["metric"]), 1)):
            host_ok = False
    psr = [i for i, j in enumerate(jobs) if j["snr"] >= 8.0]
    return {
        "files": nfiles, "samples": int(sum(j["samples"].n for j in jobs)),
        "harmonic_trials": int(ht), "slots": slots,
        "batched_s": wb, "batched_ht_s": ht / wb,
        "plain_reference_flow_s": wh, "plain_reference_flow_ht_s": ht / wh,
        "ratio_vs_plain_reference_flow": wh / wb,
        "reference_flow": "2 producers x %d host threads, 4 slots; mean host realfft+rednoise "
                          "%.3f s per file" % (threads, float(np.mean(stage_s))),
        "plain_list_mode_device_staging_s": wl, "plain_list_mode_device_staging_ht_s": ht / wl,
        "ratio_vs_plain_list_mode_device_staging": wl / wb,
        "plain_device_s": wp, "plain_device_ht_s": ht / wp,
        "ratio_vs_plain_device": wp / wb,
        "lists_identical_batched_vs_plain_device": all(_same_lists(a, b) for a, b in zip(fb, fp)),
        "lists_match_reference_flow_within_1e-4": host_ok,
        "pulsar_files_snr_ge_8_detected": "%d/%d" % (sum(_found(fb[i], jobs[i]["period"])
                                                       for i in psr), len(psr)),
        "peak_concurrent_searches": st["peak_running"],
        "peak_hbm_in_use_bytes": st["peak_in_use"],
    }


def dm_sweep_resident(nfiles, slots, N=1 << 23):
    """The DM trials of one observation: `nfiles` series of N samples, nharms
    cycling 16/32/60 (three geometries).  Batched with resident plans (each
    geometry's session and stage plan reused, staging and search replayed
    from graphs) against one file at a time with a fresh plan per file and
    one file at a time with resident plans."""
    from coherentsearch_c_amd import engine as eng
    jobs = eng.sweep_jobs(nfiles, seed=6, same_length=N)
    with eng.Engine(nslots=1, plan_cache=0) as E:        # warm: code objects, pool
        eng.run_plain_device(E, jobs[:2])
    with eng.Engine(nslots=slots, plan_cache=8) as E:
        fb, ib, wb = eng.run_batched(E, jobs)
        st = E.stats()
    ht = sum(i["harmonic_trials"] for i in ib)
    with eng.Engine(nslots=1, plan_cache=0) as E:
        fp, _, wp = eng.run_plain_device(E, jobs)
    with eng.Engine(nslots=1, plan_cache=8) as E:
        fr, _, wr = eng.run_plain_device(E, jobs)
    log("dm sweep, %d x %d samples: batched+resident %.3f s (%.4g ht/s), plain %.3f s, "
        "plain+resident %.3f s" % (nfiles, N, wb, ht / wb, wp, wr))
    return {
        "files": nfiles, "samples_per_file": N, "harmonic_trials": int(ht), "slots": slots,
        "batched_resident_s": wb, "batched_resident_ht_s": ht / wb,
        "plain_fresh_plans_s": wp, "plain_fresh_plans_ht_s": ht / wp,
        "plain_resident_s": wr, "plain_resident_ht_s": ht / wr,
        "ratio_vs_plain_fresh": wp / wb, "ratio_vs_plain_resident": wr / wb,
        "plan_hits": st["plan_hits"], "mean_plan_ms_batched": float(np.mean([i["plan_ms"] for i in ib])),
        "lists_identical": all(_same_lists(a, b) and _same_lists(a, c)
                               for a, b, c in zip(fb, fp, fr)),
    }


def serving(nreq, slots, rate=20.0):
    """The served stream: Poisson arrivals at `rate`/s, continuous batching
    on the engine; latency = arrival -> collapsed list with duty cycles (each
    request finishes on the device)."""
    from coherentsearch_c_amd import engine as eng
    t = time.perf_counter()
    jobs, arrivals = eng.serving_jobs(nreq, rate)
    log("serving: %d requests over %.1f s, generated in %.1f s"
        % (nreq, arrivals[-1], time.perf_counter() - t))
    with eng.Engine(nslots=slots) as E:
        fin, recs, wall = eng.serve(E, jobs, arrivals, post_workers=16)
        st = E.stats()
    lat = np.array([r["latency_s"] for r in recs]) * 1e3
    ht = sum(r["harmonic_trials"] for r in recs)
    psr = [i for i, j in enumerate(jobs) if j["snr"] > 0]
    log("  p50 %.1f ms, p99 %.1f ms, %.4g ht/s over %.1f s" % (np.percentile(lat, 50),
                                                             np.percentile(lat, 99),
                                                             ht / wall, wall))
    return {
        "requests": nreq, "rate_per_s": rate, "slots": slots,
        "p50_latency_ms": float(np.percentile(lat, 50)),
        "p99_latency_ms": float(np.percentile(lat, 99)),
        "max_latency_ms": float(lat.max()),
        "mean_queue_ms": float(np.mean([r["queue_ms"] for r in recs])),
        "mean_plan_ms": float(np.mean([r["plan_ms"] for r in recs])),
        "mean_stage_ms": float(np.mean([r["stage_ms"] for r in recs])),
        "mean_search_ms": float(np.mean([r["search_ms"] for r in recs])),
        "mean_kernel_ms": float(np.mean([r["kernel_ms"] for r in recs])),
        "mean_post_ms": float(np.mean([r["post_ms"] for r in recs])),
        "mean_host_return_ms": float(np.mean([1e3 * r["latency_s"] - r["latency_ms"]
                                              for r in recs])),
        "harmonic_trials": int(ht), "wall_s": wall, "throughput_ht_s": ht / wall,
        "peak_concurrent_searches": st["peak_running"],
        "pulsar_requests_detected": "%d/%d" % (sum(_found(fin[i], jobs[i]["period"])
                                                 for i in psr), len(psr)),
    }


def metric_stats_timing(spec, params, reps=3):
    """--metricstats / --normalize at the C2 size: search_file collecting the
    statistics, and normalising, beside the plain search_file on the same warm
    session (each call uploads the spectrum and collapses its list)."""
    import coherentsearch_c_amd as cs
    from coherentsearch_c_amd import abi
    cache = abi.SearchCache()
    kw = dict(blocksize=65536, threshold=8.0, cache=cache)
    base = cs.search_file(spec, params, **kw)
    ms = cs.MetricStats()
    cs.search_file(spec, params, metricstats=ms, **kw)       # warm: metric buffer
    ms.close()
    out = {}
    for name, extra in (("plain", {}), ("metricstats", None), ("normalize", {"normalize": True})):
        ts = []
        for _ in range(reps):
            m = cs.MetricStats() if extra is None else None
            t = time.perf_counter()
            c = cs.search_file(spec, params, **(dict(metricstats=m) if m else extra), **kw)
            ts.append(time.perf_counter() - t)
            if m:
                v = m.view()
                m.close()
        out[name + "_ms"] = 1e3 * min(ts)
        out[name + "_candidates"] = int(c.size)
    out["blocks"] = len(v["blocks"])
    out["windows"] = len(v["whists"])
    out["metricstats_list_equals_plain"] = bool(
        base.size == out["metricstats_candidates"])
    abi.host().search_cache_free(C.byref(cache))
    log("metric statistics: plain %.1f ms, --metricstats %.1f ms, --normalize %.1f ms"
        % (out["plain_ms"], out["metricstats_ms"], out["normalize_ms"]))
    return out


def wide_harmonics(spec, steps=5, general=True, only=None):
    """The reference README's wider configurations on the same resident
    spectrum, timed like the headline step: nharms 120 with maxdecim 16 and
    nharms 128 (the register kernels' two-word filled mask, one wave per
    SIMD); with general=True also -n 240 --maxdecim 16 and --m 64 at
    nharms 60 (the general kernel, README.md:448, 688-695)."""
    import coherentsearch_c_amd as cs
    out = {}
    cases = [(120, 16, 16, steps), (128, 6, 16, steps)]
    if general:
        cases += [(240, 16, 16, 2), (60, 6, 64, 2)]
    for nh, md, m, nst in cases:
        key = "nharms%d_maxdecim%d" % (nh, md) + ("" if m == 16 else "_m%d" % m)
        if only and key not in only:
            continue
        p = cs.make_params(nharms=nh, m=m, maxdecim=md)
        sess = cs.Session(spec, p, blocksize=65536)
        ht, trials, _ = sess.work()
        sess.upload()
        sess.search_resident(8.0)
        kms = []
        t = time.perf_counter()
        for _ in range(nst):
            sess.search_resident(8.0)
            kms.append(sess.kernel_ms())
        wall = (time.perf_counter() - t) / nst
        sess.close()
        out[key] = {"trials": trials, "harmonic_trials": ht, "ms_per_search": 1e3 * wall,
                    "kernel_ms": float(np.mean(kms)), "ht_s": ht / wall}
        log("nharms %d maxdecim %d m %d: %.2f ms per search, %.4g ht/s"
            % (nh, md, m, 1e3 * wall, ht / wall))
    return out


def _filterbank_on_device(nspec, nchan, lo, cw, dt, dm, period, duty, snr, seed):
    """uint8 [nspec, nchan] spectra generated on the GPU in slabs: round(96 + 8
    N(0, 1) + pulse) clipped to [0, 255], the pulse a von Mises pulsar delayed
    per channel by the cold-plasma law (synth.filterbank_with_pulsar's model)"""
    import torch
    dev = "cuda:0"
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    out = torch.empty((nspec, nchan), dtype=torch.uint8, device=dev)
    f = lo + torch.arange(nchan, device=dev, dtype=torch.float64) * cw
    ftop = lo + (nchan - 1) * cw
    delay = dm / (0.000241 * f * f) - dm / (0.000241 * ftop * ftop)
    kappa = np.log(2.0) / (1.0 - np.cos(np.pi * duty))
    nb = 4096
    prof = np.exp(kappa * (np.cos(2 * np.pi * (np.arange(nb) + 0.5) / nb) - 1.0))
    prof -= prof.mean()
    prof /= np.sqrt(np.mean(prof ** 2))
    amp = 8.0 * snr / (np.sqrt(nspec) * np.sqrt(nchan))
    prof_d = torch.tensor(prof * amp, dtype=torch.float32, device=dev)
    phase0 = np.random.default_rng(seed).uniform()
    slab = 1 << 15
    for r0 in range(0, nspec, slab):
        n = min(slab, nspec - r0)
        t = torch.arange(r0, r0 + n, device=dev, dtype=torch.float64)[:, None] * dt
        ph = torch.remainder((t - delay[None, :]) / period + phase0, 1.0)
        x = torch.randn((n, nchan), generator=g, device=dev, dtype=torch.float32)
        x.mul_(8.0).add_(96.0).add_(prof_d[(ph * nb).long() % nb])
        out[r0:r0 + n] = x.round_().clamp_(0, 255).to(torch.uint8)
    return out


def dedisp_timing(steps=5, ref_flow=True):
    """The prep front end's de-dispersion (PRESTO prepsubband, in memory) of a
    1024-channel 8-bit filterbank into 128 DM series, and the search of those
    series: resident in HBM (this library) against the reference's flow
    (series to the host, then host realfft + rednoise and search_file per DM,
    src/prep_search.c:44-145)."""
    import torch
    import coherentsearch_c_amd as cs
    from coherentsearch_c_amd.engine import Engine, run_plain
    nch, lo, cw, dt, P = 1024, 1200.0, 300.0 / 1024, 64e-6, 2048
    nspec = 1 << 22
    nsub, lodm, dmstep, nd, ds = 64, 0.0, 0.5, 128, 4
    period, dm0 = 0.00731, 20.0
    t0 = time.perf_counter()
    raw = _filterbank_on_device(nspec, nch, lo, cw, dt, dm0, period, 0.03, 40.0, 21)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    d = cs.Dedisp(nch, lo, cw, dt, nspec, nsub, lodm, dmstep, nd, spectra_per_subint=P,
                  downsamp=ds, clip_sigma=0.0)
    info = d.info

    class Resident:
        ptr = raw.data_ptr()
        uint8 = True

    d.run(Resident())
    walls, dev = [], []
    for _ in range(steps):
        w = time.perf_counter()
        d.run(Resident())
        walls.append(time.perf_counter() - w)
        dev.append(d.last_ms())
    dev = np.array(dev)
    written, tsub = info["written"], info["written"] + info["max_offset"]
    raw_bytes = (tsub * ds + P) * nch                      # read once (upper bound of rows)
    sub_bytes = nsub * tsub * 4                            # written, then read once
    out_bytes = nd * written * 4
    sub_b, dm_b = raw_bytes + sub_bytes, sub_bytes + out_bytes
    ms_total, ms_sub, ms_dm = (float(np.mean(dev[:, k])) for k in range(3))
    res = {"workload": "%d channels x %d spectra (uint8, %.0f s at %g us), nsub %d, %d DMs "
                       "(%g-%g), downsamp %d; spectra resident in HBM"
                       % (nch, nspec, nspec * dt, dt * 1e6, nsub, nd, lodm,
                          lodm + (nd - 1) * dmstep, ds),
           "generate_s": gen_s, "written": written, "numout": info["numout"],
           "device_ms": ms_total, "subband_ms": ms_sub, "dm_ms": ms_dm,
           "wall_ms": 1e3 * float(np.mean(walls)),
           "spectra_per_s": nspec / (ms_total * 1e-3),
           "subband_gbs": sub_b / (ms_sub * 1e-3) / 1e9, "dm_gbs": dm_b / (ms_dm * 1e-3) / 1e9,
           "roofline": {"bound": "hbm", "achieved": (sub_b + dm_b) / (ms_total * 1e-3) / 1e9,
                        "peak": 8000.0, "unit": "GB/s",
                        "frac": (sub_b + dm_b) / (ms_total * 1e-3) / 8e12,
                        "bytes": sub_b + dm_b}}
    log("dedisp: %s" % res["workload"])
    log("  device %.3f ms (subband %.3f ms, %.0f GB/s; DM %.3f ms, %.0f GB/s), wall %.2f ms"
        % (ms_total, ms_sub, res["subband_gbs"], ms_dm, res["dm_gbs"], res["wall_ms"]))
    # prepsubband's default time-domain clipping (-clip 6) on the same spectra
    dc = cs.Dedisp(nch, lo, cw, dt, nspec, nsub, lodm, dmstep, nd, spectra_per_subint=P,
                   downsamp=ds, clip_sigma=6.0)
    dc.run(Resident())
    cdev, cw_ = [], []
    for _ in range(steps):
        w = time.perf_counter()
        dc.run(Resident())
        cw_.append(time.perf_counter() - w)
        cdev.append(dc.last_ms())
    cdev = np.array(cdev)
    res["clip6_device_ms"] = float(np.mean(cdev[:, 0]))
    res["clip6_clip_ms"] = float(np.mean(cdev[:, 0] - cdev[:, 1] - cdev[:, 2]))
    res["clip6_wall_ms"] = 1e3 * float(np.mean(cw_))
    log("  with clipping (-clip 6): device %.3f ms (clip statistics %.3f ms), wall %.2f ms"
        % (res["clip6_device_ms"], res["clip6_clip_ms"], res["clip6_wall_ms"]))
    dc.close()
    # from pinned host memory (the reference reads blocks on the host)
    host = raw.cpu().pin_memory()
    w = time.perf_counter()
    d.run(host.numpy())
    res["host_input_wall_ms"] = 1e3 * (time.perf_counter() - w)
    log("  from pinned host memory: %.1f ms" % res["host_input_wall_ms"])
    del host
    # the search of every DM series on the device, then the reference's flow
    p = cs.make_params(nharms=32)
    with Engine(device=0, nslots=4) as eng:
        w = time.perf_counter()
        d.run(Resident())
        finals = cs.prep_search(d, eng, p, ncands=100, threshold=6.0)
        res["prep_search_resident_s"] = time.perf_counter() - w
    dms, _, _ = d.tables()
    best = [max([m for fr, m in zip(f["freq"], f["metric"])
                 if abs(fr * period - round(fr * period)) < 2e-3 and round(fr * period) >= 1]
                or [0.0]) for f in finals]
    res["pulsar_dm_found"] = float(dms[int(np.argmax(best))])
    res["pulsar_dm_injected"] = dm0
    log("  de-dispersion + search of %d DMs, resident: %.3f s; pulsar at DM %g (injected %g)"
        % (nd, res["prep_search_resident_s"], res["pulsar_dm_found"], dm0))
    if ref_flow:
        w = time.perf_counter()
        series = d.copy_out()
        jobs = [{"samples": series[i], "dt": d.dt_out, "params": p, "threshold": 6.0,
                 "blocksize": 65536} for i in range(nd)]
        _, wall = run_plain(jobs)
        res["prep_search_reference_flow_s"] = time.perf_counter() - w
        res["ratio_vs_reference_flow"] = (res["prep_search_reference_flow_s"]
                                          / res["prep_search_resident_s"])
        log("  reference flow (series to the host, host staging + search_file per DM): "
            "%.3f s, %.1fx" % (res["prep_search_reference_flow_s"],
                              res["ratio_vs_reference_flow"]))
    d.close()
    del raw
    torch.cuda.empty_cache()
    return res


def stage_only(sizes, reps=10):
    """Device staging alone (realfft + rednoise of a resident series): each
    size staged `reps` times after two warm-up passes; prints one JSON line
    with the mean milliseconds per staging, copy of the series included."""
    import coherentsearch_c_amd as cs
    res = []
    for N in sizes:
        x = np.random.default_rng(1).standard_normal(N).astype(np.float32)
        g = cs.Stage(N, 64e-6)
        for _ in range(2):
            g.run_host(x)
        t = time.perf_counter()
        for _ in range(reps):
            g.run_host(x)
        ms = (time.perf_counter() - t) / reps * 1e3
        g.close()
        log("stage N=%d: %.3f ms per staging (host copies included)" % (N, ms))
        res.append({"samples": N, "ms_per_staging": ms})
    print(json.dumps({"stage_only": res}), flush=True)


def calibrate_copy(nbytes=256 << 20):
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copy calibration: 3 launches of %d bytes read + %d bytes written" % (nbytes, nbytes))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--samples", type=int, default=1 << 24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-sweep", action="store_true", help="skip the 64-file DM sweep")
    ap.add_argument("--no-serve", action="store_true", help="skip the served request stream")
    ap.add_argument("--sweep-files", type=int, default=64)
    ap.add_argument("--serve-requests", type=int, default=500)
    ap.add_argument("--slots", type=int, default=4, help="engine worker streams")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (default launch shapes)")
    ap.add_argument("--sweep-only", action="store_true",
                    help="only run the 64-file DM sweep legs (prints one JSON line)")
    ap.add_argument("--wide-only", action="store_true",
                    help="only time the nharms 120 / maxdecim 16 and nharms 128 / maxdecim 6 "
                         "searches of the 2^24-sample spectrum (prints one JSON line)")
    ap.add_argument("--wide-case", action="append", default=None, metavar="KEY",
                    help="with --wide-only, time only this case (e.g. nharms240_maxdecim16); "
                         "repeatable")
    ap.add_argument("--no-general", action="store_true",
                    help="with --wide-only: only the register-kernel cases (nharms 120/128)")
    ap.add_argument("--c2-only", action="store_true",
                    help="only the headline 2^24-sample search steps and the search_file "
                         "pass, after tuning the nharms 60 class (no other legs or CPU "
                         "baseline; for profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three 256 MiB device-to-device copies (a plain copy "
                         "kernel of known size, to scale FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--no-dedisp", action="store_true",
                    help="skip the filterbank de-dispersion and its DM search")
    ap.add_argument("--dedisp-only", action="store_true",
                    help="only time the de-dispersion (prints one JSON line)")
    ap.add_argument("--stage-only", type=int, nargs="*", default=None, metavar="SAMPLES",
                    help="only time device staging (realfft + rednoise) of series of these "
                         "lengths (default 2^24 and 3*5*7*2^17)")
    args = ap.parse_args()
    if args.stage_only is not None:
        stage_only(args.stage_only or [1 << 24, 3 * 5 * 7 * (1 << 17)])
        return
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.dedisp_only:
        print(json.dumps({"dedisp": dedisp_timing(steps=max(1, args.steps),
                                                  ref_flow=False)}), flush=True)
        return
    if args.sweep_only:
        print(json.dumps({"dm_sweep": dm_sweep(args.sweep_files, args.slots)}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the search shards by request or by "
              "trial range across devices (see the design notes), --gpus %d is not run"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import coherentsearch_c_amd as cs
    from coherentsearch_c_amd import synth

    rc, name = cs.gpu_available(0)
    if rc != 1:
        log("no GPU:", name)
        sys.exit(1)
    log("device:", name)
    t = time.perf_counter()
    x, dt = synth.config2(N=args.samples)
    spec = cs.redfft(x, dt)
    del x
    log("staging (synth + realfft + rednoise): %.2f s" % (time.perf_counter() - t))
    if args.wide_only:
        print(json.dumps({"wide_harmonics": wide_harmonics(spec, steps=max(1, args.steps),
                                                           general=not args.no_general,
                                                           only=args.wide_case)}),
              flush=True)
        return
    pkw = dict(nharms=60, m=16, maxdecim=6)
    params = cs.make_params(**pkw)
    tune = []
    tune_classes = (60, 32, 16)
    if args.c2_only:
        # the headline class's tuned launch shape only
        tune_classes = (60,)
        args.no_dedisp = args.no_sweep = args.no_serve = True
        args.no_cpu_baseline = True
    if not args.no_tune:
        # launch shapes of every kernel class the bench runs (the sweep and the
        # served stream mix nharms 16/32/60); later sessions use them
        t = time.perf_counter()
        for nh in tune_classes:
            r = cs.autotune(cs.make_params(nharms=nh, m=16, maxdecim=6))
            tune.append({k: r[k] for k in ("nharms", "lane_order", "xcd_stripe", "blocksize",
                                           "ntried", "default_ms", "best_ms")})
        log("autotune: %s (%.2f s)" % (tune, time.perf_counter() - t))
    t = time.perf_counter()
    sess = cs.Session(spec, params, blocksize=65536)
    ht, trials, alg_bytes = sess.work()
    log("session: %d trials, %d chunks, %.4g harmonic-trials, create %.2f s"
        % (trials, sess.nchunks, ht, time.perf_counter() - t))
    t = time.perf_counter()
    sess.upload()
    up_s = time.perf_counter() - t

    for _ in range(args.warmup):
        c = sess.search_resident(8.0)
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        c = sess.search_resident(8.0)
        kms.append(sess.kernel_ms())
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    value = ht * args.steps / wall
    kavg = float(np.mean(kms))
    achieved = alg_bytes / (kavg * 1e-3) / 1e9
    top = cs.finalize(spec, params, cs.search_file(spec, params, blocksize=65536,
                                                   threshold=8.0))
    f0 = 1.0 / 5.3e-3
    found = bool(top.size and any(abs(fr / f0 - round(fr / f0)) < 2e-3 and round(fr / f0) >= 1
                                  for fr in top["freq"][:3]))
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_search_c2.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")
    out = {
        "metric": "harmonic-trials/s at 1 GPU vs plain form; p50/p99 search latency; "
                  "HBM GB/s frac",
        "value": value,
        "unit": "harmonic-trials/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian noise + injected von Mises pulsar)",
        "config": {"workload": "single 2^%d-sample series, dt 64 us, realfft+rednoise; "
                               "nharms 60, m 16, maxdecim 6, blocksize 65536, hifreq 125 Hz, "
                               "threshold 8" % int(np.log2(args.samples)),
                   "trials": trials, "harmonic_trials": ht, "chunks": int(sess.nchunks)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "kernel_ms": kavg,
        "candidates_per_step": int(c.size),
        "pulsar_in_top3": found,
        "upload_inclusive_ht_s": ht / (wall / args.steps + up_s),
        "upload_s": up_s,
    }
    if tune:
        out["autotune"] = {"spectrum": "2^22-sample noise, 0.1-125 Hz, m 16, f32, analytic "
                                       "sigma; lane order 0 residue classes / 1 natural, "
                                       "xcd_stripe 0 contiguous / S tiles; blocksize is "
                                       "advice (requests keep 65536)",
                           "classes": tune}
    if not args.c2_only:
        out["metric_stats"] = metric_stats_timing(spec, params)
        out["wide_harmonics"] = wide_harmonics(spec)
    if not args.no_dedisp:
        out["dedisp"] = dedisp_timing()
    if not args.no_sweep:
        out["dm_sweep"] = dm_sweep(args.sweep_files, args.slots)
        out["dm_sweep_one_observation"] = dm_sweep_resident(args.sweep_files, args.slots)
    if not args.no_serve:
        out["serving"] = serving(args.serve_requests, args.slots)
    for key, sub, ratio in (("dm_sweep", "ratio_vs_plain_device", "batched_vs_plain_device"),
                            ("dm_sweep", "ratio_vs_plain_reference_flow",
                             "batched_vs_plain_reference_list_mode"),
                            ("dm_sweep_one_observation", "ratio_vs_plain_resident",
                             "batched_vs_plain_resident_one_observation")):
        if key in out:
            out[ratio] = out[key][sub]
    if "serving" in out:
        out["serving_p50_ms"] = out["serving"]["p50_latency_ms"]
        out["serving_p99_ms"] = out["serving"]["p99_latency_ms"]
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(spec, pkw, args.cpu_budget)
            one = cpu_baseline(spec, pkw, args.cpu_budget / 2, threads=1)
            out["cpu_baseline"]["value_1_thread"] = one["value"]
            out["cpu_baseline"]["sample_1_thread"] = one["sample"]
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"error": repr(e)}
    sess.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
