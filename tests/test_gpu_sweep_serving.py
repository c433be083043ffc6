"""The batched DM sweep and the served stream at their real sizes, checked
against the oracle (GPU, through the C-ABI engine).

- Sweep: seven of the 64-file sweep's own series (lengths 2^22.1 to 2^24.95,
  the longest included; nharms 16, 32 and 60; three carry the pulsar) pushed
  through the four-slot engine at once.  Each request's raw candidate list is
  the oracle's search of the spectrum the engine staged, bit for bit (same
  Fourier bins, harmonic counts, metrics and order).  Threshold 6 instead of
  the sweep's 8, for more candidates per list (the per-trial metrics do not
  depend on it).
- Serving: the first 40 requests of the 500-request Poisson stream (20 per
  second; lengths up to 2^25; nharms 16/32/60; pulsar requests included)
  submitted at their arrival times to a four-slot engine while earlier ones
  run.  Every raw list is the oracle's, bit for bit; every list equals the
  same request run alone on a one-slot engine; the served final lists (device
  collapse, S/N cut, duty cycles) equal the host's finish of each raw list.
"""
import concurrent.futures as cf
import sys
import threading
import time

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import engine as eng
from coherentsearch_c_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits(a, f):
    return a[f].view(np.uint64 if a[f].dtype.itemsize == 8 else np.uint32)


def _same(a, b):
    return a.size == b.size and all(np.array_equal(_bits(a, f), _bits(b, f))
                                    for f in ("freq", "metric", "r", "nharm", "ducy"))


def _check_oracle(oracle, raw, spec, nh, threshold):
    """the oracle's list in chunk order (one search per chunk, chunks on
    parallel host threads) against the raw list"""
    o = oracle.Oracle(spec.amps, spec.N, spec.T,
                      oracle.config(nharms=nh, m=16, maxdecim=6, blocksize=65536))
    with cf.ThreadPoolExecutor(16) as ex:
        parts = list(ex.map(lambda c: o.search(threshold, 1, c, c + 1), range(o.nchunks)))
    want = np.concatenate(parts) if parts else o.search(threshold, 1)
    o.close()
    assert raw.size == want.size, (raw.size, want.size)
    for f in ("r", "nharm", "metric"):
        assert np.array_equal(raw[f], want[f]), f
    return want.size


def _job(m):
    x = synth.series_with_pulsar(m["N"], m["period"] if m["snr"] > 0 else 1.0, m["duty"],
                                 m["snr"], m["seed"])
    return dict(samples=x, dt=synth.DT, threshold=6.0, blocksize=65536,
                params=cs.make_params(nharms=m["nharms"], m=16, maxdecim=6))


def _submit(E, j, **kw):
    return E.submit(j["samples"], j["dt"], j["params"], threshold=j["threshold"],
                    blocksize=j["blocksize"], **kw)


def test_dm_sweep_files_match_oracle(oracle):
    meta = synth.dm_sweep(64, 4)
    pick = [51, 59, 50, 14, 32, 33, 31]
    jobs = [_job(meta[i]) for i in pick]
    say("sweep files", [(i, meta[i]["N"], meta[i]["nharms"], round(meta[i]["snr"], 1))
                        for i in pick])
    with eng.Engine(nslots=4) as E:
        tk = [_submit(E, j) for j in jobs]
        out = [E.wait(t) for t in tk]
        st = E.stats()
    say("engine", st)
    assert st["peak_running"] >= 2
    for i, j, (raw, info, spec) in zip(pick, jobs, out):
        assert info["status"] == 0
        n = _check_oracle(oracle, raw, spec, j["params"].nharms, j["threshold"])
        say("file", i, "N", spec.N, "nh", j["params"].nharms, "candidates", n,
            "kernel_ms %.2f" % info["kernel_ms"])


def test_served_stream_matches_oracle_alone_and_in_the_mix(oracle):
    meta = synth.serving_requests(500, 20.0, 5)[:40]
    jobs = [_job(m) for m in meta]
    arrivals = np.array([m["arrival"] for m in meta])
    say("stream: %d requests over %.2f s, longest %d samples"
        % (len(jobs), arrivals[-1], max(m["N"] for m in meta)))
    assert max(m["N"] for m in meta) == 1 << 25
    # the mix: each request submitted at its arrival time while others run
    mix = [None] * len(jobs)
    with eng.Engine(nslots=4) as E:
        tickets = {}
        lock = threading.Lock()
        start = time.perf_counter()

        def feeder():
            for i, j in enumerate(jobs):
                d = start + arrivals[i] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                t = _submit(E, j)
                with lock:
                    tickets[t] = i

        th = threading.Thread(target=feeder)
        th.start()
        for _ in range(len(jobs)):
            t = E.next_done()
            with lock:
                i = tickets.get(t)
            while i is None:          # the feeder has not recorded the ticket yet
                time.sleep(0.001)
                with lock:
                    i = tickets.get(t)
            mix[i] = E.wait(t)
        th.join()
        st = E.stats()
    say("mix", st)
    assert st["peak_running"] >= 2
    # alone: one request at a time
    with eng.Engine(nslots=1) as E:
        alone = [E.wait(_submit(E, j))[0] for j in jobs]
    total = 0
    for i, (j, (raw, info, spec), a) in enumerate(zip(jobs, mix, alone)):
        assert info["status"] == 0
        assert _same(raw, a), i
        total += _check_oracle(oracle, raw, spec, j["params"].nharms, j["threshold"])
    say("raw lists match the oracle and the one-slot run; %d candidates" % total)
    # the served finish (device collapse, S/N cut, duty cycles) of the same stream
    with eng.Engine(nslots=4) as E:
        finals, recs, wall = eng.serve(E, jobs, arrivals, post_workers=4, ncands=100)
    lat = np.array([r["latency_s"] for r in recs]) * 1e3
    say("served: p50 %.1f ms, p99 %.1f ms over %.2f s" % (np.percentile(lat, 50),
                                                         np.percentile(lat, 99), wall))
    for i, (j, f, (raw, _, spec)) in enumerate(zip(jobs, finals, mix)):
        want = cs.finalize(spec, j["params"], cs.collapse(raw), 100)
        assert f.size == want.size, i
        for fld in ("r", "nharm", "metric", "freq"):
            assert np.array_equal(f[fld], want[fld]), (i, fld)
        assert np.array_equal(f["ducy"], want["ducy"], equal_nan=True), i
