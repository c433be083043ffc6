"""GPU tests of the continuous-batching engine through the C-ABI.

- Each request's raw list is bit-identical to the oracle's search of the
  spectrum the engine staged (the same bar as the single-session tests).
- Results are identical alone and in the mix: the same requests run one at a
  time on a one-slot engine and all at once on a four-slot engine give the
  same lists bit for bit.
- Admission respects the HBM budget: with room for one request at a time,
  never more than one runs; a request larger than the budget still runs
  (alone).
- The served loop finishes every request and reports its latency.
"""
import sys

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import engine as eng
from coherentsearch_c_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _jobs(n=6, seed=11):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        N = synth.smooth7(int(2 ** rng.uniform(18, 20)))
        nh = (16, 32, 60)[i % 3]
        per = 3.7e-3 if i % 2 == 0 else 0.0
        x = synth.series_with_pulsar(N, per if per else 1.0, 0.04, 14.0 if per else 0.0,
                                     seed * 100 + i)
        jobs.append(dict(samples=x, dt=synth.DT, params=cs.make_params(nharms=nh, m=16,
                                                                        maxdecim=6),
                         threshold=6.0, blocksize=8192))
    return jobs


def _same(a, b):
    """field by field (the C struct's padding bytes are not part of a result)"""
    if a.size != b.size:
        return False
    return all(np.array_equal(a[f].view(np.uint64 if a[f].dtype.itemsize == 8 else np.uint32),
                              b[f].view(np.uint64 if b[f].dtype.itemsize == 8 else np.uint32))
               for f in ("freq", "metric", "r", "nharm", "ducy"))


def test_engine_requests_match_oracle(oracle):
    jobs = _jobs(4)
    with eng.Engine(nslots=4) as E:
        tk = [E.submit(j["samples"], j["dt"], j["params"], threshold=j["threshold"],
                       blocksize=j["blocksize"]) for j in jobs]
        for t, j in zip(tk, jobs):
            raw, info, spec = E.wait(t)
            nh = j["params"].nharms
            say("request", t, "N", spec.N, "nh", nh, "cands", raw.size, "ht",
                info["harmonic_trials"], "kernel_ms %.3f" % info["kernel_ms"])
            assert info["status"] == 0
            o = oracle.Oracle(spec.amps, spec.N, spec.T,
                              oracle.config(nharms=nh, m=16, blocksize=j["blocksize"]))
            want = o.search(j["threshold"], 1)
            assert raw.size == want.size
            for f in ("r", "nharm", "metric"):
                assert np.array_equal(raw[f], want[f]), f
            o.close()


def test_alone_and_in_the_mix_identical():
    jobs = _jobs(6, seed=12)
    alone = []
    with eng.Engine(nslots=1) as E:
        for j in jobs:
            t = E.submit(j["samples"], j["dt"], j["params"], threshold=j["threshold"],
                         blocksize=j["blocksize"])
            alone.append(E.wait(t)[0])
    with eng.Engine(nslots=4) as E:
        tk = [E.submit(j["samples"], j["dt"], j["params"], threshold=j["threshold"],
                       blocksize=j["blocksize"]) for j in jobs]
        mix = [E.wait(t)[0] for t in tk]
        st = E.stats()
    say("mix stats", st)
    assert st["peak_running"] >= 2
    for a, b in zip(alone, mix):
        assert _same(a, b)


def test_budget_admission():
    jobs = _jobs(4, seed=13)
    # a budget below every request's footprint: each one runs, alone
    with eng.Engine(nslots=4, hbm_budget=1) as E:
        tk = [E.submit(j["samples"], j["dt"], j["params"], threshold=j["threshold"],
                       blocksize=j["blocksize"]) for j in jobs]
        for t in tk:
            assert E.wait(t)[1]["status"] == 0
        st = E.stats()
    say("budget stats", st)
    assert st["peak_running"] == 1 and st["completed"] == len(jobs)


def test_serve_small_stream():
    jobs = _jobs(5, seed=14)
    arrivals = np.array([0.0, 0.01, 0.02, 0.2, 0.21])
    with eng.Engine(nslots=3) as E:
        finals, recs, wall = eng.serve(E, jobs, arrivals, post_workers=4)
    lat = [r["latency_s"] for r in recs]
    say("latencies", ["%.3f" % v for v in lat], "wall %.2f" % wall)
    assert all(f is not None for f in finals) and all(v > 0 for v in lat)


def _final_vs_host(spec, p, final, raw, ncands):
    """The device-finalised list against the host's finalize of the same
    collapsed list and spectrum (sort by S/N, truncate, measure_ducy), field by
    field; duty cycles are w/n of the chosen boxcar width, so they compare
    exactly."""
    want = cs.finalize(spec, p, cs.collapse(raw), ncands)
    assert final.size == want.size
    for f in ("r", "nharm", "metric", "freq"):
        assert np.array_equal(final[f], want[f]), f
    assert np.array_equal(final["ducy"], want["ducy"], equal_nan=True)
    return want


@pytest.mark.parametrize("nh,r0,nbins,hifreq", [(60, 271.3, 1 << 15, 125.0),
                                                 (16, 3333.7, 1 << 14, 2000.0),
                                                 (64, 1021.9, 1 << 15, 300.0),
                                                 (8, 16340.2, 1 << 14, 7800.0),
                                                 (120, 211.7, 1 << 15, 150.0),
                                                 (100, 431.3, 1 << 15, 150.0)])
def test_session_finalize_duty_cycles(oracle, nh, r0, nbins, hifreq):
    """cs_gpu_session_finalize: the S/N order and cut of the host's finalize
    and its duty cycles, measured on the device from the resident spectrum.
    In the middle cases only the first harmonics' 64-bin windows fit below
    the top of the spectrum (profiles of fewer harmonics than reported); in
    the last the pulsar sits at the top of the band, where no window fits."""
    amps = synth.spectrum_with_pulsar(nbins, r0, nh, 0.05, 14.0, 3)
    spec = cs.Spectrum(amps, 2 * nbins, synth.DT)
    p = cs.make_params(nharms=nh, m=16, maxdecim=6)
    sess = cs.Session(spec, p, blocksize=4096, hifreq=hifreq)
    raw = sess.run(4.0)
    col = cs.collapse(raw)
    for ncands in (100, 7):
        final = sess.finalize(col, ncands)
        say("finalize nh", nh, "r0", r0, "collapsed", col.size, "final", final.size,
            "ducy", final["ducy"][:4])
        _final_vs_host(spec, p, final, raw, ncands)
        # the oracle's duty cycle (direct transform of the same harmonics)
        od = oracle.measure_ducy(amps.view(np.float32), spec.N, final.astype(oracle.CAND_DTYPE))
        assert np.array_equal(od["ducy"], final["ducy"], equal_nan=True)
    assert sess.finalize(col[:0], 100).size == 0
    sess.close()


def test_engine_final_lists_match_host_post():
    """Requests that ask for the final list (ncands > 0) return the host's
    collapse + finalize of the same request's raw list and spectrum."""
    jobs = _jobs(4, seed=15)
    with eng.Engine(nslots=4) as E:
        raws = [E.submit(j["samples"], j["dt"], j["params"], threshold=j["threshold"],
                         blocksize=j["blocksize"]) for j in jobs]
        fins = [E.submit(j["samples"], j["dt"], j["params"], threshold=j["threshold"],
                         blocksize=j["blocksize"], ncands=50) for j in jobs]
        for tr, tf, j in zip(raws, fins, jobs):
            raw, info_r, spec = E.wait(tr)
            final, info_f, none = E.wait(tf)
            assert info_r["status"] == 0 and info_f["status"] == 0 and none is None
            say("final list", final.size, "post_ms %.3f" % info_f["post_ms"])
            _final_vs_host(spec, j["params"], final, raw, 50)


def test_engine_final_lists_wide_harmonics():
    """Requests of more than 64 harmonics that ask for the final list: duty
    cycles measured on the device over every fitting harmonic, equal to the
    host's collapse + finalize of the request's raw list."""
    jobs = []
    for i, nh in enumerate((120, 96)):
        x = synth.series_with_pulsar(1 << 19, 9.1e-3, 0.05, 15.0, 70 + i)
        jobs.append(dict(samples=x, dt=synth.DT, threshold=4.5, blocksize=16384,
                         params=cs.make_params(nharms=nh, m=16, maxdecim=6)))
    with eng.Engine(nslots=2) as E:
        for j in jobs:
            raw, info_r, spec = E.wait(E.submit(j["samples"], j["dt"], j["params"],
                                                threshold=j["threshold"],
                                                blocksize=j["blocksize"]))
            final, info_f, _ = E.wait(E.submit(j["samples"], j["dt"], j["params"],
                                               threshold=j["threshold"],
                                               blocksize=j["blocksize"], ncands=5000))
            assert info_r["status"] == 0 and info_f["status"] == 0
            say("nh", j["params"].nharms, "final list", final.size,
                "nharm", final["nharm"][:6], "ducy", final["ducy"][:6])
            assert final.size > 0 and np.any(final["nharm"] > 64)
            _final_vs_host(spec, j["params"], final, raw, 5000)


def _dm_trials(n, N=1 << 19, seed=31):
    """n series of one length (DM trials of one observation), nharms 16/32."""
    jobs = []
    for i in range(n):
        x = synth.series_with_pulsar(N, 3.3e-3, 0.05, 10.0 + i, seed + i)
        jobs.append(dict(samples=x, dt=synth.DT, threshold=6.0, blocksize=16384,
                         params=cs.make_params(nharms=(16, 32)[i % 2], m=16, maxdecim=6)))
    return jobs


def test_resident_plans_reused_with_graphs(oracle):
    """Requests of one geometry reuse the resident plan (no planning or
    uploads; staging and search replayed from graphs, zaplist and spectrum
    copies in between) and return the oracle's list of each staged spectrum,
    bit for bit, as an engine without resident plans does."""
    jobs = _dm_trials(6)
    zap = [(61.0, 0.5, 0)]
    res = {}
    for cache in (8, 0):
        with eng.Engine(nslots=1, plan_cache=cache) as E:
            out = []
            for i, j in enumerate(jobs):
                t = E.submit(j["samples"], j["dt"], j["params"], threshold=j["threshold"],
                             blocksize=j["blocksize"], zaplist=zap if i == 4 else None)
                out.append(E.wait(t))
            st = E.stats()
        say("plan cache", cache, "stats", st, "reused", [o[1]["reused"] for o in out])
        res[cache] = out
        if cache:
            assert [o[1]["reused"] for o in out] == [0, 0, 1, 1, 1, 1]
            assert st["plan_hits"] == 4 and st["plans_resident"] == 2
        else:
            assert st["plan_hits"] == 0 and st["plans_resident"] == 0
    for i, ((a, ia, sa), (b, ib, sb)) in enumerate(zip(res[8], res[0])):
        assert ia["status"] == 0 and ib["status"] == 0
        assert np.array_equal(sa.amps, sb.amps) and _same(a, b)
        p = jobs[i]["params"]
        o = oracle.Oracle(sa.amps, sa.N, sa.T,
                          oracle.config(nharms=p.nharms, m=16, blocksize=16384))
        want = o.search(6.0, 1)
        assert a.size == want.size and np.array_equal(a["metric"], want["metric"])
        o.close()


def test_resident_plans_in_the_mix():
    """Eight requests of two geometries on four slots: several plans of one
    geometry may exist at once; every list equals the one-slot, no-cache run.
    Then the same requests one at a time: after the first request of each
    geometry an idle plan of it is always resident (the cache holds four, and
    only a miss creates a plan or evicts one), so at least the last six reuse
    a plan, whatever the timing of the concurrent round."""
    jobs = _dm_trials(8, seed=41)
    with eng.Engine(nslots=1, plan_cache=0) as E:
        alone = [E.wait(_submit(E, j))[0] for j in jobs]
    with eng.Engine(nslots=4, plan_cache=4) as E:
        tk = [_submit(E, j) for j in jobs]
        mix = [E.wait(t)[0] for t in tk]
        for a, b in zip(alone, mix):
            assert _same(a, b)
        st1 = E.stats()
        serial = []
        for j in jobs:
            serial.append(E.wait(_submit(E, j)))
        st = E.stats()
    say("mix with plans:", st1, "then serial:", st, "reused", [o[1]["reused"] for o in serial])
    for a, (b, info, _) in zip(alone, serial):
        assert _same(a, b)
    assert all(o[1]["reused"] == 1 for o in serial[2:])
    assert st["plan_hits"] - st1["plan_hits"] >= 6 and st["plans_resident"] <= 4


def _submit(E, j):
    return E.submit(j["samples"], j["dt"], j["params"], threshold=j["threshold"],
                    blocksize=j["blocksize"])
