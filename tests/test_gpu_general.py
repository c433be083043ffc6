"""GPU parity of the general search kernel (the configurations the register
kernels cannot hold) through the C-ABI, against the CPU oracle.

Each case goes through cs_gpu_session_create like any other: harmonic counts
up to 256 (the reference README's -n 240 --maxdecim 16), prime counts whose
c2r is one large odd butterfly, m up to 64, double precision above 64
harmonics, and step sizes (hidr 0.75, 1.0) whose tile windows exceed the
register kernels' LDS buffer.  The interpolated stack is bit-identical to the
oracle's, and so is every per-trial metric (threshold -inf: every trial of
every decimation slot).  Requests of such sizes through the engine return the
oracle's raw list and, with ncands, the host's finish with duty cycles.
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


def _case(oracle, nh, m, maxdecim, precision, hidr, sigma="analytic", nbins=1 << 14,
          blocksize=4096, r0=41.7, hifreq=125.0):
    amps = synth.spectrum_with_pulsar(nbins, r0, min(nh, 60), 0.05, 12.0, nh + m)
    spec = cs.Spectrum(amps, 2 * nbins, synth.DT)
    p = cs.make_params(nharms=nh, m=m, hidr=hidr, maxdecim=maxdecim, precision=precision,
                       sigma=sigma)
    o = oracle.Oracle(amps.view(np.float32), spec.N, spec.T,
                      oracle.config(nharms=nh, m=m, hidr=hidr, blocksize=blocksize,
                                    maxdecim=maxdecim, hifreq=hifreq, precision=precision,
                                    sigma=sigma))
    sess = cs.Session(spec, p, blocksize=blocksize, hifreq=hifreq)
    assert sess.total == o.total and np.array_equal(sess.cstarts, o.chunk_starts())
    return spec, p, o, sess


CASES = [
    (240, 16, 16, "f32", 0.5, "analytic"),     # README.md:688-695, -n 240 --maxdecim 16
    (150, 16, 6, "f32", 0.5, "analytic"),
    (256, 16, 8, "f32", 0.5, "analytic"),
    (251, 16, 6, "f32", 0.5, "analytic"),      # a prime: one 251-point butterfly
    (127, 16, 6, "f64", 0.5, "analytic"),      # prime, double precision
    (100, 16, 6, "f64", 0.5, "analytic"),
    (100, 32, 6, "f32", 0.5, "analytic"),
    (16, 64, 6, "f32", 0.5, "analytic"),       # README.md:448, --m 64
    (60, 64, 6, "f64", 0.5, "analytic"),
    (60, 16, 6, "f32", 1.0, "analytic"),       # hidr 1.0: windows beyond the LDS tile
    (32, 16, 6, "f32", 0.75, "analytic"),
    (200, 16, 6, "f32", 0.5, "measured"),
    (90, 24, 12, "f64", 0.5, "measured"),
]


@pytest.mark.parametrize("nh,m,maxdecim,precision,hidr,sigma", CASES)
def test_general_kernel_bit_identical(oracle, nh, m, maxdecim, precision, hidr, sigma):
    spec, p, o, sess = _case(oracle, nh, m, maxdecim, precision, hidr, sigma)
    say("general nh", nh, "m", m, "maxdecim", maxdecim, precision, "hidr", hidr, sigma,
        "trials", o.total, "chunks", o.nchunks)
    for ci in sorted({0, o.nchunks // 2, o.nchunks - 1}):
        i0, n, _ = o.chunk(ci)
        want, _ = o.stack(ci)
        got = sess.dump_stack(i0, n)
        assert np.array_equal(got.transpose(1, 0, 2).view(np.uint32), want.view(np.uint32)), ci
    got = sess.run(-np.inf)
    want = o.search(-np.inf, 1)
    say("  metrics", got.size)
    assert got.size == want.size
    for f in ("r", "nharm", "metric"):
        assert np.array_equal(got[f], want[f]), f
    # the candidate list at a threshold, and the resident re-search
    assert np.array_equal(sess.run(6.0)["metric"], want["metric"][want["metric"] > 6.0])
    sess.close()
    o.close()


def test_general_engine_requests(oracle):
    """-n 240 and --m 64 requests through the engine: raw lists equal the
    oracle's search of the staged spectrum; final lists (ncands) equal the
    host's collapse + finish with duty cycles over all fitting harmonics."""
    jobs = []
    for i, (nh, m) in enumerate(((240, 16), (24, 64), (200, 16))):
        x = synth.series_with_pulsar(1 << 18, 7.7e-3, 0.05, 20.0, 90 + i)
        jobs.append(dict(samples=x, dt=synth.DT, threshold=5.0, blocksize=16384,
                         params=cs.make_params(nharms=nh, m=m, maxdecim=16 if nh == 240 else 6)))
    with eng.Engine(nslots=3) as E:
        raws = [E.submit(j["samples"], j["dt"], j["params"], threshold=j["threshold"],
                         blocksize=j["blocksize"]) for j in jobs]
        fins = [E.submit(j["samples"], j["dt"], j["params"], threshold=j["threshold"],
                         blocksize=j["blocksize"], ncands=200) for j in jobs]
        for tr, tf, j in zip(raws, fins, jobs):
            raw, ir, spec = E.wait(tr)
            final, iff, _ = E.wait(tf)
            p = j["params"]
            assert ir["status"] == 0 and iff["status"] == 0
            o = oracle.Oracle(spec.amps, spec.N, spec.T,
                              oracle.config(nharms=p.nharms, m=p.m, blocksize=16384,
                                            maxdecim=16 if p.nharms == 240 else 6))
            want = o.search(j["threshold"], 1)
            o.close()
            say("request nh", p.nharms, "m", p.m, "raw", raw.size, "final", final.size,
                "kernel_ms %.2f" % ir["kernel_ms"])
            assert raw.size == want.size
            for f in ("r", "nharm", "metric"):
                assert np.array_equal(raw[f], want[f]), f
            host = cs.finalize(spec, p, cs.collapse(raw), 200)
            assert final.size == host.size
            for f in ("r", "nharm", "metric", "freq"):
                assert np.array_equal(final[f], host[f]), f
            assert np.array_equal(final["ducy"], host["ducy"], equal_nan=True)
