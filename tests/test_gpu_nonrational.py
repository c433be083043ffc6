"""GPU search at a trial step that is no simple rational (--hidr such that
hidr/nharms has no small p/q: every plan has P = 0), through the C-ABI,
against the oracle's restatement of the reference's per-trial path
(src/interp.c:427-475, fill_row_slow).

The reference's GPU build refuses such a step and falls back to its CPU
kernel (src/search.c:1374-1383); here the stack pass computes each trial's
FP64 position, weights and rotation on the device and the usual profile
passes follow.

Bar: the device's double sin and cos (the rotation's sinpi/cospi of the
fractional offset) are not glibc's, and may differ from them in the last
bit of a double.  Rounded to the float stack such a difference shows only
when the double lies within an ulp of a float rounding boundary, so the stack
is compared exactly where it agrees (nearly everywhere) and within one float
ulp elsewhere; every metric within 1e-5 relative (the north-star tolerance is
1e-4), and the candidate lists the same away from the threshold.
"""
import math
import sys

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    (16, 16, 0.4999, "analytic", 6),           # general kernel's c2r after the stack pass
    (7, 8, 1.0 / math.sqrt(3.0), "analytic", 3),
    (60, 16, 0.5 * math.pi / 3.0, "measured", 6),
    (100, 16, 0.4999, "analytic", 12),         # register profile pass (65-128 harmonics)
    (32, 48, 0.4999, "analytic", 6),
]


@pytest.mark.parametrize("nh,m,hidr,sigma,maxdecim", CASES)
def test_nonrational_step_matches_oracle(oracle, nh, m, hidr, sigma, maxdecim):
    nbins = 1 << 14
    amps = synth.spectrum_with_pulsar(nbins, 61.7, min(nh, 60), 0.05, 12.0, 5)
    spec = cs.Spectrum(amps, 2 * nbins, synth.DT)
    p = cs.make_params(nharms=nh, m=m, hidr=hidr, maxdecim=maxdecim, sigma=sigma)
    o = oracle.Oracle(amps.view(np.float32), spec.N, spec.T,
                      oracle.config(nharms=nh, m=m, hidr=hidr, blocksize=2048,
                                    maxdecim=maxdecim, sigma=sigma))
    sess = cs.Session(spec, p, blocksize=2048)
    assert sess.total == o.total and np.array_equal(sess.cstarts, o.chunk_starts())
    say("non-rational nh", nh, "m", m, "hidr", hidr, sigma, "trials", o.total,
        "chunks", o.nchunks)
    same = total = 0
    for ci in sorted({0, o.nchunks // 2, o.nchunks - 1}):
        i0, n, _ = o.chunk(ci)
        want, _ = o.stack(ci)
        got = sess.dump_stack(i0, n).transpose(1, 0, 2)
        ulp = np.spacing(np.maximum(np.abs(want), np.float32(1e-30)))
        assert np.all(np.abs(got - want) <= ulp), ci
        same += int(np.sum(got == want))
        total += got.size
    say("  stack values identical: %d of %d" % (same, total))
    assert same >= 0.999 * total
    got = sess.run(-np.inf)
    want = o.search(-np.inf, 1)
    assert got.size == want.size
    for f in ("r", "nharm"):
        assert np.array_equal(got[f], want[f]), f
    d = np.abs(got["metric"] - want["metric"])
    say("  metrics identical: %d of %d, max |d| %.3g" % (int(np.sum(d == 0)), d.size, d.max()))
    assert np.all(d <= 1e-5 * np.maximum(np.abs(want["metric"]), 1.0))
    thr = 6.0
    a = sess.run(thr)
    b = want[want["metric"] > thr]
    away = lambda c: c[np.abs(c["metric"] - thr) > 1e-4]
    assert np.array_equal(away(a)["r"], away(b)["r"])
    sess.close()
    o.close()


def test_nonrational_f64_refused():
    """The reference keeps a double stack for a non-rational step in double
    precision (src/interp.c:464-467); the GPU search refuses that form with a
    message rather than run it differently."""
    amps = synth.spectrum_with_pulsar(1 << 13, 41.7, 8, 0.05, 8.0, 1)
    spec = cs.Spectrum(amps, 1 << 14, synth.DT)
    with pytest.raises(RuntimeError):
        cs.Session(spec, cs.make_params(nharms=8, hidr=0.4999, precision="f64"),
                   blocksize=2048)


def test_nonrational_search_file_and_engine(oracle):
    """search_file and an engine request at a non-rational step: the oracle's
    list, through the drop-in's own caller and through the batching engine."""
    from coherentsearch_c_amd import engine as eng
    x = synth.series_with_pulsar(1 << 17, 7.7e-3, 0.05, 20.0, 31)
    p = cs.make_params(nharms=24, hidr=0.4999)
    spec = cs.redfft(x, synth.DT)
    got = cs.search_file(spec, p, blocksize=8192, threshold=6.0)
    o = oracle.Oracle(spec.amps.view(np.float32), spec.N, spec.T,
                      oracle.config(nharms=24, hidr=0.4999, blocksize=8192))
    want = oracle.collapse(o.search(6.0, 1))
    say("search_file at hidr 0.4999:", got.size, "candidates")
    assert got.size == want.size and np.array_equal(got["nharm"], want["nharm"])
    assert np.all(np.abs(got["metric"] - want["metric"]) <= 1e-5 * np.abs(want["metric"]))
    with eng.Engine(nslots=2) as E:
        t = E.submit(x, synth.DT, p, threshold=6.0, blocksize=8192)
        raw, info, sp = E.wait(t)
    assert info["status"] == 0
    o2 = oracle.Oracle(sp.amps.view(np.float32), sp.N, sp.T,
                       oracle.config(nharms=24, hidr=0.4999, blocksize=8192))
    w2 = o2.search(6.0, 1)
    assert raw.size == w2.size and np.array_equal(raw["nharm"], w2["nharm"])
    o.close()
    o2.close()
