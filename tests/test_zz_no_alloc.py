"""The search that allocates nothing (cs_gpu_session_search_noalloc) through
the C-ABI, against cs_gpu_session_search_resident of the same session.

- Two calls in a row share the library's fixed record array: the first on a
  larger spectrum with many records, the second on a smaller one with fewer;
  each gives the existing entry point's candidates bit for bit, so the second
  call sees none of the first's records.
- An output array of exactly the record count works (the records are
  converted in place); one candidate short, and a search with more records
  than the fixed array holds, fail with an error and leave the next call
  correct.
"""
import sys

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _session(nbins, nh, seed, hifreq=125.0):
    amps = synth.spectrum_with_pulsar(nbins, 271.3, nh, 0.05, 12.0, seed)
    spec = cs.Spectrum(amps, 2 * nbins, synth.DT)
    sess = cs.Session(spec, cs.make_params(nharms=nh, m=16, maxdecim=6), blocksize=4096,
                      hifreq=hifreq)
    sess.upload()
    return sess


def _bits(a):
    return {f: a[f].view(np.uint64 if a[f].dtype.itemsize == 8 else np.uint32)
            for f in ("freq", "metric", "r", "nharm", "ducy")}


def _same(a, b):
    if a.size != b.size:
        return False
    ba, bb = _bits(a), _bits(b)
    return all(np.array_equal(ba[f], bb[f]) for f in ba)


@pytest.mark.parametrize("thr", [3.0, 5.0])
def test_noalloc_matches_search_resident(thr):
    for nbins, nh, seed in ((1 << 16, 32, 5), (1 << 13, 16, 6)):
        sess = _session(nbins, nh, seed)
        want = sess.search_resident(thr)
        got = sess.search_noalloc(thr, cap=want.size)
        got2 = sess.search_noalloc(thr, cap=want.size + 7)
        say("nbins", nbins, "nh", nh, "threshold", thr, "candidates", want.size)
        assert want.size > 0
        assert _same(got, want) and _same(got2, want)
        sess.close()


def test_noalloc_limits():
    sess = _session(1 << 16, 32, 7, hifreq=1000.0)
    want = sess.search_resident(4.0)
    say("candidates", want.size)
    assert want.size > 1
    with pytest.raises(RuntimeError):
        sess.search_noalloc(4.0, cap=want.size - 1)
    # every trial above -inf: more records than the fixed array holds
    with pytest.raises(RuntimeError):
        sess.search_noalloc(-np.inf, cap=1 << 20)
    assert _same(sess.search_noalloc(4.0, cap=want.size), want)
    sess.close()
