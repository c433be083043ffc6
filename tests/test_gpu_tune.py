"""GPU tests of the autotuner (cs_gpu_autotune) through the C-ABI.

- It times every candidate launch shape of a kernel class and keeps the
  fastest; the second call returns the cached choice.
- Every launch shape (lane order, tile-to-XCD mapping) gives the same raw
  candidate list bit for bit; a forced shape is loaded from a tune file.
- An engine request that asks for the tuned chunk length (blocksize < 0)
  returns the records of the default chunking (analytic sigma) in its own
  chunk order, and the same collapsed list.
"""
import sys
import time

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import abi
from coherentsearch_c_amd import engine as eng
from coherentsearch_c_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _arch():
    rc, name = cs.gpu_available(0)
    assert rc == 1
    return name.rsplit(", ", 1)[1].rstrip(")")


def _same(a, b):
    return a.size == b.size and all(np.array_equal(a[f], b[f]) for f in ("r", "nharm", "metric"))


def test_autotune_picks_and_caches():
    p = cs.make_params(nharms=32, m=16)
    t = time.perf_counter()
    r = cs.autotune(p)
    t1 = time.perf_counter() - t
    say("tune nh 32:", r, "%.2f s" % t1)
    assert r["ntried"] == 3 * 2 * 3 and r["best_ms"] <= r["default_ms"]
    assert r["blocksize"] in (16384, 65536, 262144) and r["lane_order"] in (0, 1)
    t = time.perf_counter()
    assert cs.autotune(p) == r
    assert time.perf_counter() - t < 0.05
    pm = cs.make_params(nharms=16, m=16, sigma="measured")
    rm = cs.autotune(pm)
    say("tune nh 16 measured:", rm)
    assert rm["blocksize"] == 65536 and rm["ntried"] == 2 * 3


@pytest.mark.parametrize("lane,stripe", [(1, 0), (0, 16), (1, 64), (0, 0)])
def test_every_launch_shape_gives_the_same_list(tmp_path, lane, stripe):
    nbins = 1 << 16
    amps = synth.spectrum_with_pulsar(nbins, 777.7, 60, 0.05, 13.0, 5)
    spec = cs.Spectrum(amps, 2 * nbins, synth.DT)
    p = cs.make_params(nharms=60, m=16)
    ref_sess = cs.Session(spec, p, blocksize=4096)
    ref = ref_sess.run(5.0)
    ref_sess.close()
    f = tmp_path / "tune.txt"
    f.write_text("# forced\n%s 0 60 16 0 0 %d %d 65536 1.0 1.0\n" % (_arch(), lane, stripe))
    assert abi.host().cs_gpu_tune_load(str(f).encode()) == 0
    sess = cs.Session(spec, p, blocksize=4096)
    got = sess.run(5.0)
    sess.set_graphs(False)
    eager = sess.search_resident(5.0)
    say("shape", lane, stripe, "cands", got.size)
    assert _same(got, ref) and _same(eager, ref)
    sess.close()
    out = tmp_path / "saved.txt"
    assert abi.host().cs_gpu_tune_save(str(out).encode()) == 0
    assert ("0 60 16 0 0 %d %d 65536" % (lane, stripe)) in out.read_text()


def test_engine_uses_the_tuned_chunk_length(tmp_path):
    x = synth.series_with_pulsar(1 << 20, 2.9e-3, 0.05, 14.0, 8)
    p = cs.make_params(nharms=16, m=16)
    f = tmp_path / "tune.txt"
    f.write_text("%s 0 16 16 0 0 0 0 16384 1.0 1.0\n" % _arch())
    assert abi.host().cs_gpu_tune_load(str(f).encode()) == 0
    with eng.Engine(nslots=2) as E:
        a = E.submit(x, synth.DT, p, threshold=6.0, blocksize=-1)      # tuned: 16384
        b = E.submit(x, synth.DT, p, threshold=6.0, blocksize=0)       # default: 65536
        c = E.submit(x, synth.DT, p, threshold=6.0, blocksize=16384)
        la, ia, _ = E.wait(a)
        lb, ib, _ = E.wait(b)
        lc, ic, _ = E.wait(c)
    say("tuned chunk length:", la.size, "candidates")
    assert ia["status"] == 0 and ib["status"] == 0 and ic["status"] == 0
    assert _same(la, lc)
    key = lambda v: sorted(zip(v["r"].tolist(), v["nharm"].tolist(), v["metric"].tolist()))
    assert key(la) == key(lb)
    assert _same(cs.collapse(la), cs.collapse(lb))
