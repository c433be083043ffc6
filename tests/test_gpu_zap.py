"""GPU zaplist (src/zap.c) through the C-ABI.

Bar:
- cs_gpu_session_zap leaves the resident spectrum bit-identical to the
  oracle's zap of the same spectrum (same intervals, in low-edge order, same
  lower-median levels), and the search of it returns the oracle's list;
- an engine request with a zaplist searches the zapped staged spectrum;
- the CLI with --zaplist writes the .cohout the oracle pipeline gives
  (zap, search, collapse, S/N sort, duty cycles).
"""
import ctypes as C
import sys

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


def _birds(spec, rng, n=25):
    """Birds in Hz for this spectrum: narrow, wide, overlapping, barycentric,
    near DC and past the top, plus two lines on the pulsar's harmonics."""
    nb = spec.namps
    out = [(f / spec.T, w / spec.T, int(rng.random() < 0.3))
           for f, w in zip(rng.uniform(150, nb + 40, n),
                           rng.choice([0.05, 0.9, 4.0, 25.0, 80.0], n))]
    out += [((nb - 2.0) / spec.T, 30.0 / spec.T, 0), (1000.0 / spec.T, 12.0 / spec.T, 0),
            (1004.0 / spec.T, 40.0 / spec.T, 1)]
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_session_zap_matches_oracle(oracle, seed):
    rng = np.random.default_rng(seed)
    nbins = 1 << 15
    amps = synth.spectrum_with_pulsar(nbins, 333.7, 32, 0.05, 13.0, seed)
    spec = cs.Spectrum(amps.copy(), 2 * nbins, synth.DT)
    p = cs.make_params(nharms=32, m=16)
    sess = cs.Session(spec, p, blocksize=4096, hifreq=500.0)
    sess.upload()
    birds = _birds(spec, rng) + [(2 * 333.7 / spec.T, 1.0 / spec.T, 0)]
    n = sess.zap(birds, baryv=1e-4)
    want, n_want = oracle.zap(amps.view(np.float32), spec.N, spec.T, birds, baryv=1e-4)
    say("zap seed", seed, "birds", len(birds), "applied", n)
    assert n == n_want and n > 10
    assert np.array_equal(spec.amps.view(np.float32), want)
    got = sess.search_resident(6.0)
    o = oracle.Oracle(want, spec.N, spec.T,
                      oracle.config(nharms=32, m=16, blocksize=4096, hifreq=500.0))
    ref = o.search(6.0, 1)
    assert got.size == ref.size
    for f in ("r", "nharm", "metric"):
        assert np.array_equal(got[f], ref[f]), f
    sess.close()


def test_engine_request_with_zaplist(oracle):
    x = synth.series_with_pulsar(1 << 19, 4.1e-3, 0.04, 15.0, 21)
    p = cs.make_params(nharms=16, m=16)
    T = x.size * synth.DT
    birds = [(50.0, 0.2, 0), (2 / 4.1e-3, 0.05, 0), (97.3, 1.5, 1), (110.0, 0.01, 0)]
    with eng.Engine(nslots=2) as E:
        t0 = E.submit(x, synth.DT, p, threshold=6.0, blocksize=8192)
        t1 = E.submit(x, synth.DT, p, threshold=6.0, blocksize=8192, zaplist=birds, baryv=3e-5)
        _, i0, s0 = E.wait(t0)
        raw, i1, s1 = E.wait(t1)
    assert i0["status"] == 0 and i1["status"] == 0
    want, n = oracle.zap(s0.amps, s0.N, T, birds, baryv=3e-5)
    say("engine zap applied", n, "cands", raw.size)
    assert n >= 3 and np.array_equal(s1.amps, want)
    o = oracle.Oracle(want, s0.N, T, oracle.config(nharms=16, m=16, blocksize=8192))
    ref = o.search(6.0, 1)
    assert raw.size == ref.size and np.array_equal(raw["metric"], ref["metric"])


def test_cli_zaplist_end_to_end(oracle, tmp_path):
    nbins = 1 << 16
    amps = synth.spectrum_with_pulsar(nbins, 3333.3, 60, 0.05, 14.0, 7)
    L = abi.host()
    base = tmp_path / "obs"
    L.cs_write_inf(str(base.with_suffix(".inf")).encode(), 2 * nbins, synth.DT, 0.0, b"psr")
    a = amps.view(np.float32)
    L.cs_write_floats(str(base.with_suffix(".fft")).encode(), a.ctypes.data_as(C.c_void_p),
                      a.size)
    T = 2 * nbins * synth.DT
    (tmp_path / "z.txt").write_text("# birds\n%.9f 0.05\nB %.9f 0.3\n%.9f 2.0\n"
                                    % (3 * 3333.3 / T, 7 * 3333.3 / T, 123.0))
    out = tmp_path / "out.cohout"
    argv = [b"coherent_search", b"--gpu", b"0", b"-n", b"60", b"--blocksize", b"4096",
            b"-t", b"7", b"--zaplist", str(tmp_path / "z.txt").encode(), b"--baryv", b"1e-5",
            b"-o", str(out).encode(), str(base.with_suffix(".fft")).encode()]
    arr = (C.c_char_p * len(argv))(*argv)
    assert L.cs_main(len(argv), arr) == 0
    z = cs.Zaplist.read(tmp_path / "z.txt")
    zapped, n = oracle.zap(a, 2 * nbins, T, z.tuples(), baryv=1e-5)
    assert n == 3
    o = oracle.Oracle(zapped, 2 * nbins, T, oracle.config(blocksize=4096))
    want = oracle.measure_ducy(zapped, 2 * nbins, oracle.collapse(o.search(7.0, 1)))
    text = out.read_text()
    say(text.splitlines()[:3])
    assert text == oracle.format_candidates(want)
