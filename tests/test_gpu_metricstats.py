"""--metricstats / --normalize on the GPU through the C-ABI, against the
oracle's restatement of the reference's single-threaded CPU collection
(src/search.c:409-565, 752-823, 989-1095, 1427-1444).

Bar:
- block statistics: n, min, max and the median exact; mean and std within
  1e-12 relative (the device sums in a fixed tree, the reference in trial
  order);
- window and band histograms: every count, under/overflow and total exact,
  the extrema exact, sums within 1e-12 relative;
- candidates: with --metricstats the list equals the plain search's; with
  --normalize the normalised list (scores from the histograms' quantiles,
  which depend on counts only) is bit-identical to the oracle's after the
  same collapse;
- the CLI writes the .cohout and the three report files of the oracle.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import abi, synth

pytestmark = pytest.mark.gpu

CASES = [  # (nharms, blocksize, precision, sigma)
    (32, 4096, "f32", "analytic"),
    (60, 8192, "f32", "analytic"),
    (16, 2048, "f64", "analytic"),
    (32, 4096, "f32", "measured"),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _spectrum(nh, seed):
    nbins = 1 << 16
    amps = synth.spectrum_with_pulsar(nbins, 777.3, nh, 0.05, 16.0, seed)
    return cs.Spectrum(amps, 2 * nbins, synth.DT), amps


def _close(a, b, rtol=1e-12):
    return abs(a - b) <= rtol * max(abs(a), abs(b), 1.0)


def _check_hist(g, w):
    assert (g["k"], g["Hk"], g["win"], g["total"], g["under"], g["over"]) == \
        (w.k, w.Hk, w.win, w.total, w.under, w.over)
    assert g["flo"] == w.flo and g["fhi"] == w.fhi
    assert np.array_equal(g["counts"], w.counts)
    if w.total:
        assert g["vmin"] == w.vmin and g["vmax"] == w.vmax
        assert _close(g["sum"], w.sum) and _close(g["sumsq"], w.sumsq)


@pytest.mark.parametrize("nh,bs,prec,sig", CASES)
def test_metricstats_match_oracle(oracle, nh, bs, prec, sig):
    spec, amps = _spectrum(nh, 3)
    p = cs.make_params(nharms=nh, precision=prec, sigma=sig)
    say("metricstats nh %d blocksize %d %s %s" % (nh, bs, prec, sig))
    ms = cs.MetricStats()
    got = cs.search_file(spec, p, blocksize=bs, threshold=7.0, metricstats=ms)
    plain = cs.search_file(spec, p, blocksize=bs, threshold=7.0)
    assert got.size == plain.size
    for f in ("r", "nharm", "metric"):
        assert np.array_equal(got[f], plain[f]), f
    v = ms.view()
    ms.close()
    o = oracle.Oracle(amps.view(np.float32), spec.N, spec.T,
                      oracle.config(nharms=nh, precision=prec, sigma=sig, blocksize=bs))
    want = oracle.metric_stats(o)
    say("  %d blocks, %d windows" % (len(want["blocks"]), len(want["whists"])))
    assert len(v["blocks"]) == len(want["blocks"])
    for g, w in zip(v["blocks"], want["blocks"]):
        for f in ("block", "k", "Hk", "nbins", "n", "min", "max", "median", "flo", "fhi",
                  "ngoodbins"):
            assert g[f] == w[f], (f, g, w)
        assert _close(g["mean"], w["mean"]) and _close(g["std"], w["std"]), (g, w)
    assert len(v["whists"]) == len(want["whists"]) and len(v["hists"]) == len(want["hists"])
    for g, w in zip(v["whists"] + v["hists"], want["whists"] + want["hists"]):
        _check_hist(g, w)
    o.close()


@pytest.mark.parametrize("nh,bs,prec,sig", CASES[:3])
def test_normalize_matches_oracle(oracle, nh, bs, prec, sig):
    spec, amps = _spectrum(nh, 5)
    p = cs.make_params(nharms=nh, precision=prec, sigma=sig)
    # search_file returns the collapsed list in frequency order; the oracle's
    # collapse adds the CLI's S/N sort and truncation (src/cli.c:1059-1060)
    got = oracle.collapse(cs.search_file(spec, p, blocksize=bs, threshold=4.0,
                                         normalize=True).astype(oracle.CAND_DTYPE),
                          remove=False, harm_remove=False)
    o = oracle.Oracle(amps.view(np.float32), spec.N, spec.T,
                      oracle.config(nharms=nh, precision=prec, sigma=sig, blocksize=bs))
    norm = oracle.build_metricnorm(oracle.metric_stats(o))
    raw = oracle.metric_candidates(o, 4.0, norm)
    want = oracle.collapse(raw)
    say("normalize nh %d: %d raw, %d collapsed, top %.3f" % (nh, raw.size, want.size,
                                                             want["metric"].max()))
    assert got.size == want.size
    for f in ("r", "nharm", "metric", "freq"):
        assert np.array_equal(got[f], want[f]), f
    # the pulsar stands out in noise units too
    f0 = 777.3 / spec.T
    top = got[np.argsort(-got["metric"])][:3]
    assert any(abs(f / f0 - round(f / f0)) < 1e-3 for f in top["freq"])
    o.close()


def test_cli_metricstats_and_normalize(oracle, tmp_path):
    nh = 32
    spec, amps = _spectrum(nh, 9)
    L = abi.host()
    base = tmp_path / "obs"
    L.cs_write_inf(str(base.with_suffix(".inf")).encode(), spec.N, synth.DT, 0.0, b"psr")
    a = amps.view(np.float32)
    L.cs_write_floats(str(base.with_suffix(".fft")).encode(), a.ctypes.data_as(C.c_void_p),
                      a.size)
    o = oracle.Oracle(a, spec.N, spec.T, oracle.config(nharms=nh, blocksize=4096))
    ms = oracle.metric_stats(o)
    for flag, thr in ((b"--metricstats", 7.0), (b"--normalize", 5.0)):
        out = tmp_path / ("out%s.cohout" % flag.decode().strip("-"))
        argv = [b"coherent_search", b"--gpu", b"0", b"-n", b"32", b"--blocksize", b"4096",
                b"-t", str(thr).encode(), flag, b"-o", str(out).encode(),
                str(base.with_suffix(".fft")).encode()]
        arr = (C.c_char_p * len(argv))(*argv)
        say("cli", flag)
        assert L.cs_main(len(argv), arr) == 0
        norm = oracle.build_metricnorm(ms) if flag == b"--normalize" else None
        raw = oracle.metric_candidates(o, thr, norm)
        want = oracle.measure_ducy(a, spec.N, oracle.collapse(raw))
        assert out.read_text() == oracle.format_candidates(want)
        stem = str(out)      # with -o the report base is the -o path as given (src/cli.c:1077)
        assert open(stem + "_metricstats.txt").read() == oracle.format_metricstats(ms, nh, thr)
        assert open(stem + "_metricfap.txt").read() == oracle.format_metricfap(ms, nh)
        assert open(stem + "_metrichist.txt").read() == oracle.format_metrichist(ms, nh)
    o.close()
