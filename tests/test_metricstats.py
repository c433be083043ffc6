"""Metric statistics and normalisation on the CPU (no GPU): the oracle's
restatement pinned by closed forms, and the host library's exported
functions (hist_*, log_edges, window_index, build_metricnorm,
metric_normalize, report_metricstats) equal to it.

Reference: src/search.c:409-535 (histograms, windows), 752-823 (the
normalisation model), src/cli.c:157-303 (the report).
"""
import ctypes as C

import numpy as np
import pytest

from coherentsearch_c_amd import abi


# ---- oracle pins -----------------------------------------------------------

def test_quantile_closed_forms(oracle):
    # n values in one bin: the q-quantile interpolates linearly across it
    h = oracle.Hist(1, 8, 1, 0.0, 1.0, lo=0.0, hi=64.0, nb=3200)     # bins 0.02 wide
    h.push(np.full(10, 5.011))                       # bin 250 = [5.00, 5.02)
    assert h.counts[250] == 10 and h.total == 10
    for q in (0.1, 0.5, 0.9):
        assert h.quantile(q) == pytest.approx(5.0 + q * 0.02, abs=1e-12)
    # two bins of 4 and 6 values: the median falls 1/6 of the way into the second
    h2 = oracle.Hist(1, 8, 1, 0.0, 1.0, lo=0.0, hi=1.0, nb=10)
    h2.push([0.05] * 4 + [0.15] * 6)
    assert h2.quantile(0.5) == pytest.approx(0.1 + (5 - 4) / 6 * 0.1, abs=1e-15)
    # under/overflow: a quantile inside the underflow is lo, past every bin hi
    h3 = oracle.Hist(1, 8, 1, 0.0, 1.0, lo=0.0, hi=1.0, nb=10)
    h3.push([-1.0, -2.0, 0.55, 5.0])
    assert h3.under == 2 and h3.over == 1 and h3.counts[5] == 1
    assert h3.quantile(0.5) == 0.0 and h3.quantile(0.99) == 1.0
    assert np.isnan(oracle.Hist(1, 8, 1, 0, 1).quantile(0.5))


def test_windows_closed_forms(oracle):
    e = np.array([1.0, 2.0, 4.0, 8.0])
    for f, w in [(0.5, 1), (1.0, 1), (1.5, 1), (2.0, 2), (3.9, 2), (4.0, 3), (7.9, 3),
                 (8.0, 3), (100.0, 3)]:
        assert oracle.window_index(e, 3, f) == w, f
    le = oracle.log_edges(2.0, 2000.0, 6)
    assert le[0] == pytest.approx(2.0, rel=1e-15) and le[-1] == pytest.approx(2000.0, rel=1e-15)
    assert np.allclose(le[1:] / le[:-1], 10 ** 0.5, rtol=1e-14)
    # a degenerate band is widened to a relative width of 1e-9
    d = oracle.log_edges(5.0, 5.0, 2)
    assert d[-1] == pytest.approx(5.0 * (1 + 1e-9), rel=1e-15)


def test_norm_model_falls_back_to_band(oracle):
    rng = np.random.default_rng(3)
    wa = oracle.Hist(2, 30, 1, 1.0, 2.0)
    wb = oracle.Hist(2, 30, 2, 2.0, 4.0)
    wa.push(rng.normal(10.0, 1.0, 5000))
    wb.push(rng.normal(12.0, 2.0, 50))            # fewer than 200 trials: band values
    g = oracle.Hist(2, 30, 0, 1.0, 4.0)
    g.merge(wa)
    g.merge(wb)
    norm = oracle.build_metricnorm(dict(hists=[g], whists=[wa, wb]))
    e, loc, sc = norm[2]
    assert list(e) == [1.0, 2.0, 4.0]
    assert loc[0] == wa.quantile(0.5) and sc[0] == wa.quantile(0.8413) - wa.quantile(0.5)
    assert loc[1] == g.quantile(0.5) and sc[1] == g.quantile(0.8413) - g.quantile(0.5)
    # a Gaussian window: loc ~ mean, scale ~ sigma
    assert loc[0] == pytest.approx(10.0, abs=0.05) and sc[0] == pytest.approx(1.0, abs=0.06)


def test_oracle_stats_invariants(oracle):
    from coherentsearch_c_amd import synth
    amps = synth.spectrum_with_pulsar(1 << 13, 211.3, 16, 0.05, 14.0, 2)
    N = 1 << 14
    o = oracle.Oracle(amps.view(np.float32), N, N * synth.DT,
                      oracle.config(nharms=16, blocksize=1024))
    ms = oracle.metric_stats(o)
    n_trials = {}
    for b in ms["blocks"]:
        n_trials[b["k"]] = n_trials.get(b["k"], 0) + b["n"]
        assert b["min"] <= b["median"] <= b["max"]
    for h in ms["hists"]:
        assert h.total == n_trials[h.k]
        assert h.under + h.over + int(h.counts.sum()) == h.total
    assert n_trials[1] == o.total


# ---- host library against the oracle ------------------------------------------

def _host_hist(L, h):
    hh = abi.MetricHistogram()
    L.hist_init(C.byref(hh), h.k, h.Hk, h.win, h.flo, h.fhi, h.lo, h.hi, h.nb)
    return hh


def test_host_histograms_equal_oracle(oracle):
    L = abi.host()
    rng = np.random.default_rng(7)
    v = np.concatenate([rng.normal(8.0, 3.0, 20000), [-1.0, 64.0, 70.0, 0.0, 63.99999]])
    h = oracle.Hist(3, 20, 2, 1.5, 3.0)
    h.push(v)
    hh = _host_hist(L, h)
    for x in v:
        L.hist_push(C.byref(hh), float(x))
    counts = np.ctypeslib.as_array(hh.counts, shape=(hh.nb,))
    assert np.array_equal(counts, h.counts)
    assert (hh.under, hh.over, hh.total) == (h.under, h.over, h.total)
    assert hh.sum == h.sum and hh.sumsq == h.sumsq and hh.vmin == h.vmin and hh.vmax == h.vmax
    for q in (0.0, 0.1, 0.5, 0.8413, 0.9, 1 - 1e-4, 1.0):
        assert L.hist_quantile(C.byref(hh), q) == h.quantile(q), q
    L.hist_free(C.byref(hh))
    e = (C.c_double * 17)()
    L.log_edges(0.123, 456.7, 16, e)
    assert np.array_equal(np.array(e[:]), oracle.log_edges(0.123, 456.7, 16))
    for f in np.concatenate([rng.uniform(0.0, 500.0, 200), np.array(e[:])]):
        assert L.window_index(e, 16, float(f)) == oracle.window_index(np.array(e[:]), 16, f)


def _host_stats(L, ms):
    """a host MetricStats holding the oracle's histograms and blocks"""
    hs = abi.MetricStats()
    L.metricstats_init(C.byref(hs), ms["lo"], ms["hi"], ms["nb"], ms["nwin"])
    for h in ms["whists"] + ms["hists"]:
        hh = _host_hist(L, h)
        counts = np.ctypeslib.as_array(hh.counts, shape=(hh.nb,))
        counts[:] = h.counts
        hh.under, hh.over, hh.total = h.under, h.over, h.total
        hh.sum, hh.sumsq, hh.vmin, hh.vmax = h.sum, h.sumsq, h.vmin, h.vmax
        (L.metricstats_append_whist if h.win > 0 else L.metricstats_append_hist)(
            C.byref(hs), hh)
    for b in ms["blocks"]:
        bb = abi.BlockMetricStats(**b)
        L.metricstats_append_block(C.byref(hs), bb)
    return hs


def test_host_norm_and_report_equal_oracle(oracle, tmp_path):
    from coherentsearch_c_amd import make_params, synth
    L = abi.host()
    amps = synth.spectrum_with_pulsar(1 << 14, 1211.7, 32, 0.04, 15.0, 4)
    N = 1 << 15
    o = oracle.Oracle(amps.view(np.float32), N, N * synth.DT,
                      oracle.config(nharms=32, blocksize=2048))
    ms = oracle.metric_stats(o)
    hs = _host_stats(L, ms)
    nm = L.build_metricnorm(C.byref(hs))
    want = oracle.build_metricnorm(ms)
    assert nm.contents.n == len(want)
    for i in range(nm.contents.n):
        nk = nm.contents.ks[i]
        e, loc, sc = want[nk.k]
        assert np.array_equal(np.array(nk.edges[:nk.nwin + 1]), e)
        assert np.array_equal(np.array(nk.loc[:nk.nwin]), loc)
        assert np.array_equal(np.array(nk.scale[:nk.nwin]), sc)
        for f in np.linspace(e[0] * 0.9, e[-1] * 1.1, 50):
            w = oracle.window_index(e, len(loc), f) - 1
            assert L.metric_normalize(nm, nk.k, f, 9.5) == (9.5 - loc[w]) / sc[w]
    assert L.metric_normalize(nm, 99, 1.0, 9.5) == 9.5     # no model for k: unchanged
    L.metricnorm_free(nm)
    base = tmp_path / "obs"
    p = make_params(nharms=32)
    L.report_metricstats(C.byref(hs), str(base).encode(), C.byref(p), 8.0)
    assert (tmp_path / "obs_metricstats.txt").read_text() == \
        oracle.format_metricstats(ms, 32, 8.0)
    assert (tmp_path / "obs_metricfap.txt").read_text() == oracle.format_metricfap(ms, 32)
    assert (tmp_path / "obs_metrichist.txt").read_text() == oracle.format_metrichist(ms, 32)
    L.metricstats_free(C.byref(hs))
