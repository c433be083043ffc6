"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Bar: the interpolated harmonic stack is bit-identical to the oracle's (same
f32 FMA chains), and since the kernels run the oracle's c2r op for op and the
boxcar in the reference's order, every per-trial metric and therefore every
candidate list is bit-identical too.  Against the FP64 oracle the f32 metric
error is stated and bounded (|d| <= 1e-4 * max(|m|, 1)).
"""
import os
import sys

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(oracle, nh, m, lobin=100, nbins=1 << 15, blocksize=2048, r0=271.3, seed=1,
          maxdecim=6, hifreq=125.0, precision="f32", sigma="analytic"):
    amps = synth.spectrum_with_pulsar(nbins, r0, nh, 0.05, 12.0, seed)
    spec = cs.Spectrum(amps, 2 * nbins, synth.DT)
    p = cs.make_params(nharms=nh, m=m, maxdecim=maxdecim, precision=precision, sigma=sigma)
    o = oracle.Oracle(amps.view(np.float32), spec.N, spec.T,
                      oracle.config(nharms=nh, m=m, lobin=lobin, blocksize=blocksize,
                                    maxdecim=maxdecim, hifreq=hifreq, precision=precision,
                                    sigma=sigma))
    sess = cs.Session(spec, p, lobin=lobin, blocksize=blocksize, hifreq=hifreq)
    assert sess.total == o.total and np.array_equal(sess.cstarts, o.chunk_starts())
    return spec, p, o, sess


CASES = [(16, 16, 100), (16, 32, 3), (32, 16, 100), (32, 32, 2), (60, 16, 100), (60, 32, 100)]


@pytest.mark.parametrize("nh,m,lobin", CASES)
def test_stack_bit_identical(oracle, nh, m, lobin):
    spec, p, o, sess = _case(oracle, nh, m, lobin)
    for ci in sorted({0, 1, o.nchunks // 2, o.nchunks - 1}):
        i0, n, _ = o.chunk(ci)
        say("stack", nh, m, lobin, "chunk", ci, i0, n)
        want, filled = o.stack(ci)
        got = sess.dump_stack(i0, n)
        assert np.array_equal(got.transpose(1, 0, 2).view(np.uint32), want.view(np.uint32))
    sess.close()


@pytest.mark.parametrize("nh,m,lobin", CASES)
def test_every_trial_metric_bit_identical(oracle, nh, m, lobin):
    spec, p, o, sess = _case(oracle, nh, m, lobin)
    say("all metrics", nh, m, lobin, "trials", o.total)
    got = sess.run(-np.inf)
    want = o.search(-np.inf, 1)
    assert got.size == want.size
    for f in ("r", "freq", "nharm"):
        assert np.array_equal(got[f], want[f]), f
    assert np.array_equal(got["metric"], want["metric"])
    sess.close()


# every harmonic count has a kernel (M = 16 or 32 padded taps, m <= M front-padded
# with zero weights; decimations up to 16): odd and prime counts, small and
# large m, long decimation lists
OTHER = [(1, 16, 1), (7, 16, 6), (13, 8, 6), (20, 12, 10), (37, 24, 12), (45, 16, 6),
         (64, 32, 16), (60, 16, 16), (17, 2, 8)]


@pytest.mark.parametrize("nh,m,maxdecim", OTHER)
def test_other_configurations_bit_identical(oracle, nh, m, maxdecim):
    spec, p, o, sess = _case(oracle, nh, m, lobin=100, nbins=1 << 14, maxdecim=maxdecim)
    ci = o.nchunks // 2
    i0, n, _ = o.chunk(ci)
    say("config nh", nh, "m", m, "maxdecim", maxdecim, "trials", o.total)
    want, _ = o.stack(ci)
    got = sess.dump_stack(i0, n)
    assert np.array_equal(got.transpose(1, 0, 2).view(np.uint32), want.view(np.uint32))
    got = sess.run(-np.inf)
    want = o.search(-np.inf, 1)
    assert got.size == want.size
    for f in ("r", "nharm", "metric"):
        assert np.array_equal(got[f], want[f]), f
    sess.close()


@pytest.mark.parametrize("nh,maxdecim,sigma,blocksize", [(65, 6, "analytic", 2048),
                                                         (72, 16, "analytic", 4096),
                                                         (100, 6, "measured", 2048),
                                                         (120, 16, "analytic", 8192),
                                                         (128, 8, "analytic", 2048)])
def test_above_64_harmonics_bit_identical(oracle, nh, maxdecim, sigma, blocksize):
    """nharms 65-128 (m 16, f32): the filled-harmonic mask spans two words and
    the upper harmonics stop fitting chunk by chunk in this band; the stack
    and every per-trial metric are bit-identical to the oracle's."""
    spec, p, o, sess = _case(oracle, nh, 16, nbins=1 << 15, blocksize=blocksize,
                             maxdecim=maxdecim, sigma=sigma, r0=61.7)
    say("nh", nh, "maxdecim", maxdecim, sigma, "trials", o.total, "chunks", o.nchunks)
    for ci in sorted({0, o.nchunks // 2, o.nchunks - 1}):
        i0, n, _ = o.chunk(ci)
        want, filled = o.stack(ci)
        got = sess.dump_stack(i0, n)
        assert np.array_equal(got.transpose(1, 0, 2).view(np.uint32), want.view(np.uint32)), ci
    got = sess.run(-np.inf)
    want = o.search(-np.inf, 1)
    assert got.size == want.size
    for f in ("r", "nharm", "metric"):
        assert np.array_equal(got[f], want[f]), f
    sess.close()


@pytest.mark.parametrize("nh,m,maxdecim,precision,slice_", [
    (120, 16, 16, "f32", 5), (100, 64, 6, "f32", 3), (77, 28, 12, "f32", 0),
    (60, 64, 6, "f32", 4), (240, 16, 6, "f32", 6), (90, 16, 6, "f64", 7)])
def test_two_pass_slices_bit_identical(oracle, nh, m, maxdecim, precision, slice_, monkeypatch):
    """The two-pass search (stack pass, then the register profile pass above
    64 harmonics in f32, the general kernel's c2r otherwise) over slices of a
    few tiles, so that every slice boundary is crossed: the stack and every
    per-trial metric bit-identical to the oracle's."""
    if slice_:
        monkeypatch.setenv("CS_SPLIT_SLICE", str(slice_))
    spec, p, o, sess = _case(oracle, nh, m, nbins=1 << 14, blocksize=2048, maxdecim=maxdecim,
                             precision=precision, r0=61.7)
    say("two-pass nh", nh, "m", m, "maxdecim", maxdecim, precision, "slice", slice_,
        "trials", o.total, "chunks", o.nchunks)
    for ci in sorted({0, o.nchunks // 2, o.nchunks - 1}):
        i0, n, _ = o.chunk(ci)
        want, _ = o.stack(ci)
        got = sess.dump_stack(i0, n)
        assert np.array_equal(got.transpose(1, 0, 2).view(np.uint32), want.view(np.uint32)), ci
    got = sess.run(-np.inf)
    want = o.search(-np.inf, 1)
    assert got.size == want.size
    for f in ("r", "nharm", "metric"):
        assert np.array_equal(got[f], want[f]), f
    assert np.array_equal(sess.run(6.0)["metric"], want["metric"][want["metric"] > 6.0])
    sess.close()


@pytest.mark.parametrize("nh,m,maxdecim", [(16, 16, 6), (32, 32, 6), (60, 16, 6), (7, 8, 3),
                                            (45, 16, 12)])
def test_f64_every_trial_metric_bit_identical(oracle, nh, m, maxdecim):
    """--precision f64: the f32 stack widened to double, c2r, fold and boxcar
    in double -- bit-identical to the oracle's f64 path."""
    spec, p, o, sess = _case(oracle, nh, m, nbins=1 << 14, maxdecim=maxdecim, precision="f64")
    say("f64 nh", nh, "m", m, "trials", o.total)
    got = sess.run(-np.inf)
    want = o.search(-np.inf, 1)
    assert got.size == want.size
    for f in ("r", "nharm", "metric"):
        assert np.array_equal(got[f], want[f]), f
    sess.close()


@pytest.mark.parametrize("nh,m,blocksize,precision", [
    (16, 16, 128, "f32"),      # small chunks: every row of every chunk sampled
    (32, 32, 2048, "f32"),     # large chunks: rows c*nv/ncol
    (60, 16, 2048, "f32"),
    (60, 16, 8192, "f64"),
    (21, 8, 100, "f64")])
def test_measured_sigma_every_trial_metric_bit_identical(oracle, nh, m, blocksize, precision):
    """--sigma measured: per chunk and decimation, 1.4826 * median |profile|
    over the sampled rows (exact order statistic), then the same boxcar."""
    spec, p, o, sess = _case(oracle, nh, m, nbins=1 << 14, blocksize=blocksize,
                             precision=precision, sigma="measured")
    say("measured sigma nh", nh, "m", m, "blocksize", blocksize, precision, "chunks", o.nchunks)
    got = sess.run(-np.inf)
    want = o.search(-np.inf, 1)
    assert got.size == want.size
    for f in ("r", "nharm", "metric"):
        assert np.array_equal(got[f], want[f]), f
    sess.close()


@pytest.mark.parametrize("sigma", ["analytic", "measured"])
def test_graph_replays_match_eager(oracle, sigma):
    """The captured search graph, replayed several times with eager launches,
    host copies and synchronisations in between, returns the eager result
    every time (the graph holds kernels only: the counter reset is a fill
    kernel)."""
    spec, p, o, sess = _case(oracle, 32, 16, nbins=1 << 14, sigma=sigma)
    sess.set_graphs(False)
    eager = sess.run(5.0)
    sess.set_graphs(True)
    for i in range(4):
        got = sess.search_resident(5.0)
        say("graph replay", i, sigma, got.size)
        for f in ("r", "nharm", "metric"):
            assert np.array_equal(got[f], eager[f]), (i, f)
        i0, n, _ = o.chunk(i)
        sess.dump_stack(i0, min(n, 64))            # an eager launch in between
        sess.upload()                               # and a host-to-device copy
    # a threshold change re-captures
    lo = sess.search_resident(4.0)
    sess.set_graphs(False)
    assert np.array_equal(lo["metric"], sess.run(4.0)["metric"])
    sess.close()


def test_config1_candidate_list_identical(oracle):
    """BASELINE config 1 (2^20-bin .fft, pulsar at r0 = 12345.37, 60 harmonics,
    S/N 15), blocksize 2048, threshold 8: raw lists identical in order and
    value; after collapse the pulsar leads."""
    amps, N, dt = synth.config1()
    spec = cs.Spectrum(amps, N, dt)
    p = cs.make_params()
    o = oracle.Oracle(amps.view(np.float32), N, N * dt, oracle.config(blocksize=2048))
    sess = cs.Session(spec, p, blocksize=2048)
    say("config1 trials", sess.total, "chunks", sess.nchunks)
    got = sess.run(8.0)
    want = o.search(8.0, os.cpu_count() or 1)
    # CPU order with several threads is thread-major; compare as sets then
    # against the one-thread order, which is the GPU's
    key = lambda a: list(zip(a["r"], a["nharm"], a["metric"]))
    assert sorted(key(got)) == sorted(key(want))
    top = oracle.collapse(got.astype(oracle.CAND_DTYPE))
    assert abs(top["r"][0] - 12345.37) < 0.05 and top["nharm"][0] == 60
    assert top["metric"][0] > 12.0
    sess.close()


def test_fp64_error_bounded(oracle):
    spec, p, o, sess = _case(oracle, 32, 16, nbins=1 << 14)
    got = sess.run(-np.inf)
    ci = o.nchunks // 2
    f = o.metrics_fp64(ci)
    i0, n, rs = o.chunk(ci)
    # pick this chunk's base-slot records out of the ordered list
    chunk_of = np.searchsorted(o.chunk_starts(), np.round((got["r"] - o.r_lo) / o.lodr),
                               side="right") - 1
    sel = (chunk_of == ci) & (got["nharm"] == 32)
    m = got["metric"][sel]
    ref = f[0][:n]
    err = np.abs(m - ref) / np.maximum(np.abs(ref), 1.0)
    say("fp64 max rel error", float(err.max()))
    assert err.max() < 1e-4
    sess.close()


def test_search_file_end_to_end(oracle, tmp_path):
    amps = synth.spectrum_with_pulsar(1 << 16, 3333.3, 60, 0.05, 14.0, 7)
    spec = cs.Spectrum(amps, 1 << 17, synth.DT)
    p = cs.make_params()
    got = cs.finalize(spec, p, cs.search_file(spec, p, blocksize=4096, threshold=7.0))
    o = oracle.Oracle(amps.view(np.float32), spec.N, spec.T, oracle.config(blocksize=4096))
    want = oracle.measure_ducy(amps.view(np.float32), spec.N,
                               oracle.collapse(o.search(7.0, 1)))
    assert oracle.format_candidates(got.astype(oracle.CAND_DTYPE)) == \
        oracle.format_candidates(want)


def test_session_reuse_and_mismatch(oracle):
    spec, p, o, sess = _case(oracle, 16, 16)
    a = sess.run(6.0)
    b = sess.run(6.0)
    for f in ("r", "freq", "nharm", "metric"):
        assert np.array_equal(a[f], b[f]), f
    import ctypes as C
    # same geometry, new data: the resident session serves it (a DM sweep)
    amps2 = synth.spectrum_with_pulsar(1 << 15, 500.1, 16, 0.05, 12.0, 9)
    spec2 = cs.Spectrum(amps2, spec.N, spec.dt)
    c = cs.take_list(sess.L.cs_gpu_session_run(sess.h, C.byref(spec2.ft), 6.0))
    o2 = oracle.Oracle(amps2.view(np.float32), spec.N, spec.T,
                       oracle.config(nharms=16, blocksize=2048))
    assert np.array_equal(c["metric"], o2.search(6.0, 1)["metric"])
    # a spectrum of another length is refused with an empty list
    small = cs.Spectrum(amps2[:1000], 2000, spec.dt)
    e = cs.take_list(sess.L.cs_gpu_session_run(sess.h, C.byref(small.ft), 6.0))
    assert e.size == 0
    sess.close()


@pytest.mark.parametrize("nh,m,precision", [(16, 16, "f32"), (60, 32, "f32"), (32, 16, "f64"),
                                             (72, 32, "f32")])
def test_search_region_one_shot(oracle, nh, m, precision):
    """cs_gpu_search_region (src/cuda.cu:1238-1255): create, run and free in
    one call; r_hi is ignored.  Its list is the oracle's, trial for trial, and
    the resident session's."""
    import ctypes as C
    spec, p, o, sess = _case(oracle, nh, m, nbins=1 << 14, precision=precision)
    say("search_region nh", nh, "m", m, precision, "trials", o.total)
    L = sess.L
    cl = L.cs_gpu_search_region(C.byref(spec.ft), C.byref(p), C.byref(sess.plans),
                                sess._cs.ctypes.data_as(C.POINTER(C.c_longlong)), sess.total,
                                sess.Nprof, sess.nchunks, sess.r_lo, -12345.0, sess.lodr,
                                -np.inf, 0)
    got = cs.take_list(cl)
    want = o.search(-np.inf, 1)
    assert got.size == want.size
    for f in ("r", "freq", "nharm", "metric"):
        assert np.array_equal(got[f], want[f]), f
    # a device that does not exist: an empty list, as a failed create, and
    # the library's later searches are unaffected
    bad = cs.take_list(L.cs_gpu_search_region(
        C.byref(spec.ft), C.byref(p), C.byref(sess.plans),
        sess._cs.ctypes.data_as(C.POINTER(C.c_longlong)), sess.total, sess.Nprof,
        sess.nchunks, sess.r_lo, 0.0, sess.lodr, 8.0, 4096))
    assert bad.size == 0
    again = sess.run(-np.inf)
    assert np.array_equal(again["metric"], got["metric"])
    sess.close()


def test_unsupported_configuration_fails_loudly():
    """Outside the GPU search's range cs_gpu_session_create returns NULL with
    a message: more than 512 harmonics, m above 64, decimations above 16;
    inside it the general kernel takes what the register kernels cannot
    hold."""
    amps = synth.spectrum_with_pulsar(1 << 14, 300.1, 4, 0.05, 5.0, 1)
    spec = cs.Spectrum(amps, 1 << 15, synth.DT)
    for kw in (dict(nharms=513), dict(nharms=16, m=66), dict(nharms=60, maxdecim=20)):
        with pytest.raises(RuntimeError):
            cs.Session(spec, cs.make_params(**kw), blocksize=2048)
    for kw in (dict(nharms=16, m=34), dict(nharms=129), dict(nharms=72, m=32),
               dict(nharms=72, precision="f64"), dict(nharms=16, hidr=0.9)):
        cs.Session(spec, cs.make_params(**kw), blocksize=2048).close()
