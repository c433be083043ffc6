"""Precision and kernel-width sweep on one 2^24-sample series (dt 64 us, a
5.3 ms von Mises pulsar of duty 0.03 at S/N 12, staged on the device):
m in {16, 32} x nharms in {16, 32, 60} x sigma in {analytic, measured},
maxdecim 6, blocksize 65536, hifreq 125 Hz, threshold 8.

For each of the 12 runs, through the C-ABI:
- the raw candidate list is identical (same records, bit-equal metrics) to
  the f32 oracle's list of the same spectrum and chunk table, and the
  --precision f64 list to the f64 oracle's;
- the per-trial metric error against the FP64 oracle (direct sinc
  interpolation in double, direct DFT profiles, double boxcar,
  src/interp.c:317-344) is measured on two row windows -- around the
  strongest candidate and in the middle chunk -- over every decimation, for
  the f32 path and for the f64 path, and bounded by 1e-4 * max(|m|, 1);
- the collapsed f32 and f64 candidate lists are compared and their
  difference is reported.

Each run prints one `PRECISION {...}` JSON line with these figures.
"""
import json
import os
import sys

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import synth

pytestmark = pytest.mark.gpu

BLOCK = 65536
HIFREQ = 125.0
WIN = 256


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


@pytest.fixture(scope="module")
def c2spec():
    x, dt = synth.config2()
    st = cs.Stage(x.size, dt)
    amps = st.run_host(x)
    st.close()
    return cs.Spectrum(amps, x.size, dt)


def _key(a):
    return sorted(zip(a["r"].tolist(), a["nharm"].tolist(), a["metric"].tolist()))


def _locate(o, rec, slot_of_H):
    """(chunk, row, slot) of a raw record: decimated records carry k * r."""
    s = slot_of_H[int(rec["nharm"])]
    r = float(rec["r"]) / o.ks[s]
    t = int(round((r - o.r_lo) / o.lodr))
    starts = o.chunk_starts()
    ci = int(np.searchsorted(starts, t, side="right") - 1)
    return ci, t - int(starts[ci]), s


def _window_error(o32, o64, ofp, ci, j0, nr):
    """max |m - m64| / max(|m64|, 1) over rows [j0, j0+nr) of chunk ci and
    every decimation, for the f32 and the f64 oracle paths (the GPU's metrics
    are bit-identical to these)."""
    ref = ofp.metrics_fp64_rows(ci, j0, nr)[:, j0:j0 + nr]
    a = o32.metrics(ci)[:, j0:j0 + nr]
    b = o64.metrics(ci)[:, j0:j0 + nr]
    ok = np.isfinite(ref)
    assert np.array_equal(ok, np.isfinite(a)) and np.array_equal(ok, np.isfinite(b))
    den = np.maximum(np.abs(ref[ok]), 1.0)
    return (float((np.abs(a[ok] - ref[ok]) / den).max()),
            float((np.abs(b[ok] - ref[ok]) / den).max()), int(ok.sum()))


def _collapsed_set(oracle, raw):
    c = oracle.collapse(raw.astype(oracle.CAND_DTYPE))
    return c, {(round(float(r), 3), int(h)) for r, h in zip(c["r"], c["nharm"])}


@pytest.mark.parametrize("sigma", ["analytic", "measured"])
@pytest.mark.parametrize("nh", [16, 32, 60])
@pytest.mark.parametrize("m", [16, 32])
def test_precision_sweep(oracle, c2spec, m, nh, sigma):
    spec = c2spec
    amps = spec.amps.view(np.float32)
    say("precision sweep m", m, "nharms", nh, "sigma", sigma)
    res = {"m": m, "nharms": nh, "sigma": sigma}
    lists = {}
    ors = {}
    for prec in ("f32", "f64"):
        p = cs.make_params(nharms=nh, m=m, maxdecim=6, precision=prec, sigma=sigma)
        sess = cs.Session(spec, p, blocksize=BLOCK, hifreq=HIFREQ)
        got = sess.run(8.0)
        o = oracle.Oracle(amps, spec.N, spec.T,
                          oracle.config(nharms=nh, m=m, blocksize=BLOCK, maxdecim=6,
                                        hifreq=HIFREQ, precision=prec, sigma=sigma))
        assert sess.total == o.total and np.array_equal(sess.cstarts, o.chunk_starts())
        want = o.search(8.0, _threads())
        same = _key(got) == _key(want)
        res["raw_" + prec] = int(got.size)
        res["identical_to_oracle_" + prec] = same
        assert same, (prec, got.size, want.size)
        lists[prec] = got
        ors[prec] = o
        sess.close()

    o32, o64 = ors["f32"], ors["f64"]
    slot_of_H = {h: s for s, h in enumerate(o32.hks)}
    c32, s32 = _collapsed_set(oracle, lists["f32"])
    c64, s64 = _collapsed_set(oracle, lists["f64"])
    res["collapsed_f32"] = len(s32)
    res["collapsed_f64"] = len(s64)
    res["collapsed_only_f32"] = len(s32 - s64)
    res["collapsed_only_f64"] = len(s64 - s32)

    ofp = oracle.Oracle(amps, spec.N, spec.T,
                        oracle.config(nharms=nh, m=m, blocksize=BLOCK, maxdecim=6,
                                      hifreq=HIFREQ, precision="f32", sigma=sigma))
    windows = []
    raw = lists["f32"]
    if raw.size:
        top = raw[int(np.argmax(raw["metric"]))]
        ci, j, s = _locate(o32, top, slot_of_H)
        n = o32.chunk(ci)[1]
        j0 = max(0, min(j - WIN // 2, n - WIN))
        windows.append((ci, j0, min(WIN, n)))
        tm = ofp.metrics_fp64_rows(ci, j, 1)[s, j]
        res["top"] = {"r": float(top["r"]), "nharm": int(top["nharm"]),
                      "metric": float(top["metric"]), "metric_fp64": float(tm),
                      "rel_err": float(abs(top["metric"] - tm) / max(abs(tm), 1.0))}
    mid = o32.nchunks // 2
    windows.append((mid, 0, min(WIN, o32.chunk(mid)[1])))
    e32 = e64 = 0.0
    ncmp = 0
    for ci, j0, nr in windows:
        a, b, k = _window_error(o32, o64, ofp, ci, j0, nr)
        e32, e64, ncmp = max(e32, a), max(e64, b), ncmp + k
    res["fp64_max_rel_err_f32"] = e32
    res["fp64_max_rel_err_f64"] = e64
    res["fp64_compared_metrics"] = ncmp
    say("PRECISION " + json.dumps(res))
    assert e32 <= 1e-4 and e64 <= 1e-4
    for o in (o32, o64, ofp):
        o.close()
