"""Candidate lists at a threshold against the full metric list.

The f32 search kernels skip the windows of a boxcar width whose bound
cannot exceed the candidate threshold (the bound itself is checked in
tests/test_boxcar_bound.py).  The list at a threshold must therefore be
exactly the records of the unthresholded run (every trial's metric, where
nothing is skipped, checked against the oracle by the parity tests) whose
metric exceeds it: same trials, same values, bit for bit.  Thresholds from
below zero to above the pulsar, at harmonic counts of every kernel form
(one-pass, two-pass with the register profile pass, f64, and the general
kernel, which keeps its own boxcar).
"""
import sys

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import synth

pytestmark = pytest.mark.gpu

THRESHOLDS = [-2.0, 0.0, 2.0, 3.5, 5.0, 6.5, 8.0, 11.0]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def key(c):
    return sorted(zip(c["r"].tolist(), c["nharm"].tolist(), c["metric"].tolist()))


@pytest.mark.parametrize("nh,m,maxdecim,precision", [
    (60, 16, 6, "f32"), (32, 16, 6, "f32"), (16, 32, 6, "f32"), (7, 16, 3, "f32"),
    (45, 16, 12, "f32"), (60, 16, 16, "f32"), (100, 16, 12, "f32"), (24, 48, 6, "f32"),
    (32, 16, 6, "f64")])
def test_threshold_list_is_the_full_list_above_it(nh, m, maxdecim, precision):
    nbins = 1 << 15
    amps = synth.spectrum_with_pulsar(nbins, 271.3, min(nh, 60), 0.05, 9.0, 3)
    spec = cs.Spectrum(amps, 2 * nbins, synth.DT)
    p = cs.make_params(nharms=nh, m=m, maxdecim=maxdecim, precision=precision)
    sess = cs.Session(spec, p, blocksize=4096)
    say("boxcar skip", nh, m, maxdecim, precision, "trials", sess.total)
    full = sess.run(-np.inf)
    assert full.size > 0
    for thr in THRESHOLDS:
        got = sess.run(thr)
        want = full[full["metric"] > thr]
        say("  threshold", thr, "records", got.size)
        assert got.size == want.size, thr
        assert key(got) == key(want), thr
    sess.close()
