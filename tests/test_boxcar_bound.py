"""The f32 boxcar's width bound, checked in float32 arithmetic on the CPU.

The search kernels compute the maximum window sum exactly for the anchor
widths 1, 4, 13 and 42 and bound every other width w by the sum of the anchor
maxima of a split of w, plus a margin of 4u(k+2)(NB+WC+2k+4) * max|P|
(u = 2^-24, k parts, P the wrapped prefix sums): a width whose bound cannot
exceed the candidate threshold is not computed.  These tests check that
inequality, m_w <= bound, on profiles built to stress the rounding (large
offsets against small signals, spikes, alternating signs, tiny values), for
every profile length the kernels use, with the prefix sums accumulated in
float32 in the kernels' order; and that on noise profiles the bound lies far
enough below a threshold of 8 for the skip to pay.
"""
import numpy as np
import pytest

U = 2.0 ** -24


def wseq(i):
    w = 1
    for _ in range(i):
        g = int(1.5 * w)
        w = g if g > w + 1 else w + 1
    return w


def wcap(nb):
    return max(1, min(int(0.3 * nb), nb - 1))


def widths(nb):
    out = []
    i = 0
    while i < 16 and wseq(i) <= wcap(nb):
        out.append(wseq(i))
        i += 1
    return out


def split(w):
    n42, r = divmod(w, 42)
    n13, r = divmod(r, 13)
    n4, n1 = divmod(r, 4)
    return [(42, n42), (13, n13), (4, n4), (1, n1)]


def prefix(x, wc):
    """float32 prefix sums over the profile extended by its first wc values,
    accumulated one value at a time"""
    nb = x.shape[-1]
    ext = np.concatenate([x, x[..., :wc]], axis=-1).astype(np.float32)
    P = np.zeros(ext.shape[:-1] + (nb + wc + 1,), np.float32)
    acc = np.zeros(ext.shape[:-1], np.float32)
    for i in range(nb + wc):
        acc = (acc + ext[..., i]).astype(np.float32)
        P[..., i + 1] = acc
    return P


def winmax(P, nb, w):
    d = (P[..., w:w + nb] - P[..., 0:nb]).astype(np.float32)
    return np.nanmax(d, axis=-1)


def check(x):
    nb = x.shape[-1]
    ws = widths(nb)
    wc = ws[-1]
    P = prefix(x, wc)
    pmax = np.nanmax(np.abs(P), axis=-1).astype(np.float32)
    m = {w: winmax(P, nb, w) for w in ws}
    anchors = {a: m.get(a) for a in (1, 4, 13, 42)}
    worst = 0.0
    for w in ws:
        if w in (1, 4, 13, 42):
            continue
        parts = split(w)
        k = sum(n for _, n in parts)
        cm = np.float32(4 * (k + 2) * (nb + wc + 2 * k + 4) * U)
        sb = np.zeros_like(pmax)
        for a, n in parts:
            for _ in range(n):
                sb = (sb + anchors[a]).astype(np.float32)
        bnd = (sb.astype(np.float64) + pmax.astype(np.float64) * np.float64(cm)).astype(np.float32)
        assert np.all(m[w] <= bnd), (nb, w)
        worst = max(worst, float(np.max((m[w] - sb) / np.maximum(pmax, 1e-30))))
    return worst


NBS = [2, 6, 8, 14, 20, 24, 30, 40, 60, 120, 128, 200, 240, 256]


@pytest.mark.parametrize("nb", NBS)
def test_bound_holds_on_rounding_stress_profiles(nb):
    rng = np.random.default_rng(nb)
    n = 400
    noise = rng.standard_normal((n, nb))
    cases = [
        noise,
        1e4 + noise,                         # offset: prefix sums dwarf the windows
        -3e5 + 1e-2 * noise,
        np.where(np.arange(nb) % 2, 1.0, -1.0) * (1e3 + noise),
        noise * 1e-30,                       # near the subnormal range
        np.cumsum(noise, axis=-1),           # a random walk
    ]
    spikes = noise.copy()
    spikes[:, rng.integers(0, nb, n)] += 1e6
    cases.append(spikes)
    for x in cases:
        check(x.astype(np.float32))


@pytest.mark.parametrize("nb", [20, 40, 60, 120, 240])
def test_bound_is_tight_on_noise(nb):
    """On unit-variance noise the bounded S/N of every skipped width stays
    below 8 for nearly all profiles (the skip is taken for a whole wave of 64
    trials only when all of them qualify)."""
    rng = np.random.default_rng(7)
    x = rng.standard_normal((4096, nb)).astype(np.float32)
    x -= x.mean(axis=-1, keepdims=True)
    ws = widths(nb)
    wc = ws[-1]
    P = prefix(x, wc)
    m = {w: winmax(P, nb, w) for w in ws}
    for w in ws:
        if w in (1, 4, 13, 42):
            continue
        sb = sum(n * m[a] for a, n in split(w) if n)
        snr = sb / np.sqrt(w * (1.0 - w / nb))
        assert np.mean(snr > 8.0) < 1e-3, (nb, w, float(np.max(snr)))
