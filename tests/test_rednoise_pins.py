"""Pins of the oracle's staging arithmetic that the oracle's own code cannot
satisfy by construction: rednoise (src/datpipe.c:143-253), its helpers, and
the FP contraction of the reference's build of datpipe.c and zap.c
(src/Makefile:129-138).

- The inverse square root against an independent bit-level implementation
  of the published method (magic constant 0x5f3759df, then one Newton step),
  and its error bound: one Newton step leaves at most 0.18 % relative error
  (the raw bit trick, 3.4 %).
- The quickselect median against np.sort: element (n-1)/2 ascending.
- The block layout: a probe spectrum that is constant within each block of
  the layout PRESTO's rednoise documents (first block startwidth = 6 bins;
  then startwidth * ln(bins so far) while the bin count over T stays below
  endfreq = 6 Hz; then endwidth = 100 bins; the last block what remains).
  Every block's lower median is then its own constant, so at the middle bin
  of each block the output is the input times invsqrtf(power / ln 2), and
  across the first half block it is that of block 0; between two middles the
  normalising level is the straight line from one block's mean to the next.
  A different start width, end width, end frequency or a slope of the wrong
  sign breaks these.
- FP contraction: the oracle writes out the fused forms gcc makes of the
  reference's expressions under -ffp-contract=fast.  The oracle's source
  compiled with the reference's unfused expressions (ORC_SOURCE_FORM) and
  -ffp-contract=fast must give the oracle's rednoise, unpack and zap outputs
  bit for bit, and the same source without contraction must not.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LN2 = np.log(2.0)


def invsqrtf_ref(x):
    """0x5f3759df - (bits >> 1), then y * (1.5 - 0.5 x y y) in float32 with
    the last multiply-subtract fused (one rounding, the reference build's
    contraction); the fused step is exact in long double (both terms are
    near 1 and the product has 48 bits)."""
    x = np.asarray(x, np.float32)
    y = (np.int32(0x5f3759df) - (x.view(np.int32) >> 1)).view(np.float32)
    hxy = (np.float32(0.5) * x) * y
    t = (np.longdouble(1.5) - hxy.astype(np.longdouble) * y.astype(np.longdouble))
    return (y * t.astype(np.float32)).astype(np.float32)


def test_invsqrtf_bit_level_and_error_bound(oracle):
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(1e-6, 10.0, 40000), 10.0 ** rng.uniform(-30, 30, 40000),
                        [1.0, 2.0, 4.0, 0.25, 1e-38, 3e38]]).astype(np.float32)
    got = oracle.presto_invsqrtf(x)
    assert np.array_equal(got.view(np.uint32), invsqrtf_ref(x).view(np.uint32))
    normal = x >= np.finfo(np.float32).tiny     # the bit trick assumes a normal float
    rel = np.abs(got[normal].astype(np.float64) * np.sqrt(x[normal].astype(np.float64)) - 1.0)
    assert rel.max() < 1.8e-3           # one Newton step; none would leave 3.4 %
    assert rel.max() > 1.0e-3           # and it is the approximation, not 1/sqrt


def test_presto_median_is_lower_median(oracle):
    rng = np.random.default_rng(2)
    for n in list(range(1, 40)) + [99, 100, 101, 255, 1024]:
        for kind in range(3):
            if kind == 0:
                v = rng.standard_normal(n)
            elif kind == 1:
                v = rng.integers(0, 4, n)          # many ties
            else:
                v = np.sort(rng.exponential(1.0, n))[::-1]
            v = v.astype(np.float32)
            assert oracle.presto_median(v) == np.sort(v)[(n - 1) // 2]


def layout(nc, T, startwidth=6, endwidth=100, endfreq=6.0):
    """block (start, length) list of PRESTO's rednoise over bins 1..nc-1"""
    blocks, start, binnum = [], 1, 1
    length = startwidth
    first = True
    while start < nc:
        n = min(length, nc - start)
        blocks.append((start, n))
        start += n
        binnum += n
        if first or np.float32(binnum) / np.float32(T) < endfreq:
            length = int(startwidth * np.log(binnum))
        else:
            length = endwidth
        first = False
    return blocks


@pytest.mark.parametrize("N,dt", [(1 << 16, 1e-3), (50000, 2e-3), (1 << 14, 64e-6)])
def test_rednoise_block_layout_probe(oracle, N, dt):
    nc = N // 2
    T = float(np.float32(N * dt))
    blocks = layout(nc, T)
    # one amplitude per block, exact squares, neighbours at least 20 % apart in power
    rng = np.random.default_rng(N)
    amp = np.zeros(nc, np.float32)
    levels = rng.permutation(np.arange(8, 40)).astype(np.float32) / 8.0
    a_of = []
    for j, (s, n) in enumerate(blocks):
        a = levels[j % levels.size]
        if j and abs(a * a / (a_of[-1] ** 2) - 1.0) < 0.2:
            a = np.float32(a_of[-1] * 1.5 if a_of[-1] < 3.0 else a_of[-1] / 1.5)
        a_of.append(np.float32(a))
        amp[s:s + n] = a
    packed = np.zeros(N, np.float32)
    packed[2::2] = amp[1:]
    out = oracle.rednoise(packed, dt).view(np.complex64)
    re = out.real
    assert out[0] == 1.0 + 0.0j and np.all(out.imag[1:] == 0.0)
    mean = [np.float32(np.float64(np.float32(a * a)) / LN2) for a in a_of]
    norm = [invsqrtf_ref(np.float32(m))[()] for m in mean]
    s0, n0 = blocks[0]
    assert np.all(re[1:1 + n0 // 2] == a_of[0] * norm[0])
    for j, (s, n) in enumerate(blocks):
        mid = s + n // 2
        assert re[mid] == np.float32(a_of[j] * norm[j]), (j, s, n)
    # between two middles the level is the straight line from one block's mean
    # to the next, reaching it after half the two blocks' lengths; after the
    # last middle it goes on with the last slope.  The recovered level has the
    # inverse square root's error (< 0.18 % each way, pinned above), twice.
    p_out = re.astype(np.float64) ** 2
    p_in = amp.astype(np.float64) ** 2
    mean64 = np.array(mean, np.float64)
    for j in range(len(blocks)):
        s, n = blocks[j]
        jj = min(j, len(blocks) - 2)
        (sa, na), (sb, nb) = blocks[jj], blocks[jj + 1]
        slope = (mean64[jj + 1] - mean64[jj]) / (0.5 * (na + nb))
        end = blocks[j + 1][0] + blocks[j + 1][1] // 2 if j + 1 < len(blocks) else nc
        k = np.arange(s + n // 2, end)
        line = mean64[j] + slope * (k - (s + n // 2))
        lev = p_in[k] / p_out[k]
        assert np.all(np.abs(lev / line - 1.0) < 4e-3), (j, s, n)


def _build(tmp_path, name, flags):
    so = str(tmp_path / name)
    subprocess.run(["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-std=gnu11",
                    "-shared", "-o", so, os.path.join(ROOT, "oracle", "oracle.c"), "-lm"] + flags,
                   check=True)
    L = C.CDLL(so)
    L.orc_rednoise.argtypes = [C.c_void_p, C.c_longlong, C.c_double, C.c_void_p]
    L.orc_nr_unpack.argtypes = [C.c_void_p, C.c_longlong]
    L.orc_zap.argtypes = [C.c_void_p, C.c_longlong, C.c_longlong, C.c_double, C.c_void_p,
                          C.c_void_p, C.c_void_p, C.c_int, C.c_double]
    return L


def _outputs(L, seed):
    rng = np.random.default_rng(seed)
    N, dt = 1 << 18, 64e-6
    x = (rng.standard_normal(N) * 40.0).astype(np.float32)
    red = np.zeros(N, np.float32)
    assert L.orc_rednoise(x.ctypes.data, N, dt, red.ctypes.data) == 0
    unp = x.copy()
    assert L.orc_nr_unpack(unp.ctypes.data, N) == 0
    amps = (rng.standard_normal(N) * 3.0).astype(np.float32)
    T = N * dt
    f = np.ascontiguousarray(rng.uniform(30, N / 2 / T - 5, 60))
    w = np.ascontiguousarray(rng.choice([0.01, 0.07, 0.3, 2.0], 60) * (1 + rng.uniform(0, 1, 60)))
    b = np.ascontiguousarray((rng.uniform(0, 1, 60) < 0.3).astype(np.int32))
    L.orc_zap(amps.ctypes.data, N // 2, N, T, f.ctypes.data, w.ctypes.data, b.ctypes.data, 60,
              3.3e-4)
    return red, unp, amps


def test_fp_contraction_of_the_reference_build(oracle, tmp_path):
    fast = _build(tmp_path, "src_fast.so", ["-DORC_SOURCE_FORM", "-ffp-contract=fast"])
    off = _build(tmp_path, "src_off.so", ["-DORC_SOURCE_FORM", "-ffp-contract=off"])
    mine = oracle.lib()
    differs = 0
    for seed in (1, 2, 3):
        a, b, c = _outputs(mine, seed), _outputs(fast, seed), _outputs(off, seed)
        for x, y in zip(a, b):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
        differs += int(not np.array_equal(a[0].view(np.uint32), c[0].view(np.uint32)))
    assert differs == 3      # unfused rednoise differs in the last bits
