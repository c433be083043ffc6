"""De-dispersion (the prep front end's prepsubband loop) on the CPU, no GPU.

The oracle restates the reference's blocked CPU loop
(src/prep_dedisp.c:485-530, 745-790, 1104-1290; PRESTO dispersion.c,
backend_common.c, misc_utils.c).  It is pinned here by:
- the direct closed form (no blocks) the loop must equal bit for bit, over
  downsampling factors, subband counts and block lengths;
- the written length derived from the loop's structure;
- a dispersed pulse that lines up only at the injected DM;
- the textbook dispersion constant;
- choose_good_N as the smallest power of two or good factor times a power of
  ten at or above N (brute force).
The library's tables (tables-only plan, device -1) and helpers equal the
oracle's.
"""
import math

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import synth

CFGS = [  # numchan, lofreq, chan_wid, dt, P, nsub, lodm, dmstep, numdms, downsamp
    (64, 1200.0, 2.0, 64e-6, 256, 16, 40.0, 2.0, 21, 2),
    (32, 400.0, 0.25, 1e-4, 128, 32, 0.0, 1.0, 7, 1),
    (48, 700.0, 1.0, 5e-5, 192, 4, 10.0, 3.0, 12, 4),
    (16, 1400.0, 8.0, 1e-4, 256, 1, 0.0, 10.0, 5, 8),
]


def _raw(cfg, nblocks, seed, u8=False):
    nch, P = cfg[0], cfg[4]
    x = synth.filterbank_with_pulsar(P * nblocks, nch, cfg[1], cfg[2], cfg[3], 45.0, 0.031,
                                     0.05, 30.0, seed, dtype=np.uint8 if u8 else np.float32)
    return x.astype(np.float32)


@pytest.mark.parametrize("cfg", CFGS)
def test_blocked_loop_equals_closed_form(oracle, cfg):
    c = oracle.dedisp_config(*cfg)
    raw = _raw(cfg, 48, 3, u8=cfg[0] == 48)
    out, n, pad = oracle.dedisp(c, raw)
    bpr, dms, idisp, offs, numout = oracle.dedisp_tables(c, raw.shape[0])
    W = cfg[4] * bpr
    # written: work block m goes out once block m + 1 was read whole, and the
    # first read takes two raw blocks
    m = max(0, (raw.shape[0] - cfg[4]) // W - 1)
    assert n == min(m * W // cfg[9], numout) and n > 0
    direct = oracle.dedisp_direct(c, raw, n)
    assert np.array_equal(out[:, :n].view(np.uint32), direct.view(np.uint32))
    # padding: DM 0's mean, on every DM
    assert np.all(out[:, n:] == np.float32(pad))
    assert pad == pytest.approx(out[0, :n].astype(np.float64).mean(), rel=1e-6, abs=1e-6)


def test_tables_closed_forms(oracle):
    c = oracle.dedisp_config(*CFGS[0])
    bpr, dms, idisp, offs, numout = oracle.dedisp_tables(c, 256 * 48)
    assert np.allclose(dms, 40.0 + 2.0 * np.arange(21))
    cps = 4
    # within a subband the top channel has no delay and delays grow downwards;
    # across DMs the subband offsets grow with DM and vanish for the top subband
    assert np.all(idisp.reshape(16, cps)[:, -1] == 0)
    assert np.all(np.diff(idisp.reshape(16, cps), axis=1) <= 0)
    assert np.all(offs[:, -1] == 0) and np.all(np.diff(offs[:, 0]) >= 0)
    # offsets are the subband top channels' delays in downsampled samples
    f = 1200.0 + (np.arange(16) * cps + cps - 1) * 2.0
    want = np.array([[math.floor((d / (0.000241 * fi * fi) - d / (0.000241 * f[-1] ** 2))
                                 / 128e-6 + 0.5) for fi in f] for d in dms])
    assert np.array_equal(offs, want)


def test_dispersion_constant(oracle):
    # PRESTO's 1/0.000241 s MHz^2 is the textbook 4.148808e3 to 1.4e-4
    for f in (150.0, 1400.0):
        k = oracle.delay_from_dm(1.0, f) * f * f
        assert abs(k - 4148.808) / 4148.808 < 2e-4
    assert oracle.delay_from_dm(10.0, 400.0) == pytest.approx(
        10.0 * oracle.delay_from_dm(1.0, 400.0), rel=1e-15)
    assert oracle.delay_from_dm(1.0, 0.0) == 0.0


def test_pulse_lines_up_at_its_dm(oracle):
    nch, lo, cw, dt, P = 64, 1200.0, 4.0, 64e-6, 512
    c = oracle.dedisp_config(nch, lo, cw, dt, P, 16, 0.0, 20.0, 10, 1)
    period = 0.02
    raw = synth.filterbank_with_pulsar(P * 64, nch, lo, cw, dt, 100.0, period, 0.02, 40.0, 5)
    out, n, _ = oracle.dedisp(c, raw)
    nb = 64
    ph = (np.arange(n) * dt / period % 1.0 * nb).astype(int)
    peak = []
    for i in range(10):
        prof = np.bincount(ph, weights=out[i, :n].astype(np.float64), minlength=nb)
        prof /= np.bincount(ph, minlength=nb)
        peak.append((prof.max() - prof.mean()) / prof.std())
    assert int(np.argmax(peak)) == 5, peak          # DM 100 of 0, 20, ..., 180


def test_choose_good_N_is_the_smallest_good_length(oracle):
    good = [1000] + [v for v in range(1001, 10001) if v % 16 == 0 and _smooth11(v)]
    assert len(good) == 114
    rng = np.random.default_rng(1)
    ns = [1, 5, 100, 999, 1000, 1001, 1008, 12000, 12001, 2 ** 20, 2 ** 20 + 1, 851968,
          6815744] + list(rng.integers(1, 10 ** 9, 300))
    for n in ns:
        n = int(n)
        cands = [1 << k for k in range(1, 40) if (1 << k) >= n]
        cands += [g * 10 ** k for g in good for k in range(0, 8) if g * 10 ** k >= n]
        assert oracle.choose_good_N(n) == min(cands), n
        assert cs.abi.host().cs_choose_good_N(n) == oracle.choose_good_N(n)


def _smooth11(v):
    for p in (2, 3, 5, 7, 11):
        while v % p == 0:
            v //= p
    return v == 1


@pytest.mark.parametrize("cfg", CFGS)
def test_library_tables_equal_oracle(oracle, cfg):
    c = oracle.dedisp_config(*cfg)
    nspec = cfg[4] * 48
    bpr, dms, idisp, offs, numout = oracle.dedisp_tables(c, nspec)
    d = cs.Dedisp(cfg[0], cfg[1], cfg[2], cfg[3], nspec, cfg[5], cfg[6], cfg[7], cfg[8],
                  spectra_per_subint=cfg[4], downsamp=cfg[9], clip_sigma=0.0, device=-1)
    gd, gi, go = d.tables()
    info = d.info
    assert info["blocksperread"] == bpr and info["numout"] == numout
    assert np.array_equal(gd, dms) and np.array_equal(gi, idisp) and np.array_equal(go, offs)
    raw = _raw(cfg, 48, 3)
    _, n, _ = oracle.dedisp(c, raw)
    assert info["written"] == n
    for f in (150.0, 1400.0):
        assert cs.abi.host().cs_delay_from_dm(30.0, f) == oracle.delay_from_dm(30.0, f)
    d.close()


def test_rejections(oracle):
    base = dict(numchan=64, lofreq=1200.0, chan_wid=2.0, dt=64e-6, nspec=256 * 8, nsub=16,
                lodm=0.0, dmstep=1.0, numdms=4, spectra_per_subint=256, downsamp=1,
                clip_sigma=0.0)
    for bad in (dict(nsub=24),                         # does not divide 64 channels
                dict(downsamp=3),                      # not a factor of the block
                dict(nspec=256 * 8 + 5),               # not whole blocks
                dict(lodm=5000.0, spectra_per_subint=16, nspec=16 * 64)):  # subband smear > block
        with pytest.raises(ValueError):
            cs.Dedisp(**dict(base, **bad), device=-1)
        args = dict(base, **bad)
        c = oracle.dedisp_config(args["numchan"], args["lofreq"], args["chan_wid"], args["dt"],
                                 args["spectra_per_subint"], args["nsub"], args["lodm"],
                                 args["dmstep"], args["numdms"], args["downsamp"])
        if args["nspec"] % args["spectra_per_subint"] == 0:
            assert oracle.dedisp_tables(c, args["nspec"])[0] == -1
    d = cs.Dedisp(**base, device=-1)
    with pytest.raises(RuntimeError):
        d.run(np.zeros((256 * 8, 64), np.float32))
    d.close()


# ---- clip_times (clipping.c:41-215) ------------------------------------------

def _rfi(raw, nrows, seed, level=40.0):
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(raw.shape[0], nrows, replace=False))
    raw = raw.copy()
    raw[rows] += level
    return raw, rows


def test_clip_statistics_closed_forms(oracle):
    rng = np.random.default_rng(4)
    x = rng.normal(3.0, 2.0, 4097).astype(np.float32)
    m, v = oracle.avg_var(x)
    assert m == pytest.approx(x.astype(np.float64).mean(), rel=1e-13)
    assert v == pytest.approx(x.astype(np.float64).var(ddof=1), rel=1e-12)
    for n in (1, 2, 3, 4, 2048, 2049):
        y = rng.normal(size=n).astype(np.float32)
        assert oracle.lower_median(y) == np.sort(y)[(n - 1) // 2]


def test_clip_replaces_exactly_the_rfi_rows(oracle):
    cfg = CFGS[0]
    c = oracle.dedisp_config(*cfg, clip_sigma=6.0)
    raw = _raw(cfg, 48, 3)
    dirty, rows = _rfi(raw, 40, 1)
    x, levels, n = oracle.clip_all(c, dirty)
    changed = np.where(np.any(x != dirty, axis=1))[0]
    assert n == len(rows) and np.array_equal(changed, rows)
    # a replaced row holds its block's channel levels; the first block's
    # levels are the channel means of its good (unflagged) rows
    P = cfg[4]
    for r in rows:
        assert np.array_equal(x[r], levels[r // P])
    b0 = dirty[:P].astype(np.float64)
    good = np.setdiff1d(np.arange(P), rows)
    assert np.allclose(levels[0], b0[good].mean(axis=0), rtol=1e-6)
    # clip_sigma 0 leaves the data alone
    c0 = oracle.dedisp_config(*cfg, clip_sigma=0.0)
    out0, _, _ = oracle.dedisp(c0, dirty)
    outc, nw, _ = oracle.dedisp(c, dirty)
    assert not np.array_equal(out0, outc)


@pytest.mark.parametrize("cfg", CFGS[:3])
def test_clipped_blocked_loop_equals_closed_form(oracle, cfg):
    c = oracle.dedisp_config(*cfg, clip_sigma=6.0)
    dirty, _ = _rfi(_raw(cfg, 48, 3), 60, 2)
    out, n, pad = oracle.dedisp(c, dirty)
    x, _, _ = oracle.clip_all(c, dirty)
    direct = oracle.dedisp_direct(c, x, n)
    assert np.array_equal(out[:, :n].view(np.uint32), direct.view(np.uint32))


MARKSTEIN_C = r"""
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
/* the clipping kernels' quotient: q = RN(d y), r = d - q a (fused), RN(q + r y) */
static double div_rcp(double d, double a, double y)
{
    const double q = d * y;
    const double r = fma(-q, a, d);
    return fma(r, y, q);
}
static uint64_t s = 0x9e3779b97f4a7c15ull;
static uint64_t next(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double unit(void) { return (double)(next() >> 11) / 9007199254740992.0; }
typedef unsigned __int128 u128;
/* a dividend whose quotient by a lies 1/(2 a') of an ulp from a rounding
 * midpoint (a' the odd part of a), the hardest case of division by a: with
 * quotients K in [2^52, 2^53) and d = a' (2K + 1) / 2 -+ 1/2 a multiple of
 * d's ulp u, which a' (2K + 1) == +-1 (mod 2u) gives.  Its correctly
 * rounded quotient goes to *want; 0 if the construction left the binade. */
static int hard_case(int a, double *d, double *want)
{
    int sh = 0, ao = a;
    while (!(ao & 1)) { ao >>= 1; sh++; }
    const uint64_t K0 = (1ull << 52) + (next() >> 12);
    int t = 0;
    while (((u128)ao * K0) >> (53 + t)) t++;          /* d's ulp 2^t */
    const uint64_t m2u = 2ull << t;                    /* 2u */
    uint64_t inv = 1;                                  /* ao^-1 mod 2^64 */
    for (int i = 0; i < 6; i++) inv *= 2 - (uint64_t)ao * inv;
    const int up = (int)(next() & 1);
    const uint64_t r = (up ? inv : (0 - inv)) & (m2u - 1);   /* odd residue of 2K + 1 */
    const uint64_t u = m2u / 2;
    uint64_t K = (K0 / u) * u + (r - 1) / 2;
    if (K < (1ull << 52) || K >= (1ull << 53)) return 0;
    const u128 T = (u128)ao * (2 * (u128)K + 1);          /* == +-1 mod 2u */
    const u128 D = up ? (T - 1) / 2 : (T + 1) / 2;      /* quotient K + 1/2 -+ 1/(2 ao) */
    if ((D >> (53 + t)) || !(D >> (52 + t)) || (D & (u - 1))) return 0;
    const int e = (int)(next() % 120) - 60;
    *d = ldexp((double)(uint64_t)(D >> t), t + e);
    *want = ldexp((double)(up ? K : K + 1), e - sh);
    return 1;
}
int main(int argc, char **argv)
{
    const int amax = atoi(argv[1]), per = atoi(argv[2]);
    long long bad = 0, n = 0, gen_bad = 0;
    for (int a = 1; a <= amax; a++) {
        const double A = (double)a, y = 1.0 / A;
        for (int j = 0; j < per; j++) {
            double d, want;
            if (j & 1) {
                if (!hard_case(a, &d, &want)) continue;
                if (a > 1 && (a & (a - 1)) && d / A != want) gen_bad++;
            } else {
                d = ldexp(1.0 + unit(), (int)(next() % 200) - 100);
            }
            if (next() & 1) d = -d;
            n++;
            if (div_rcp(d, A, y) != d / A) bad++;
        }
    }
    printf("%lld %lld %lld\n", n, bad, gen_bad);
    return 0;
}
"""


def test_clip_reciprocal_division_is_correctly_rounded(tmp_path):
    """The clipping kernels divide AS 52's differences by the step count with
    a tabulated reciprocal and one fused correction (Markstein).  For every
    divisor of the table (1 .. 4096) that equals IEEE division, on random
    dividends over 60 decades and on the hardest dividends for each divisor
    (quotients 1/(2a) of an ulp from a rounding midpoint, whose rounding the
    construction fixes independently); the uncorrected product d * RN(1/a)
    fails on them."""
    import subprocess

    def run(source, per):
        src = tmp_path / "markstein.c"
        src.write_text(source)
        exe = str(tmp_path / "markstein")
        subprocess.run(["gcc", "-O2", "-march=x86-64-v3", "-ffp-contract=off", "-o", exe,
                        str(src), "-lm"], check=True)
        out = subprocess.run([exe, "4096", str(per)], check=True, capture_output=True, text=True)
        return map(int, out.stdout.split())

    n, bad, gen_bad = run(MARKSTEIN_C, 2000)
    assert n > 4096 * 1900 and bad == 0 and gen_bad == 0
    n, bad, _ = run(MARKSTEIN_C.replace("return fma(r, y, q);", "(void)r; return q;"), 200)
    assert bad > 10000
