"""De-dispersion on the GPU through the C-ABI (cs_gpu_dedisp_*), against the
oracle's restatement of the reference's blocked prepsubband loop
(src/prep_dedisp.c:485-530, 745-790, 1104-1427).

Bar: every DM series bit-identical to the oracle's (the written samples and
the padding), for float32 and uint8 spectra, from host memory and from HBM,
in one device pass and in several.  Then the prep front end's search
(src/prep_search.c:44-145) on the resident series: the final lists equal
those of the same series searched from host memory, and the pulsar is found
at the DM it was injected at.
"""
import sys

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import synth
from coherentsearch_c_amd.engine import Engine

pytestmark = pytest.mark.gpu

# numchan, lofreq, chan_wid, dt, P, nsub, lodm, dmstep, numdms, downsamp, nblocks, uint8
CASES = [
    (64, 1200.0, 2.0, 64e-6, 256, 16, 40.0, 2.0, 21, 2, 48, False),
    (48, 700.0, 1.0, 5e-5, 192, 4, 10.0, 3.0, 12, 4, 40, True),
    (16, 1400.0, 8.0, 1e-4, 256, 1, 0.0, 10.0, 5, 8, 30, False),
    # 2^20 spectra: one device pass, and 16 with passes of 2^16 (CHUNKED)
    (64, 1300.0, 3.0, 64e-6, 1024, 16, 0.0, 4.0, 32, 1, 1024, True),
    # bytes loaded one by one (3-channel subbands), three DM tile groups
    (30, 1100.0, 2.0, 64e-6, 256, 10, 5.0, 1.5, 70, 2, 64, True),
    # a downsampling factor that does not divide a wave (untiled subbands)
    (32, 900.0, 1.0, 64e-6, 192, 8, 0.0, 2.0, 9, 3, 48, False),
    # raw blocks of 16384 spectra (clipping: the radix-select median)
    (16, 1400.0, 8.0, 1e-4, 16384, 4, 0.0, 10.0, 5, 1, 8, True),
    # raw blocks of 200 spectra: clipping's group loops end in a partial group
    (24, 1250.0, 2.0, 64e-6, 200, 8, 10.0, 2.0, 11, 2, 60, True),
    # 1024 and 2048 uint8 channels (clipping's one-read statistics, a spectrum
    # over one wave and over two)
    (1024, 1200.0, 0.29296875, 64e-6, 2048, 64, 0.0, 2.0, 8, 4, 8, True),
    (2048, 1200.0, 0.146484375, 64e-6, 512, 32, 0.0, 2.0, 6, 2, 8, True),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _plan(c, nspec, numout=0):
    return cs.Dedisp(c[0], c[1], c[2], c[3], nspec, c[5], c[6], c[7], c[8],
                     spectra_per_subint=c[4], downsamp=c[9], numout=numout, clip_sigma=0.0)


CHUNKED = [(CASES[3], "tiled"), (CASES[3], "untiled"), (CASES[4], "tiled")]


# every case in one device pass with both kernel sets; the CHUNKED ones also
# in passes of 2^16 spectra
SERIES = [pytest.param(c, k, 0, id="c%d-%s-0" % (i, k))
          for i, c in enumerate(CASES) for k in ("tiled", "untiled")] + \
         [pytest.param(c, k, 1 << 16, id="c%d-%s-65536" % (CASES.index(c), k))
          for c, k in CHUNKED]


@pytest.mark.parametrize("c,kernels,chunk", SERIES)
def test_series_match_oracle(oracle, c, kernels, chunk, monkeypatch):
    # the untiled kernels serve shapes the LDS tiles do not fit
    monkeypatch.setenv("CS_DEDISP_SIMPLE", "1" if kernels == "untiled" else "0")
    monkeypatch.setenv("CS_DEDISP_CHUNK", str(chunk))
    nspec = c[4] * c[10]
    u8 = c[11]
    raw = synth.filterbank_with_pulsar(nspec, c[0], c[1], c[2], c[3], 37.0, 0.0213, 0.04, 25.0,
                                       7, dtype=np.uint8 if u8 else np.float32)
    say("dedisp %s: %d chans x %d spectra (%s), nsub %d, %d DMs, downsamp %d"
        % (kernels, c[0], nspec, raw.dtype, c[5], c[8], c[9]))
    want, n, pad = oracle.dedisp(oracle.dedisp_config(*c[:10]), raw.astype(np.float32))
    d = _plan(c, nspec)
    info = d.info
    say("  written %d of %d, %d passes" % (n, info["numout"], -(-n * c[9] // info["chunk"])))
    assert info["written"] == n and info["numout"] == want.shape[1]
    d.run(raw)
    got = d.copy_out()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert d.info["pad"] == pad
    # the same spectra already resident in HBM
    import torch
    t = torch.from_numpy(raw).to("cuda:0")

    class Resident:
        ptr = t.data_ptr()
        uint8 = u8

    d.run(Resident())
    torch.cuda.synchronize()
    assert np.array_equal(d.copy_out().view(np.uint32), want.view(np.uint32))
    total, ksub, kdm = d.last_ms()
    say("  device %.3f ms (subband %.3f, DM %.3f)" % (total, ksub, kdm))
    d.close()


CLIPPED = [  # (case, passes of 2^16 spectra or one pass)
    (CASES[0], 0), (CASES[1], 0), (CASES[3], 1 << 16), (CASES[5], 0), (CASES[6], 0),
    (CASES[7], 0), (CASES[8], 0), (CASES[9], 0)]


@pytest.mark.parametrize("kernels", ["tiled", "untiled"])
@pytest.mark.parametrize("case", range(len(CLIPPED)))
def test_clipped_series_match_oracle(oracle, case, kernels, monkeypatch):
    """prepsubband's default time-domain clipping (clip_times at 6 sigma) of
    every raw block, on spectra with broadband RFI bursts"""
    c, chunk = CLIPPED[case]
    monkeypatch.setenv("CS_DEDISP_SIMPLE", "1" if kernels == "untiled" else "0")
    monkeypatch.setenv("CS_DEDISP_CHUNK", str(chunk))
    nspec = c[4] * c[10]
    u8 = c[11]
    raw = synth.filterbank_with_pulsar(nspec, c[0], c[1], c[2], c[3], 37.0, 0.0213, 0.04, 25.0,
                                       13, dtype=np.uint8 if u8 else np.float32)
    rng = np.random.default_rng(5)
    rows = rng.choice(nspec, max(8, nspec // 500), replace=False)
    if u8:
        raw[rows] = np.minimum(raw[rows].astype(np.int32) + 60, 255).astype(np.uint8)
    else:
        raw[rows] += 40.0
    cfg = oracle.dedisp_config(*c[:10], clip_sigma=6.0)
    want, n, pad = oracle.dedisp(cfg, raw.astype(np.float32))
    _, _, nclip = oracle.clip_all(cfg, raw.astype(np.float32))
    say("clip %s: %d chans x %d spectra (%s), %d rows clipped, written %d"
        % (kernels, c[0], nspec, raw.dtype, nclip, n))
    assert nclip > 0
    d = cs.Dedisp(c[0], c[1], c[2], c[3], nspec, c[5], c[6], c[7], c[8],
                  spectra_per_subint=c[4], downsamp=c[9], clip_sigma=6.0)
    d.run(raw)
    assert np.array_equal(d.copy_out().view(np.uint32), want.view(np.uint32))
    assert d.info["pad"] == pad
    import torch
    t = torch.from_numpy(raw).to("cuda:0")

    class Resident:
        ptr = t.data_ptr()
        uint8 = u8

    d.run(Resident())
    torch.cuda.synchronize()
    assert np.array_equal(d.copy_out().view(np.uint32), want.view(np.uint32))
    d.close()


def test_numout_shorter_and_longer(oracle):
    c = CASES[0]
    nspec = c[4] * c[10]
    raw = synth.filterbank_with_pulsar(nspec, c[0], c[1], c[2], c[3], 50.0, 0.0213, 0.04, 25.0,
                                       9)
    for numout in (1000, 9000):
        cfg = oracle.dedisp_config(*c[:10], numout=numout)
        want, n, pad = oracle.dedisp(cfg, raw)
        d = _plan(c, nspec, numout=numout)
        d.run(raw)
        say("numout %d: written %d" % (numout, n))
        assert d.info["written"] == n and want.shape[1] == numout
        assert np.array_equal(d.copy_out().view(np.uint32), want.view(np.uint32))
        d.close()


def test_prep_search_on_resident_series(oracle, tmp_path):
    """de-dispersion -> device staging -> search -> duty cycles, the series
    never leaving HBM; against the same series searched from host memory"""
    nch, lo, cw, dt, P = 64, 1200.0, 4.0, 64e-6, 2048
    nspec = P * 512                         # 2^20 spectra, 67 s
    period, dm0 = 0.0113, 60.0
    raw = synth.filterbank_with_pulsar(nspec, nch, lo, cw, dt, dm0, period, 0.03, 60.0, 11,
                                       dtype=np.uint8)
    d = cs.Dedisp(nch, lo, cw, dt, nspec, 16, 0.0, 10.0, 13, spectra_per_subint=P,
                  downsamp=2)
    d.run(raw)
    dms, _, _ = d.tables()
    p = cs.make_params(nharms=32)
    say("prep search: %d DMs x %d samples" % (d.numdms, d.numout))
    with Engine(device=0, nslots=4) as eng:
        dev = cs.prep_search(d, eng, p, ncands=20, threshold=6.0, outfile=str(tmp_path / "OUT"))
        host_series = d.copy_out()
        tickets = [eng.submit(host_series[i], d.dt_out, p, blocksize=65536, threshold=6.0,
                              ncands=20) for i in range(d.numdms)]
        host = [eng.wait(t)[0] for t in tickets]
    for i in range(d.numdms):
        assert dev[i].size == host[i].size, i
        for f in ("r", "nharm", "metric", "ducy"):
            assert np.array_equal(dev[i][f], host[i][f], equal_nan=True), (i, f)
        assert (tmp_path / ("OUT_DM%.2f.cohout" % dms[i])).exists()
    # the pulsar: strongest at its DM
    f0 = 1.0 / period
    best = []
    for i in range(d.numdms):
        hit = [m for fr, m in zip(dev[i]["freq"], dev[i]["metric"])
               if abs(fr / f0 - round(fr / f0)) < 2e-3 and round(fr / f0) >= 1]
        best.append(max(hit) if hit else 0.0)
    say("  best pulsar metric per DM:", ["%.1f" % b for b in best])
    assert dms[int(np.argmax(best))] == dm0
    d.close()
