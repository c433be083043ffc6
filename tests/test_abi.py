"""CPU checks of the C-ABI libraries: they load, export every symbol the
public headers declare, keep the reference's struct layouts, and their host
geometry, collapse, staging and output agree with the oracle.  No GPU calls
except cs_gpu_available's no-device answer."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import abi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"CS_API\s+[\w\s\*]+?\b(\w+)\s*\(", txt))


def test_headers_match_exports():
    hip_decl = _declared("cs_gpu.h") | _declared("cs_dedisp.h")
    host_decl = _declared("cs_host.h")
    assert hip_decl == set(abi.HIP_SYMBOLS)
    assert host_decl == set(abi.HOST_SYMBOLS)
    hip = abi.hip()
    host = abi.host()
    for s in hip_decl:
        assert hasattr(hip, s), s
    for s in host_decl:
        assert hasattr(host, s), s


def test_only_declared_symbols_exported():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", abi.HIP_LIB], capture_output=True,
                         text=True, check=True).stdout
    funcs = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert funcs == set(abi.HIP_SYMBOLS)


def test_struct_layouts():
    assert C.sizeof(abi.FFTFile) == 168 and abi.FFTFile.amps.offset == 128
    assert abi.FFTFile.namps.offset == 136 and abi.FFTFile.T.offset == 96
    assert abi.SearchParams.decimations.offset == 60 and abi.SearchParams.sigma.offset == 320
    assert abi.DirectPlan.row.offset == 48 and abi.DirectPlan.gA_im.offset == 136
    assert abi.Candidate.ducy.offset == 32
    assert C.sizeof(abi.CsSearchRequest) == 120 and abi.CsSearchRequest.amps_out.offset == 64
    assert abi.CsSearchRequest.zaplist.offset == 104 and abi.CsSearchRequest.baryv.offset == 112
    assert C.sizeof(abi.ZapBird) == 24 and C.sizeof(abi.ZapList) == 16
    assert C.sizeof(abi.CsTuneResult) == 48 and abi.CsTuneResult.default_ms.offset == 32
    assert abi.CsSearchRequest.lobin.offset == 48 and abi.CsSearchRequest.threshold.offset == 56
    assert abi.CsSearchRequest.ncands.offset == 72 and abi.CsSearchRequest.harm_remove.offset == 84
    assert abi.CsSearchRequest.dr_tol.offset == 88 and abi.CsSearchRequest.harm_tol.offset == 96
    assert C.sizeof(abi.CsSearchInfo) == 96 and abi.CsSearchInfo.kernel_ms.offset == 56
    assert abi.CsSearchInfo.search_ms.offset == 80 and abi.CsSearchInfo.post_ms.offset == 88
    assert C.sizeof(abi.CsEngineStats) == 72 and abi.CsEngineStats.completed.offset == 48
    assert C.sizeof(abi.CsDedispConfig) == 80 and abi.CsDedispConfig.numout.offset == 64
    assert abi.CsDedispConfig.clip_sigma.offset == 72
    assert abi.CsDedispConfig.lodm.offset == 40 and abi.CsDedispConfig.downsamp.offset == 60
    assert C.sizeof(abi.CsDedispInfo) == 56 and abi.CsDedispInfo.pad.offset == 48
    assert abi.CsEngineStats.plans_resident.offset == 64 and abi.CsSearchInfo.reused.offset == 4


def test_no_device_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    rc, msg = cs.gpu_available(0)
    assert rc == -1 and msg


@pytest.mark.parametrize("nh,m,lobin,blocksize", [(60, 16, 100, 2048), (16, 32, 3, 512),
                                                  (32, 16, 100, 65536), (37, 8, 100, 1000)])
def test_plans_and_chunks_match_oracle(oracle, nh, m, lobin, blocksize):
    L = abi.host()
    nbins = 1 << 16
    amps = synth.spectrum_with_pulsar(nbins, 1000.3, 4, 0.1, 5.0, 1)
    spec = cs.Spectrum(amps, 2 * nbins, synth.DT)
    p = cs.make_params(nharms=nh, m=m)
    o = oracle.Oracle(amps.view(np.float32), spec.N, spec.T,
                      oracle.config(nharms=nh, m=m, lobin=lobin, blocksize=blocksize))
    r_lo = o.r_lo
    plans = abi.DirectPlanSet()
    assert L.direct_plans_build(C.byref(plans), C.byref(p), r_lo) == 0
    csp = C.POINTER(C.c_longlong)()
    nc = C.c_longlong()
    assert L.chunk_starts(C.byref(plans), C.byref(spec.ft), C.byref(p), o.total, blocksize,
                          C.byref(csp), C.byref(nc)) == 0
    got = np.ctypeslib.as_array(csp, shape=(nc.value,)).copy()
    L.free_(C.cast(csp, C.c_void_p))
    assert np.array_equal(got, o.chunk_starts())
    for h in range(1, nh + 1):
        pl = plans.plans[h - 1]
        W = np.zeros(pl.P * m, np.float32)
        A = np.zeros(pl.P * 2, np.float32)
        carry = np.zeros(pl.P, np.uint8)
        assert oracle.lib().orc_plan_tables(o.h, h, W.ctypes.data, A.ctypes.data,
                                            carry.ctypes.data, None) == pl.P
        assert np.array_equal(W.view(np.uint32),
                              np.ctypeslib.as_array(pl.W, (pl.P * m,)).view(np.uint32))
        assert np.array_equal(A[0::2], np.ctypeslib.as_array(pl.A_re, (pl.P,)))
        assert np.array_equal(A[1::2], np.ctypeslib.as_array(pl.A_im, (pl.P,)))
        assert np.array_equal(carry, np.ctypeslib.as_array(pl.carry, (pl.P,)))
    L.direct_plans_free(C.byref(plans))


def _rand_cands(rng, n):
    r = np.sort(rng.uniform(100, 3000, n))
    r[::7] = r[::7] * 2.0            # harmonic pairs
    r[1::5] = r[:-1:5] + rng.uniform(0, 1.5, r[1::5].size)   # near duplicates
    m = np.round(rng.uniform(8, 30, n), 2)
    a = np.zeros(n, dtype=cs.CAND_DTYPE)
    a["r"], a["freq"], a["metric"] = r, r / 100.0, m
    a["nharm"] = rng.choice([16, 32, 60], n)
    a["ducy"] = np.nan
    return a


def test_collapse_matches_oracle(oracle):
    L = abi.host()
    rng = np.random.default_rng(1)
    for _ in range(5):
        a = _rand_cands(rng, 300)
        cl = cs.to_list(a)
        d = L.remove_duplicates(C.byref(cl), 1.0)
        h = L.remove_harmonics(C.byref(d), 16, 1.0)
        L.candlist_sort_metric_desc(C.byref(h))
        got = cs.take_list(h)
        L.candlist_free(C.byref(cl))
        L.candlist_free(C.byref(d))
        want = oracle.collapse(a.astype(oracle.CAND_DTYPE), ncands=10 ** 9)
        assert np.array_equal(got[["r", "metric", "nharm"]], want[["r", "metric", "nharm"]])


def test_ducy_and_cohout_match_oracle(oracle, tmp_path):
    nbins = 1 << 15
    amps = synth.spectrum_with_pulsar(nbins, 1234.5, 32, 0.1, 40.0, 3)
    spec = cs.Spectrum(amps, 2 * nbins, synth.DT)
    a = np.zeros(3, dtype=cs.CAND_DTYPE)
    a["r"] = [1234.5, 617.25, 2469.0]
    a["freq"] = a["r"] / spec.T
    a["metric"] = [30.0, 12.0, 9.5]
    a["nharm"] = [32, 60, 16]
    p = cs.make_params()
    got = cs.finalize(spec, p, a)
    want = oracle.measure_ducy(amps.view(np.float32), spec.N, a.astype(oracle.CAND_DTYPE))
    np.testing.assert_allclose(got["ducy"], want["ducy"], rtol=1e-12)
    out = tmp_path / "x.cohout"
    cl = cs.to_list(got)
    assert abi.host().write_candidates(C.byref(cl), str(out).encode(), 8.0) == 0
    abi.host().candlist_free(C.byref(cl))
    assert out.read_text() == oracle.format_candidates(want.astype(oracle.CAND_DTYPE))


def test_redfft_matches_oracle(oracle):
    x, dt = synth.config2(N=1 << 18)
    spec = cs.redfft(x, dt)
    ref = oracle.realfft_rednoise(x, dt)
    d = np.abs(spec.amps - ref)
    assert np.max(d) <= 1e-5 * np.max(np.abs(ref))
    assert np.mean(d == 0) > 0.9


def test_fftfile_roundtrip(tmp_path):
    L = abi.host()
    nbins = 4096
    amps = synth.spectrum_with_pulsar(nbins, 100.5, 4, 0.1, 10.0, 1)
    base = tmp_path / "obs"
    L.cs_write_inf(str(base.with_suffix(".inf")).encode(), 2 * nbins, synth.DT, 12.5, b"psr")
    a = amps.view(np.float32)
    L.cs_write_floats(str(base.with_suffix(".fft")).encode(), a.ctypes.data_as(C.c_void_p), a.size)
    ft = abi.FFTFile()
    assert L.fftfile_open(C.byref(ft), str(base.with_suffix(".fft")).encode()) == 0
    assert ft.N == 2 * nbins and ft.namps == nbins and ft.T == pytest.approx(2 * nbins * synth.DT)
    assert ft.inf.DM == 12.5
    got = np.ctypeslib.as_array(ft.amps, (2 * nbins,))
    assert np.array_equal(got, a)
    L.fftfile_close(C.byref(ft))
    assert L.fftfile_open(C.byref(ft), str(tmp_path / "missing.fft").encode()) == -1


def _birds(rng, n, namps):
    out = []
    for i in range(n):
        f = rng.uniform(10, namps + 50)
        w = rng.choice([0.05, 0.8, 3.0, 17.0, 60.0])
        out.append((f, w, int(rng.random() < 0.3)))
    out += [(namps - 3.0, 20.0, 0), (320.0, 30.0, 0), (301.2, 0.2, 1)]
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_zaplist_apply_matches_oracle(oracle, seed):
    """zaplist_apply (host library) against the oracle: same bins, same
    levels, bit for bit, with overlapping, narrow, barycentric, near-DC and
    past-the-top birds."""
    rng = np.random.default_rng(seed)
    nbins = 1 << 14
    amps = synth.spectrum_with_pulsar(nbins, 123.4, 8, 0.1, 10.0, seed)
    spec = cs.Spectrum(amps.copy(), 2 * nbins, synth.DT)
    birds = _birds(rng, 40, nbins)
    # frequencies in Hz for this T
    birds = [(f / spec.T, w / spec.T, b) for f, w, b in birds]
    n = cs.zap(spec, birds, baryv=2e-4)
    want, n_want = oracle.zap(amps.view(np.float32), spec.N, spec.T, birds, baryv=2e-4)
    assert n == n_want and n > 10
    assert np.array_equal(spec.amps.view(np.float32), want)


def test_zaplist_read(tmp_path):
    p = tmp_path / "birds.zaplist"
    p.write_text("# freq width\n\n60.0 0.5\nB 100.5\t0.25\n   \nabc def\n50 -1\n0 2\n"
                 "B0.5 0.1\n120.25 1e-2 trailing words\n")
    z = cs.Zaplist.read(p)
    assert z.tuples() == [(60.0, 0.5, 0), (100.5, 0.25, 1), (0.5, 0.1, 1), (120.25, 0.01, 0)]
    with pytest.raises(OSError):
        cs.Zaplist.read(tmp_path / "missing")


def test_inf_fields_and_errors(tmp_path, capfd):
    """.inf rules of the reference's reader (src/fileio.c:38-72): a key matches
    at the start of the raw line (leading space included), the value is what
    follows the last '=' trimmed, a later line wins, a line without '=' gives
    the whole trimmed line, other keys and unindented keys are ignored; a
    missing 'Number of bins' or 'Width' fails fftfile_open with its message."""
    L = abi.host()
    L.inf_read.argtypes = [C.c_char_p, C.POINTER(abi.SimpleInf)]
    L.inf_read.restype = C.c_int
    L.inf_free.argtypes = [C.POINTER(abi.SimpleInf)]
    p = tmp_path / "x.inf"
    p.write_bytes(b" Object being observed   =  J0000+00 = B1937 \r\n"
                  b" Number of bins in the time series = 1000\n"
                  b"Number of bins in the time series = 7\n"
                  b" Number of bins in the time series = 4096\n"
                  b" Epoch of observation (MJD) = 59000.5\n"
                  b" Width of each time series bin (sec)    =   6.4e-05   \n"
                  b" Dispersion measure  12.5")
    inf = abi.SimpleInf()
    assert L.inf_read(str(p).encode(), C.byref(inf)) == 0
    assert inf.object == b"B1937" and inf.N == 4096 and inf.has_N == 1
    assert inf.dt == 6.4e-05 and inf.has_dt == 1 and inf.epoch == 59000.5
    assert inf.has_DM == 1 and inf.DM == 0.0     # no '=': strtod of the whole line
    L.inf_free(C.byref(inf))
    assert L.inf_read(str(tmp_path / "none.inf").encode(), C.byref(inf)) == -1
    L.inf_free(C.byref(inf))
    for text, msg in ((b" Width of each time series bin = 1e-4\n", "Missing 'Number of bins'"),
                      (b" Number of bins = 8\n", "Missing 'Width of each time series bin'")):
        (tmp_path / "y.inf").write_bytes(text)
        (tmp_path / "y.fft").write_bytes(np.zeros(8, np.float32).tobytes())
        ft = abi.FFTFile()
        capfd.readouterr()
        assert L.fftfile_open(C.byref(ft), str(tmp_path / "y.fft").encode()) == -1
        assert msg in capfd.readouterr().err
        L.fftfile_close(C.byref(ft))
    # a .fft path shorter than its suffix: the .inf is ".inf" itself
    ft = abi.FFTFile()
    assert L.fftfile_open(C.byref(ft), b"abc") == -1
    assert "The .inf file '.inf' was not found." in capfd.readouterr().err
    L.fftfile_close(C.byref(ft))
