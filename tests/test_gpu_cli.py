"""coherent_search_list through the C-ABI (src/cli.c:556-1113): a .dat list
prepared by the prefetch producers while the GPU searches.

Bar: every .cohout the list run writes is byte-identical to the one-file path
of the same library run on that input alone -- realfft + rednoise on the GPU
(lengths whose half is 7-smooth) or on the host (the others, and every input
with --host-staging), then search_file, sort, truncate, duty cycles --
whatever order the producers finish in; an input that cannot be opened sets
exit status 1 and the others are still written.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import coherentsearch_c_amd as cs
from coherentsearch_c_amd import abi, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# 7-smooth halves on the multi-level host transform, one half with a prime
# factor 11 (the chirp-z path), a repeated length (the session is reused)
SIZES = [1 << 17, 2 * 3 * 5 * 7 * 2**9, 2 * 11 * 2**13, 1 << 17, 2 * 3**4 * 2**10]


def test_list_mode_with_prefetch(oracle, tmp_path):
    L = abi.host()
    dt = synth.DT
    names, series = [], []
    for i, N in enumerate(SIZES):
        if i == 1:
            x = synth.series_with_pulsar(N, 7.3e-3, 0.04, 25.0, 11 + i)
        else:
            x = np.random.default_rng(100 + i).standard_normal(N).astype(np.float32)
        base = tmp_path / ("beam_DM%d" % i)
        L.cs_write_inf(str(base.with_suffix(".inf")).encode(), N, dt, float(i), b"psr")
        L.cs_write_floats(str(base.with_suffix(".dat")).encode(),
                          x.ctypes.data_as(C.c_void_p), x.size)
        names.append(base.with_suffix(".dat").name)
        series.append(x)
    # an input that cannot be opened, in the middle of the list
    lst = tmp_path / "sweep.dat.list"
    lst.write_text("\n".join(names[:3] + ["missing_DM9.dat"] + names[3:]) + "\n")
    p = cs.make_params(nharms=32)

    def smooth7(n):
        for f in (2, 3, 5, 7):
            while n % f == 0:
                n //= f
        return n == 1

    for host in (False, True):
        outdir = tmp_path / ("out_host" if host else "out_dev")
        argv = [b"coherent_search_list", b"--gpu", b"0", b"-n", b"32", b"--blocksize", b"8192",
                b"-t", b"7", b"--outdir", str(outdir).encode(), str(lst).encode()]
        if host:
            argv.insert(1, b"--host-staging")
        arr = (C.c_char_p * len(argv))(*argv)
        say("list of", len(names) + 1, "inputs, host staging" if host else "device staging")
        assert L.cs_main_list(len(argv), arr) == 1
        assert not (outdir / "missing_DM9.cohout").exists()
        for i, (name, x) in enumerate(zip(names, series)):
            if host or not smooth7(x.size // 2):
                spec = cs.redfft(x, dt)
            else:
                g = cs.Stage(x.size, dt)
                spec = cs.Spectrum(g.run_host(x), x.size, dt)
                g.close()
            want = cs.finalize(spec, p, cs.search_file(spec, p, blocksize=8192, threshold=7.0))
            text = (outdir / name.replace(".dat", ".cohout")).read_text()
            say(name, want.size, "candidates")
            assert text == oracle.format_candidates(want.astype(oracle.CAND_DTYPE)), name
            if i == 1:
                f0 = 1.0 / 7.3e-3
                assert want.size and any(abs(f / f0 - round(f / f0)) < 2e-3
                                         for f in want["freq"][:3])


def test_reference_flow_equals_one_at_a_time():
    """The prefetching form of the reference's list mode (host staging on
    producer threads, inputs searched in order with one shared cache) gives
    each input the list it gets alone."""
    from coherentsearch_c_amd import engine as eng
    jobs = []
    for i, N in enumerate(SIZES):
        x = synth.series_with_pulsar(N, 4.1e-3 * (1 + i), 0.05, 20.0, 30 + i)
        jobs.append(dict(samples=x, dt=synth.DT, params=cs.make_params(nharms=(16, 32, 60)[i % 3]),
                         blocksize=8192, threshold=7.0))
    got, wall, stage_s = eng.run_reference_flow(jobs, producers=2, slots=2, threads=2)
    want, _ = eng.run_plain(jobs)
    say("reference flow %.3f s, staging %s" % (wall, ["%.3f" % v for v in stage_s]))
    for i, (a, b) in enumerate(zip(got, want)):
        assert a.size == b.size, i
        for f in ("r", "nharm", "metric", "ducy"):
            assert np.array_equal(a[f], b[f], equal_nan=True), (i, f)
    # producers staging on the GPU: the lists of the device-staged inputs
    # agree with the host-staged ones within the metric tolerance
    dev, wall_d, _ = eng.run_reference_flow(jobs, producers=2, slots=2, device_staging=True)
    say("device-staging producers %.3f s" % wall_d)
    for i, (a, b) in enumerate(zip(dev, want)):
        assert a.size == b.size and np.array_equal(a["nharm"], b["nharm"]), i
        assert np.all(np.abs(a["metric"] - b["metric"]) <= 1e-4 * np.maximum(np.abs(b["metric"]), 1))
