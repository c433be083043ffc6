"""Pins for the CPU oracle (oracle/), which restates the reference's CPU path.

The reference ships no golden vectors or known-answer tests for this path, and
it cannot be built here (every translation unit includes <fftw3.h>), so each
oracle function is pinned by mathematics chosen so that a plausible mistake
(dropped term, wrong sign or index, transposed operand) fails: numpy's FFT,
brute-force sums on small inputs, closed forms, exact special cases and
Monte-Carlo checks of the noise model.
"""
import math

import numpy as np
import pytest

from coherentsearch_c_amd import synth


def test_c2r_matches_numpy_irfft(oracle):
    rng = np.random.default_rng(3)
    # prime powers (Cooley-Tukey with twiddles), coprime products (prime-factor
    # split), and mixtures of both
    for H in [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 15, 16, 20, 24, 25, 27, 30, 32,
              36, 37, 45, 48, 49, 56, 60, 63, 64]:
        X = rng.standard_normal(H + 1) + 1j * rng.standard_normal(H + 1)
        ref = np.fft.irfft(X, 2 * H) * (2 * H)
        x64 = oracle.c2r_f64(X)
        assert np.max(np.abs(x64 - ref)) <= 1e-12 * np.max(np.abs(ref)), H
        x32 = oracle.c2r_f32(X.astype(np.complex64))
        assert np.max(np.abs(x32 - ref)) <= 3e-6 * np.max(np.abs(ref)), H


def test_c2r_f64_follows_the_written_operation_sequence(oracle):
    """Bit-for-bit against an independent replay of the definition (exact
    fma, separate libm cos/sin): prime-factor splits, odd-prime butterflies,
    radix-4/2 stages and the half-length pre-processing all in the order the
    definition states.  A merged sincos twiddle or a reordered sum fails."""
    import c2r_spec_replay as sp
    rng = np.random.default_rng(9)
    for H in (3, 5, 6, 12, 15, 16, 30, 45, 60):
        for _ in range(6):
            X = (rng.standard_normal(H + 1) + 1j * rng.standard_normal(H + 1))
            X = X.astype(np.complex64).astype(np.complex128)
            want = sp.c2r([(float(v.real), float(v.imag)) for v in X])
            got = oracle.c2r_f64(X)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), H


def test_c2r_ignores_dc_nyquist_imag_and_brute_force(oracle):
    rng = np.random.default_rng(4)
    for H in range(1, 7):
        X = rng.standard_normal(H + 1) + 1j * rng.standard_normal(H + 1)
        n = 2 * H
        brute = np.zeros(n)
        for t in range(n):
            s = X[0].real + X[H].real * (-1) ** t
            for h in range(1, H):
                s += 2 * (X[h] * np.exp(2j * np.pi * h * t / n)).real
            brute[t] = s
        x = oracle.c2r_f64(X)
        np.testing.assert_allclose(x, brute, rtol=0, atol=1e-12 * np.max(np.abs(brute)))
        Y = X.copy()
        Y[0] += 5j
        Y[H] -= 7j
        assert np.array_equal(oracle.c2r_f64(Y), x)


def test_rational_grid(oracle):
    import ctypes as C
    L = oracle.lib()
    p, q = C.c_int(), C.c_int()
    for lodr, pq in [(0.5 / 60, (1, 120)), (0.5 / 32, (1, 64)), (0.5 / 16, (1, 32)),
                     (0.5 / 37, (1, 74)), (0.3 / 60, (1, 200)), (0.75 / 10, (3, 40))]:
        assert L.orc_trial_grid_rational(lodr, C.byref(p), C.byref(q)) == 1
        assert (p.value, q.value) == pq
    assert L.orc_trial_grid_rational(1.0 / math.pi, C.byref(p), C.byref(q)) == 0
    assert L.orc_trial_grid_rational(-1.0, C.byref(p), C.byref(q)) == 0


def _small_oracle(oracle, nh=16, m=16, nbins=1 << 15, lobin=100, blocksize=2048, seed=1,
                  maxdecim=6, sigma="analytic", precision="f32", r0=271.3):
    amps = synth.spectrum_with_pulsar(nbins, r0, nh, 0.05, 12.0, seed)
    N = 2 * nbins
    cfg = oracle.config(nharms=nh, m=m, lobin=lobin, blocksize=blocksize, maxdecim=maxdecim,
                        sigma=sigma, precision=precision)
    return oracle.Oracle(amps.view(np.float32), N, N * synth.DT, cfg), amps


def test_plan_closed_forms(oracle):
    import ctypes as C
    o, _ = _small_oracle(oracle, nh=60)
    for h in range(1, 61):
        rf, dr0, P, s, ba = C.c_int(), C.c_double(), C.c_int(), C.c_int(), C.c_int()
        oracle.lib().orc_plan_info(o.h, h, C.byref(rf), C.byref(dr0), C.byref(P), C.byref(s),
                                   C.byref(ba))
        assert P.value == 120 // math.gcd(h, 120)
        assert (s.value, ba.value) == (h % 120, h // 120)
        W = np.zeros(P.value * 16, np.float32)
        A = np.zeros(P.value * 2, np.float32)
        carry = np.zeros(P.value, np.uint8)
        oracle.lib().orc_plan_tables(o.h, h, W.ctypes.data, A.ctypes.data, carry.ctypes.data,
                                     None)
        W = W.reshape(P.value, 16)
        # the weights' sum of sinc kernel magnitudes is bounded and finite
        assert np.all(np.isfinite(W)) and np.all(np.isfinite(A))
        # |A| = |sin(pi dr)| / pi <= 1/pi
        assert np.all(np.hypot(A[0::2], A[1::2]) <= 1 / np.pi + 1e-7) or \
            np.any((A[0::2] == 1) & (A[1::2] == 0))


def test_on_bin_trials_are_exact_bins(oracle):
    """A trial whose harmonic lands exactly on a bin returns that bin."""
    nbins = 1 << 14
    rng = np.random.default_rng(7)
    amps = ((rng.standard_normal(nbins) + 1j * rng.standard_normal(nbins)) / 2).astype(np.complex64)
    N = 2 * nbins
    # lobin=200 with lodr = 1/32: trial t sits at r = 200 + t/32, on-bin when t % 32 == 0
    cfg = oracle.config(nharms=16, lobin=200, blocksize=4096, hifreq=190.0)
    o = oracle.Oracle(amps.view(np.float32), N, N * synth.DT, cfg)
    st, filled = o.stack(0)
    i0, n, _ = o.chunk(0)
    for j in range(0, n, 32):
        r = 200 + (i0 + j) / 32
        for h in range(1, 17):
            if not filled[h - 1]:
                continue
            b = int(round(h * r))
            assert st[h, j, 0] == amps[b].real and st[h, j, 1] == amps[b].imag


@pytest.mark.parametrize("nh,m,lobin", [(16, 16, 100), (32, 16, 3), (60, 16, 100),
                                        (60, 32, 100), (7, 8, 2)])
def test_group_form_equals_direct_form(oracle, nh, m, lobin):
    o, _ = _small_oracle(oracle, nh=nh, m=m, lobin=lobin)
    for ci in sorted({0, 1, o.nchunks // 3, o.nchunks - 1}):
        g, fg = o.stack(ci, 0)
        d, fd = o.stack(ci, 1)
        assert np.array_equal(fg, fd)
        assert np.array_equal(g.view(np.uint32), d.view(np.uint32)), ci


def test_interpolation_vs_exact_sinc(oracle):
    """f32 tabulated interpolation vs the FP64 direct-sinc formula."""
    import ctypes as C
    o, amps = _small_oracle(oracle, nh=32, m=16)
    a = amps.view(np.float32)
    re, im = C.c_double(), C.c_double()
    ci = o.nchunks // 2
    st, filled = o.stack(ci)
    i0, n, _ = o.chunk(ci)
    worst = 0.0
    for j in range(0, n, max(n // 40, 1)):
        r = o.r_lo + (i0 + j) * o.lodr
        for h in range(1, 33):
            if not filled[h - 1]:
                continue
            oracle.lib().orc_fourier_interp_complex(h * r, a.ctypes.data, amps.size, 16,
                                                    C.byref(re), C.byref(im))
            ref = complex(re.value, im.value)
            got = complex(st[h, j, 0], st[h, j, 1])
            worst = max(worst, abs(got - ref) / max(abs(ref), 1e-3))
    assert worst < 2e-5


def _nonrational_oracle(oracle, nh=16, m=16, hidr=0.4999, nbins=1 << 14, lobin=100,
                        blocksize=2048, precision="f32"):
    amps = synth.spectrum_with_pulsar(nbins, 271.3, nh, 0.05, 12.0, 4)
    N = 2 * nbins
    cfg = oracle.config(nharms=nh, m=m, hidr=hidr, lobin=lobin, blocksize=blocksize,
                        precision=precision)
    return oracle.Oracle(amps.view(np.float32), N, N * synth.DT, cfg), amps


@pytest.mark.parametrize("nh,m,hidr", [(16, 16, 0.4999), (7, 8, 1.0 / math.sqrt(3.0)),
                                       (32, 32, 0.5 * math.pi / 3.0)])
def test_nonrational_step_is_the_fp64_sinc_rounded(oracle, nh, m, hidr):
    """A trial step that is no simple rational (P = 0): every trial's position
    in FP64 and the m-tap kernel 1/(dr - k) times (sin cos, -sin^2)/pi in
    double.  That is sinc(off) e^{-i pi off} summed over the same taps, written
    differently, so the stack equals the FP64 direct-sinc formula (a separate
    function: sin and cos of pi*off per tap) rounded to float, up to the
    double rounding of the two forms: within 1 ulp of the float, exact for
    nearly every value.  A wrong tap offset, sign or position fails by orders
    of magnitude."""
    import ctypes as C
    o, amps = _nonrational_oracle(oracle, nh=nh, m=m, hidr=hidr)
    assert not o.L.orc_trial_grid_rational(hidr / nh, C.byref(C.c_int()), C.byref(C.c_int()))
    a = amps.view(np.float32)
    re, im = C.c_double(), C.c_double()
    exact = total = 0
    for ci in sorted({0, o.nchunks // 2, o.nchunks - 1}):
        st, filled = o.stack(ci)
        i0, n, _ = o.chunk(ci)
        assert not np.any(st[0])                      # the DC row
        for h in range(1, nh + 1):
            rf, dr0, P, sdum, ba = C.c_int(), C.c_double(), C.c_int(), C.c_int(), C.c_int()
            o.L.orc_plan_info(o.h, h, C.byref(rf), C.byref(dr0), C.byref(P), C.byref(sdum),
                              C.byref(ba))
            assert P.value == 0
            if not filled[h - 1]:
                assert not np.any(st[h])
                continue
            lodr = hidr / nh
            for j in range(0, n, max(n // 50, 1)):
                pos = (rf.value + dr0.value) + (i0 + j) * (h * lodr)
                o.L.orc_fourier_interp_complex(pos, a.ctypes.data, amps.size, m,
                                               C.byref(re), C.byref(im))
                want = np.array([re.value, im.value], dtype=np.float32)
                got = st[h, j]
                ulp = np.spacing(np.maximum(np.abs(want), np.float32(1e-30)))
                assert np.all(np.abs(got - want) <= ulp), (h, j, got, want)
                exact += int(np.array_equal(got, want))
                total += 1
    assert total > 100 and exact >= 0.98 * total


def test_nonrational_chunks_and_fits(oracle):
    """P = 0 fit sets from FP64 positions: constant inside every chunk, chunk
    starts at the change points and every blocksize, and the same chunk table
    as the host library's own restatement (core.c) of the same rule."""
    import ctypes as C
    from coherentsearch_c_amd import abi
    o, amps = _nonrational_oracle(oracle, nh=16, lobin=2, nbins=1 << 12, blocksize=512)
    cs = list(o.chunk_starts()) + [o.total]
    fits = np.zeros((16, o.total), dtype=bool)
    for h in range(1, 17):
        for t in range(o.total):
            fits[h - 1, t] = o.L.orc_harmonic_fits(o.h, h, t)
    assert not fits.all() and fits.any()
    for a_, b_ in zip(cs[:-1], cs[1:]):
        assert b_ > a_
        assert np.all(fits[:, a_:b_] == fits[:, a_:a_ + 1])
    import coherentsearch_c_amd as cs_
    spec = cs_.Spectrum(amps, 1 << 13, synth.DT)
    p = cs_.make_params(nharms=16, m=16, hidr=0.4999)
    L = abi.host()
    plans = abi.DirectPlanSet()
    assert L.direct_plans_build(C.byref(plans), C.byref(p), 2.0) == 0
    assert plans.plans[0].P == 0
    ptr = C.POINTER(C.c_longlong)()
    nc = C.c_longlong()
    assert L.chunk_starts(C.byref(plans), C.byref(spec.ft), C.byref(p), o.total, 512,
                          C.byref(ptr), C.byref(nc)) == 0
    host = [ptr[i] for i in range(nc.value)]
    L.free_(C.cast(ptr, C.c_void_p))
    L.direct_plans_free(C.byref(plans))
    assert host == list(o.chunk_starts())


def test_nonrational_f64_is_not_restated(oracle):
    """The reference keeps a double stack for a non-rational step in double
    precision (src/interp.c:464-467); that form is not restated."""
    with pytest.raises(ValueError):
        _nonrational_oracle(oracle, precision="f64")


def test_pure_tone_recovered(oracle):
    """A noiseless fractional-bin harmonic train is recovered at its trial with
    the m-tap truncation loss only (about 0.2/m, 1.3% at m=16)."""
    nbins = 1 << 14
    r0 = 123.0 + 1.0 / 3.0   # a multiple of lodr = 1/96 (nh=48, hidr 0.5)
    amps = synth.spectrum_with_pulsar(nbins, r0, 4, 0.5, 10.0, 1, noise=False)
    N = 2 * nbins
    cfg = oracle.config(nharms=48, lobin=100, blocksize=1 << 16, hifreq=200.0)
    o = oracle.Oracle(amps.view(np.float32), N, N * synth.DT, cfg)
    t = int(round((r0 - o.r_lo) / o.lodr))
    ci = int(np.searchsorted(o.chunk_starts(), t, side="right") - 1)
    i0, n, _ = o.chunk(ci)
    st, filled = o.stack(ci)
    a, _ = synth.vonmises_harmonics(0.5, 4)
    scale = 10.0 / np.sqrt(2 * np.sum(a ** 2))
    for h in range(1, 5):
        got = np.hypot(*st[h, t - i0])
        assert abs(got / (scale * a[h - 1]) - 1) < 0.02


def test_chunks_have_constant_fit_sets(oracle):
    o, _ = _small_oracle(oracle, nh=16, lobin=2, nbins=1 << 12, blocksize=512)
    cs = list(o.chunk_starts()) + [o.total]
    fits = np.zeros((16, o.total), dtype=bool)
    for h in range(1, 17):
        for t in range(o.total):
            fits[h - 1, t] = oracle.lib().orc_harmonic_fits(o.h, h, t)
    for a, b in zip(cs[:-1], cs[1:]):
        assert b > a
        assert np.all(fits[:, a:b] == fits[:, a:a + 1])
    assert all(x % 512 == 0 or x in cs for x in range(0, o.total, 512))


def test_widths_and_ladder(oracle):
    import ctypes as C
    L = oracle.lib()
    out = (C.c_int * 256)()
    n = L.orc_boxcar_widths(120, 1.5, 0.3, out, 256)
    assert list(out[:n]) == [1, 2, 3, 4, 6, 9, 13, 19, 28]
    cfg = oracle.config(nharms=60, maxdecim=6)
    got = {}
    for k in range(1, 7):
        n = L.orc_ladder_widths(2 * (60 // k), k, C.byref(cfg), out, 256)
        got[k] = list(out[:n])
    assert got[1] == [1, 2, 3, 4, 6] and got[6] == [2, 3, 4, 6]
    assert all(got[k] == [2, 3, 4, 6] for k in range(2, 6))
    cfg = oracle.config(nharms=16, maxdecim=6)
    want = {1: [1, 2, 3, 4, 6], 2: [2, 3, 4], 3: [2, 3], 4: [2], 5: [1], 6: [1]}
    for k, w in want.items():
        n = L.orc_ladder_widths(2 * (16 // k), k, C.byref(cfg), out, 256)
        assert list(out[:n]) == w, k
    cfg = oracle.config(nharms=60, maxdecim=1)
    n = L.orc_ladder_widths(120, 1, C.byref(cfg), out, 256)
    assert list(out[:n]) == [1, 2, 3, 4, 6, 9, 13, 19, 28]


def test_analytic_sigma_closed_form(oracle):
    L = oracle.lib()
    for Hk in [2, 8, 30, 60]:
        f = np.ones(Hk, np.uint8)
        assert L.orc_analytic_sigma(f.ctypes.data, 1, Hk) == pytest.approx(
            math.sqrt(2 * (Hk - 1) + 0.5), rel=1e-15)
        f[-1] = 0
        assert L.orc_analytic_sigma(f.ctypes.data, 1, Hk) == pytest.approx(
            math.sqrt(2 * (Hk - 1)), rel=1e-15)
    f = np.zeros(60, np.uint8)
    assert L.orc_analytic_sigma(f.ctypes.data, 1, 60) == 0.0


def test_boxcar_vs_brute_force(oracle):
    rng = np.random.default_rng(11)
    L = oracle.lib()
    for nb in [8, 20, 40, 120]:
        widths = np.array([w for w in [1, 2, 3, 4, 6, 9, 13] if w <= 0.3 * nb or w == 1],
                          np.int32)
        shape = rng.uniform(0.5, 2.0, widths.size)
        for _ in range(20):
            x = rng.standard_normal(nb)
            invs = 0.37
            got = L.orc_profile_boxcar_f64(x.ctypes.data, nb, widths.ctypes.data, widths.size,
                                           invs, shape.ctypes.data)
            best = -np.inf
            for i, w in enumerate(widths):
                sums = [sum(x[(p + u) % nb] for u in range(w)) for p in range(nb)]
                best = max(best, (max(sums) - w / nb * x.sum()) * invs * shape[i])
            assert got == pytest.approx(best, rel=1e-12, abs=1e-12)


def test_boxcar_noise_model_montecarlo(oracle):
    """For white harmonics with E|X|^2 = 1, Var(window sum of width w) equals
    sigma^2/shape_w^2 (the normalisation boxcar_shape encodes)."""
    import ctypes as C
    L = oracle.lib()
    rng = np.random.default_rng(5)
    H, nb, ntr = 30, 60, 40000
    widths = np.array([1, 2, 3, 6, 9], np.int32)
    filled = np.ones(H, np.uint8)
    filled[[3, 11]] = 0
    shape = np.zeros(widths.size)
    L.orc_boxcar_shape(shape.ctypes.data, nb, widths.ctypes.data, widths.size,
                       filled.ctypes.data, 1, H)
    sig = L.orc_analytic_sigma(filled.ctypes.data, 1, H)
    X = (rng.standard_normal((ntr, H + 1)) + 1j * rng.standard_normal((ntr, H + 1))) * np.sqrt(0.5)
    X[:, 0] = 0
    X[:, 1:][:, filled == 0] = 0
    prof = np.fft.irfft(X, nb, axis=1) * nb
    assert np.std(prof) == pytest.approx(sig, rel=0.02)
    for i, w in enumerate(widths):
        s = prof[:, :w].sum(axis=1)
        assert np.var(s) == pytest.approx((sig / shape[i]) ** 2, rel=0.03), w


def test_measured_sigma_order_statistic(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(2)
    for n, nb in [(7, 3), (40, 20), (68, 120)]:    # all samples when n*nb <= 8192
        M = rng.standard_normal((n, nb)).astype(np.float32)
        got = L.orc_block_sigma_f32(M.ctypes.data, nb, n, 8192)
        a = np.sort(np.abs(M).ravel())
        k = a.size
        med = a[k // 2] if k % 2 else np.float32(0.5) * (a[k // 2 - 1] + a[k // 2])
        assert got == np.float32(np.float32(1.4826) * med)
    # rows c*n/ncol when the chunk exceeds 8192 samples
    M = rng.standard_normal((1000, 120)).astype(np.float32)
    ncol = 8192 // 120
    rows = [(c * 1000) // ncol for c in range(ncol)]
    a = np.sort(np.abs(M[rows]).ravel())
    med = np.float32(0.5) * (a[a.size // 2 - 1] + a[a.size // 2])
    assert L.orc_block_sigma_f32(M.ctypes.data, 120, 1000, 8192) == np.float32(np.float32(1.4826) * med)


def test_metrics_vs_fp64_oracle(oracle):
    o, _ = _small_oracle(oracle, nh=16, nbins=1 << 14)
    for ci in [o.nchunks // 2, o.nchunks - 1]:
        m = o.metrics(ci)
        f = o.metrics_fp64(ci)
        ok = np.isfinite(m)
        assert np.array_equal(ok, np.isfinite(f))
        assert np.max(np.abs(m[ok] - f[ok]) / np.maximum(np.abs(f[ok]), 1.0)) < 2e-5


def test_search_threads_round_robin(oracle):
    o, _ = _small_oracle(oracle, nh=32, nbins=1 << 15)
    c1 = o.search(6.0, 1)
    c4 = o.search(6.0, 4)
    assert c1.size > 0 and c1.size == c4.size
    k = lambda a: sorted(zip(a["r"], a["nharm"], a["metric"]))
    assert k(c1) == k(c4)
    # one thread: chunk order, base then decimations, ascending trial
    assert np.all(np.diff(c1["r"][c1["nharm"] == 32]) > 0)


def test_injected_pulsar_found_with_harmonics(oracle):
    o, _ = _small_oracle(oracle, nh=16, nbins=1 << 15, r0=271.3)
    c = o.search(8.0, 4)
    top = oracle.collapse(c)
    # the strongest survivor is the pulsar (base or a decimated grid point)
    assert abs(top["r"][0] - 271.3) < 0.1 and top["metric"][0] > 9.0


def test_collapse_rules(oracle):
    def C(r, m, nh=16):
        return (r / 100.0, m, r, nh, 0, np.nan)
    a = np.array([C(100.0, 9), C(100.5, 12), C(101.2, 10), C(300.0, 11), C(200.02, 20),
                  C(50.0, 8.5), C(1000.0, 30)], dtype=oracle.CAND_DTYPE)
    d = a.copy()
    n = oracle.lib().orc_remove_duplicates(d.ctypes.data, d.size, 1.0)
    d = d[:n]
    # 100, 100.5, 101.2 chain within 1 bin of each predecessor -> one, metric 12
    assert sorted(d["r"]) == [50.0, 100.5, 200.02, 300.0, 1000.0]
    n = oracle.lib().orc_remove_harmonics(d.ctypes.data, d.size, 16, 1.0)
    h = d[:n]
    # by metric: 1000 kept; 200.02 removed (|1000 - 5*200.02| = 0.1 <= 1);
    # 100.5 kept (|1000 - 10*100.5| = 5 > 1, other ratios need n > 16);
    # 300 removed (|3*1000 - 10*300| = 0 <= 3); 50 removed by 100.5
    # (|100.5 - 2*50| = 0.5 <= 1) although 1000/50 = 20 > numharm
    assert sorted(h["r"]) == [100.5, 1000.0]
    assert list(h["r"]) == sorted(h["r"])   # sorted by frequency


def test_cohout_format(oracle):
    a = np.array([(91.97534, 15.234, 12345.0, 60, 0, 0.0333), (1.5, 8.0049, 201.3, 16, 0, np.nan)],
                 dtype=oracle.CAND_DTYPE)
    txt = oracle.format_candidates(a)
    lines = txt.splitlines()
    assert lines[0] == "#Num    'S/N'      Frequency (Hz)        Period (ms)    #Harm  Ducy(%)"
    assert lines[1] == "1        15.23     91.975340000000     10.872479514618    60     3.33"
    assert lines[2] == "2         8.00      1.500000000000    666.666666666667    16        -"


def test_realfft_packing_vs_numpy(oracle):
    rng = np.random.default_rng(9)
    for N in [64, 1000, 4096, 6000]:
        x = rng.standard_normal(N).astype(np.float32)
        got = oracle.realfft(x).view(np.complex64)
        ref = np.fft.rfft(x.astype(np.float64))
        assert got[0].real == pytest.approx(ref[0].real, abs=1e-3)
        assert got[0].imag == pytest.approx(ref[-1].real, abs=1e-3)
        np.testing.assert_allclose(got[1:], ref[1:N // 2], rtol=0, atol=2e-4 * np.sqrt(N))


def test_rednoise_normalises_white_noise(oracle):
    rng = np.random.default_rng(10)
    N = 1 << 16
    x = (3.0 * rng.standard_normal(N)).astype(np.float32)
    s = oracle.realfft_rednoise(x, synth.DT).view(np.complex64)
    assert s[0] == 1.0 + 0.0j
    p = np.abs(s[1:]) ** 2
    assert np.mean(p) == pytest.approx(1.0, rel=0.05)


def test_measure_ducy_of_injected_profile(oracle):
    nbins = 1 << 15
    amps = synth.spectrum_with_pulsar(nbins, 1234.5, 32, 0.1, 60.0, 3, noise=False)
    c = np.array([(1234.5 / (2 * nbins * synth.DT), 50.0, 1234.5, 32, 0, np.nan)],
                 dtype=oracle.CAND_DTYPE)
    d = oracle.measure_ducy(amps.view(np.float32), 2 * nbins, c)
    assert 0.05 <= d["ducy"][0] <= 0.15


# ---- zaplist (src/zap.c) ---------------------------------------------------

LN2 = -np.log(0.5)


def _flat(namps, value=(1.0, 1.0)):
    a = np.zeros(2 * namps, dtype=np.float32)
    a[0::2], a[1::2] = value
    return a


def test_zap_closed_forms(oracle):
    """T = 1 s, so Hz are bins.  A flat spectrum of power 2: each zapped bin
    becomes f32(sqrt(0.5 * (2 + 2) / ln 2)); bins floor(lo) .. ceil(hi) - 1
    change and nothing else."""
    namps, N = 4096, 8192
    a = _flat(namps)
    avg = np.float32(np.sqrt(0.5 * np.float32(2.0 + 2.0) / LN2))
    z, n = oracle.zap(a, N, 1.0, [(1000.25, 10.0, 0)])
    assert n == 1
    want = a.copy()
    want[2 * 995:2 * 1006:2], want[2 * 995 + 1:2 * 1006:2] = avg, 0.0
    assert np.array_equal(z, want)
    # narrower than a bin: widened to one bin around its centre (1999.8 .. 2000.8)
    z, n = oracle.zap(a, N, 1.0, [(2000.3, 0.1, 0)])
    assert n == 1 and np.flatnonzero(z != a).tolist() == [2 * 1999, 2 * 1999 + 1, 4000, 4001]
    # within 301 bins of DC: skipped (lo - 300 must exceed 1)
    for f, ok in ((250.0, 0), (301.5, 0), (302.5, 1)):
        z, n = oracle.zap(a, N, 1.0, [(f, 1.0, 0)])
        assert n == ok and np.array_equal(z, a) == (not ok)
    # the top: hi clipped to N/2 - 1; a bird starting at N/2 and everything
    # after it (in low-edge order) is dropped
    z, n = oracle.zap(a, N, 1.0, [(4097.0, 2.0, 0), (4095.0, 10.0, 0)])
    assert n == 1
    changed = np.flatnonzero(z[0::2] != a[0::2])
    assert changed.tolist() == list(range(4090, 4095))


def test_zap_medians(oracle):
    """The level is the lower median of the 200 powers at lo-300 .. lo-101
    and at hi-100 .. hi+99: with powers k^2 (k = 1..200 and 201..400,
    shuffled) those are 100^2 and 300^2."""
    namps, N = 4096, 8192
    rng = np.random.default_rng(3)
    a = np.zeros(2 * namps, dtype=np.float32)
    a[2 * 700:2 * 900:2] = rng.permutation(np.arange(1, 201)).astype(np.float32)
    a[2 * 920:2 * 1120:2] = rng.permutation(np.arange(201, 401)).astype(np.float32)
    z, n = oracle.zap(a, N, 1.0, [(1010.0, 20.0, 0)])
    avg = np.float32(np.sqrt(0.5 * np.float32(100.0 ** 2 + 300.0 ** 2) / LN2))
    assert n == 1
    assert np.all(z[2 * 1000:2 * 1020:2] == avg) and np.all(z[2 * 1000 + 1:2 * 1020:2] == 0)
    assert z[2 * 999] == a[2 * 999] and z[2 * 1020] == a[2 * 1020]


def test_zap_order_and_scaling(oracle):
    """Intervals apply in low-edge order whatever the input order, each one
    after the previous (a later median window sees earlier zapped bins);
    non-'B' lines scale by 1 + baryv, 'B' lines do not."""
    namps, N = 1 << 14, 1 << 15
    rng = np.random.default_rng(4)
    a = (rng.standard_normal(2 * namps) * 0.7).astype(np.float32)
    b1, b2 = (1010.0, 20.0, 0), (1260.0, 20.0, 0)   # b2's low window covers b1's bins
    both, n = oracle.zap(a, N, 1.0, [b2, b1])
    assert n == 2
    step, _ = oracle.zap(oracle.zap(a, N, 1.0, [b1])[0], N, 1.0, [b2])
    assert np.array_equal(both, step)
    alone, _ = oracle.zap(a, N, 1.0, [b2])
    assert not np.array_equal(alone[2 * 1250:2 * 1270], both[2 * 1250:2 * 1270])
    v = 1e-3
    scaled, _ = oracle.zap(a, N, 1.0, [(3000.0, 4.0, 0)], baryv=v)
    bary, _ = oracle.zap(a, N, 1.0, [(3000.0 * (1 + v), 4.0 * (1 + v), 1)], baryv=0.0)
    assert np.array_equal(scaled, bary)
    same, _ = oracle.zap(a, N, 1.0, [(3000.0, 4.0, 1)], baryv=v)
    plain, _ = oracle.zap(a, N, 1.0, [(3000.0, 4.0, 0)])
    assert np.array_equal(same, plain)
