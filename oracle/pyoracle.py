"""ORACLE -- test infrastructure only.

ctypes view of oracle/liboracle.so, the CPU restatement of the reference's
CPU search path (see oracle/oracle.h for the file:line map).  Imported only by
tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always
as the checker, never as the thing measured or shipped.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class OrcCand(C.Structure):
    _fields_ = [("freq", C.c_double), ("metric", C.c_double), ("r", C.c_double),
                ("nharm", C.c_int), ("ducy", C.c_double)]


class OrcConfig(C.Structure):
    _fields_ = [("nharms", C.c_int), ("m", C.c_int), ("hidr", C.c_double),
                ("ndecim", C.c_int), ("decimations", C.c_int * 64),
                ("precision", C.c_int), ("sigma", C.c_int),
                ("fsp", C.c_double), ("maxfrac", C.c_double),
                ("lofreq", C.c_double), ("hifreq", C.c_double),
                ("lobin", C.c_int), ("blocksize", C.c_int)]


CAND_DTYPE = np.dtype([("freq", "f8"), ("metric", "f8"), ("r", "f8"),
                       ("nharm", "i4"), ("pad", "i4"), ("ducy", "f8")])

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, I, D, LL = C.c_void_p, C.c_int, C.c_double, C.c_longlong
        L.orc_config_defaults.argtypes = [C.POINTER(OrcConfig)]
        L.orc_decimation_set.argtypes = [I, I, C.POINTER(C.c_int), I]
        L.orc_decimation_set.restype = I
        L.orc_ctx_create.argtypes = [P, LL, LL, D, C.POINTER(OrcConfig)]
        L.orc_ctx_create.restype = P
        L.orc_ctx_free.argtypes = [P]
        L.orc_ctx_total.argtypes = [P]; L.orc_ctx_total.restype = LL
        L.orc_ctx_nchunks.argtypes = [P]; L.orc_ctx_nchunks.restype = LL
        L.orc_ctx_chunk.argtypes = [P, LL, C.POINTER(LL), C.POINTER(I), C.POINTER(D)]
        L.orc_ctx_chunk.restype = I
        L.orc_ctx_r_lo.argtypes = [P]; L.orc_ctx_r_lo.restype = D
        L.orc_ctx_lodr.argtypes = [P]; L.orc_ctx_lodr.restype = D
        L.orc_ctx_rational.argtypes = [P, C.POINTER(I), C.POINTER(I)]
        L.orc_ctx_ndec.argtypes = [P, C.POINTER(I), C.POINTER(I)]; L.orc_ctx_ndec.restype = I
        L.orc_ctx_harmonic_trials.argtypes = [P]; L.orc_ctx_harmonic_trials.restype = LL
        L.orc_plan_info.argtypes = [P, I, C.POINTER(I), C.POINTER(D), C.POINTER(I),
                                    C.POINTER(I), C.POINTER(I)]
        L.orc_plan_tables.argtypes = [P, I, P, P, P, P]; L.orc_plan_tables.restype = I
        L.orc_harmonic_fits.argtypes = [P, I, LL]; L.orc_harmonic_fits.restype = I
        L.orc_chunk_stack.argtypes = [P, LL, I, P, P]; L.orc_chunk_stack.restype = I
        L.orc_chunk_profiles.argtypes = [P, LL, I, P]; L.orc_chunk_profiles.restype = I
        L.orc_chunk_metrics.argtypes = [P, LL, P]; L.orc_chunk_metrics.restype = I
        L.orc_chunk_metrics_fp64.argtypes = [P, LL, P]; L.orc_chunk_metrics_fp64.restype = I
        L.orc_chunk_metrics_fp64_rows.argtypes = [P, LL, I, I, P]
        L.orc_chunk_metrics_fp64_rows.restype = I
        L.orc_search.argtypes = [P, D, I, LL, LL, C.POINTER(C.POINTER(OrcCand))]
        L.orc_search.restype = LL
        L.orc_c2r_f32.argtypes = [I, P, P]
        L.orc_c2r_f64.argtypes = [I, P, P]
        L.orc_boxcar_widths.argtypes = [I, D, D, P, I]; L.orc_boxcar_widths.restype = I
        L.orc_ladder_widths.argtypes = [I, I, C.POINTER(OrcConfig), P, I]
        L.orc_ladder_widths.restype = I
        L.orc_analytic_sigma.argtypes = [P, I, I]; L.orc_analytic_sigma.restype = D
        L.orc_boxcar_shape.argtypes = [P, I, P, I, P, I, I]
        L.orc_block_sigma_f32.argtypes = [P, I, I, I]; L.orc_block_sigma_f32.restype = C.c_float
        L.orc_profile_boxcar_f32.argtypes = [P, I, P, I, C.c_float, P]
        L.orc_profile_boxcar_f32.restype = C.c_float
        L.orc_profile_boxcar_f64.argtypes = [P, I, P, I, D, P]
        L.orc_profile_boxcar_f64.restype = D
        L.orc_trial_grid_rational.argtypes = [D, C.POINTER(I), C.POINTER(I)]
        L.orc_trial_grid_rational.restype = I
        L.orc_sinpi.argtypes = [D]; L.orc_sinpi.restype = D
        L.orc_cospi.argtypes = [D]; L.orc_cospi.restype = D
        L.orc_fourier_interp_complex.argtypes = [D, P, LL, I, C.POINTER(D), C.POINTER(D)]
        L.orc_remove_duplicates.argtypes = [P, C.c_size_t, D]
        L.orc_remove_duplicates.restype = C.c_size_t
        L.orc_remove_harmonics.argtypes = [P, C.c_size_t, I, D]
        L.orc_remove_harmonics.restype = C.c_size_t
        L.orc_sort_metric_desc.argtypes = [P, C.c_size_t]
        L.orc_sort_freq.argtypes = [P, C.c_size_t]
        L.orc_harmonically_related.argtypes = [D, D, I, D]
        L.orc_harmonically_related.restype = I
        L.orc_measure_ducy.argtypes = [P, LL, LL, P, C.c_size_t, D, D]
        L.orc_boxcar_best_width.argtypes = [P, I, D, D, C.POINTER(I), C.POINTER(D)]
        L.orc_boxcar_best_width.restype = I
        L.orc_format_candidates.argtypes = [P, C.c_size_t, C.c_char_p, C.c_size_t]
        L.orc_format_candidates.restype = C.c_size_t
        L.orc_realfft_rednoise.argtypes = [P, LL, D, P]; L.orc_realfft_rednoise.restype = I
        L.orc_realfft.argtypes = [P, LL, P]; L.orc_realfft.restype = I
        L.orc_rednoise.argtypes = [P, LL, D, P]; L.orc_rednoise.restype = I
        L.orc_nr_unpack.argtypes = [P, LL]; L.orc_nr_unpack.restype = I
        L.orc_presto_invsqrtf.argtypes = [C.c_float]; L.orc_presto_invsqrtf.restype = C.c_float
        L.orc_presto_median.argtypes = [P, I]; L.orc_presto_median.restype = C.c_float
        L.orc_zap.argtypes = [P, LL, LL, D, P, P, P, I, D]; L.orc_zap.restype = I
        L.orc_delay_from_dm.argtypes = [D, D]; L.orc_delay_from_dm.restype = D
        L.orc_choose_good_N.argtypes = [LL]; L.orc_choose_good_N.restype = LL
        L.orc_dedisp_tables.argtypes = [C.POINTER(OrcDedispConfig), LL, P, P, P,
                                        C.POINTER(LL)]
        L.orc_dedisp_tables.restype = I
        L.orc_dedisp.argtypes = [C.POINTER(OrcDedispConfig), P, LL, P, C.POINTER(C.c_float)]
        L.orc_dedisp.restype = LL
        L.orc_dedisp_direct.argtypes = [C.POINTER(OrcDedispConfig), P, LL, LL, P]
        L.orc_dedisp_direct.restype = I
        L.orc_clip_all.argtypes = [C.POINTER(OrcDedispConfig), P, LL, P]
        L.orc_clip_all.restype = LL
        L.orc_avg_var.argtypes = [P, I, C.POINTER(D), C.POINTER(D)]
        L.orc_lower_median.argtypes = [P, I]; L.orc_lower_median.restype = C.c_float
        L.free = C.CDLL(None).free
        L.free.argtypes = [P]
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def config(nharms=60, m=16, hidr=0.5, maxdecim=6, precision="f32", sigma="analytic",
           lofreq=0.1, hifreq=125.0, lobin=100, blocksize=0, decimations=None,
           fsp=1.5, maxfrac=0.3):
    L = lib()
    c = OrcConfig()
    L.orc_config_defaults(C.byref(c))
    c.nharms, c.m, c.hidr = nharms, m, hidr
    if decimations is None:
        c.ndecim = L.orc_decimation_set(nharms, maxdecim, c.decimations, 64)
    else:
        c.ndecim = len(decimations)
        for i, k in enumerate(decimations):
            c.decimations[i] = k
    c.precision = 1 if precision == "f64" else 0
    c.sigma = 1 if sigma == "measured" else 0
    c.fsp, c.maxfrac = fsp, maxfrac
    c.lofreq, c.hifreq, c.lobin, c.blocksize = lofreq, hifreq, lobin, blocksize
    return c


def cands_to_array(p, n):
    buf = (C.c_char * (n * 40)).from_address(C.cast(p, C.c_void_p).value) if n else b""
    arr = np.frombuffer(bytes(buf), dtype=CAND_DTYPE).copy() if n else \
        np.zeros(0, dtype=CAND_DTYPE)
    return arr


class Oracle:
    """One spectrum + search configuration (plans and chunk table)."""

    def __init__(self, amps, N, T, cfg):
        self.L = lib()
        self.amps = np.ascontiguousarray(amps, dtype=np.float32).reshape(-1)
        self.namps = self.amps.size // 2
        self.N, self.T, self.cfg = int(N), float(T), cfg
        self.h = self.L.orc_ctx_create(ptr(self.amps), self.namps, self.N, self.T,
                                       C.byref(cfg))
        if not self.h:
            raise ValueError("oracle context creation failed")
        ks = (C.c_int * 64)()
        hk = (C.c_int * 64)()
        nd = self.L.orc_ctx_ndec(self.h, ks, hk)
        self.ks = [1] + [ks[i] for i in range(nd)]
        self.hks = [cfg.nharms] + [hk[i] for i in range(nd)]

    def close(self):
        if self.h:
            self.L.orc_ctx_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def total(self):
        return self.L.orc_ctx_total(self.h)

    @property
    def nchunks(self):
        return self.L.orc_ctx_nchunks(self.h)

    @property
    def r_lo(self):
        return self.L.orc_ctx_r_lo(self.h)

    @property
    def lodr(self):
        return self.L.orc_ctx_lodr(self.h)

    def harmonic_trials(self):
        return self.L.orc_ctx_harmonic_trials(self.h)

    def chunk(self, ci):
        i0, n, rs = C.c_longlong(), C.c_int(), C.c_double()
        if self.L.orc_ctx_chunk(self.h, ci, C.byref(i0), C.byref(n), C.byref(rs)) != 0:
            raise IndexError(ci)
        return i0.value, n.value, rs.value

    def chunk_starts(self):
        return np.array([self.chunk(c)[0] for c in range(self.nchunks)], dtype=np.int64)

    def stack(self, ci, form=0):
        i0, n, _ = self.chunk(ci)
        out = np.zeros((self.cfg.nharms + 1, n, 2), dtype=np.float32)
        filled = np.zeros(self.cfg.nharms, dtype=np.uint8)
        self.L.orc_chunk_stack(self.h, ci, form, ptr(out), ptr(filled))
        return out, filled

    def profiles(self, ci, slot=0):
        i0, n, _ = self.chunk(ci)
        nb = 2 * self.hks[slot]
        dt = np.float64 if self.cfg.precision == 1 else np.float32
        out = np.zeros((n, nb), dtype=dt)
        self.L.orc_chunk_profiles(self.h, ci, slot, ptr(out))
        return out

    def metrics(self, ci):
        i0, n, _ = self.chunk(ci)
        out = np.zeros((len(self.ks), n), dtype=np.float64)
        self.L.orc_chunk_metrics(self.h, ci, ptr(out))
        return out

    def metrics_fp64(self, ci):
        i0, n, _ = self.chunk(ci)
        out = np.zeros((len(self.ks), n), dtype=np.float64)
        self.L.orc_chunk_metrics_fp64(self.h, ci, ptr(out))
        return out

    def metrics_fp64_rows(self, ci, j0, nr):
        """FP64 metrics of rows [j0, j0+nr) of chunk ci, NaN at other rows."""
        i0, n, _ = self.chunk(ci)
        out = np.zeros((len(self.ks), n), dtype=np.float64)
        if self.L.orc_chunk_metrics_fp64_rows(self.h, ci, j0, nr, ptr(out)) < 0:
            raise ValueError("row window outside the chunk")
        return out

    def search(self, threshold, nthreads=1, chunk_lo=0, chunk_hi=-1):
        p = C.POINTER(OrcCand)()
        n = self.L.orc_search(self.h, threshold, nthreads, chunk_lo, chunk_hi, C.byref(p))
        arr = cands_to_array(p, n)
        self.L.free(C.cast(p, C.c_void_p))
        return arr


def c2r_f32(X):
    X = np.ascontiguousarray(X, dtype=np.complex64)
    H = X.size - 1
    x = np.zeros(2 * H, dtype=np.float32)
    lib().orc_c2r_f32(H, ptr(X), ptr(x))
    return x


def c2r_f64(X):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    H = X.size - 1
    x = np.zeros(2 * H, dtype=np.float64)
    lib().orc_c2r_f64(H, ptr(X), ptr(x))
    return x


def collapse(cands, dr_tol=1.0, numharm=16, harm_tol=1.0, remove=True, harm_remove=True,
             ncands=100):
    """remove_duplicates -> remove_harmonics -> metric sort -> truncate
    (search.c:1467-1478, cli.c:1059-1060)."""
    L = lib()
    a = np.ascontiguousarray(cands.copy())
    n = a.size
    if remove:
        n = L.orc_remove_duplicates(ptr(a), n, dr_tol)
    else:
        L.orc_sort_freq(ptr(a), n)
    if harm_remove:
        n = L.orc_remove_harmonics(ptr(a), n, numharm, harm_tol)
    a = a[:n].copy()
    L.orc_sort_metric_desc(ptr(a), n)
    return a[:ncands].copy()


def measure_ducy(amps, N, cands, fsp=1.5, maxfrac=0.3):
    a = np.ascontiguousarray(cands.copy())
    amps = np.ascontiguousarray(amps, dtype=np.float32).reshape(-1)
    lib().orc_measure_ducy(ptr(amps), amps.size // 2, int(N), ptr(a), a.size, fsp, maxfrac)
    return a


def format_candidates(cands):
    a = np.ascontiguousarray(cands)
    n = lib().orc_format_candidates(ptr(a), a.size, None, 0)
    buf = C.create_string_buffer(n)
    lib().orc_format_candidates(ptr(a), a.size, buf, n)
    return buf.value.decode()


def realfft_rednoise(samples, dt):
    s = np.ascontiguousarray(samples, dtype=np.float32)
    out = np.zeros(s.size, dtype=np.float32)
    rc = lib().orc_realfft_rednoise(ptr(s), s.size, dt, ptr(out))
    if rc != 0:
        raise ValueError("realfft/rednoise failed")
    return out


def realfft(samples):
    s = np.ascontiguousarray(samples, dtype=np.float32)
    out = np.zeros(s.size, dtype=np.float32)
    if lib().orc_realfft(ptr(s), s.size, ptr(out)) != 0:
        raise ValueError("realfft failed")
    return out


def zap(amps, N, T, birds, baryv=0.0):
    """Zaplist applied to a copy of an interleaved spectrum; birds: list of
    (freq Hz, width Hz, bary).  Returns (zapped copy, intervals applied)."""
    a = np.ascontiguousarray(amps, dtype=np.float32).reshape(-1).copy()
    f = np.ascontiguousarray([b[0] for b in birds], dtype=np.float64)
    w = np.ascontiguousarray([b[1] for b in birds], dtype=np.float64)
    y = np.ascontiguousarray([int(b[2]) for b in birds], dtype=np.int32)
    n = lib().orc_zap(ptr(a), a.size // 2, int(N), float(T), ptr(f), ptr(w), ptr(y), len(birds),
                      float(baryv))
    return a, n


def rednoise(packed, dt):
    """rednoise alone of a packed realfft output (N floats)."""
    s = np.ascontiguousarray(packed, dtype=np.float32)
    out = np.zeros(s.size, dtype=np.float32)
    if lib().orc_rednoise(ptr(s), s.size, dt, ptr(out)) != 0:
        raise ValueError("rednoise failed")
    return out


# ---------------------------------------------------------------------------
# Metric statistics and normalisation (--metricstats / --normalize), the
# reference's single-threaded order: chunks in order, the base then the
# decimations in configuration order, trials ascending.
# ---------------------------------------------------------------------------

def log_edges(a, b, nwin):
    """src/search.c:471-478: nwin log-spaced windows from a to b."""
    import math
    a = max(a, np.finfo(np.float64).tiny)
    if b < a * (1.0 + 1e-9):
        b = a * (1.0 + 1e-9)
    la, lb = math.log10(a), math.log10(b)      # libm, as the reference
    return np.array([math.pow(10.0, la + (lb - la) * i / nwin) for i in range(nwin + 1)])


def window_index(edges, nwin, f):
    """src/search.c:479-492: 1 + the last window whose left edge is <= f,
    clamped to 1..nwin (the count of edges <= f)."""
    cnt = int(np.sum(np.asarray(edges[:nwin + 1]) <= f))
    return min(max(cnt, 1), nwin)


class Hist:
    """src/search.c:412-453: nb equal bins over [lo, hi), under/over counts,
    running sum, sum of squares, minimum and maximum."""

    def __init__(self, k, Hk, win, flo, fhi, lo=0.0, hi=64.0, nb=3200):
        self.k, self.Hk, self.nbins, self.win = k, Hk, 2 * Hk, win
        self.flo, self.fhi, self.lo, self.hi, self.nb = flo, fhi, lo, hi, nb
        self.invw = nb / (hi - lo)
        self.counts = np.zeros(nb, dtype=np.int64)
        self.under = self.over = self.total = 0
        self.sum = self.sumsq = 0.0
        self.vmin, self.vmax = np.inf, -np.inf

    def push(self, v):
        """hist_push over v in order (sequential double sums)."""
        v = np.asarray(v, dtype=np.float64)
        if v.size == 0:
            return
        self.total += v.size
        self.sum = float(np.add.accumulate(np.concatenate([[self.sum], v]))[-1])
        self.sumsq = float(np.add.accumulate(np.concatenate([[self.sumsq], v * v]))[-1])
        self.vmin = min(self.vmin, float(v.min()))
        self.vmax = max(self.vmax, float(v.max()))
        under = v < self.lo
        over = ~under & (v >= self.hi)
        self.under += int(under.sum())
        self.over += int(over.sum())
        mid = v[~under & ~over]
        idx = np.clip(np.floor((mid - self.lo) * self.invw).astype(np.int64), 0, self.nb - 1)
        np.add.at(self.counts, idx, 1)

    def merge(self, b):
        self.counts += b.counts
        self.under += b.under
        self.over += b.over
        self.total += b.total
        self.sum += b.sum
        self.sumsq += b.sumsq
        self.vmin = min(self.vmin, b.vmin)
        self.vmax = max(self.vmax, b.vmax)

    def quantile(self, q):
        """src/search.c:454-470: linear inside the bin crossing q*total."""
        if self.total == 0:
            return np.nan
        target = q * self.total
        cum = self.under
        if cum >= target:
            return self.lo
        w = 1.0 / self.invw
        for i in range(self.nb):
            c = int(self.counts[i])
            if cum + c >= target:
                edge = self.lo + i * w
                return edge if c == 0 else edge + (target - cum) / c * w
            cum += c
        return self.hi


def _median(v):
    """median_f64 (src/boxcar.c:62-71)."""
    n = v.size
    s = np.sort(v)
    h = n // 2
    return float(s[h]) if n & 1 else 0.5 * (float(s[h - 1]) + float(s[h]))


def metric_stats(o, nwin=16, lo=0.0, hi=64.0, nb=3200):
    """The statistics of src/search.c:989-1095 for the oracle's search o:
    per (chunk, decimation) block statistics (src/search.c:537-565), the
    per-(k, window) histograms (a chunk's metrics go to the window of its mean
    frequency, src/search.c:864-868, 924-929) and the band-wide histogram per
    k.  Returns dict(blocks=[...], whists=[Hist], hists=[Hist], statks=[k],
    edges={k: array})."""
    cfg = o.cfg
    nh = cfg.nharms
    r_lo, lodr, T, N = o.r_lo, o.lodr, o.T, o.N
    r_hi = cfg.hifreq * T
    statks = sorted(set(o.ks))
    edges = {k: log_edges(k * r_lo / T, k * r_hi / T, nwin) for k in statks}
    wh = {k: [Hist(k, nh // k, w + 1, edges[k][w], edges[k][w + 1], lo, hi, nb)
              for w in range(nwin)] for k in statks}
    blocks = []
    nyq = N / 2.0
    for c in range(o.nchunks):
        i0, n, rs = o.chunk(c)
        rmean = rs + (n - 1) * lodr / 2.0
        M = o.metrics(c)
        for s, k in enumerate(o.ks):
            Hk = nh // k
            if s == 0:
                v = M[0, :n]
            else:
                j = np.arange(n, dtype=np.float64)
                nv = int(np.sum(k * (rs + j * lodr) < nyq))
                v = M[s, :nv]
            w = window_index(edges[k], nwin, k * rmean / T)
            wh[k][w - 1].push(v)
            if v.size == 0:
                continue
            sm = float(np.add.accumulate(v)[-1])
            mean = sm / v.size
            ss = float(np.add.accumulate((v - mean) ** 2)[-1])
            if s == 0:
                ngood = min(N / 2.0 / rmean, float(nh))
                flo, fhi = rs / T, (rs + (v.size - 1) * lodr) / T
            else:
                ngood = min(N / 2.0 / (k * rmean), float(Hk))
                flo, fhi = k * rs / T, k * (rs + (v.size - 1) * lodr) / T
            blocks.append(dict(block=c + 1, k=k, Hk=Hk, nbins=2 * Hk, ngoodbins=ngood,
                               flo=flo, fhi=fhi, n=v.size, min=float(v.min()),
                               median=_median(v), mean=mean,
                               std=float(np.sqrt(ss / (v.size - 1))) if v.size > 1 else 0.0,
                               max=float(v.max())))
    blocks.sort(key=lambda b: (b["block"], b["k"]))
    whists, hists = [], []
    for k in statks:
        g = Hist(k, nh // k, 0, edges[k][0], edges[k][nwin], lo, hi, nb)
        for h in wh[k]:
            g.merge(h)
            whists.append(h)
        hists.append(g)
    return dict(blocks=blocks, whists=whists, hists=hists, statks=statks, edges=edges,
                nwin=nwin, lo=lo, hi=hi, nb=nb)


def build_metricnorm(ms, min_win_trials=200):
    """src/search.c:752-801: per k and window, loc = median and scale =
    84.13 % quantile - median, the band-wide values where a window has fewer
    than min_win_trials trials or no spread (scale 1 if neither has)."""
    def ls(h):
        loc = h.quantile(0.5)
        sc = h.quantile(0.8413) - loc
        return loc, (sc if sc > 0.0 else np.nan)
    norm = {}
    for g in ms["hists"]:
        gloc, gsc = ls(g)
        if np.isnan(gsc):
            gsc = 1.0
        wins = [h for h in ms["whists"] if h.k == g.k]
        if not wins:
            continue
        e = [h.flo for h in wins] + [wins[-1].fhi]
        loc, sc = [], []
        for h in wins:
            l, s = ls(h)
            if h.total < min_win_trials or np.isnan(s):
                l, s = gloc, gsc
            loc.append(l)
            sc.append(s)
        norm[g.k] = (np.array(e), np.array(loc), np.array(sc))
    return norm


def metric_candidates(o, threshold, norm=None):
    """The candidates of src/search.c:605-747, 830-945 in the reference's
    order with the metric, or with --normalize its score (M - loc)/scale in
    the window of the trial's frequency (src/search.c:814-823)."""
    cfg = o.cfg
    nh, lodr, T, N = cfg.nharms, o.lodr, o.T, o.N
    parts = []
    for c in range(o.nchunks):
        i0, n, rs = o.chunk(c)
        M = o.metrics(c)
        j = np.arange(n, dtype=np.float64)
        for s, k in enumerate(o.ks):
            r = rs + j * lodr if s == 0 else float(k) * (rs + j * lodr)
            ok = np.ones(n, bool) if s == 0 else r < N / 2.0
            m = M[s, :n].astype(np.float64)
            if norm is not None and k in norm:
                e, loc, sc = norm[k]
                nwin = len(loc)
                w = np.clip(np.searchsorted(e[:nwin + 1], r / T, side="right"), 1, nwin) - 1
                m = (m - loc[w]) / sc[w]
            keep = ok & (m > threshold)
            if np.any(keep):
                a = np.zeros(int(keep.sum()), dtype=CAND_DTYPE)
                a["r"] = r[keep]
                a["freq"] = r[keep] / T
                a["metric"] = m[keep]
                a["nharm"] = nh // k
                a["ducy"] = np.nan
                parts.append(a)
    return np.concatenate(parts) if parts else np.zeros(0, dtype=CAND_DTYPE)


def _fmt_threshold(x):
    """cs_format_double (src/candidate.c:336-349): the shortest %g form that
    reads back exactly, with ".0" appended to a bare integer."""
    if np.isnan(x):
        return "NaN"
    if np.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    s = "%.17g" % x
    for p in range(1, 18):
        s = "%.*g" % (p, x)
        if float(s) == x:
            break
    if not any(ch in s for ch in ".eE"):
        s += ".0"
    return s


def _hist_row(h, faps):
    """src/cli.c:157-168."""
    N = h.total
    mean = h.sum / N if N > 0 else np.nan
    std = np.sqrt(max(0.0, (h.sumsq - N * mean * mean) / (N - 1))) if N > 1 else 0.0
    med = h.quantile(0.5)
    fv = [h.quantile(1.0 - f) for f in faps]
    ov = h.over / N if N > 0 else 0.0
    return mean, std, med, fv, ov


FAPS = [0.1, 0.01, 1e-3, 1e-4, 1e-5]


def format_metricstats(ms, nharms, threshold):
    """<base>_metricstats.txt (src/cli.c:238-253)."""
    L = ["# Per-block, per-decimation metric statistics\n",
         "# metric=boxcar nharms=%d threshold=%s\n" % (nharms, _fmt_threshold(threshold)),
         "# nbins = 2*Hk; the pure-noise metric floor scales ~sqrt(nbins).\n",
         "#%-7s %3s %4s %5s %10s %14s %14s %8s %9s %9s %9s %9s %9s\n"
         % ("block", "k", "Hk", "nbins", "ngoodbins", "f_lo(Hz)", "f_hi(Hz)", "n", "min",
            "median", "mean", "std", "max")]
    for b in ms["blocks"]:
        L.append("%-8d %3d %4d %5d %10.3f %14.8f %14.8f %8d %9.3f %9.3f %9.3f %9.3f %9.3f\n"
                 % (b["block"], b["k"], b["Hk"], b["nbins"], b["ngoodbins"], b["flo"], b["fhi"],
                    b["n"], b["min"], b["median"], b["mean"], b["std"], b["max"]))
    return "".join(L)


def format_metricfap(ms, nharms):
    """<base>_metricfap.txt (src/cli.c:255-276)."""
    L = ["# Per-(decimation, frequency-window) empirical false-alarm thresholds\n",
         "# metric=boxcar nharms=%d\n" % nharms,
         "# fap_X = metric value whose single-trial, single-decimation false-alarm prob is X "
         "(in-window)\n",
         "#%-3s %4s %5s %4s %13s %13s %10s %8s %8s %8s | %s\n"
         % ("k", "Hk", "nbins", "win", "f_lo(Hz)", "f_hi(Hz)", "ntrials", "mean", "std",
            "median", "fap1e-1 fap1e-2 fap1e-3 fap1e-4 fap1e-5")]
    for h in ms["whists"]:
        mean, std, med, fv, ov = _hist_row(h, FAPS)
        s = "%-4d %4d %5d %4d %13.7f %13.7f %10d %8.3f %8.3f %8.3f |" % (
            h.k, h.Hk, h.nbins, h.win, h.flo, h.fhi, h.total, mean, std, med)
        s += "".join(" %7.3f" % v for v in fv)
        if ov > FAPS[4]:
            s += "  # overflow %.1e" % ov
        L.append(s + "\n")
    return "".join(L)


def format_metrichist(ms, nharms):
    """<base>_metrichist.txt (src/cli.c:278-302)."""
    bw = (ms["hi"] - ms["lo"]) / ms["nb"]
    L = ["# Per-(decimation, frequency-window) metric histograms (one streaming pass)\n",
         "# metric=boxcar nharms=%d\n" % nharms,
         "# bin = left edge of a bin of width %g; count = trials in [bin, bin+width)\n" % bw,
         "#%-3s %4s %4s %13s %13s %12s %14s\n"
         % ("k", "Hk", "win", "f_lo(Hz)", "f_hi(Hz)", "bin", "count")]
    for h in ms["whists"]:
        nz = np.nonzero(h.counts)[0]
        if nz.size == 0:
            continue
        for b in range(nz[0], nz[-1] + 1):
            L.append("%-4d %4d %4d %13.7f %13.7f %12.5f %14d\n"
                     % (h.k, h.Hk, h.win, h.flo, h.fhi, ms["lo"] + b * bw, h.counts[b]))
    return "".join(L)


# ---- PRESTO prepsubband de-dispersion (src/prep_dedisp.c) ----------------------

class OrcDedispConfig(C.Structure):
    _fields_ = [("numchan", C.c_int), ("lofreq", C.c_double), ("chan_wid", C.c_double),
                ("dt", C.c_double), ("P", C.c_int), ("nsub", C.c_int),
                ("lodm", C.c_double), ("dmstep", C.c_double), ("numdms", C.c_int),
                ("downsamp", C.c_int), ("numout", C.c_longlong), ("clip_sigma", C.c_float)]


def dedisp_config(numchan, lofreq, chan_wid, dt, P, nsub, lodm, dmstep, numdms, downsamp=1,
                  numout=0, clip_sigma=0.0):
    return OrcDedispConfig(numchan, lofreq, chan_wid, dt, P, nsub, lodm, dmstep, numdms,
                           downsamp, numout, clip_sigma)


def clip_all(cfg, raw):
    """clip_times over every raw block, in order: (clipped copy, levels
    [nblocks, numchan], rows replaced)"""
    x = np.array(raw, dtype=np.float32, order="C")
    lv = np.zeros((x.shape[0] // cfg.P, cfg.numchan), np.float32)
    n = lib().orc_clip_all(C.byref(cfg), ptr(x), x.shape[0], ptr(lv))
    if n < 0:
        raise ValueError("not whole blocks")
    return x, lv, n


def nr_unpack(fft_out):
    """the Numerical Recipes unpack alone of an N/2-point complex FFT (N floats)"""
    d = np.array(fft_out, dtype=np.float32, copy=True)
    if lib().orc_nr_unpack(ptr(d), d.size) != 0:
        raise ValueError("unpack failed")
    return d


def presto_invsqrtf(x):
    f = lib().orc_presto_invsqrtf
    return np.array([f(float(v)) for v in np.asarray(x, np.float32).ravel()], dtype=np.float32)


def presto_median(x):
    x = np.ascontiguousarray(x, np.float32)
    return lib().orc_presto_median(ptr(x), x.size)


def avg_var(x):
    x = np.ascontiguousarray(x, np.float32)
    m, v = C.c_double(), C.c_double()
    lib().orc_avg_var(ptr(x), x.size, C.byref(m), C.byref(v))
    return m.value, v.value


def lower_median(x):
    x = np.ascontiguousarray(x, np.float32)
    return lib().orc_lower_median(ptr(x), x.size)


def delay_from_dm(dm, f):
    return lib().orc_delay_from_dm(dm, f)


def choose_good_N(n):
    return lib().orc_choose_good_N(n)


def dedisp_tables(cfg, nspec):
    """(blocksperread, dms, idispdt, offsets[numdms, nsub], numout); blocksperread
    is -1 for a configuration the reference rejects"""
    dms = np.zeros(cfg.numdms)
    idisp = np.zeros(cfg.numchan, np.int32)
    offs = np.zeros((cfg.numdms, cfg.nsub), np.int32)
    nout = C.c_longlong(0)
    bpr = lib().orc_dedisp_tables(C.byref(cfg), nspec, ptr(dms), ptr(idisp), ptr(offs),
                                  C.byref(nout))
    return bpr, dms, idisp, offs, nout.value


def dedisp(cfg, raw):
    """the prepsubband write loop over raw float spectra [nspec, numchan]:
    (series[numdms, numout], samples written per DM, padding value)"""
    raw = np.ascontiguousarray(raw, np.float32)
    nspec = raw.shape[0]
    bpr, _, _, _, numout = dedisp_tables(cfg, nspec)
    if bpr < 0:
        raise ValueError("configuration rejected")
    out = np.zeros((cfg.numdms, numout), np.float32)
    pad = C.c_float(0)
    n = lib().orc_dedisp(C.byref(cfg), ptr(raw), nspec, ptr(out), C.byref(pad))
    if n < 0:
        raise ValueError("de-dispersion failed")
    return out, n, pad.value


def dedisp_direct(cfg, raw, ntime):
    raw = np.ascontiguousarray(raw, np.float32)
    out = np.zeros((cfg.numdms, ntime), np.float32)
    if lib().orc_dedisp_direct(C.byref(cfg), ptr(raw), raw.shape[0], ntime, ptr(out)) != 0:
        raise ValueError("direct de-dispersion out of range")
    return out
