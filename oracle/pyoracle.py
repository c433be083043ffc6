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
        L.orc_zap.argtypes = [P, LL, LL, D, P, P, P, I, D]; L.orc_zap.restype = I
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
