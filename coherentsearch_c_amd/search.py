"""Python view of the host interface (reference src/search.c search_file,
src/interp.c plans/chunks, src/cuda.cu sessions), over the C-ABI libraries.

Everything here forwards to libcs_host.so / libcs_hip.so; the numerical work
happens on the GPU.
"""
import ctypes as C

import numpy as np

from . import abi

CAND_DTYPE = np.dtype([("freq", "f8"), ("metric", "f8"), ("r", "f8"),
                       ("nharm", "i4"), ("pad", "i4"), ("ducy", "f8")])


def make_params(nharms=60, m=16, hidr=0.5, maxdecim=6, precision="f32",
                sigma="analytic", threshold=8.0, decimations=None):
    """SearchParams as cli.c:943-951 builds them."""
    L = abi.host()
    p = abi.SearchParams()
    L.params_defaults(C.byref(p))
    p.nharms, p.m, p.hidr, p.threshold = nharms, m, hidr, threshold
    p.precision = 1 if precision == "f64" else 0
    p.sigma = 1 if sigma == "measured" else 0
    if decimations is None:
        p.ndecim = L.decimation_set(nharms, maxdecim, p.decimations, 64)
    else:
        p.ndecim = len(decimations)
        for i, k in enumerate(decimations):
            p.decimations[i] = k
    return p


class Spectrum:
    """An in-memory .fft: namps complex amplitudes (interleaved float32),
    N time samples of width dt (T = N*dt)."""

    def __init__(self, amps, N, dt, path="memory"):
        a = np.asarray(amps)
        if np.iscomplexobj(a):
            a = a.astype(np.complex64).view(np.float32)
        self.amps = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
        self.N, self.dt = int(N), float(dt)
        self.T = self.N * self.dt
        self._path = path.encode()
        f = abi.FFTFile()
        f.path = self._path
        f.N = self.N
        f.T = self.T
        f.df = 1.0 / self.T
        f.amps = self.amps.ctypes.data_as(C.POINTER(C.c_float))
        f.namps = self.amps.size // 2
        f.inf.N, f.inf.has_N, f.inf.dt, f.inf.has_dt = self.N, 1, self.dt, 1
        f.DC = float(self.amps[0]) if self.amps.size else 0.0
        f.Nyquist = float(self.amps[1]) if self.amps.size > 1 else 0.0
        self.ft = f

    @property
    def namps(self):
        return self.amps.size // 2


class Stage:
    """Device realfft + rednoise plan for series of N samples at dt
    (cs_gpu_stage_create; datpipe_create's widths)."""

    def __init__(self, N, dt, device=0, startwidth=6, endwidth=100, endfreq=6.0):
        self.L = abi.host()
        self.N, self.dt = int(N), float(dt)
        self.h = self.L.cs_gpu_stage_create(self.N, self.dt, startwidth, endwidth, endfreq,
                                            device)
        if not self.h:
            raise RuntimeError("cs_gpu_stage_create failed (see stderr)")

    def run_host(self, samples, packed=False):
        """Stage on the device and copy the spectrum (and optionally the
        realfft output before rednoise) back: float32 arrays of N values."""
        s = np.ascontiguousarray(samples, dtype=np.float32)
        if s.size != self.N:
            raise ValueError("series of %d samples for a plan of %d" % (s.size, self.N))
        amps = np.zeros(self.N, dtype=np.float32)
        pk = np.zeros(self.N, dtype=np.float32) if packed else None
        rc = self.L.cs_gpu_stage_run_host(self.h, s.ctypes.data, amps.ctypes.data,
                                          pk.ctypes.data if packed else None)
        if rc != 0:
            raise RuntimeError("device staging failed")
        return (amps, pk) if packed else amps

    def close(self):
        if self.h:
            self.L.cs_gpu_stage_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def empty_spectrum(N, dt, path="memory"):
    """A Spectrum of N samples whose amplitudes are filled later by device
    staging (bin 0 is (1, 0) after rednoise)."""
    sp = Spectrum(np.zeros(int(N), dtype=np.float32), N, dt, path)
    sp.ft.DC, sp.ft.Nyquist = 1.0, 0.0
    return sp


def redfft(samples, dt, path="memory"):
    """realfft + rednoise of a float32 series in memory (datpipe_redfft_mem)."""
    L = abi.host()
    s = np.ascontiguousarray(samples, dtype=np.float32)
    pipe = L.datpipe_create(6, 100, 6.0)
    ft = abi.FFTFile()
    rc = L.datpipe_redfft_mem(pipe, s.ctypes.data_as(C.c_void_p), s.size, dt,
                              path.encode(), C.byref(ft))
    if rc != 0:
        L.datpipe_free(pipe)
        raise ValueError("realfft/rednoise failed (N must be even and > 2*startwidth)")
    amps = np.ctypeslib.as_array(ft.amps, shape=(2 * ft.namps,)).copy()
    L.fftfile_close(C.byref(ft))
    L.datpipe_free(pipe)
    return Spectrum(amps, s.size, dt, path)


class Zaplist:
    """A zaplist (ZapList) built from (freq Hz, width Hz, bary) tuples or
    read from a file (zaplist_read); keeps its birds alive for the C calls."""

    def __init__(self, birds=()):
        birds = list(birds)
        self.birds = (abi.ZapBird * max(len(birds), 1))()
        for i, (f, w, b) in enumerate(birds):
            self.birds[i].freq, self.birds[i].width, self.birds[i].bary = float(f), float(w), int(b)
        self.zl = abi.ZapList()
        self.zl.v = C.cast(self.birds, C.POINTER(abi.ZapBird))
        self.zl.n = len(birds)

    @classmethod
    def read(cls, path):
        L = abi.host()
        z = abi.ZapList()
        if L.zaplist_read(str(path).encode(), C.byref(z)) != 0:
            raise OSError("cannot read zaplist '%s'" % path)
        birds = [(z.v[i].freq, z.v[i].width, z.v[i].bary) for i in range(z.n)]
        L.zaplist_free(C.byref(z))
        return cls(birds)

    def tuples(self):
        return [(self.birds[i].freq, self.birds[i].width, self.birds[i].bary)
                for i in range(self.zl.n)]

    def __len__(self):
        return self.zl.n


def zap(spec, zaplist, baryv=0.0):
    """zaplist_apply on the host spectrum in place (zap.c:155-187); returns
    the number of intervals zapped."""
    z = zaplist if isinstance(zaplist, Zaplist) else Zaplist(zaplist)
    n = abi.host().zaplist_apply(C.byref(z.zl), C.byref(spec.ft), float(baryv))
    if n < 0:
        raise RuntimeError("zaplist_apply failed")
    spec.ft.DC = float(spec.amps[0])
    spec.ft.Nyquist = float(spec.amps[1])
    return n


def take_list(cl):
    """Copy a CandidateList into a numpy array and free it."""
    L = abi.host()
    n = int(cl.n)
    if n:
        buf = C.string_at(C.cast(cl.v, C.c_void_p), n * C.sizeof(abi.Candidate))
        arr = np.frombuffer(buf, dtype=CAND_DTYPE).copy()
    else:
        arr = np.zeros(0, dtype=CAND_DTYPE)
    L.candlist_free(C.byref(cl))
    return arr


def to_list(arr):
    """numpy candidates -> CandidateList owning malloc'ed memory (free with
    candlist_free)."""
    arr = np.ascontiguousarray(arr, dtype=CAND_DTYPE)
    cl = abi.CandidateList()
    n = arr.size
    libc = C.CDLL(None)
    libc.malloc.restype = C.c_void_p
    p = libc.malloc(max(n, 1) * C.sizeof(abi.Candidate))
    C.memmove(p, arr.ctypes.data, n * C.sizeof(abi.Candidate))
    cl.v = C.cast(p, C.POINTER(abi.Candidate))
    cl.n = n
    cl.cap = n
    return cl


class Session:
    """Plans + chunk table (host) and a resident GPU session for one
    spectrum geometry (search.c:1257-1341 then cuda.cu session_create)."""

    def __init__(self, spec, params, lofreq=0.1, hifreq=125.0, lobin=100, blocksize=65536,
                 device=0):
        L = abi.host()
        self.L, self.spec, self.params = L, spec, params
        self.lodr = params.hidr / params.nharms
        self.r_lo = lofreq * spec.T if lobin == 100 else float(lobin)
        r_hi = hifreq * spec.T
        self.total = max(int(np.floor((r_hi - self.r_lo) / self.lodr)) + 1, 0)
        self.Nprof = blocksize if blocksize > 0 else 65536
        self.plans = abi.DirectPlanSet()
        if L.direct_plans_build(C.byref(self.plans), C.byref(params), self.r_lo) != 0:
            raise ValueError("plan build failed")
        cs = C.POINTER(C.c_longlong)()
        nc = C.c_longlong()
        if L.chunk_starts(C.byref(self.plans), C.byref(spec.ft), C.byref(params), self.total,
                          self.Nprof, C.byref(cs), C.byref(nc)) != 0:
            raise ValueError("chunk table failed")
        self.nchunks = nc.value
        self.cstarts = np.ctypeslib.as_array(cs, shape=(max(self.nchunks, 1),))[:self.nchunks].copy()
        L.free_(C.cast(cs, C.c_void_p))
        self._cs = np.ascontiguousarray(self.cstarts, dtype=np.int64)
        self.h = L.cs_gpu_session_create(
            C.byref(spec.ft), C.byref(params), C.byref(self.plans),
            self._cs.ctypes.data_as(C.POINTER(C.c_longlong)), self.total, self.Nprof,
            self.nchunks, self.r_lo, self.lodr, device)
        if not self.h:
            L.direct_plans_free(C.byref(self.plans))
            raise RuntimeError("cs_gpu_session_create failed (see stderr)")

    def run(self, threshold):
        return take_list(self.L.cs_gpu_session_run(self.h, C.byref(self.spec.ft), threshold))

    def upload(self, spec=None):
        spec = spec or self.spec
        if self.L.cs_gpu_session_upload(self.h, C.byref(spec.ft)) != 0:
            raise RuntimeError("spectrum upload failed")

    def stage(self, stage, samples, host_copy=True):
        """realfft + rednoise of the raw series on the device, straight into
        the resident spectrum; with host_copy the session's Spectrum receives
        the staged amplitudes for host post-processing."""
        s = np.ascontiguousarray(samples, dtype=np.float32)
        if s.size != self.spec.N:
            raise ValueError("series of %d samples for a session of N=%d" % (s.size, self.spec.N))
        rc = self.L.cs_gpu_session_stage(self.h, stage.h, s.ctypes.data,
                                         self.spec.amps.ctypes.data if host_copy else None)
        if rc != 0:
            raise RuntimeError("device staging failed")
        if host_copy:
            self.spec.ft.DC = float(self.spec.amps[0])
            self.spec.ft.Nyquist = float(self.spec.amps[1])

    def set_graphs(self, enable):
        """replay the search from a captured hipGraph (default) or launch eagerly"""
        if self.L.cs_gpu_session_set_graphs(self.h, int(bool(enable))) != 0:
            raise RuntimeError("graph mode not available for this session")

    def search_resident(self, threshold):
        return take_list(self.L.cs_gpu_session_search_resident(self.h, threshold))

    def search_noalloc(self, threshold, cap):
        """cs_gpu_session_search_noalloc into an array of `cap` candidates:
        the candidates, or RuntimeError when the search finds more records
        than cap or the library's fixed record array holds."""
        buf = np.zeros(max(int(cap), 0), dtype=CAND_DTYPE)
        n = self.L.cs_gpu_session_search_noalloc(self.h, threshold,
                                                 buf.ctypes.data if buf.size else None, int(cap))
        if n < 0:
            raise RuntimeError("cs_gpu_session_search_noalloc failed (see stderr)")
        return buf[:n].copy()

    def zap(self, zaplist, baryv=0.0, host_copy=True):
        """Apply a zaplist to the resident spectrum (cs_gpu_session_zap); with
        host_copy the session's Spectrum receives the zapped amplitudes.
        Returns the number of intervals zapped."""
        z = zaplist if isinstance(zaplist, Zaplist) else Zaplist(zaplist)
        n = self.L.cs_gpu_session_zap(self.h, C.byref(z.zl), float(baryv),
                                      self.spec.amps.ctypes.data if host_copy else None)
        if n < 0:
            raise RuntimeError("device zaplist failed (see stderr)")
        return n

    def finalize(self, cands, ncands=100):
        """cli.c:1059-1061 on the device: sort by S/N, keep ncands, duty cycles
        from the resident spectrum (cs_gpu_session_finalize)."""
        cl = to_list(cands)
        out = self.L.cs_gpu_session_finalize(self.h, C.byref(cl), int(ncands))
        n_in = int(cl.n)
        self.L.candlist_free(C.byref(cl))
        res = take_list(out)
        if n_in and ncands != 0 and res.size == 0:
            raise RuntimeError("device duty-cycle pass failed (see stderr)")
        return res

    def dump_stack(self, t0, n):
        out = np.zeros((n, self.params.nharms + 1, 2), dtype=np.float32)
        rc = self.L.cs_gpu_session_dump_stack(self.h, C.byref(self.spec.ft), t0, n,
                                              out.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError("stack dump failed")
        return out

    def kernel_ms(self):
        return self.L.cs_gpu_session_last_kernel_ms(self.h)

    def work(self):
        ht, tr, b = C.c_longlong(), C.c_longlong(), C.c_double()
        self.L.cs_gpu_session_work(self.h, C.byref(ht), C.byref(tr), C.byref(b))
        return ht.value, tr.value, b.value

    def close(self):
        if getattr(self, "h", None):
            self.L.cs_gpu_session_free(self.h)
            self.h = None
            self.L.direct_plans_free(C.byref(self.plans))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MetricStats:
    """A MetricStats (src/search.c:409-535) for search_file's --metricstats
    collection, with the histogram range and window count of the CLI
    (src/cli.c:1050); view() copies the statistics out."""

    def __init__(self, lo=0.0, hi=64.0, nb=3200, nwin=16):
        self.L = abi.host()
        self.ms = abi.MetricStats()
        self.L.metricstats_init(C.byref(self.ms), lo, hi, nb, nwin)

    def view(self):
        m = self.ms
        blocks = [{f: getattr(m.blocks[i], f) for f, _ in abi.BlockMetricStats._fields_}
                  for i in range(m.nblocks)]

        def hist(h):
            d = {f: getattr(h, f) for f, _ in abi.MetricHistogram._fields_ if f != "counts"}
            d["counts"] = np.ctypeslib.as_array(h.counts, shape=(h.nb,)).copy()
            return d
        return dict(blocks=blocks, whists=[hist(m.whists[i]) for i in range(m.nwhists)],
                    hists=[hist(m.hists[i]) for i in range(m.nhists)],
                    nwin=m.nwin, lo=m.hist_lo, hi=m.hist_hi, nb=m.hist_nb)

    def report(self, base, params, threshold):
        """report_metricstats (src/cli.c:170-303)."""
        self.L.report_metricstats(C.byref(self.ms), str(base).encode(), C.byref(params),
                                  threshold)

    def close(self):
        if self.ms is not None:
            self.L.metricstats_free(C.byref(self.ms))
            self.ms = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def search_file(spec, params, lofreq=0.1, hifreq=125.0, lobin=100, blocksize=0,
                threshold=8.0, remove=True, dr_tol=1.0, harm_remove=True, numharm=16,
                device=0, cache=None, harm_tol=1.0, metricstats=None, normalize=False):
    """search.c:1247-1486 (GPU path): raw search then collapse.  metricstats
    (a MetricStats) collects the per-block and per-window statistics;
    normalize scores the candidates per (k, frequency window)."""
    L = abi.host()
    cl = L.search_file(C.byref(spec.ft), C.byref(params), lofreq, hifreq, lobin, blocksize,
                       threshold, int(remove), dr_tol, int(harm_remove), numharm,
                       C.byref(metricstats.ms) if metricstats is not None else None,
                       int(bool(normalize)), 0, 0, None,
                       C.byref(cache) if cache is not None else None, harm_tol, device)
    return take_list(cl)


def collapse(cands, dr_tol=1.0, numharm=16, harm_tol=1.0):
    """search.c:1467-1478: remove_duplicates (runs within dr_tol keep the
    strongest) then remove_harmonics of a raw candidate list."""
    L = abi.host()
    cl = to_list(cands)
    a = L.remove_duplicates(C.byref(cl), dr_tol)
    L.candlist_free(C.byref(cl))
    b = L.remove_harmonics(C.byref(a), numharm, harm_tol)
    L.candlist_free(C.byref(a))
    return take_list(b)


def finalize(spec, params, cands, ncands=100):
    """cli.c:1059-1061: sort by S/N, truncate, measure duty cycles."""
    L = abi.host()
    cl = to_list(cands)
    L.candlist_sort_metric_desc(C.byref(cl))
    if cl.n > ncands:
        cl.n = ncands
    d = L.measure_ducy(C.byref(spec.ft), C.byref(cl), C.byref(params))
    L.candlist_free(C.byref(cl))
    return take_list(d)


def autotune(params, device=0):
    """cs_gpu_autotune: the fastest launch shape (and, for analytic sigma,
    chunk length) of this kernel class, measured on a synthetic spectrum and
    used by later sessions and engine requests.  Returns a dict."""
    r = abi.CsTuneResult()
    if abi.host().cs_gpu_autotune(int(device), C.byref(params), C.byref(r)) != 0:
        raise RuntimeError("no GPU kernel serves these parameters")
    return {k: getattr(r, k) for k, _ in abi.CsTuneResult._fields_}


def gpu_available(device=0):
    buf = C.create_string_buffer(256)
    rc = abi.host().cs_gpu_available(device, buf, 256)
    return rc, buf.value.decode()
