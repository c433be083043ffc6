"""De-dispersion of raw filterbank spectra into DM series resident in HBM
(include/cs_dedisp.h), and the search of those series without a round trip
through the host.

The reference's prep front end (src/prep_dedisp.c) de-disperses with PRESTO
prepsubband, copies every DM series back to the host and hands the set to
prep_search_run (src/prep_search.c:44-145), which stages and searches each
series in turn.  Here `Dedisp.run` leaves the series on the device and
`prep_search` submits pointers into them to the engine: device staging,
search and duty cycles run on the resident series.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .search import to_list


class DeviceSeries:
    """N float32 samples in device memory (a pointer into a resident array
    the owner keeps alive): what the engine accepts in place of a host
    series."""

    def __init__(self, ptr, n, owner=None):
        self.ptr = int(ptr)
        self.n = int(n)
        self.owner = owner


class Dedisp:
    """A prepsubband plan on one GPU for `nspec` spectra of `numchan`
    channels (lowest frequency first): the reference's delay tables, device
    rings and the resident series [numdms][numout].  Arguments follow
    prepsubband's options (-nsub, -lodm, -dmstep, -numdms, -downsamp,
    -numout, -clip: 0 is -noclip; the default 6 is prepsubband's) and the raw
    header (lofreq = centre of the lowest channel in MHz, chan_wid in MHz, dt
    in s, the reader's block length)."""

    def __init__(self, numchan, lofreq, chan_wid, dt, nspec, nsub, lodm, dmstep, numdms,
                 spectra_per_subint=2048, downsamp=1, numout=0, clip_sigma=6.0, device=0):
        self.L = abi.host()
        self.cfg = abi.CsDedispConfig(int(numchan), float(lofreq), float(chan_wid), float(dt),
                                      int(spectra_per_subint), int(nsub), float(lodm),
                                      float(dmstep), int(numdms), int(downsamp), int(numout),
                                      float(clip_sigma))
        self.nspec = int(nspec)
        self.h = self.L.cs_gpu_dedisp_create(C.byref(self.cfg), self.nspec, int(device))
        if not self.h:
            raise ValueError("de-dispersion configuration rejected (see stderr)")
        self.numdms = int(numdms)
        self.numchan = int(numchan)

    @property
    def info(self):
        i = abi.CsDedispInfo()
        self.L.cs_gpu_dedisp_info(self.h, C.byref(i))
        return {k: getattr(i, k) for k, _ in abi.CsDedispInfo._fields_ if k != "pad_"}

    @property
    def numout(self):
        return self.info["numout"]

    @property
    def dt_out(self):
        return self.info["dsdt"]

    def tables(self):
        """(dms, idispdt[numchan], offsets[numdms, nsub])"""
        dms = np.zeros(self.numdms)
        idisp = np.zeros(self.numchan, np.int32)
        offs = np.zeros((self.numdms, self.cfg.nsub), np.int32)
        self.L.cs_gpu_dedisp_tables(self.h, dms.ctypes.data, idisp.ctypes.data,
                                    offs.ctypes.data)
        return dms, idisp, offs

    def run(self, spectra):
        """De-disperse [nspec, numchan] float32 or uint8 spectra (a numpy
        array, or a DeviceSeries-like object with .ptr for spectra already in
        HBM, then `uint8` tells the type).  The series stay on the device."""
        if isinstance(spectra, np.ndarray):
            if spectra.dtype not in (np.float32, np.uint8):
                raise ValueError("spectra must be float32 or uint8")
            if spectra.shape != (self.nspec, self.numchan) or not spectra.flags.c_contiguous:
                raise ValueError("spectra must be C-contiguous [%d, %d]"
                                 % (self.nspec, self.numchan))
            ptr, u8 = spectra.ctypes.data, spectra.dtype == np.uint8
        else:
            ptr, u8 = int(spectra.ptr), bool(getattr(spectra, "uint8", False))
        if self.L.cs_gpu_dedisp_run(self.h, C.c_void_p(ptr), int(u8)) != 0:
            raise RuntimeError("cs_gpu_dedisp_run failed (see stderr)")
        return self

    def series(self, i):
        p = self.L.cs_gpu_dedisp_series(self.h, int(i))
        if not p:
            raise IndexError(i)
        return DeviceSeries(p, self.numout, self)

    def copy_out(self, i0=0, n=None):
        n = self.numdms - i0 if n is None else n
        out = np.empty((n, self.numout), np.float32)
        if self.L.cs_gpu_dedisp_copy_out(self.h, int(i0), int(n), out.ctypes.data) != 0:
            raise RuntimeError("cs_gpu_dedisp_copy_out failed")
        return out

    def last_ms(self):
        """(device ms of the last run, subband kernels ms, DM kernels ms)"""
        a, b = C.c_double(), C.c_double()
        t = self.L.cs_gpu_dedisp_last_ms(self.h, C.byref(a), C.byref(b))
        return t, a.value, b.value

    def close(self):
        if self.h:
            self.L.cs_gpu_dedisp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def candidate_path(outfile, dms, i, dmprecision=2, outdir=None):
    """<outfile>_DM<dm>.cohout, or that name under outdir
    (src/prep_search.c:30-42)"""
    base = outfile
    if outdir:
        return os.path.join(outdir, "%s_DM%.*f.cohout" % (os.path.basename(outfile),
                                                           dmprecision, dms[i]))
    return "%s_DM%.*f.cohout" % (base, dmprecision, dms[i])


def prep_search(dd, engine, params, ncands=100, threshold=6.0, lofreq=0.1, hifreq=125.0,
                lobin=100, blocksize=65536, outfile=None, dmprecision=2, outdir=None,
                zaplist=None, baryv=0.0):
    """Search every DM series of a run plan (prep_search_run,
    src/prep_search.c:44-145): each series is staged (realfft + rednoise),
    searched, collapsed, sorted by S/N, cut to ncands and given duty cycles
    on the device, all requests in flight together on the engine.  Returns
    the final candidate arrays in DM order; with `outfile`, also writes each
    one's .cohout under the reference's name."""
    dms, _, _ = dd.tables()
    tickets = {}
    for i in range(dd.numdms):
        t = engine.submit(dd.series(i), dd.dt_out, params, lofreq=lofreq, hifreq=hifreq,
                          lobin=lobin, blocksize=blocksize, threshold=threshold,
                          ncands=ncands, zaplist=zaplist, baryv=baryv)
        tickets[t] = i
    finals = [None] * dd.numdms
    for _ in range(dd.numdms):
        t = engine.next_done()
        lst, info, _ = engine.wait(t)
        if info["status"] != 0:
            raise RuntimeError("search of DM %g failed" % dms[tickets[t]])
        finals[tickets[t]] = lst
    if outfile:
        L = abi.host()
        if outdir:
            os.makedirs(outdir, exist_ok=True)
        for i, lst in enumerate(finals):
            cl = to_list(lst)
            rc = L.write_candidates(C.byref(cl), candidate_path(outfile, dms, i, dmprecision,
                                                                outdir).encode(),
                                    float(threshold))
            L.candlist_free(C.byref(cl))
            if rc != 0:
                raise RuntimeError("cannot write the candidates of DM %g" % dms[i])
    return finals
