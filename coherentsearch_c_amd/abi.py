"""ctypes mirror of the C-ABI (include/cs_types.h, cs_gpu.h, cs_host.h).

The structs keep the reference's layouts (include/coherent_search.h:48-195);
sizes are asserted at import.  The libraries are loaded from this package's
directory and a missing or unbuildable library raises -- there is no fallback
implementation.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(HERE, "libcs_hip.so")
HOST_LIB = os.path.join(HERE, "libcs_host.so")


class SimpleInf(C.Structure):
    _fields_ = [("path", C.c_char_p), ("object", C.c_char_p), ("has_epoch", C.c_int),
                ("epoch", C.c_double), ("N", C.c_longlong), ("has_N", C.c_int),
                ("dt", C.c_double), ("has_dt", C.c_int), ("DM", C.c_double),
                ("has_DM", C.c_int)]


class FFTFile(C.Structure):
    _fields_ = [("path", C.c_char_p), ("inf", SimpleInf), ("N", C.c_longlong),
                ("T", C.c_double), ("df", C.c_double), ("dereddened", C.c_int),
                ("detrended", C.c_int), ("DC", C.c_float), ("Nyquist", C.c_float),
                ("amps", C.POINTER(C.c_float)), ("namps", C.c_longlong),
                ("map", C.c_void_p), ("mapsize", C.c_size_t), ("owns_amps", C.c_int)]


class SearchParams(C.Structure):
    _fields_ = [("nharms", C.c_int), ("m", C.c_int), ("numbetween", C.c_int),
                ("hidr", C.c_double), ("threshold", C.c_double),
                ("boxcar_fsp", C.c_double), ("boxcar_maxfrac", C.c_double),
                ("boxcar_gatemargin", C.c_double), ("ndecim", C.c_int),
                ("decimations", C.c_int * 64), ("precision", C.c_int), ("sigma", C.c_int)]


class DirectPlan(C.Structure):
    _fields_ = [("h", C.c_int), ("m", C.c_int), ("q", C.c_int), ("s", C.c_int),
                ("base_adv", C.c_int), ("pnum", C.c_int), ("rfloor0", C.c_int),
                ("dr0", C.c_double), ("P", C.c_int),
                ("row", C.POINTER(C.c_int32)), ("W", C.POINTER(C.c_float)),
                ("A_re", C.POINTER(C.c_float)), ("A_im", C.POINTER(C.c_float)),
                ("carry", C.POINTER(C.c_uint8)), ("ngrp", C.c_int),
                ("grow", C.POINTER(C.c_int32)), ("goff", C.POINTER(C.c_int32)),
                ("gnj", C.POINTER(C.c_int32)), ("gW", C.POINTER(C.c_float)),
                ("gA_re", C.POINTER(C.c_float)), ("gA_im", C.POINTER(C.c_float))]


class DirectPlanSet(C.Structure):
    _fields_ = [("plans", C.POINTER(DirectPlan)), ("n", C.c_int)]


class Candidate(C.Structure):
    _fields_ = [("freq", C.c_double), ("metric", C.c_double), ("r", C.c_double),
                ("nharm", C.c_int), ("ducy", C.c_double)]


class CandidateList(C.Structure):
    _fields_ = [("v", C.POINTER(Candidate)), ("n", C.c_size_t), ("cap", C.c_size_t)]


class ZapBird(C.Structure):
    _fields_ = [("freq", C.c_double), ("width", C.c_double), ("bary", C.c_int)]


class ZapList(C.Structure):
    _fields_ = [("v", C.POINTER(ZapBird)), ("n", C.c_int)]


class SearchCache(C.Structure):
    _fields_ = [("Nprof", C.c_int), ("plan_valid", C.c_int),
                ("params", C.POINTER(SearchParams)), ("plan_T", C.c_double),
                ("plan_N", C.c_longlong), ("plan_namps", C.c_longlong),
                ("plan_lofreq", C.c_double), ("plan_hifreq", C.c_double),
                ("plan_lobin", C.c_int), ("plan_Nprof", C.c_int),
                ("plan_total", C.c_longlong), ("dplans", DirectPlanSet),
                ("cstarts", C.POINTER(C.c_longlong)), ("nchunks", C.c_longlong),
                ("gpu_session", C.c_void_p)]


class CsSearchRequest(C.Structure):
    _fields_ = [("samples", C.c_void_p), ("N", C.c_longlong), ("dt", C.c_double),
                ("p", C.POINTER(SearchParams)), ("lofreq", C.c_double), ("hifreq", C.c_double),
                ("lobin", C.c_int), ("blocksize", C.c_int), ("threshold", C.c_double),
                ("amps_out", C.c_void_p), ("ncands", C.c_int), ("numharm", C.c_int),
                ("remove", C.c_int), ("harm_remove", C.c_int), ("dr_tol", C.c_double),
                ("harm_tol", C.c_double), ("zaplist", C.POINTER(ZapList)),
                ("baryv", C.c_double)]


class CsSearchInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("reused", C.c_int), ("harmonic_trials", C.c_longlong),
                ("trials", C.c_longlong), ("footprint", C.c_longlong),
                ("queue_ms", C.c_double), ("run_ms", C.c_double), ("latency_ms", C.c_double),
                ("kernel_ms", C.c_double), ("plan_ms", C.c_double), ("stage_ms", C.c_double),
                ("search_ms", C.c_double), ("post_ms", C.c_double)]


class CsTuneResult(C.Structure):
    _fields_ = [("nharms", C.c_int), ("m_padded", C.c_int), ("precision", C.c_int),
                ("sigma", C.c_int), ("lane_order", C.c_int), ("xcd_stripe", C.c_int),
                ("blocksize", C.c_int), ("ntried", C.c_int), ("default_ms", C.c_double),
                ("best_ms", C.c_double)]


class CsEngineStats(C.Structure):
    _fields_ = [("budget", C.c_longlong), ("in_use", C.c_longlong),
                ("peak_in_use", C.c_longlong), ("nslots", C.c_int), ("running", C.c_int),
                ("peak_running", C.c_int), ("pad", C.c_int), ("queued", C.c_longlong),
                ("completed", C.c_longlong), ("plan_hits", C.c_longlong),
                ("plans_resident", C.c_longlong)]


class BlockMetricStats(C.Structure):
    _fields_ = [("block", C.c_int), ("k", C.c_int), ("Hk", C.c_int), ("nbins", C.c_int),
                ("ngoodbins", C.c_double), ("flo", C.c_double), ("fhi", C.c_double),
                ("n", C.c_longlong), ("min", C.c_double), ("median", C.c_double),
                ("mean", C.c_double), ("std", C.c_double), ("max", C.c_double)]


class MetricHistogram(C.Structure):
    _fields_ = [("k", C.c_int), ("Hk", C.c_int), ("nbins", C.c_int), ("win", C.c_int),
                ("flo", C.c_double), ("fhi", C.c_double), ("lo", C.c_double),
                ("hi", C.c_double), ("invw", C.c_double),
                ("counts", C.POINTER(C.c_longlong)), ("nb", C.c_int),
                ("under", C.c_longlong), ("over", C.c_longlong), ("total", C.c_longlong),
                ("sum", C.c_double), ("sumsq", C.c_double), ("vmin", C.c_double),
                ("vmax", C.c_double)]


class MetricStats(C.Structure):
    _fields_ = [("hist_lo", C.c_double), ("hist_hi", C.c_double), ("hist_nb", C.c_int),
                ("nwin", C.c_int), ("blocks", C.POINTER(BlockMetricStats)),
                ("nblocks", C.c_size_t), ("capblocks", C.c_size_t),
                ("hists", C.POINTER(MetricHistogram)), ("nhists", C.c_size_t),
                ("caphists", C.c_size_t), ("whists", C.POINTER(MetricHistogram)),
                ("nwhists", C.c_size_t), ("capwhists", C.c_size_t)]


class NormK(C.Structure):
    _fields_ = [("k", C.c_int), ("edges", C.POINTER(C.c_double)),
                ("loc", C.POINTER(C.c_double)), ("scale", C.POINTER(C.c_double)),
                ("nwin", C.c_int)]


class MetricNorm(C.Structure):
    _fields_ = [("ks", C.POINTER(NormK)), ("n", C.c_int)]


class CsMetricBlock(C.Structure):
    _fields_ = [("n", C.c_longlong), ("min", C.c_double), ("max", C.c_double),
                ("sum", C.c_double), ("sumsq", C.c_double), ("ss", C.c_double),
                ("median", C.c_double), ("under", C.c_longlong), ("over", C.c_longlong)]


assert C.sizeof(SimpleInf) == 80 and C.sizeof(FFTFile) == 168
assert C.sizeof(SearchParams) == 328 and C.sizeof(DirectPlan) == 144
assert C.sizeof(Candidate) == 40 and C.sizeof(CandidateList) == 24

class CsDedispConfig(C.Structure):
    _fields_ = [("numchan", C.c_int), ("lofreq", C.c_double), ("chan_wid", C.c_double),
                ("dt", C.c_double), ("spectra_per_subint", C.c_int), ("nsub", C.c_int),
                ("lodm", C.c_double), ("dmstep", C.c_double), ("numdms", C.c_int),
                ("downsamp", C.c_int), ("numout", C.c_longlong), ("clip_sigma", C.c_float)]


class CsDedispInfo(C.Structure):
    _fields_ = [("blocksperread", C.c_int), ("worklen", C.c_int), ("nsub", C.c_int),
                ("max_offset", C.c_int), ("numout", C.c_longlong), ("written", C.c_longlong),
                ("chunk", C.c_longlong), ("dsdt", C.c_double), ("pad", C.c_float),
                ("pad_", C.c_int)]


# Every symbol the public headers declare, per library.
HIP_SYMBOLS = ["cs_gpu_available", "cs_gpu_session_create", "cs_gpu_session_matches",
               "cs_gpu_session_run", "cs_gpu_session_free", "cs_gpu_search_region",
               "cs_gpu_session_dump_stack", "cs_gpu_session_last_kernel_ms",
               "cs_gpu_session_upload", "cs_gpu_session_search_resident",
               "cs_gpu_session_search_noalloc",
               "cs_gpu_session_work", "cs_gpu_stage_create", "cs_gpu_stage_free",
               "cs_gpu_stage_run_host", "cs_gpu_session_stage", "cs_gpu_session_set_graphs",
               "cs_gpu_session_finalize", "cs_gpu_session_zap", "cs_gpu_autotune",
               "cs_gpu_tune_save", "cs_gpu_tune_load", "cs_engine_create",
               "cs_engine_submit", "cs_engine_poll", "cs_engine_next_done", "cs_engine_wait",
               "cs_engine_stats", "cs_engine_free", "cs_engine_set_plan_cache", "cs_host_alloc",
               "cs_host_free", "cs_gpu_session_metric_stats",
               "cs_gpu_session_metric_candidates",
               # include/cs_dedisp.h
               "cs_gpu_dedisp_create", "cs_gpu_dedisp_free", "cs_gpu_dedisp_info",
               "cs_gpu_dedisp_tables", "cs_gpu_dedisp_run", "cs_gpu_dedisp_series",
               "cs_gpu_dedisp_copy_out", "cs_gpu_dedisp_last_ms", "cs_choose_good_N",
               "cs_delay_from_dm"]
HOST_SYMBOLS = ["params_defaults", "decimation_set", "trial_grid_rational",
                "direct_plans_build", "direct_plans_free", "cs_chunk_state",
                "cs_harmonic_fits", "harmonic_change_trials", "chunk_starts", "cs_sinpi",
                "cs_cospi", "cs_eps", "fourier_interp_complex", "boxcar_widths",
                "ladder_boxcar_widths", "analytic_sigma", "boxcar_shape",
                "boxcar_best_width", "candlist_init", "candlist_free", "candlist_push",
                "candlist_append", "candlist_sort_metric_desc", "candlist_sort_freq",
                "candlist_sort_r", "candlist_sort_metric_desc_stable",
                "candlist_sort_freq_stable", "candlist_sort_r_stable", "remove_duplicates",
                "remove_harmonics", "harmonically_related", "candidate_profile",
                "measure_ducy", "write_candidates", "cs_format_double",
                "read_candidates_file", "inf_read", "inf_free", "zaplist_read",
                "zaplist_free", "zaplist_apply", "fftfile_open",
                "fftfile_close", "datpipe_create", "datpipe_free", "datpipe_set_threads", "datpipe_set_device",
                "datpipe_redfft",
                "datpipe_redfft_mem", "cs_open_input", "cs_write_inf", "cs_write_floats",
                "metricstats_init", "metricstats_free", "metricstats_append_block",
                "metricstats_append_hist", "metricstats_append_whist", "hist_init", "hist_free",
                "hist_push", "hist_merge", "hist_quantile", "log_edges", "window_index",
                "build_metricnorm", "metricnorm_free", "metric_normalize", "report_metricstats",
                "search_file", "search_cache_free", "cs_main", "cs_main_list"]

_host = None
_hip = None


def _declare(L):
    P, I, D, LL = C.c_void_p, C.c_int, C.c_double, C.c_longlong
    Ps = C.POINTER
    L.params_defaults.argtypes = [Ps(SearchParams)]
    L.decimation_set.argtypes = [I, I, Ps(C.c_int), I]
    L.decimation_set.restype = I
    L.trial_grid_rational.argtypes = [D, Ps(I), Ps(I)]
    L.trial_grid_rational.restype = I
    L.direct_plans_build.argtypes = [Ps(DirectPlanSet), Ps(SearchParams), D]
    L.direct_plans_build.restype = I
    L.direct_plans_free.argtypes = [Ps(DirectPlanSet)]
    L.chunk_starts.argtypes = [Ps(DirectPlanSet), Ps(FFTFile), Ps(SearchParams), LL, I,
                               Ps(Ps(LL)), Ps(LL)]
    L.chunk_starts.restype = I
    L.cs_harmonic_fits.argtypes = [Ps(DirectPlan), Ps(FFTFile), Ps(SearchParams), LL]
    L.cs_harmonic_fits.restype = I
    L.ladder_boxcar_widths.argtypes = [I, I, Ps(SearchParams), Ps(C.c_int), I]
    L.ladder_boxcar_widths.restype = I
    L.boxcar_widths.argtypes = [I, D, D, Ps(C.c_int), I]
    L.boxcar_widths.restype = I
    L.analytic_sigma.argtypes = [P, I, I]
    L.analytic_sigma.restype = D
    L.candlist_free.argtypes = [Ps(CandidateList)]
    L.remove_duplicates.argtypes = [Ps(CandidateList), D]
    L.remove_duplicates.restype = CandidateList
    L.remove_harmonics.argtypes = [Ps(CandidateList), I, D]
    L.remove_harmonics.restype = CandidateList
    L.candlist_sort_metric_desc.argtypes = [Ps(CandidateList)]
    L.candlist_sort_freq_stable.argtypes = [Ps(CandidateList)]
    L.measure_ducy.argtypes = [Ps(FFTFile), Ps(CandidateList), Ps(SearchParams)]
    L.measure_ducy.restype = CandidateList
    L.write_candidates.argtypes = [Ps(CandidateList), C.c_char_p, D]
    L.write_candidates.restype = I
    L.fftfile_open.argtypes = [Ps(FFTFile), C.c_char_p]
    L.zaplist_read.argtypes = [C.c_char_p, Ps(ZapList)]
    L.zaplist_read.restype = I
    L.zaplist_free.argtypes = [Ps(ZapList)]
    L.zaplist_free.restype = None
    L.zaplist_apply.argtypes = [Ps(ZapList), Ps(FFTFile), D]
    L.zaplist_apply.restype = I
    L.fftfile_open.restype = I
    L.fftfile_close.argtypes = [Ps(FFTFile)]
    L.datpipe_create.argtypes = [I, I, D]
    L.datpipe_create.restype = P
    L.datpipe_free.argtypes = [P]
    L.datpipe_set_threads.argtypes = [P, I]
    L.datpipe_set_threads.restype = None
    L.datpipe_set_device.argtypes = [P, I]
    L.datpipe_set_device.restype = None
    L.datpipe_redfft_mem.argtypes = [P, P, LL, D, C.c_char_p, Ps(FFTFile)]
    L.datpipe_redfft_mem.restype = I
    L.cs_write_inf.argtypes = [C.c_char_p, LL, D, D, C.c_char_p]
    L.cs_write_inf.restype = I
    L.cs_write_floats.argtypes = [C.c_char_p, P, LL]
    L.cs_write_floats.restype = I
    L.search_file.argtypes = [Ps(FFTFile), Ps(SearchParams), D, D, I, I, D, I, D, I, I,
                              Ps(MetricStats), I, I, I, C.c_char_p, Ps(SearchCache), D, I]
    L.metricstats_init.argtypes = [Ps(MetricStats), D, D, I, I]
    L.metricstats_init.restype = None
    L.metricstats_free.argtypes = [Ps(MetricStats)]
    L.metricstats_free.restype = None
    L.hist_init.argtypes = [Ps(MetricHistogram), I, I, I, D, D, D, D, I]
    L.hist_init.restype = None
    L.hist_free.argtypes = [Ps(MetricHistogram)]
    L.hist_free.restype = None
    L.hist_push.argtypes = [Ps(MetricHistogram), D]
    L.hist_push.restype = None
    L.hist_merge.argtypes = [Ps(MetricHistogram), Ps(MetricHistogram)]
    L.hist_merge.restype = None
    L.hist_quantile.argtypes = [Ps(MetricHistogram), D]
    L.hist_quantile.restype = D
    L.log_edges.argtypes = [D, D, I, Ps(D)]
    L.log_edges.restype = None
    L.window_index.argtypes = [Ps(D), I, D]
    L.window_index.restype = I
    L.metricstats_append_whist.argtypes = [Ps(MetricStats), MetricHistogram]
    L.metricstats_append_whist.restype = None
    L.metricstats_append_hist.argtypes = [Ps(MetricStats), MetricHistogram]
    L.metricstats_append_hist.restype = None
    L.metricstats_append_block.argtypes = [Ps(MetricStats), BlockMetricStats]
    L.metricstats_append_block.restype = None
    L.build_metricnorm.argtypes = [Ps(MetricStats)]
    L.build_metricnorm.restype = Ps(MetricNorm)
    L.metricnorm_free.argtypes = [Ps(MetricNorm)]
    L.metricnorm_free.restype = None
    L.metric_normalize.argtypes = [Ps(MetricNorm), I, D, D]
    L.metric_normalize.restype = D
    L.report_metricstats.argtypes = [Ps(MetricStats), C.c_char_p, Ps(SearchParams), D]
    L.report_metricstats.restype = None
    L.search_file.restype = CandidateList
    L.search_cache_free.argtypes = [Ps(SearchCache)]
    L.cs_main.argtypes = [I, Ps(C.c_char_p)]
    L.cs_main.restype = I
    L.cs_main_list.argtypes = [I, Ps(C.c_char_p)]
    L.cs_main_list.restype = I
    # GPU backend
    L.cs_gpu_available.argtypes = [I, C.c_char_p, C.c_size_t]
    L.cs_gpu_available.restype = I
    L.cs_gpu_session_create.argtypes = [Ps(FFTFile), Ps(SearchParams), Ps(DirectPlanSet),
                                        Ps(LL), LL, I, LL, D, D, I]
    L.cs_gpu_session_create.restype = P
    L.cs_gpu_session_matches.argtypes = [P, Ps(FFTFile), Ps(SearchParams), I, LL, D, D, I]
    L.cs_gpu_session_matches.restype = I
    L.cs_gpu_session_run.argtypes = [P, Ps(FFTFile), D]
    L.cs_gpu_session_run.restype = CandidateList
    L.cs_gpu_session_free.argtypes = [P]
    L.cs_gpu_search_region.argtypes = [Ps(FFTFile), Ps(SearchParams), Ps(DirectPlanSet),
                                       Ps(LL), LL, I, LL, D, D, D, D, I]
    L.cs_gpu_search_region.restype = CandidateList
    L.cs_gpu_session_dump_stack.argtypes = [P, Ps(FFTFile), LL, LL, P]
    L.cs_gpu_session_dump_stack.restype = I
    L.cs_gpu_session_upload.argtypes = [P, Ps(FFTFile)]
    L.cs_gpu_session_upload.restype = I
    L.cs_gpu_session_search_resident.argtypes = [P, D]
    L.cs_gpu_session_search_resident.restype = CandidateList
    L.cs_gpu_session_search_noalloc.argtypes = [P, D, P, LL]
    L.cs_gpu_session_search_noalloc.restype = LL
    L.cs_gpu_session_last_kernel_ms.argtypes = [P]
    L.cs_gpu_session_last_kernel_ms.restype = D
    L.cs_gpu_session_work.argtypes = [P, Ps(LL), Ps(LL), Ps(D)]
    L.cs_gpu_session_work.restype = I
    L.cs_gpu_stage_create.argtypes = [LL, D, I, I, D, I]
    L.cs_gpu_stage_create.restype = P
    L.cs_gpu_stage_free.argtypes = [P]
    L.cs_gpu_stage_free.restype = None
    L.cs_gpu_stage_run_host.argtypes = [P, P, P, P]
    L.cs_gpu_stage_run_host.restype = I
    L.cs_gpu_session_stage.argtypes = [P, P, P, P]
    L.cs_gpu_session_stage.restype = I
    L.cs_gpu_session_set_graphs.argtypes = [P, I]
    L.cs_gpu_session_set_graphs.restype = I
    L.cs_gpu_autotune.argtypes = [I, Ps(SearchParams), Ps(CsTuneResult)]
    L.cs_gpu_autotune.restype = I
    L.cs_gpu_tune_save.argtypes = [C.c_char_p]
    L.cs_gpu_tune_save.restype = I
    L.cs_gpu_tune_load.argtypes = [C.c_char_p]
    L.cs_gpu_tune_load.restype = I
    L.cs_gpu_session_zap.argtypes = [P, Ps(ZapList), D, P]
    L.cs_gpu_session_zap.restype = I
    L.cs_gpu_session_metric_stats.argtypes = [P, Ps(FFTFile), Ps(I), I, D, D, I,
                                              Ps(CsMetricBlock), Ps(LL)]
    L.cs_gpu_session_metric_stats.restype = I
    L.cs_gpu_session_metric_candidates.argtypes = [P, D, I, Ps(I), Ps(D), Ps(D), Ps(D)]
    L.cs_gpu_session_metric_candidates.restype = CandidateList
    L.cs_gpu_session_finalize.argtypes = [P, Ps(CandidateList), I]
    L.cs_gpu_session_finalize.restype = CandidateList
    L.cs_engine_create.argtypes = [I, LL, I]
    L.cs_engine_create.restype = P
    L.cs_engine_submit.argtypes = [P, Ps(CsSearchRequest)]
    L.cs_engine_submit.restype = LL
    L.cs_engine_poll.argtypes = [P, LL]
    L.cs_engine_poll.restype = I
    L.cs_engine_next_done.argtypes = [P, D]
    L.cs_engine_next_done.restype = LL
    L.cs_engine_wait.argtypes = [P, LL, Ps(CandidateList), Ps(CsSearchInfo)]
    L.cs_engine_wait.restype = I
    L.cs_engine_stats.argtypes = [P, Ps(CsEngineStats)]
    L.cs_engine_stats.restype = None
    L.cs_engine_set_plan_cache.argtypes = [P, I]
    L.cs_engine_set_plan_cache.restype = None
    L.cs_engine_free.argtypes = [P]
    L.cs_engine_free.restype = None
    L.cs_host_alloc.argtypes = [LL]
    L.cs_host_alloc.restype = P
    L.cs_host_free.argtypes = [P]
    L.cs_host_free.restype = None
    L.cs_gpu_dedisp_create.argtypes = [Ps(CsDedispConfig), LL, I]
    L.cs_gpu_dedisp_create.restype = P
    L.cs_gpu_dedisp_free.argtypes = [P]
    L.cs_gpu_dedisp_free.restype = None
    L.cs_gpu_dedisp_info.argtypes = [P, Ps(CsDedispInfo)]
    L.cs_gpu_dedisp_info.restype = I
    L.cs_gpu_dedisp_tables.argtypes = [P, P, P, P]
    L.cs_gpu_dedisp_tables.restype = I
    L.cs_gpu_dedisp_run.argtypes = [P, P, I]
    L.cs_gpu_dedisp_run.restype = I
    L.cs_gpu_dedisp_series.argtypes = [P, I]
    L.cs_gpu_dedisp_series.restype = P
    L.cs_gpu_dedisp_copy_out.argtypes = [P, I, I, P]
    L.cs_gpu_dedisp_copy_out.restype = I
    L.cs_gpu_dedisp_last_ms.argtypes = [P, Ps(D), Ps(D)]
    L.cs_gpu_dedisp_last_ms.restype = D
    L.cs_choose_good_N.argtypes = [LL]
    L.cs_choose_good_N.restype = LL
    L.cs_delay_from_dm.argtypes = [D, D]
    L.cs_delay_from_dm.restype = D
    L.free_ = C.CDLL(None).free
    L.free_.argtypes = [P]


def host():
    """libcs_host.so (also pulls in libcs_hip.so).  Raises if not built."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB) or not os.path.exists(HIP_LIB):
            raise RuntimeError(
                "coherentsearch_c_amd native libraries are not built (%s); run "
                "__graft_entry__.build() or make -C coherentsearch_c_amd/csrc" % HERE)
        L = C.CDLL(HOST_LIB, mode=C.RTLD_GLOBAL)
        _declare(L)
        _host = L
    return _host


def hip():
    """libcs_hip.so on its own (for symbol checks)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError("libcs_hip.so is not built")
        _hip = C.CDLL(HIP_LIB)
    return _hip
