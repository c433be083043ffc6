"""MI355X-native coherent harmonic-summing search (gfx950 HIP kernels behind
the reference's cs_gpu_* C-ABI, with a C host library mirroring the
reference's search_file / plan / candidate interface)."""
from . import abi
from .dedisp import Dedisp, DeviceSeries, prep_search
from .search import (CAND_DTYPE, MetricStats, Session, Spectrum, Stage, Zaplist, autotune,
                     collapse,
                     empty_spectrum,
                     finalize, gpu_available, make_params, redfft, search_file, take_list,
                     to_list, zap)

__all__ = ["abi", "Dedisp", "DeviceSeries", "prep_search", "CAND_DTYPE", "MetricStats", "Session", "Spectrum", "Stage", "collapse", "empty_spectrum",
           "finalize", "gpu_available", "make_params", "redfft", "search_file", "take_list",
           "to_list", "Zaplist", "zap", "autotune"]
