"""gcorr_amd: MI355X-native FX correlation (the fxkernel/gcorr per-subintegration path).

The compute runs in gcorr_amd/libgxcorr.so (hand-written gfx950 HIP kernels
behind the C-ABI of include/gxcorr.h); this package is its Python host-side
mirror of fxkernel's interface.
"""
from ._lib import GxcError, load  # noqa: F401
from .batch import Batch  # noqa: F401
from .config import CorrConfig, parse_config, write_config  # noqa: F401
from .fxkernel import (FxKernel, PinnedBuffer, lag_snr, num_rows, requantise_2bit,  # noqa: F401
                       requantise_8bit, station_bytes, xengine)

__all__ = ["Batch", "FxKernel", "PinnedBuffer", "GxcError", "CorrConfig", "parse_config", "write_config", "num_rows",
           "station_bytes", "xengine", "lag_snr", "requantise_2bit", "requantise_8bit", "load"]
