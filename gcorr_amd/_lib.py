"""ctypes bindings of include/gxcorr.h (gcorr_amd/libgxcorr.so).

The library is the product: there is no fallback.  If it is missing or fails
to load, every entry point raises.
"""
import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "libgxcorr.so")
HEADER = os.path.join(ROOT, "include", "gxcorr.h")

GXC_OK = 0
STATUS = {0: "ok", 1: "invalid argument", 2: "unsupported", 3: "no device", 4: "HIP error",
          5: "out of memory", 6: "bad state", 7: "I/O error"}
PRECISION = {"fp32": 0, "fp16": 1, "bf16": 2, "fp8": 3}


class GxcError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        super().__init__("%s (%d): %s" % (STATUS.get(status, "?"), status, message))


class Config(ctypes.Structure):
    _fields_ = [
        ("nant", ctypes.c_int), ("nchan", ctypes.c_int), ("numffts", ctypes.c_int),
        ("nbit", ctypes.c_int), ("npol", ctypes.c_int), ("is_complex", ctypes.c_int),
        ("lo", ctypes.c_double), ("bw", ctypes.c_double),
        ("precision", ctypes.c_int), ("autos", ctypes.c_int), ("exact_phase", ctypes.c_int),
        ("chunk_ffts", ctypes.c_int), ("overlap", ctypes.c_int), ("graph", ctypes.c_int),
        ("fblock", ctypes.c_int), ("xtile", ctypes.c_int), ("xsplits", ctypes.c_int),
        ("reserved", ctypes.c_int * 1),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("fblock", ctypes.c_int), ("xtile", ctypes.c_int), ("xsplits", ctypes.c_int),
        ("graph", ctypes.c_int), ("fstage_ms", ctypes.c_double), ("xstage_ms", ctypes.c_double),
        ("process_ms", ctypes.c_double), ("default_ms", ctypes.c_double),
        ("candidates", ctypes.c_int),
    ]


class VdifSpec(ctypes.Structure):
    _fields_ = [
        ("frame_bytes", ctypes.c_int), ("thread_id", ctypes.c_int),
        ("frames_per_second", ctypes.c_int), ("ref_epoch", ctypes.c_int),
        ("ref_second", ctypes.c_longlong), ("ref_frame", ctypes.c_int),
        ("offset_binary8", ctypes.c_int), ("header_bytes", ctypes.c_int),
        ("reserved", ctypes.c_int),
    ]


class VdifStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_longlong) for k in
                ("frames", "placed", "missing", "duplicate", "invalid", "other_thread",
                 "out_of_window", "malformed")]


class DifxMeta(ctypes.Structure):
    _fields_ = [
        ("mjd", ctypes.c_int), ("seconds", ctypes.c_double), ("config_index", ctypes.c_int),
        ("source_index", ctypes.c_int), ("freq_index", ctypes.c_int), ("pols", ctypes.c_char * 2),
        ("uvw", ctypes.POINTER(ctypes.c_double)),
    ]


class Timings(ctypes.Structure):
    _fields_ = [
        ("h2d_ms", ctypes.c_double), ("fstage_ms", ctypes.c_double),
        ("xstage_ms", ctypes.c_double), ("finalise_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double), ("fstage_launches", ctypes.c_int),
        ("xstage_launches", ctypes.c_int),
    ]


class ConfFile(ctypes.Structure):
    _fields_ = [
        ("cfg", Config), ("nant_read", ctypes.c_int),
        ("names", ctypes.POINTER(ctypes.c_char_p)), ("files", ctypes.POINTER(ctypes.c_char_p)),
        ("delays", ctypes.POINTER(ctypes.c_double)), ("offsets", ctypes.POINTER(ctypes.c_double)),
    ]


class Subband(ctypes.Structure):
    _fields_ = [("nchan", ctypes.c_int), ("numffts", ctypes.c_int),
                ("lo", ctypes.c_double), ("bw", ctypes.c_double)]


class BatchConfig(ctypes.Structure):
    _fields_ = [
        ("nant", ctypes.c_int), ("nbit", ctypes.c_int), ("npol", ctypes.c_int),
        ("is_complex", ctypes.c_int), ("precision", ctypes.c_int), ("autos", ctypes.c_int),
        ("exact_phase", ctypes.c_int), ("nsub", ctypes.c_int),
        ("subbands", ctypes.POINTER(Subband)), ("chunks", ctypes.c_int),
        ("overlap", ctypes.c_int), ("graph", ctypes.c_int), ("fstage_mfma", ctypes.c_int),
        ("reserved", ctypes.c_int * 3),
    ]


class CF32(ctypes.Structure):
    _fields_ = [("re", ctypes.c_float), ("im", ctypes.c_float)]


_lib = None


def declared_functions():
    """Names of the functions include/gxcorr.h declares."""
    with open(HEADER) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gxc_[a-z0-9_]+)\s*\(", text)))


def load():
    """Load libgxcorr.so and declare argument types.  Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GxcError(4, "libgxcorr.so not built; run `python -m gcorr_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp, u8p, dp, ip = ctypes.c_void_p, P(ctypes.c_uint8), P(ctypes.c_double), P(ctypes.c_int)
    H = vp
    sig = {
        "gxc_config_init": (None, [P(Config)]),
        "gxc_create": (ctypes.c_int, [P(Config), P(vp)]),
        "gxc_destroy": (None, [H]),
        "gxc_last_error": (ctypes.c_char_p, [H]),
        "gxc_status_string": (ctypes.c_char_p, [ctypes.c_int]),
        "gxc_set_input": (ctypes.c_int, [H, P(u8p)]),
        "gxc_set_input_device": (ctypes.c_int, [H, P(vp)]),
        "gxc_set_delays": (ctypes.c_int, [H, P(dp), dp]),
        "gxc_set_pol_conversion": (ctypes.c_int, [H, vp]),
        "gxc_requantise_2bit": (ctypes.c_int, [vp, vp, ctypes.c_size_t, ctypes.c_float, vp, vp]),
        "gxc_requantise_8bit": (ctypes.c_int, [vp, vp, ctypes.c_size_t, ctypes.c_float, vp, vp]),
        "gxc_vdif_spec_init": (None, [P(VdifSpec)]),
        "gxc_vdif_slots": (ctypes.c_longlong, [P(Config), ctypes.c_int, ctypes.c_int]),
        "gxc_set_input_vdif": (ctypes.c_int, [H, ctypes.c_int, vp, ctypes.c_size_t, P(VdifSpec),
                                              P(VdifStats)]),
        "gxc_get_frame_flags": (ctypes.c_int, [H, ctypes.c_int, u8p, ctypes.c_size_t,
                                               P(ctypes.c_longlong)]),
        "gxc_process": (ctypes.c_int, [H]),
        "gxc_process_async": (ctypes.c_int, [H]),
        "gxc_synchronize": (ctypes.c_int, [H]),
        "gxc_accumulate": (ctypes.c_int, [H, vp]),
        "gxc_get_visibilities": (ctypes.c_int, [H, vp, ctypes.c_size_t]),
        "gxc_get_row_counts": (ctypes.c_int, [H, ip, ctypes.c_size_t]),
        "gxc_save_visibilities": (ctypes.c_int, [H, ctypes.c_char_p]),
        "gxc_visibilities_device": (vp, [H]),
        "gxc_save_visibilities_binary": (ctypes.c_int, [H, ctypes.c_char_p]),
        "gxc_save_difx": (ctypes.c_int, [H, ctypes.c_char_p, P(DifxMeta)]),
        "gxc_lag_snr": (ctypes.c_int, [H, P(ctypes.c_float), ip, ctypes.c_size_t]),
        "gxc_lag_snr_series": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_float), ip]),
        "gxc_station_bytes": (ctypes.c_size_t, [P(Config)]),
        "gxc_num_rows": (ctypes.c_int, [P(Config)]),
        "gxc_fstage": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, vp, u8p]),
        "gxc_xengine": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, vp]),
        "gxc_xengine_noalloc": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int, vp, vp]),
        "gxc_host_alloc": (vp, [ctypes.c_size_t]),
        "gxc_host_free": (None, [vp]),
        "gxc_integrate_begin": (ctypes.c_int, [H]),
        "gxc_submit": (ctypes.c_int, [H, P(u8p), P(dp), dp]),
        "gxc_integrate_end": (ctypes.c_int, [H, vp, ctypes.c_size_t, ip]),
        "gxc_set_profiling": (ctypes.c_int, [H, ctypes.c_int]),
        "gxc_autotune": (ctypes.c_int, [H, P(Tuning)]),
        "gxc_get_timings": (ctypes.c_int, [H, P(Timings)]),
        "gxc_parse_config": (ctypes.c_int, [ctypes.c_char_p, P(ConfFile)]),
        "gxc_conffile_free": (None, [P(ConfFile)]),
        "gxc_batch_config_init": (None, [P(BatchConfig)]),
        "gxc_batch_create": (ctypes.c_int, [P(BatchConfig), P(vp)]),
        "gxc_batch_destroy": (None, [H]),
        "gxc_batch_last_error": (ctypes.c_char_p, [H]),
        "gxc_batch_set_input": (ctypes.c_int, [H, ctypes.c_int, P(u8p)]),
        "gxc_batch_set_input_device": (ctypes.c_int, [H, ctypes.c_int, P(vp)]),
        "gxc_batch_set_delays": (ctypes.c_int, [H, ctypes.c_int, P(dp), dp]),
        "gxc_batch_set_pol_conversion": (ctypes.c_int, [H, ctypes.c_int, vp]),
        "gxc_batch_set_input_vdif": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, vp, ctypes.c_size_t,
                                                    P(VdifSpec), P(VdifStats)]),
        "gxc_batch_get_frame_flags": (ctypes.c_int, [H, ctypes.c_int, ctypes.c_int, u8p, ctypes.c_size_t,
                                                     P(ctypes.c_longlong)]),
        "gxc_batch_process": (ctypes.c_int, [H]),
        "gxc_batch_process_async": (ctypes.c_int, [H]),
        "gxc_batch_synchronize": (ctypes.c_int, [H]),
        "gxc_batch_get_visibilities": (ctypes.c_int, [H, ctypes.c_int, vp, ctypes.c_size_t]),
        "gxc_batch_get_row_counts": (ctypes.c_int, [H, ctypes.c_int, ip, ctypes.c_size_t]),
        "gxc_batch_set_profiling": (ctypes.c_int, [H, ctypes.c_int]),
        "gxc_batch_autotune": (ctypes.c_int, [H, P(Tuning)]),
        "gxc_batch_get_timings": (ctypes.c_int, [H, P(Timings)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status, handle=None):
    if status != GXC_OK:
        msg = load().gxc_last_error(handle)
        raise GxcError(status, msg.decode() if msg else "")


def check_batch(status, handle=None):
    if status != GXC_OK:
        msg = load().gxc_batch_last_error(handle)
        raise GxcError(status, msg.decode() if msg else "")
