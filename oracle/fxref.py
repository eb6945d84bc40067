"""ctypes bindings for the fxkernel restatement (oracle/libfxref.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, as the checker or as the timed CPU
baseline.  The product package (gcorr_amd) never imports it.

Every function mirrors one in fxref.h, which cites the fxkernel lines it
restates (fxkernel/src/fxkernel.cpp).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfxref.so")


class Config(ctypes.Structure):
    _fields_ = [
        ("nant", ctypes.c_int),
        ("nchan", ctypes.c_int),
        ("numffts", ctypes.c_int),
        ("nbit", ctypes.c_int),
        ("is_complex", ctypes.c_int),
        ("lo", ctypes.c_double),
        ("bw", ctypes.c_double),
        ("autos", ctypes.c_int),
        ("quirk", ctypes.c_int),
        ("frame_ok", ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8))),
        ("frame_samples", ctypes.POINTER(ctypes.c_int)),
        ("polgains", ctypes.POINTER(ctypes.c_float)),
    ]


class VdifSpec(ctypes.Structure):
    _fields_ = [
        ("frame_bytes", ctypes.c_int),
        ("thread_id", ctypes.c_int),
        ("frames_per_second", ctypes.c_int),
        ("ref_epoch", ctypes.c_int),
        ("ref_second", ctypes.c_longlong),
        ("ref_frame", ctypes.c_int),
        ("offset_binary8", ctypes.c_int),
        ("header_bytes", ctypes.c_int),
    ]


class VdifStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_longlong) for k in
                ("frames", "placed", "missing", "duplicate", "invalid", "other_thread",
                 "out_of_window", "malformed")]


class Derived(ctypes.Structure):
    _fields_ = [
        ("fftchannels", ctypes.c_int),
        ("stride", ctypes.c_int),
        ("substride", ctypes.c_int),
        ("nsteps", ctypes.c_int),
        ("fracsteps", ctypes.c_int),
        ("maxoffset", ctypes.c_int),
        ("nbaselines", ctypes.c_int),
        ("nrows", ctypes.c_int),
        ("sampletime", ctypes.c_double),
        ("bytes_per_station", ctypes.c_size_t),
    ]


_lib = None
_lib_path = None  # another build of the same source (use_library), or None


def use_library(path):
    """Load the oracle from `path` (a build of the same source with other
    compiler flags, e.g. the Makefile's `native` target) from now on; None
    returns to oracle/libfxref.so."""
    global _lib, _lib_path
    _lib, _lib_path = None, path


def compile_flags(native=False):
    """The compiler flags of oracle/Makefile's portable build, or of its
    `native` target."""
    flags = None
    with open(os.path.join(_HERE, "Makefile")) as f:
        for line in f:
            if line.startswith("CFLAGS ?="):
                flags = line.split("=", 1)[1].strip()
    if flags and native:
        flags = flags.replace("-march=x86-64-v3", "-march=native")
    return flags


def build():
    """Compile libfxref.so with its committed Makefile."""
    import subprocess
    subprocess.run(["make", "-B", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        path = _lib_path or LIB_PATH
        if _lib_path is None and not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        dp, fp, ip, u8p = (P(ctypes.c_double), P(ctypes.c_float),
                           P(ctypes.c_int), P(ctypes.c_uint8))
        cfgp = P(Config)
        L.fxref_derive.argtypes = [cfgp, P(Derived)]
        L.fxref_lut2bit.argtypes = [fp]
        L.fxref_station_delay.argtypes = [dp, ctypes.c_int, dp, dp, dp]
        L.fxref_delay_split.argtypes = [cfgp, dp, ctypes.c_double, ctypes.c_int,
                                        ip, dp, dp, dp]
        for suf, rp in (("f32", fp), ("f64", dp)):
            getattr(L, "fxref_unpack_" + suf).argtypes = [cfgp, u8p, ctypes.c_int, rp]
            getattr(L, "fxref_rotator_" + suf).argtypes = [cfgp, ctypes.c_double,
                                                           ctypes.c_double, rp]
            getattr(L, "fxref_fracrot_" + suf).argtypes = [cfgp, ctypes.c_double, rp]
            getattr(L, "fxref_fft_" + suf).argtypes = [ctypes.c_int, rp, rp]
            getattr(L, "fxref_station_spectrum_" + suf).argtypes = [
                cfgp, u8p, dp, ctypes.c_double, ctypes.c_int, rp]
            getattr(L, "fxref_process_" + suf).argtypes = [
                cfgp, P(u8p), dp, dp, rp, ip]
        L.fxref_accumulate_range_f32.argtypes = [cfgp, P(u8p), dp, dp, ctypes.c_int,
                                                 ctypes.c_int, fp, ip]
        L.fxref_xmac_f32.argtypes = [fp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, fp]
        L.fxref_bench_f32.argtypes = [cfgp, P(u8p), dp, dp, ctypes.c_int, ctypes.c_int]
        L.fxref_bench_f32.restype = ctypes.c_double
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def make_config(nant, nchan, numffts, lo, bw, nbit=2, is_complex=False, autos=False,
                quirk=True, frame_ok=None, frame_samples=None, polgains=None):
    """frame_ok: per-station uint8 flags of a VDIF ingest, with
    frame_samples[station] time samples per frame (None: no missing data).
    polgains: complex [nant][2][nchan] polarisation-conversion gains (None: none)."""
    c = Config(int(nant), int(nchan), int(numffts), int(nbit), int(bool(is_complex)),
               float(lo), float(bw), int(bool(autos)), int(bool(quirk)))
    keep = []
    if polgains is not None:
        g = np.ascontiguousarray(np.asarray(polgains, np.complex64).reshape(int(nant), 2, int(nchan)))
        gf = g.view(np.float32)
        c.polgains = _ptr(gf, ctypes.c_float)
        keep.append(gf)
        c._keep = (keep,)
    if frame_ok is not None:
        flags = [np.ascontiguousarray(f, np.uint8) for f in frame_ok]
        spf = np.ascontiguousarray(frame_samples, np.int32).reshape(int(nant))
        ptrs = (ctypes.POINTER(ctypes.c_uint8) * int(nant))(*[_ptr(f, ctypes.c_uint8) for f in flags])
        c.frame_ok = ctypes.cast(ptrs, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)))
        c.frame_samples = _ptr(spf, ctypes.c_int)
        c._keep = (flags, spf, ptrs, keep)
    return c


def derive(cfg):
    d = Derived()
    rc = lib().fxref_derive(ctypes.byref(cfg), ctypes.byref(d))
    if rc:
        raise ValueError("invalid oracle config (code %d)" % rc)
    return d


def lut2bit():
    out = np.zeros((256, 4), np.float32)
    lib().fxref_lut2bit(_ptr(out, ctypes.c_float))
    return out


def station_delay(poly, fftindex):
    poly = np.ascontiguousarray(poly, np.float64)
    m, a, b = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().fxref_station_delay(_ptr(poly, ctypes.c_double), int(fftindex),
                              ctypes.byref(m), ctypes.byref(a), ctypes.byref(b))
    return m.value, a.value, b.value


def delay_split(cfg, poly, fileoffset, fftindex):
    """Returns (valid, offset, fracdelay_seconds, a, b)."""
    poly = np.ascontiguousarray(poly, np.float64)
    off, fr, a, b = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    v = lib().fxref_delay_split(ctypes.byref(cfg), _ptr(poly, ctypes.c_double),
                                float(fileoffset), int(fftindex), ctypes.byref(off),
                                ctypes.byref(fr), ctypes.byref(a), ctypes.byref(b))
    return bool(v), off.value, fr.value, a.value, b.value


def _rt(mode):
    return (np.float32, ctypes.c_float) if mode == "f32" else (np.float64, ctypes.c_double)


def unpack(cfg, data, offset, mode="f32"):
    dt, ct = _rt(mode)
    L = derive(cfg).fftchannels
    out = np.zeros((2, L, 2), dt)
    data = np.ascontiguousarray(data, np.uint8)
    getattr(lib(), "fxref_unpack_" + mode)(ctypes.byref(cfg), _ptr(data, ctypes.c_uint8),
                                          int(offset), _ptr(out, ct))
    return out[..., 0] + 1j * out[..., 1]


def rotator(cfg, a, b, mode="f32"):
    dt, ct = _rt(mode)
    L = derive(cfg).fftchannels
    out = np.zeros((L, 2), dt)
    getattr(lib(), "fxref_rotator_" + mode)(ctypes.byref(cfg), float(a), float(b),
                                           _ptr(out, ct))
    return out[:, 0] + 1j * out[:, 1]


def fracrot(cfg, frac, mode="f32"):
    dt, ct = _rt(mode)
    out = np.zeros((cfg.nchan, 2), dt)
    getattr(lib(), "fxref_fracrot_" + mode)(ctypes.byref(cfg), float(frac), _ptr(out, ct))
    return out[:, 0] + 1j * out[:, 1]


def fft(x, mode="f32"):
    dt, ct = _rt(mode)
    x = np.asarray(x)
    L = x.shape[-1]
    inp = np.ascontiguousarray(np.stack([x.real, x.imag], -1).astype(dt))
    out = np.zeros_like(inp)
    getattr(lib(), "fxref_fft_" + mode)(int(L), _ptr(inp, ct), _ptr(out, ct))
    return out[..., 0] + 1j * out[..., 1]


def station_spectrum(cfg, data, poly, fileoffset, fftindex, mode="f32"):
    """[2][nchan] complex spectrum of one (station, FFT), and its validity."""
    dt, ct = _rt(mode)
    out = np.zeros((2, cfg.nchan, 2), dt)
    data = np.ascontiguousarray(data, np.uint8)
    poly = np.ascontiguousarray(poly, np.float64)
    v = getattr(lib(), "fxref_station_spectrum_" + mode)(
        ctypes.byref(cfg), _ptr(data, ctypes.c_uint8), _ptr(poly, ctypes.c_double),
        float(fileoffset), int(fftindex), _ptr(out, ct))
    return out[..., 0] + 1j * out[..., 1], bool(v)


def _station_ptrs(data):
    arrs = [np.ascontiguousarray(d, np.uint8) for d in data]
    ptrs = (ctypes.POINTER(ctypes.c_uint8) * len(arrs))(
        *[_ptr(a, ctypes.c_uint8) for a in arrs])
    return arrs, ptrs


def process(cfg, data, delays, offsets, mode="f32"):
    """FxKernel::process on the CPU.  Returns (vis[nrows,4,nchan] complex, counts)."""
    dt, ct = _rt(mode)
    d = derive(cfg)
    for a in data:
        if len(a) < d.bytes_per_station:
            raise ValueError("station buffer shorter than %d bytes" % d.bytes_per_station)
    arrs, ptrs = _station_ptrs(data)
    dl = np.ascontiguousarray(delays, np.float64).reshape(cfg.nant, 3)
    of = np.ascontiguousarray(offsets, np.float64).reshape(cfg.nant)
    vis = np.zeros((d.nrows, 4, cfg.nchan, 2), dt)
    counts = np.zeros(d.nrows, np.int32)
    rc = getattr(lib(), "fxref_process_" + mode)(
        ctypes.byref(cfg), ptrs, _ptr(dl, ctypes.c_double), _ptr(of, ctypes.c_double),
        _ptr(vis, ct), _ptr(counts, ctypes.c_int))
    if rc:
        raise RuntimeError("oracle process failed (%d)" % rc)
    return vis[..., 0] + 1j * vis[..., 1], counts


def vdif_place(cfg, stream, frame_bytes=0, thread_id=0, frames_per_second=1, ref_epoch=0,
               ref_second=0, ref_frame=0, offset_binary8=True, header_bytes=0):
    """Place a VDIF byte stream into one station buffer (gcorr ingest,
    ingest/src/ingest.c:267-355).  Returns (station bytes, frame_ok flags,
    samples per frame, stats dict)."""
    d = derive(cfg)
    stream = np.ascontiguousarray(stream, np.uint8)
    spec = VdifSpec(int(frame_bytes), int(thread_id), int(frames_per_second), int(ref_epoch),
                    int(ref_second), int(ref_frame), int(bool(offset_binary8)), int(header_bytes))
    station = np.empty(d.bytes_per_station, np.uint8)
    flags = np.empty(d.bytes_per_station // 8 + 1, np.uint8)
    spf = ctypes.c_int(0)
    st = VdifStats()
    L = lib()
    L.fxref_vdif_place.restype = ctypes.c_longlong
    L.fxref_vdif_place.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_uint8),
                                   ctypes.c_size_t, ctypes.POINTER(VdifSpec),
                                   ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint8),
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(VdifStats)]
    n = L.fxref_vdif_place(ctypes.byref(cfg), _ptr(stream, ctypes.c_uint8), stream.size,
                           ctypes.byref(spec), _ptr(station, ctypes.c_uint8),
                           _ptr(flags, ctypes.c_uint8), ctypes.byref(spf), ctypes.byref(st))
    if n < 0:
        raise ValueError("unusable VDIF stream")
    return station, flags[:n].copy(), spf.value, {k: getattr(st, k) for k, _ in VdifStats._fields_}


def accumulate_range(cfg, data, delays, offsets, t0, t1):
    """Unnormalised FP32 sums over FFTs [t0, t1).  Returns (sums, counts)."""
    d = derive(cfg)
    arrs, ptrs = _station_ptrs(data)
    dl = np.ascontiguousarray(delays, np.float64).reshape(cfg.nant, 3)
    of = np.ascontiguousarray(offsets, np.float64).reshape(cfg.nant)
    vis = np.zeros((d.nrows, 4, cfg.nchan, 2), np.float32)
    counts = np.zeros(d.nrows, np.int32)
    lib().fxref_accumulate_range_f32(
        ctypes.byref(cfg), ptrs, _ptr(dl, ctypes.c_double), _ptr(of, ctypes.c_double),
        int(t0), int(t1), _ptr(vis, ctypes.c_float), _ptr(counts, ctypes.c_int))
    return vis[..., 0] + 1j * vis[..., 1], counts


def xmac(spectra, nant, autos=False):
    """Unnormalised sequential FP32 X sums over spectra [nfft][2*nant][nchan]."""
    spectra = np.asarray(spectra)
    nfft, ninp, nchan = spectra.shape
    sp = np.ascontiguousarray(np.stack([spectra.real, spectra.imag], -1).astype(np.float32))
    nrows = nant * (nant - 1) // 2 + (nant if autos else 0)
    vis = np.zeros((nrows, 4, nchan, 2), np.float32)
    lib().fxref_xmac_f32(_ptr(sp, ctypes.c_float), int(nant), int(nchan), int(nfft),
                         int(bool(autos)), _ptr(vis, ctypes.c_float))
    return vis[..., 0] + 1j * vis[..., 1]


def bench(cfg, data, delays, offsets, nthreads, nloop=1):
    """Wall seconds for nthreads x nloop FP32 process() calls (bench_fxkernel)."""
    arrs, ptrs = _station_ptrs(data)
    dl = np.ascontiguousarray(delays, np.float64).reshape(cfg.nant, 3)
    of = np.ascontiguousarray(offsets, np.float64).reshape(cfg.nant)
    t = lib().fxref_bench_f32(ctypes.byref(cfg), ptrs, _ptr(dl, ctypes.c_double),
                              _ptr(of, ctypes.c_double), int(nthreads), int(nloop))
    if t < 0:
        raise RuntimeError("oracle bench failed (%g)" % t)
    return t
