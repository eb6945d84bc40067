"""Python mirror of fxkernel's FxKernel over the gxcorr C-ABI.

Same constructor arguments and methods as fxkernel/src/fxkernel.h:18-64
(setInputData, setDelays, process, accumulate, saveVisibilities), plus
access to the visibilities as a numpy array and the stage entry points used
by tests and benchmarks.  Visibilities are [row][product][channel] complex64:
cross baselines in fxkernel order (fxkernel.cpp:328-355), product p = 2l+m
= X_j,l * conj(X_k,m), then (with autos=True) one row per station.
"""
import collections
import ctypes

import numpy as np

from . import _lib
from ._lib import check


def _config(nant, nchan, nfft, nbit, lo, bw, is_complex, precision, autos, exact_phase,
            chunk_ffts, overlap=False, graph=True, fblock=0, xtile=0, xsplits=0):
    c = _lib.Config()
    _lib.load().gxc_config_init(ctypes.byref(c))
    c.nant, c.nchan, c.numffts, c.nbit = int(nant), int(nchan), int(nfft), int(nbit)
    c.npol = 2
    c.is_complex = int(bool(is_complex))
    c.lo, c.bw = float(lo), float(bw)
    c.precision = _lib.PRECISION[precision] if isinstance(precision, str) else int(precision)
    c.autos = int(bool(autos))
    c.exact_phase = int(bool(exact_phase))
    c.chunk_ffts = int(chunk_ffts)
    c.overlap = int(bool(overlap))
    c.graph = int(bool(graph))
    c.fblock, c.xtile, c.xsplits = int(fblock), int(xtile), int(xsplits)
    return c


def station_bytes(nant, nchan, nfft, nbit=2, is_complex=False):
    """allocDataHost's per-station size (fxkernel/src/common.cpp:16)."""
    c = _config(nant, nchan, nfft, nbit, 0.0, 1.0, is_complex, 0, False, False, 0)
    return int(_lib.load().gxc_station_bytes(ctypes.byref(c)))


def num_rows(nant, autos=False):
    return nant * (nant - 1) // 2 + (nant if autos else 0)


class PinnedBuffer:
    """Page-locked host memory from gxc_host_alloc, viewed as a uint8 array.

    Station buffers in pinned memory let gxc_submit copy them to the device
    asynchronously; pageable buffers are copied synchronously."""

    def __init__(self, nbytes):
        self._L = _lib.load()
        self.nbytes = int(nbytes)
        self._p = self._L.gxc_host_alloc(self.nbytes)
        if not self._p:
            raise MemoryError("gxc_host_alloc(%d) failed" % self.nbytes)
        self.array = np.ctypeslib.as_array(ctypes.cast(self._p, ctypes.POINTER(ctypes.c_uint8)),
                                           shape=(self.nbytes,))

    def free(self):
        if getattr(self, "_p", None):
            self.array = None
            self._L.gxc_host_free(self._p)
        self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class FxKernel:
    """FxKernel(nant, nchan, nfft, nbit, lo, bw, is_complex) on the MI355X path."""

    def __init__(self, nant, nchan, nfft, nbit, lo, bw, is_complex=False, *, precision="fp32",
                 autos=False, exact_phase=False, chunk_ffts=0, overlap=False, graph=True,
                 fblock=0, xtile=0, xsplits=0):
        self._L = _lib.load()
        self.cfg = _config(nant, nchan, nfft, nbit, lo, bw, is_complex, precision, autos,
                           exact_phase, chunk_ffts, overlap, graph, fblock, xtile, xsplits)
        self.nant, self.nchan, self.numffts = int(nant), int(nchan), int(nfft)
        self.nrows = int(self._L.gxc_num_rows(ctypes.byref(self.cfg)))
        self.station_bytes = int(self._L.gxc_station_bytes(ctypes.byref(self.cfg)))
        h = ctypes.c_void_p()
        check(self._L.gxc_create(ctypes.byref(self.cfg), ctypes.byref(h)))
        self._h = h
        self._keep = None
        self._inflight = collections.deque(maxlen=2)
        self._vdif_keep = {}

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.gxc_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- fxkernel interface ---------------------------------------------------
    def setInputData(self, idata):
        """Host buffers, one per station (fxkernel.cpp:242-245).  Kept alive here."""
        arrs = [np.ascontiguousarray(a, dtype=np.uint8) for a in idata]
        if len(arrs) != self.nant:
            raise ValueError("need %d station buffers" % self.nant)
        for a in arrs:
            if a.size < self.station_bytes:
                raise ValueError("station buffer has %d bytes, need %d" % (a.size, self.station_bytes))
        ptrs = (ctypes.POINTER(ctypes.c_uint8) * self.nant)(
            *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)) for a in arrs])
        check(self._L.gxc_set_input(self._h, ptrs), self._h)
        self._keep = (arrs, ptrs)

    def setInputDevice(self, dev_ptrs, keepalive=None):
        """Device-resident station buffers (integers, e.g. tensor.data_ptr())."""
        if len(dev_ptrs) != self.nant:
            raise ValueError("need %d station buffers" % self.nant)
        ptrs = (ctypes.c_void_p * self.nant)(*[int(p) for p in dev_ptrs])
        check(self._L.gxc_set_input_device(self._h, ptrs), self._h)
        self._keep = (keepalive, ptrs)

    def setInputVDIF(self, station, dev_ptr, nbytes, *, frames_per_second, ref_second=0,
                     ref_frame=0, ref_epoch=0, thread_id=0, frame_bytes=0, offset_binary8=True,
                     header_bytes=0, keepalive=None, stats=True):
        """Station `station`'s input from VDIF frames in device memory
        (gxc_set_input_vdif).  header_bytes: 32, 16 (legacy) or 0 (from the
        first frame).  Returns the ingest counts as a dict (stats=True waits
        for the ingest)."""
        spec = _lib.VdifSpec()
        self._L.gxc_vdif_spec_init(ctypes.byref(spec))
        spec.frame_bytes, spec.thread_id = int(frame_bytes), int(thread_id)
        spec.frames_per_second, spec.ref_epoch = int(frames_per_second), int(ref_epoch)
        spec.ref_second, spec.ref_frame = int(ref_second), int(ref_frame)
        spec.offset_binary8 = int(bool(offset_binary8))
        spec.header_bytes = int(header_bytes)
        st = _lib.VdifStats()
        check(self._L.gxc_set_input_vdif(self._h, int(station), ctypes.c_void_p(int(dev_ptr)),
                                         int(nbytes), ctypes.byref(spec),
                                         ctypes.byref(st) if stats else None), self._h)
        # the ingest is queued: the frames must live until it has run
        self._vdif_keep[int(station)] = keepalive
        return {k: getattr(st, k) for k, _ in _lib.VdifStats._fields_} if stats else None

    def frame_flags(self, station):
        """Per-slot presence flags of the station's last VDIF ingest."""
        n = ctypes.c_longlong(0)
        check(self._L.gxc_get_frame_flags(self._h, int(station), None, 0, ctypes.byref(n)), self._h)
        out = np.empty(n.value, np.uint8)
        check(self._L.gxc_get_frame_flags(self._h, int(station),
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), out.size,
                                          None), self._h)
        return out

    def set_pol_conversion(self, gains):
        """gains: complex [nant][2][nchan] (g0, g1 per station and channel) or
        None to turn the conversion off (gxc_set_pol_conversion)."""
        if gains is None:
            check(self._L.gxc_set_pol_conversion(self._h, None), self._h)
            return
        g = np.ascontiguousarray(gains, np.complex64).reshape(self.nant, 2, self.nchan)
        check(self._L.gxc_set_pol_conversion(self._h, g.ctypes.data_as(ctypes.c_void_p)), self._h)

    def setDelays(self, d, f):
        """d: [nant][3] polynomial (s, t in FFTs), f: [nant] file offsets (s)."""
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(self.nant, 3)
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.nant)
        rows = (ctypes.POINTER(ctypes.c_double) * self.nant)(
            *[d[a].ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in range(self.nant)])
        check(self._L.gxc_set_delays(self._h, rows, f.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
              self._h)

    def process(self):
        check(self._L.gxc_process(self._h), self._h)

    def process_async(self):
        check(self._L.gxc_process_async(self._h), self._h)

    def synchronize(self):
        check(self._L.gxc_synchronize(self._h), self._h)

    def accumulate(self, odata):
        """odata: complex64 array [rows][4][nchan], updated in place (fxkernel.cpp:414-423)."""
        if odata.dtype != np.complex64 or odata.shape != (self.nrows, 4, self.nchan) \
                or not odata.flags.c_contiguous:
            raise ValueError("odata must be C-contiguous complex64 of shape %s"
                             % ((self.nrows, 4, self.nchan),))
        base = odata.ctypes.data
        row_t = ctypes.POINTER(_lib.CF32) * 4
        rows = (ctypes.POINTER(ctypes.POINTER(_lib.CF32)) * self.nrows)()
        keep = []
        for r in range(self.nrows):
            prods = row_t(*[ctypes.cast(base + ((r * 4 + p) * self.nchan) * 8,
                                        ctypes.POINTER(_lib.CF32)) for p in range(4)])
            keep.append(prods)
            rows[r] = ctypes.cast(prods, ctypes.POINTER(ctypes.POINTER(_lib.CF32)))
        check(self._L.gxc_accumulate(self._h, ctypes.cast(rows, ctypes.c_void_p)), self._h)

    def saveVisibilities(self, path):
        check(self._L.gxc_save_visibilities(self._h, str(path).encode()), self._h)

    def save_binary(self, path):
        """gxc_save_visibilities_binary (read back with gcorr_amd.visio.read_binary)."""
        check(self._L.gxc_save_visibilities_binary(self._h, str(path).encode()), self._h)

    def save_difx(self, path, mjd=0, seconds=0.0, pols="XY", uvw=None, config_index=0,
                  source_index=0, freq_index=0):
        """gxc_save_difx: DiFX binary visibility records, one per row and product."""
        m = _lib.DifxMeta()
        m.mjd, m.seconds = int(mjd), float(seconds)
        m.config_index, m.source_index, m.freq_index = int(config_index), int(source_index), int(freq_index)
        m.pols = pols.encode()[:2]
        keep = None
        if uvw is not None:
            keep = np.ascontiguousarray(uvw, np.float64).reshape(self.nrows, 3)
            m.uvw = keep.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        check(self._L.gxc_save_difx(self._h, str(path).encode(), ctypes.byref(m)), self._h)
        del keep

    def lag_snr(self):
        """gxc_lag_snr: (snr, peak) arrays [rows][4] of the last process()."""
        n = self.nrows * 4
        snr = np.empty(n, np.float32)
        peak = np.empty(n, np.int32)
        check(self._L.gxc_lag_snr(self._h, snr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                  peak.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n), self._h)
        return snr.reshape(self.nrows, 4), peak.reshape(self.nrows, 4)

    # -- extras -----------------------------------------------------------------
    def visibilities(self):
        out = np.empty((self.nrows, 4, self.nchan), np.complex64)
        check(self._L.gxc_get_visibilities(self._h, out.ctypes.data_as(ctypes.c_void_p), out.size),
              self._h)
        return out

    def row_counts(self):
        out = np.empty(self.nrows, np.int32)
        check(self._L.gxc_get_row_counts(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                         out.size), self._h)
        return out

    def fstage(self, t0, nfft):
        """Spectra [nfft][2*nant][nchan] complex64 and validity [nant][nfft] of FFTs t0..t0+nfft."""
        spec = np.empty((nfft, 2 * self.nant, self.nchan), np.complex64)
        valid = np.empty((self.nant, nfft), np.uint8)
        check(self._L.gxc_fstage(self._h, int(t0), int(nfft), spec.ctypes.data_as(ctypes.c_void_p),
                                 valid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))), self._h)
        return spec, valid.astype(bool)

    def integrate_begin(self):
        """Start a streamed sum of subintegrations (gxc_integrate_begin)."""
        check(self._L.gxc_integrate_begin(self._h), self._h)
        self._inflight.clear()

    def submit(self, idata, d, f):
        """Queue one subintegration from host station buffers (numpy uint8 or
        PinnedBuffer arrays) with its delays; its visibilities are added to
        the sum.  Returns before the work is done; the previous submission's
        buffers may be refilled once this returns.  Buffers are kept alive
        here until they have been read."""
        arrs = [a if isinstance(a, np.ndarray) and a.dtype == np.uint8 and a.flags.c_contiguous
                else np.ascontiguousarray(a, dtype=np.uint8) for a in idata]
        if len(arrs) != self.nant:
            raise ValueError("need %d station buffers" % self.nant)
        for a in arrs:
            if a.size < self.station_bytes:
                raise ValueError("station buffer has %d bytes, need %d" % (a.size, self.station_bytes))
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(self.nant, 3)
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.nant)
        ptrs = (ctypes.POINTER(ctypes.c_uint8) * self.nant)(
            *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)) for a in arrs])
        rows = (ctypes.POINTER(ctypes.c_double) * self.nant)(
            *[d[a].ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in range(self.nant)])
        check(self._L.gxc_submit(self._h, ptrs, rows, f.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
              self._h)
        self._inflight.append(arrs)

    def integrate_end(self):
        """Wait for the submitted subintegrations; returns (sum, count), the
        sum a complex64 array [rows][4][nchan]."""
        out = np.empty((self.nrows, 4, self.nchan), np.complex64)
        n = ctypes.c_int(0)
        check(self._L.gxc_integrate_end(self._h, out.ctypes.data_as(ctypes.c_void_p), out.size,
                                        ctypes.byref(n)), self._h)
        self._inflight.clear()
        return out, n.value

    def autotune(self):
        """gxc_autotune: time the launch shapes on this handle's input and keep
        the fastest.  Returns the choice and its timings as a dict."""
        t = _lib.Tuning()
        check(self._L.gxc_autotune(self._h, ctypes.byref(t)), self._h)
        return {k: getattr(t, k) for k, _ in _lib.Tuning._fields_}

    def set_profiling(self, on=True):
        check(self._L.gxc_set_profiling(self._h, int(bool(on))), self._h)

    def timings(self):
        t = _lib.Timings()
        check(self._L.gxc_get_timings(self._h, ctypes.byref(t)), self._h)
        return {k: getattr(t, k) for k, _ in _lib.Timings._fields_}

    # snake_case aliases
    set_input_data = setInputData
    set_input_device = setInputDevice
    set_input_vdif = setInputVDIF
    set_delays = setDelays
    save_visibilities = saveVisibilities


def requantise_2bit(pol0_ptr, pol1_ptr, nsamples, threshold, out_ptr, stream=0):
    """gxc_requantise_2bit on device pointers (integers, e.g. tensor.data_ptr());
    stream: a hipStream_t as an integer (0: the default stream)."""
    check(_lib.load().gxc_requantise_2bit(ctypes.c_void_p(int(pol0_ptr)), ctypes.c_void_p(int(pol1_ptr)),
                                          int(nsamples), float(threshold), ctypes.c_void_p(int(out_ptr)),
                                          ctypes.c_void_p(int(stream)) if stream else None))


def requantise_8bit(pol0_ptr, pol1_ptr, nsamples, scale, out_ptr, stream=0):
    """gxc_requantise_8bit on device pointers (complex samples, re/im interleaved)."""
    check(_lib.load().gxc_requantise_8bit(ctypes.c_void_p(int(pol0_ptr)), ctypes.c_void_p(int(pol1_ptr)),
                                          int(nsamples), float(scale), ctypes.c_void_p(int(out_ptr)),
                                          ctypes.c_void_p(int(stream)) if stream else None))


def lag_snr(vis):
    """gxc_lag_snr_series on the GPU: vis [..., nchan] complex (nchan a power
    of two); returns (snr, peak) of shape vis.shape[:-1]."""
    L = _lib.load()
    v = np.ascontiguousarray(vis, np.complex64)
    nchan = v.shape[-1]
    n = v.size // nchan
    snr = np.empty(n, np.float32)
    peak = np.empty(n, np.int32)
    check(L.gxc_lag_snr_series(v.ctypes.data_as(ctypes.c_void_p), n, nchan,
                               snr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               peak.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    return snr.reshape(v.shape[:-1]), peak.reshape(v.shape[:-1])


def xengine(spectra, nant, autos=False, precision="fp32"):
    """X-stage alone over spectra [nfft][2*nant][nchan]; result / nfft (gcorr CrossCorr semantics)."""
    L = _lib.load()
    spectra = np.ascontiguousarray(spectra, np.complex64)
    nfft, ninp, nchan = spectra.shape
    if ninp != 2 * nant:
        raise ValueError("spectra must have 2*nant inputs")
    out = np.empty((num_rows(nant, autos), 4, nchan), np.complex64)
    prec = _lib.PRECISION[precision] if isinstance(precision, str) else int(precision)
    check(L.gxc_xengine(spectra.ctypes.data_as(ctypes.c_void_p), int(nant), int(nchan), int(nfft),
                        int(bool(autos)), prec, out.ctypes.data_as(ctypes.c_void_p)))
    return out
