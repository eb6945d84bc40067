"""Several subbands of one subintegration through one launch per stage.

Python mirror of the gxc_batch_* entry points of include/gxcorr.h.  fxkernel
correlates one band per FxKernel object (fxkernel/src/fxkernel.h:29); a
Batch holds every subband (same stations; own channel count, bandwidth, LO,
FFT count, delays and input) and returns per subband what an FxKernel with
the same settings returns.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check_batch
from .fxkernel import num_rows, station_bytes


class Batch:
    """Batch(nant, [(nchan, numffts, lo, bw), ...]) over gxc_batch_*.

    fstage_mfma: 0 (False) keeps every subband in the multi-subband F launch;
    1 (True) gives the subbands of 2048 points and more a launch each on the
    matrix-core F-stage where it covers them, else the register F-stage; 2
    gives them a launch each on the register F-stage.  autotune() chooses.
    """

    def __init__(self, nant, subbands, *, precision="fp32", autos=False, exact_phase=False,
                 chunks=0, overlap=False, graph=True, fstage_mfma=False):
        self._L = _lib.load()
        c = _lib.BatchConfig()
        self._L.gxc_batch_config_init(ctypes.byref(c))
        self.subbands = [(int(n), int(f), float(lo), float(bw)) for n, f, lo, bw in subbands]
        arr = (_lib.Subband * len(self.subbands))(
            *[_lib.Subband(n, f, lo, bw) for n, f, lo, bw in self.subbands])
        c.nant = int(nant)
        c.precision = _lib.PRECISION[precision] if isinstance(precision, str) else int(precision)
        c.autos = int(bool(autos))
        c.exact_phase = int(bool(exact_phase))
        c.nsub = len(self.subbands)
        c.subbands = arr
        c.chunks = int(chunks)
        c.overlap = int(bool(overlap))
        c.graph = int(bool(graph))
        c.fstage_mfma = 2 if fstage_mfma == 2 else int(bool(fstage_mfma))
        self.nant = int(nant)
        self.nrows = num_rows(nant, autos)
        h = ctypes.c_void_p()
        check_batch(self._L.gxc_batch_create(ctypes.byref(c), ctypes.byref(h)))
        self._h = h
        self._keep = {}

    def station_bytes(self, sub):
        n, f, _, _ = self.subbands[sub]
        return station_bytes(self.nant, n, f)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.gxc_batch_destroy(self._h)
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

    def set_input(self, sub, idata):
        arrs = [np.ascontiguousarray(a, dtype=np.uint8) for a in idata]
        if len(arrs) != self.nant:
            raise ValueError("need %d station buffers" % self.nant)
        need = self.station_bytes(sub)
        for a in arrs:
            if a.size < need:
                raise ValueError("station buffer has %d bytes, need %d" % (a.size, need))
        ptrs = (ctypes.POINTER(ctypes.c_uint8) * self.nant)(
            *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)) for a in arrs])
        check_batch(self._L.gxc_batch_set_input(self._h, int(sub), ptrs), self._h)
        self._keep[sub] = (arrs, ptrs)

    def set_input_device(self, sub, dev_ptrs, keepalive=None):
        if len(dev_ptrs) != self.nant:
            raise ValueError("need %d station buffers" % self.nant)
        ptrs = (ctypes.c_void_p * self.nant)(*[int(p) for p in dev_ptrs])
        check_batch(self._L.gxc_batch_set_input_device(self._h, int(sub), ptrs), self._h)
        self._keep[sub] = (keepalive, ptrs)

    def set_delays(self, sub, d, f):
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(self.nant, 3)
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.nant)
        rows = (ctypes.POINTER(ctypes.c_double) * self.nant)(
            *[d[a].ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in range(self.nant)])
        check_batch(self._L.gxc_batch_set_delays(self._h, int(sub), rows,
                                                 f.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    self._h)

    def set_input_vdif(self, sub, station, dev_ptr, nbytes, *, frames_per_second, ref_second=0,
                       ref_frame=0, ref_epoch=0, thread_id=0, frame_bytes=0, offset_binary8=True,
                       header_bytes=0, keepalive=None, stats=True):
        """Station `station` of subband `sub` from VDIF frames in device memory
        (gxc_batch_set_input_vdif); as FxKernel.setInputVDIF."""
        spec = _lib.VdifSpec()
        self._L.gxc_vdif_spec_init(ctypes.byref(spec))
        spec.frame_bytes, spec.thread_id = int(frame_bytes), int(thread_id)
        spec.frames_per_second, spec.ref_epoch = int(frames_per_second), int(ref_epoch)
        spec.ref_second, spec.ref_frame = int(ref_second), int(ref_frame)
        spec.offset_binary8 = int(bool(offset_binary8))
        spec.header_bytes = int(header_bytes)
        st = _lib.VdifStats()
        check_batch(self._L.gxc_batch_set_input_vdif(self._h, int(sub), int(station), ctypes.c_void_p(int(dev_ptr)),
                                                     int(nbytes), ctypes.byref(spec),
                                                     ctypes.byref(st) if stats else None), self._h)
        # the ingest is queued: the frames must live until it has run
        self._keep[("vdif", int(sub), int(station))] = keepalive
        return {k: getattr(st, k) for k, _ in _lib.VdifStats._fields_} if stats else None

    def frame_flags(self, sub, station):
        """Per-slot presence flags of the last VDIF ingest of (sub, station)."""
        n = ctypes.c_longlong(0)
        check_batch(self._L.gxc_batch_get_frame_flags(self._h, int(sub), int(station), None, 0, ctypes.byref(n)),
                    self._h)
        out = np.empty(n.value, np.uint8)
        check_batch(self._L.gxc_batch_get_frame_flags(self._h, int(sub), int(station),
                                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), out.size,
                                                      None), self._h)
        return out

    def set_pol_conversion(self, sub, gains):
        """gains: complex [nant][2][nchan of subband sub], or None to turn the
        conversion off (gxc_batch_set_pol_conversion)."""
        if gains is None:
            check_batch(self._L.gxc_batch_set_pol_conversion(self._h, int(sub), None), self._h)
            return
        n = self.subbands[sub][0]
        g = np.ascontiguousarray(gains, np.complex64).reshape(self.nant, 2, n)
        check_batch(self._L.gxc_batch_set_pol_conversion(self._h, int(sub), g.ctypes.data_as(ctypes.c_void_p)),
                    self._h)

    def process(self):
        check_batch(self._L.gxc_batch_process(self._h), self._h)

    def process_async(self):
        check_batch(self._L.gxc_batch_process_async(self._h), self._h)

    def synchronize(self):
        check_batch(self._L.gxc_batch_synchronize(self._h), self._h)

    def visibilities(self, sub):
        out = np.empty((self.nrows, 4, self.subbands[sub][0]), np.complex64)
        check_batch(self._L.gxc_batch_get_visibilities(self._h, int(sub),
                                                       out.ctypes.data_as(ctypes.c_void_p), out.size),
                    self._h)
        return out

    def row_counts(self, sub):
        out = np.empty(self.nrows, np.int32)
        check_batch(self._L.gxc_batch_get_row_counts(
            self._h, int(sub), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), out.size), self._h)
        return out

    def autotune(self):
        """gxc_batch_autotune: F workgroup size, X unit count (splits) and graph
        replay timed on this batch's input; the fastest is kept."""
        t = _lib.Tuning()
        check_batch(self._L.gxc_batch_autotune(self._h, ctypes.byref(t)), self._h)
        return {k: getattr(t, k) for k, _ in _lib.Tuning._fields_}

    def set_profiling(self, on=True):
        check_batch(self._L.gxc_batch_set_profiling(self._h, int(bool(on))), self._h)

    def timings(self):
        t = _lib.Timings()
        check_batch(self._L.gxc_batch_get_timings(self._h, ctypes.byref(t)), self._h)
        return {k: getattr(t, k) for k, _ in _lib.Timings._fields_}
