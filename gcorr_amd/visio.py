"""Readers for the visibility files the library writes.

read_binary: gxc_save_visibilities_binary ("GXCVIS01", layout in
include/gxcorr.h).  read_difx: gxc_save_difx, DiFX binary visibility
records (74-byte header + nchan complex64).  read_text: the vis.out text of
gxc_save_visibilities (fxkernel/src/fxkernel.cpp:389-399).
"""
import numpy as np

DIFX_SYNC = 0xFF00FF00
_DIFX_HDR = np.dtype([("sync", "<u4"), ("version", "<i4"), ("baseline", "<i4"), ("mjd", "<i4"),
                      ("seconds", "<f8"), ("config_index", "<i4"), ("source_index", "<i4"),
                      ("freq_index", "<i4"), ("polpair", "S2"), ("pulsar_bin", "<i4"),
                      ("weight", "<f8"), ("u", "<f8"), ("v", "<f8"), ("w", "<f8")])
assert _DIFX_HDR.itemsize == 74


def read_binary(path):
    raw = np.fromfile(path, np.uint8)
    if raw[:8].tobytes() != b"GXCVIS01":
        raise ValueError("not a GXCVIS01 file")
    hdr = raw[8:40].view("<i4")
    version, nant, nchan, nrows, nprod, autos, numffts, precision = (int(x) for x in hdr)
    bw, lo = raw[40:56].view("<f8")
    off = 56
    stations = raw[off:off + 8 * nrows].view("<i4").reshape(nrows, 2)
    off += 8 * nrows
    counts = raw[off:off + 4 * nrows].view("<i4")
    off += 4 * nrows
    vis = raw[off:off + 8 * nrows * nprod * nchan].view("<c8").reshape(nrows, nprod, nchan)
    return {"version": version, "nant": nant, "nchan": nchan, "autos": bool(autos),
            "numffts": numffts, "precision": precision, "bandwidth": float(bw), "lo": float(lo),
            "stations": stations.copy(), "counts": counts.copy(), "vis": vis.copy()}


def read_difx(path, nchan):
    """Records of a DiFX binary visibility file: list of (header dict, complex64[nchan])."""
    rec = np.dtype([("h", _DIFX_HDR), ("vis", "<c8", (nchan,))])
    arr = np.fromfile(path, rec)
    out = []
    for r in arr:
        h = {k: (r["h"][k].decode() if k == "polpair" else r["h"][k].item()) for k in _DIFX_HDR.names}
        if h["sync"] != DIFX_SYNC:
            raise ValueError("bad DiFX sync word")
        out.append((h, r["vis"].copy()))
    return out
