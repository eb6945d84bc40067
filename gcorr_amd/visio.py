"""Readers for the visibility files the library writes.

read_binary: gxc_save_visibilities_binary ("GXCVIS01", layout in
include/gxcorr.h).  read_difx: gxc_save_difx, DiFX binary visibility
records (74-byte header + nchan complex64).  read_text: the vis.out text of
gxc_save_visibilities (fxkernel/src/fxkernel.cpp:389-399).  compare_text:
two vis.out files compared as fxkernel/scripts/verifyfxkernel.py intends
(the command-line form is gcorr_amd/verify.py).
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


def read_text(path, nprod=4):
    """A vis.out text file of gxc_save_visibilities / fxkernel's
    saveVisibilities (fxkernel/src/fxkernel.cpp:389-399): one line per
    channel, the channel index and centre frequency in MHz, then for each row
    and product re, im, amplitude and phase.

    Returns {"channel": int[nchan], "freq_mhz": float[nchan],
    "vis": complex64[nrows][nprod][nchan], "amp", "phase": float[nrows][nprod][nchan]}.
    """
    data = np.loadtxt(path, dtype=np.float64, ndmin=2)
    ncol = data.shape[1]
    if ncol < 2 or (ncol - 2) % (4 * nprod):
        raise ValueError("%s: %d columns is not 2 + 4 x %d x rows" % (path, ncol, nprod))
    nrows = (ncol - 2) // (4 * nprod)
    body = data[:, 2:].reshape(data.shape[0], nrows, nprod, 4).transpose(1, 2, 3, 0)
    return {"channel": data[:, 0].astype(np.int64), "freq_mhz": data[:, 1],
            "vis": (body[:, :, 0] + 1j * body[:, :, 1]).astype(np.complex64),
            "amp": body[:, :, 2], "phase": body[:, :, 3]}


def compare_text(ref_path, new_path, nrows, maxepsilon=1e-7, nprod=4):
    """Compare two vis.out files as fxkernel/scripts/verifyfxkernel.py:57-77
    intends: for every row and product, the largest difference over channels
    as a fraction of that series' typical visibility.

    The reference script indexes its columns from 1 (the frequency column),
    takes the signed maximum of the differences and divides by the signed
    median of one component, so a negative median or a difference of either
    sign can pass or fail by accident.  Here the measure is
    max_c |V_new - V_ref| / median_c |V_ref| on the complex values (a series
    whose reference is zero everywhere counts any difference as infinite).

    Returns {"pass", "max_abs_diff", "worst": (row, product, fraction),
    "fail": [(row, product, fraction)], "fraction": float[nrows][nprod]}.
    """
    a, b = read_text(ref_path, nprod), read_text(new_path, nprod)
    if a["vis"].shape != b["vis"].shape:
        raise ValueError("visibility shapes differ: %s, %s" % (a["vis"].shape, b["vis"].shape))
    if a["vis"].shape[0] != nrows:
        raise ValueError("%d rows in the files, %d expected" % (a["vis"].shape[0], nrows))
    ref = a["vis"].astype(np.complex128)
    diff = np.abs(b["vis"].astype(np.complex128) - ref).max(axis=-1)
    typ = np.median(np.abs(ref), axis=-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        frac = np.where(typ > 0, diff / np.where(typ > 0, typ, 1.0), np.where(diff > 0, np.inf, 0.0))
    fail = [(int(r), int(p), float(frac[r, p])) for r, p in zip(*np.nonzero(frac > maxepsilon))]
    r, p = np.unravel_index(int(np.argmax(frac)), frac.shape)
    return {"pass": not fail, "max_abs_diff": float(diff.max()),
            "worst": (int(r), int(p), float(frac[r, p])), "fail": fail, "fraction": frac}
