"""Restatement of util/lagsnr.py's fringe check (lines 33-36).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of gxc_lag_snr.

For each series of channel visibilities: lags = |ifft(vis)|,
snr = max(lags) / std(lags) (numpy's population standard deviation) and the
peak = argmax(lags) (the first index of the maximum).
"""
import numpy as np


def lag_snr(vis):
    v = np.asarray(vis, np.complex128)
    lags = np.abs(np.fft.ifft(v, axis=-1))
    return lags.max(axis=-1) / lags.std(axis=-1), lags.argmax(axis=-1)


def read_vis_out(path):
    """Channel visibilities [nblpol][nchan] of a vis.out text file, with the
    column rule of util/lagsnr.py:14-26 (8 leading columns when the rest
    splits into groups of 4, else 2)."""
    rows = [ln.split() for ln in open(path)]
    ncols = len(rows[0])
    skip = 8 if (ncols - 8) % 4 == 0 else 2
    if (ncols - skip) % 4:
        raise ValueError("unknown vis.out layout")
    data = np.array([[float(x) for x in r[skip:]] for r in rows])
    nbl = (ncols - skip) // 4
    return np.stack([data[:, 4 * i] + 1j * data[:, 4 * i + 1] for i in range(nbl)])
