"""Visibility comparison metrics.

eps: per (row, product) ||V - V_ref||_2 / ||V_ref||_2 over channels, maximum
over rows and products (the parity measure).  Amplitude and phase errors per
polarisation product (AA, AB, BA, BB): RMS over rows and channels of
|V|/|V_ref| - 1 and of arg(V conj(V_ref)), plus eps restricted to the product.
"""
import numpy as np

PRODUCTS = ("AA", "AB", "BA", "BB")


def eps(v, ref):
    num = np.linalg.norm(v - ref, axis=-1)
    den = np.linalg.norm(ref, axis=-1)
    ok = den > 0
    if np.any(num[~ok] != 0):
        return float("inf")
    return float((num[ok] / den[ok]).max()) if ok.any() else 0.0


def amp_phase_errors(v, ref, rows=None):
    """{product: {"amp_rms", "phase_rms_rad", "eps"}} over the given rows."""
    v = np.asarray(v, np.complex128)
    ref = np.asarray(ref, np.complex128)
    if rows is not None:
        v, ref = v[rows], ref[rows]
    out = {}
    for p, name in enumerate(PRODUCTS):
        a, r = v[:, p], ref[:, p]
        mask = np.abs(r) > 0
        amp = np.abs(a[mask]) / np.abs(r[mask]) - 1.0
        ph = np.angle(a[mask] * np.conj(r[mask]))
        out[name] = {"amp_rms": float(np.sqrt(np.mean(amp ** 2))),
                     "phase_rms_rad": float(np.sqrt(np.mean(ph ** 2))),
                     "eps": eps(a, r)}
    return out
