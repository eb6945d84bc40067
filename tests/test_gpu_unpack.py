"""Bit-exact 2-bit unpack in every F-stage kernel, seen through closed-form spectra.

fxkernel's layout (fxkernel/src/fxkernel.cpp:443-476, levels 15-28): time
sample s is nibble s & 1 of byte s >> 1, polarisation A in bits 0-1 of the
nibble and B in bits 2-3; codes 0..3 are -3.3359, -1, +1, +3.3359.

With zero delays and zero phase the rotator and the fractional-sample
correction are exactly the identity, so the F-stage returns the FFT of the
unpacked samples.  Every FFT window here holds a pattern of period 4 samples
x0..x3 built from two byte values, so its only non-zero kept bins are
  X[0]   = (L/4) (x0 + x1 + x2 + x3)
  X[L/4] = (L/4) ((x0 - x2) + i (x3 - x1)),
and the order of the four samples (which byte, which nibble, which two bits)
shows in X[L/4].  Station 0 reads its windows from even sample offsets and
station 1 (delayed by one sample) from odd ones; over the windows every byte
value 0..255 is read at both parities.  The two bins must equal the closed
form to within a few FP32 ulps (the FP32 matrix-core form, which carries
samples as f16 hi + lo pairs, to 22 bits), while a wrongly unpacked sample
pattern moves X[0] or X[L/4] by at least L/4 times the smallest level step
(2), and breaks the period (other kept bins non-zero).
"""
import numpy as np
import pytest

import gcorr_amd

pytestmark = pytest.mark.gpu

LEVELS = np.array([-3.3359, -1.0, 1.0, 3.3359], np.float32).astype(np.float64)


def code(byte, nibble, pol):
    return (int(byte) >> (4 * nibble + 2 * pol)) & 3


def build(L, numffts):
    """Station buffers and, per (station, window), the four bytes of the
    sample pattern as ((byte, nibble) for x0..x3)."""
    nbytes = (numffts + 1) * L // 2  # (numffts+1) L time samples, two per byte
    s0 = np.zeros(nbytes, np.uint8)
    s1 = np.zeros(nbytes, np.uint8)
    pat = {}
    half = L // 2  # bytes per window
    # station 0, even offsets tL: bytes B0, B1 alternate; window t reads
    # x = (lo B0, hi B0, lo B1, hi B1)
    for t in range(numffts):
        b0, b1 = (2 * t) % 256, (2 * t + 1) % 256
        seg = np.empty(half, np.uint8)
        seg[0::2], seg[1::2] = b0, b1
        s0[t * half:(t + 1) * half] = seg
        pat[(0, t)] = ((b0, 0), (b0, 1), (b1, 0), (b1, 1))
    # station 1, odd offsets tL - 1 (even t >= 2): bytes tL/2 - 1 .. tL/2 +
    # L/2 - 1 alternate D1, D0, .., D1; window t reads x = (hi D1, lo D0,
    # hi D0, lo D1)
    for i, t in enumerate(range(2, numffts, 2)):
        d0, d1 = (2 * i) % 256, (2 * i + 1) % 256
        seg = np.empty(half + 1, np.uint8)
        seg[0::2], seg[1::2] = d1, d0
        s1[t * half - 1:t * half + half] = seg
        pat[(1, t)] = ((d1, 1), (d0, 0), (d0, 1), (d1, 0))
    return [s0, s1], pat


def expected(pattern, pol, L):
    x = [LEVELS[code(b, n, pol)] for b, n in pattern]
    return (L / 4) * (x[0] + x[1] + x[2] + x[3]), (L / 4) * complex(x[0] - x[2], x[3] - x[1])


@pytest.mark.parametrize("L,fblock", [(256, 0), (2048, 0), (2048, 64), (2048, -1), (4096, -1), (8192, 0),
                                      (8192, 256)])
def test_unpack_bit_exact_through_closed_form_spectra(gpu, L, fblock):
    nchan, numffts, bw = L // 2, 258, 64e6
    bufs, pat = build(L, numffts)
    assert all(b.size == gcorr_amd.station_bytes(2, nchan, numffts) for b in bufs)
    ts = 1.0 / (2 * bw)
    k = gcorr_amd.FxKernel(2, nchan, numffts, 2, 1.65e9, bw, fblock=fblock)
    k.setInputData(bufs)
    k.setDelays(np.zeros((2, 3)), np.array([0.0, -ts]))
    spec, valid = k.fstage(0, numffts)
    k.close()
    assert valid.all()
    mfma = fblock == -1
    seen = {(0, 0): set(), (0, 1): set(), (1, 0): set(), (1, 1): set()}  # (parity, nibble) -> bytes
    other = np.ones(nchan, bool)
    other[[0, L // 4]] = False
    for (st, t), p in pat.items():
        for b, n in p:
            seen[(st, n)].add(b)
        for pol in range(2):
            X = spec[t, 2 * st + pol].astype(np.complex128)
            dc, q = expected(p, pol, L)
            tol = (lambda v: 1e-6 * abs(v) + 1e-3) if mfma else (lambda v: 4 * np.spacing(np.float32(abs(v))) + 1e-6)
            assert abs(X[0].real - dc) <= tol(dc) and abs(X[0].imag) <= tol(0.0), (st, t, pol, X[0], dc)
            assert abs(X[L // 4].real - q.real) <= tol(q.real), (st, t, pol, X[L // 4], q)
            assert abs(X[L // 4].imag - q.imag) <= tol(q.imag), (st, t, pol, X[L // 4], q)
            # the other kept bins hold only rounding residue of the twiddles
            assert np.abs(X[other]).max() <= 1e-5 * L, (st, t, pol)
    # every byte value read through both nibbles at even and at odd sample offsets
    for key, s in seen.items():
        assert s == set(range(256)), key
