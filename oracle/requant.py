"""Restatement of the requantisers (gxc_requantise_2bit / _8bit).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.

2-bit: gpubench packs float filterbank output to 2 bits
(examples/chris/bench/gpubench.cu:64-130) with its own code order; the codes
here are the ones fxkernel's LUT decodes (fxkernel/src/fxkernel.cpp:15-28):
x < -t -> 0, -t <= x < 0 -> 1, 0 <= x < t -> 2, x >= t -> 3 (NaN -> 1), two
polarisations packed LSB first as [A_t, B_t, A_t+1, B_t+1]
(fxkernel.cpp:443-476).  8-bit: complex samples to
clamp(round-half-even(x * scale), -127, 127) (NaN -> 0) as
[Are, Aim, Bre, Bim] int8 per time sample.
"""
import numpy as np


def codes2(x, t):
    x = np.asarray(x, np.float32)
    c = np.ones(x.shape, np.uint8)
    c[x >= 0] = 2
    c[x >= np.float32(t)] = 3
    c[x < -np.float32(t)] = 0
    return c


def requant2(a, b, t):
    ca, cb = codes2(a, t), codes2(b, t)
    return (ca[0::2] | (cb[0::2] << 2) | (ca[1::2] << 4) | (cb[1::2] << 6)).astype(np.uint8)


def q8(x, scale):
    y = np.rint(np.asarray(x, np.float32) * np.float32(scale))
    y = np.where(np.isnan(y), 0, np.clip(y, -127, 127))
    return y.astype(np.int8)


def requant8(a, b, scale):
    """a, b: complex arrays of one polarisation each."""
    a = np.asarray(a, np.complex64)
    b = np.asarray(b, np.complex64)
    out = np.stack([q8(a.real, scale), q8(a.imag, scale), q8(b.real, scale), q8(b.imag, scale)], axis=1)
    return out.reshape(-1).view(np.uint8)
