"""Requantiser and polarisation-conversion restatements, on the CPU.

Requantisers (oracle/requant.py) are pinned by threshold edge cases, by the
data generator's own quantiser and packer, and by decoding their output
through the fxkernel unpacker restatement.  The per-spectrum polarisation
conversion of the oracle (gpubench.cu:141-166) is pinned by exact identities
(g = (1, 0) copies pol 0 into both outputs; g = (0, 1) gives -x1, x1) and by
the FP64 equality with the conversion applied to the visibilities,
V' = J_j V J_k^H, which is what the GPU computes.
"""
import math

import numpy as np
import pytest
import torch

from gcorr_amd import synth
from oracle import fxref
from oracle.requant import codes2, requant2, requant8


def test_codes_at_the_edges():
    t = 0.98
    x = np.array([-np.inf, -1.0, -0.98, -0.5, -0.0, 0.0, 0.5, 0.98, 1.0, np.inf, np.nan], np.float32)
    assert list(codes2(x, t)) == [0, 0, 1, 1, 2, 2, 2, 3, 3, 3, 1]


def test_matches_the_generator_quantiser():
    g = torch.Generator().manual_seed(3)
    a = torch.randn(4096, generator=g)
    b = torch.randn(4096, generator=g)
    want = synth.pack2bit(synth._quantise2(a, 1.0), synth._quantise2(b, 1.0)).numpy()
    np.testing.assert_array_equal(requant2(a.numpy(), b.numpy(), synth.THRESH), want)


def test_decodes_through_the_unpacker():
    rng = np.random.default_rng(4)
    n = 2 * 64 * 3
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    packed = requant2(a, b, 0.98)
    cfg = fxref.make_config(2, 64, 2, 1.65e9, 16e6)
    u = fxref.unpack(cfg, packed, 0)  # [2][L] complex levels of FFT 0
    lev = np.array([-fxref.lut2bit()[0b11111111][0], -1.0, 1.0, fxref.lut2bit()[0b11111111][0]])
    L = 128
    np.testing.assert_array_equal(u[0].real, lev[codes2(a[:L], 0.98)])
    np.testing.assert_array_equal(u[1].real, lev[codes2(b[:L], 0.98)])


def test_8bit_rounding_and_clamp():
    a = np.array([0.5 + 1.5j, 2.5 - 2.5j, 200 - 300j, np.nan + 1j], np.complex64)
    b = np.array([-0.5 - 1.5j, 3.5 + 0j, -127.4 + 126.6j, 0j], np.complex64)
    q = requant8(a, b, 1.0).view(np.int8).reshape(4, 4)
    assert q.tolist() == [[0, 2, 0, -2], [2, -2, 4, 0], [127, -127, -127, 127], [0, 1, 0, 0]]


def jones_apply(v, gains, nant, autos=True):
    rows = [(j, k) for j in range(nant) for k in range(j + 1, nant)] + \
        ([(a, a) for a in range(nant)] if autos else [])
    out = np.empty_like(v)
    for r, (j, k) in enumerate(rows):
        Jj = np.array([[gains[j, 0], -gains[j, 1]], [gains[j, 0], gains[j, 1]]])  # [2][2][nchan]
        Jk = np.array([[gains[k, 0], -gains[k, 1]], [gains[k, 0], gains[k, 1]]])
        V = v[r].reshape(2, 2, -1)
        out[r] = np.einsum("lpc,pqc,mqc->lmc", Jj, V, np.conj(Jk)).reshape(4, -1)
    return out


def case(nant=3, nchan=64, numffts=40, seed=5):
    w = synth.Workload("p", nant=nant, nchan=nchan, numffts=numffts, bw=32e6, lo=1.65e9, seed=seed)
    rng = np.random.default_rng(seed)
    data = [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(nant)]
    poly, offs = synth.make_delays(w)
    g = (rng.standard_normal((nant, 2, nchan)) + 1j * rng.standard_normal((nant, 2, nchan))) / math.sqrt(2)
    return w, data, poly, offs, g.astype(np.complex64)


def test_polconv_per_spectrum_equals_jones_on_visibilities():
    w, data, poly, offs, g = case()
    base = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True)
    conv = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True, polgains=g)
    v0, c0 = fxref.process(base, data, poly, offs, "f64")
    v1, c1 = fxref.process(conv, data, poly, offs, "f64")
    np.testing.assert_array_equal(c0, c1)
    np.testing.assert_allclose(v1, jones_apply(v0, g.astype(np.complex128), w.nant), rtol=1e-10,
                               atol=1e-10 * np.abs(v0).max())


def test_polconv_exact_identities():
    w, data, poly, offs, _ = case(seed=6)
    base = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True)
    v0, _ = fxref.process(base, data, poly, offs)
    g = np.zeros((w.nant, 2, w.nchan), np.complex64)
    g[:, 0] = 1.0
    v1, _ = fxref.process(fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True,
                                            polgains=g), data, poly, offs)
    for p in range(4):  # y0 = y1 = x0: every product is the (0, 0) product
        np.testing.assert_array_equal(v1[:, p], v0[:, 0])
    g[:] = 0
    g[:, 1] = 1.0
    v2, _ = fxref.process(fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True,
                                            polgains=g), data, poly, offs)
    np.testing.assert_array_equal(v2[:, 0], v0[:, 3])   # (-x1)(-x1)* = x1 x1*
    np.testing.assert_array_equal(v2[:, 1], -v0[:, 3])
    np.testing.assert_array_equal(v2[:, 2], -v0[:, 3])
    np.testing.assert_array_equal(v2[:, 3], v0[:, 3])
