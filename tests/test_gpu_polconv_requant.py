"""Polarisation conversion and requantisers on the GPU.

gxc_set_pol_conversion must give the oracle's per-spectrum conversion
(gpubench.cu:141-166) within the FP32 bound, with unchanged row counts;
gxc_requantise_2bit/_8bit must equal the restatement bit for bit, and their
output must correlate like the oracle on the same bytes.
"""
import math

import numpy as np
import pytest
import torch

import gcorr_amd
from gcorr_amd import synth
from oracle import fxref
from oracle.requant import requant2, requant8

pytestmark = pytest.mark.gpu

EPS = 1e-5


def eps(v, ref):
    num = np.linalg.norm(v - ref, axis=-1)
    den = np.linalg.norm(ref, axis=-1)
    mask = den > 0
    assert np.all(num[~mask] == 0)
    return float((num[mask] / den[mask]).max()) if mask.any() else 0.0


def case(nant, nchan, numffts, seed):
    w = synth.Workload("p", nant=nant, nchan=nchan, numffts=numffts, bw=64e6, lo=1.65e9, seed=seed)
    rng = np.random.default_rng(seed)
    data = [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(nant)]
    poly, offs = synth.make_delays(w)
    g = (rng.standard_normal((nant, 2, nchan)) + 1j * rng.standard_normal((nant, 2, nchan))) / math.sqrt(2)
    return w, data, poly, offs, g.astype(np.complex64)


@pytest.mark.parametrize("graph", [True, False])
def test_pol_conversion_matches_oracle(gpu, graph):
    w, data, poly, offs, g = case(4, 512, 120, seed=51)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True, graph=graph)
    k.setInputData(data)
    k.setDelays(poly, offs)
    k.process()
    plain, counts = k.visibilities(), k.row_counts()
    k.set_pol_conversion(g)
    k.process()
    conv = k.visibilities()
    np.testing.assert_array_equal(k.row_counts(), counts)
    cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True, polgains=g)
    ref, rc = fxref.process(cfg, data, poly, offs)
    np.testing.assert_array_equal(rc, counts)
    assert eps(conv, ref) < EPS
    # circular from linear: g0 = 1/sqrt2, g1 = i/sqrt2 for every station
    circ = np.zeros_like(g)
    circ[:, 0] = 1 / math.sqrt(2)
    circ[:, 1] = 1j / math.sqrt(2)
    k.set_pol_conversion(circ)
    k.process()
    ref2, _ = fxref.process(fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True,
                                              polgains=circ), data, poly, offs)
    assert eps(k.visibilities(), ref2) < EPS
    k.set_pol_conversion(None)
    k.process()
    np.testing.assert_array_equal(k.visibilities(), plain)
    k.close()


def test_pol_conversion_reduced_precision(gpu):
    w, data, poly, offs, g = case(4, 1024, 200, seed=52)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True, precision="fp16")
    k.setInputData(data)
    k.setDelays(poly, offs)
    k.set_pol_conversion(g)
    k.process()
    ref, _ = fxref.process(fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True,
                                             polgains=g), data, poly, offs, "f64")
    assert eps(k.visibilities(), ref) < 1e-3
    k.close()


@pytest.mark.parametrize("n", [16, 2 * 7, 4096 * 3 + 10, 1 << 20])
def test_requantise_2bit_bit_exact(gpu, n):
    gen = torch.Generator(device="cuda").manual_seed(n)
    a = torch.randn(n, device="cuda", generator=gen) * 1.3
    b = torch.randn(n, device="cuda", generator=gen)
    a[:8] = torch.tensor([0.98, -0.98, 0.0, -0.0, float("nan"), float("inf"), -float("inf"), 0.9799999],
                         device="cuda")[: min(8, n)]
    out = torch.empty(n // 2, dtype=torch.uint8, device="cuda")
    gcorr_amd.requantise_2bit(a.data_ptr(), b.data_ptr(), n, 0.98, out.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), requant2(a.cpu().numpy(), b.cpu().numpy(), 0.98))


@pytest.mark.parametrize("n", [4, 7, 1001, 1 << 18])
def test_requantise_8bit_bit_exact(gpu, n):
    gen = torch.Generator(device="cuda").manual_seed(n + 1)
    a = torch.randn(n, dtype=torch.complex64, device="cuda", generator=gen) * 60
    b = torch.randn(n, dtype=torch.complex64, device="cuda", generator=gen) * 60
    a[0] = complex(2.5, -2.5)
    a[1] = complex(float("nan"), 300.0)
    out = torch.empty(4 * n, dtype=torch.uint8, device="cuda")
    gcorr_amd.requantise_8bit(a.data_ptr(), b.data_ptr(), n, 1.0, out.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), requant8(a.cpu().numpy(), b.cpu().numpy(), 1.0))


def test_requantised_input_correlates_like_the_oracle(gpu):
    w = synth.Workload("q", nant=3, nchan=256, numffts=100, bw=64e6, lo=1.65e9, seed=53)
    n = w.station_bytes * 2  # time samples per pol
    gen = torch.Generator(device="cuda").manual_seed(53)
    common = torch.randn(2, n, device="cuda", generator=gen)
    bufs = []
    for a in range(w.nant):
        x = 0.5 * common + torch.randn(2, n, device="cuda", generator=gen)
        out = torch.empty(w.station_bytes, dtype=torch.uint8, device="cuda")
        gcorr_amd.requantise_2bit(x[0].data_ptr(), x[1].data_ptr(), n, 0.98 * math.sqrt(1.25), out.data_ptr())
        bufs.append(out)
    torch.cuda.synchronize()
    poly, offs = synth.make_delays(w)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True)
    k.setInputDevice([b.data_ptr() for b in bufs], keepalive=bufs)
    k.setDelays(poly, offs)
    k.process()
    host = [b.cpu().numpy() for b in bufs]
    ref, _ = fxref.process(fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True),
                           host, poly, offs)
    assert eps(k.visibilities(), ref) < EPS
    k.close()


def test_requantiser_argument_errors(gpu):
    a = torch.zeros(64, device="cuda")
    out = torch.empty(64, dtype=torch.uint8, device="cuda")
    with pytest.raises(gcorr_amd.GxcError):
        gcorr_amd.requantise_2bit(a.data_ptr(), a.data_ptr(), 63, 0.98, out.data_ptr())  # odd count
    with pytest.raises(gcorr_amd.GxcError):
        gcorr_amd.requantise_2bit(a.data_ptr() + 4, a.data_ptr(), 32, 0.98, out.data_ptr())  # misaligned
    with pytest.raises(gcorr_amd.GxcError):
        gcorr_amd.requantise_2bit(a.data_ptr(), a.data_ptr(), 32, 0.0, out.data_ptr())
    with pytest.raises(gcorr_amd.GxcError):
        gcorr_amd.requantise_8bit(a.data_ptr(), a.data_ptr(), 8, 1.0, out.data_ptr() + 8)
