"""Streamed subintegrations: gxc_integrate_begin / gxc_submit / gxc_integrate_end.

fxkernel's caller processes one subintegration at a time and sums them with
accumulate() (fxkernel/src/fxkernel.cpp:414-423); gcorr's driver overlaps the
host copies with compute (gcorr/src/testgpukernel.cu:448-486).  The streamed
sum must equal, bit for bit, the float32 sum in submission order of what
gxc_process gives for each subintegration alone, and match the oracle's sum
within the FP32 bound.
"""
import numpy as np
import pytest

import gcorr_amd
from gcorr_amd import synth
from oracle import fxref

pytestmark = pytest.mark.gpu

EPS = 1e-5


def eps(v, ref):
    num = np.linalg.norm(v - ref, axis=-1)
    den = np.linalg.norm(ref, axis=-1)
    mask = den > 0
    assert np.all(num[~mask] == 0)
    return float((num[mask] / den[mask]).max()) if mask.any() else 0.0


def subints(w, n, seed):
    """n subintegrations of random bytes, each with its own delays."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        data = [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(w.nant)]
        poly, offs = synth.make_delays(w)
        offs = offs + k * 3.0 * w.sampletime
        poly = poly.copy()
        poly[:, 2] += k * 1e-7
        out.append((data, poly, offs))
    return out


def individual(k, subs):
    vis = []
    for data, poly, offs in subs:
        k.setInputData(data)
        k.setDelays(poly, offs)
        k.process()
        vis.append(k.visibilities())
    return vis


def ordered_sum(vis):
    acc = np.zeros_like(vis[0])
    for v in vis:
        acc = acc + v  # complex64 + complex64: per-component float32 adds
    return acc


@pytest.mark.parametrize("precision,overlap", [("fp32", False), ("fp32", True), ("fp16", False)])
def test_stream_equals_sum_of_process(gpu, precision, overlap):
    w = synth.Workload("s", nant=4, nchan=512, numffts=200, bw=64e6, lo=1.65e9, seed=3)
    subs = subints(w, 5, seed=21)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True,
                           precision=precision, overlap=overlap, chunk_ffts=64)
    expect = ordered_sum(individual(k, subs))
    k.integrate_begin()
    for data, poly, offs in subs:
        k.submit(data, poly, offs)
    got, n = k.integrate_end()
    k.close()
    assert n == len(subs)
    np.testing.assert_array_equal(got, expect)


def test_stream_matches_oracle_sum_from_pinned_buffers(gpu):
    w = synth.Workload("s", nant=3, nchan=256, numffts=150, bw=32e6, lo=1.65e9, seed=4)
    subs = subints(w, 4, seed=5)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True)
    # two sets of pinned station buffers, refilled as a capture loop would:
    # set i&1 was last submitted with subintegration i-2, and the submit of
    # i-1 has returned, so it may be overwritten
    pins = [[gcorr_amd.PinnedBuffer(w.station_bytes) for _ in range(w.nant)] for _ in range(2)]
    k.integrate_begin()
    for i, (data, poly, offs) in enumerate(subs):
        bufs = pins[i & 1]
        for b, d in zip(bufs, data):
            b.array[:] = d
        k.submit([b.array for b in bufs], poly, offs)
    got, n = k.integrate_end()
    k.close()
    assert n == 4
    cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True)
    ref = None
    for data, poly, offs in subs:
        v, _ = fxref.process(cfg, data, poly, offs, "f32")
        ref = v.astype(np.complex128) if ref is None else ref + v
    assert eps(got, ref) < EPS
    for row in pins:
        for b in row:
            b.free()


def test_stream_restarts_and_mixes_with_process(gpu):
    w = synth.Workload("s", nant=3, nchan=128, numffts=64, bw=16e6, lo=1.65e9, seed=6)
    subs = subints(w, 3, seed=8)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw)
    vis = individual(k, subs)
    k.integrate_begin()
    k.submit(*subs[0])
    k.submit(*subs[1])
    first, n1 = k.integrate_end()
    assert n1 == 2
    np.testing.assert_array_equal(first, ordered_sum(vis[:2]))
    # a plain process() between integrations still gives its own result
    k.setInputData(subs[2][0])
    k.setDelays(subs[2][1], subs[2][2])
    k.process()
    np.testing.assert_array_equal(k.visibilities(), vis[2])
    # a new integration starts from zero
    k.integrate_begin()
    k.submit(*subs[2])
    second, n2 = k.integrate_end()
    assert n2 == 1
    np.testing.assert_array_equal(second, vis[2])
    k.close()


def test_stream_state_errors(gpu):
    w = synth.Workload("s", nant=3, nchan=64, numffts=16, bw=16e6, lo=1.65e9, seed=7)
    (data, poly, offs), = subints(w, 1, seed=9)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw)
    with pytest.raises(gcorr_amd.GxcError):
        k.submit(data, poly, offs)  # before integrate_begin
    with pytest.raises(gcorr_amd.GxcError):
        k.integrate_end()
    k.integrate_begin()
    with pytest.raises(ValueError):
        k.submit(data[:2], poly, offs)
    with pytest.raises(ValueError):
        k.submit([d[:-1] for d in data], poly, offs)
    got, n = k.integrate_end()  # nothing submitted: zero sum
    assert n == 0 and not np.any(got)
    k.close()


def test_stream_with_pol_conversion(gpu):
    w = synth.Workload("s", nant=3, nchan=256, numffts=64, bw=32e6, lo=1.65e9, seed=12)
    subs = subints(w, 3, seed=13)
    rng = np.random.default_rng(14)
    g = ((rng.standard_normal((3, 2, 256)) + 1j * rng.standard_normal((3, 2, 256))) / 2).astype(np.complex64)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True)
    k.set_pol_conversion(g)
    expect = ordered_sum(individual(k, subs))
    k.integrate_begin()
    for data, poly, offs in subs:
        k.submit(data, poly, offs)
    got, n = k.integrate_end()
    k.close()
    assert n == 3
    np.testing.assert_array_equal(got, expect)


def test_stream_after_vdif_input(gpu):
    """A VDIF ingest's frame flags do not leak into submitted subintegrations,
    and the handle needs new input afterwards."""
    import torch
    from gcorr_amd import vdif
    w = synth.Workload("s", nant=3, nchan=128, numffts=64, bw=16e6, lo=1.65e9, seed=15)
    subs = subints(w, 2, seed=16)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw)
    expect = ordered_sum(individual(k, subs))
    data0, poly0, offs0 = subs[0]
    devs = []
    for a in range(w.nant):
        fr = vdif.frames_from_station(data0[a], 256, frames_per_second=10).reshape(-1)
        fr = fr[: fr.size - fr.size // 2]  # drop the second half of the frames
        dev = torch.from_numpy(np.ascontiguousarray(fr)).to("cuda")
        devs.append(dev)
        k.setInputVDIF(a, dev.data_ptr(), fr.size, frames_per_second=10, keepalive=dev)
    k.setDelays(poly0, offs0)
    k.process()
    assert k.row_counts().max() < w.numffts  # the missing frames took FFTs out
    k.integrate_begin()
    for data, poly, offs in subs:
        k.submit(data, poly, offs)
    got, n = k.integrate_end()
    np.testing.assert_array_equal(got, expect)
    with pytest.raises(gcorr_amd.GxcError):  # the VDIF input was released
        k.process()
    k.close()
