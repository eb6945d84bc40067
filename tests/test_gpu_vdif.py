"""VDIF ingest on the GPU (gxc_set_input_vdif) against the oracle's placement.

Frames are shuffled, dropped, duplicated, flagged invalid, mixed with
another thread and with frames outside the subintegration.  The GPU's
counts and per-slot flags must equal the oracle's exactly, and the
visibilities and row counts of the ingested subintegration must match the
oracle run on the oracle's placed buffers with its missing-frame flags
(FP32 within 1e-5, counts exact).  With every frame present the result is
bit-identical to handing the original buffers to gxc_set_input.
"""
import numpy as np
import pytest
import torch

import gcorr_amd
from gcorr_amd import synth, vdif
from oracle import fxref

pytestmark = pytest.mark.gpu

EPS = 1e-5


def eps(v, ref):
    num = np.linalg.norm(v - ref, axis=-1)
    den = np.linalg.norm(ref, axis=-1)
    mask = den > 0
    assert np.all(num[~mask] == 0)
    return float((num[mask] / den[mask]).max()) if mask.any() else 0.0


def case(nant, nchan, numffts, seed, is_complex=False):
    w = synth.Workload("v", nant=nant, nchan=nchan, numffts=numffts, bw=64e6, lo=1.65e9,
                       is_complex=is_complex, seed=seed)
    rng = np.random.default_rng(seed)
    data = [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(nant)]
    poly, offs = synth.make_delays(w)
    return w, data, poly, offs, rng


def damaged_stream(buf, payload, rng, nbit, is_complex, fps, ref_second, ref_frame, drop):
    fr = vdif.frames_from_station(buf, payload, nbit, is_complex, frames_per_second=fps,
                                  ref_second=ref_second, ref_frame=ref_frame)
    n = fr.shape[0]
    keep = [k for k in range(n) if k not in drop]
    extra = []
    for k in rng.choice(n, 3, replace=False):
        d = fr[k].copy()
        d[40:] ^= 0x5A
        extra.append(d)  # later duplicates: ignored
    inv = fr[rng.integers(n)].copy()
    vdif.set_header(inv, invalid=1)
    extra.append(inv)
    oth = fr[rng.integers(n)].copy()
    vdif.set_header(oth, thread=5)
    extra.append(oth)
    early = fr[0].copy()
    vdif.set_header(early, seconds=ref_second - 1)
    extra.append(early)
    allf = np.concatenate([fr[keep]] + [e[None] for e in extra])
    rng.shuffle(allf[:len(keep)])  # originals in any order, before their duplicates
    return allf.reshape(-1)


def ingest_all(k, streams, fps, ref_second, ref_frame, frame_bytes=0):
    stats = []
    devs = []
    for a, s in enumerate(streams):
        dev = torch.from_numpy(s).to("cuda")
        devs.append(dev)
        stats.append(k.setInputVDIF(a, dev.data_ptr(), s.size, frames_per_second=fps,
                                    ref_second=ref_second, ref_frame=ref_frame,
                                    frame_bytes=frame_bytes, keepalive=dev))
    torch.cuda.synchronize()
    return stats


@pytest.mark.parametrize("graph", [True, False])
def test_vdif_ingest_matches_oracle(gpu, graph):
    w, data, poly, offs, rng = case(4, 256, 150, seed=31)
    payload, fps, s0, f0 = 1000, 400, 1234, 17
    drops = [[], [3, 4], [0, 11], [19]]
    streams = [damaged_stream(data[a], payload, rng, 2, False, fps, s0, f0, drops[a])
               for a in range(w.nant)]
    cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True)
    placed, flags, spfs, ostats = [], [], [], []
    for s in streams:
        st, ok, spf, stt = fxref.vdif_place(cfg, s, frames_per_second=fps, ref_second=s0, ref_frame=f0)
        placed.append(st)
        flags.append(ok)
        spfs.append(spf)
        ostats.append(stt)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True, graph=graph)
    # with the frame size given the ingest does not read the header on the host
    gstats = ingest_all(k, streams, fps, s0, f0, frame_bytes=32 + payload if graph else 0)
    assert gstats == ostats
    for a in range(w.nant):
        np.testing.assert_array_equal(k.frame_flags(a), flags[a])
    k.setDelays(poly, offs)
    k.process()
    v, c = k.visibilities(), k.row_counts()
    cfgm = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True, frame_ok=flags,
                             frame_samples=spfs)
    ref, rc = fxref.process(cfgm, placed, poly, offs)
    np.testing.assert_array_equal(c, rc)
    assert c.min() < w.numffts  # the dropped frames removed FFTs
    assert eps(v, ref) < EPS
    # replay with the same input gives the same result
    k.process()
    np.testing.assert_array_equal(k.visibilities(), v)
    k.close()


@pytest.mark.parametrize("precision,bound,mode", [("fp16", 1e-3, "f64"), ("fp32", EPS, "f32")])
def test_vdif_missing_frames_matrix_core_fstage(gpu, precision, bound, mode):
    """Missing frames invalidate FFTs on the matrix-core F-stage too (its
    delay splits, frame checks included, are computed a batch of units at a
    time)."""
    w, data, poly, offs, rng = case(3, 1024, 60, seed=33)
    payload, fps, s0, f0 = 4000, 100, 55, 2
    drops = [[1], [], [6, 7]]
    streams = [damaged_stream(data[a], payload, rng, 2, False, fps, s0, f0, drops[a])
               for a in range(w.nant)]
    cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True)
    placed, flags, spfs = [], [], []
    for st in streams:
        buf, ok, spf, _ = fxref.vdif_place(cfg, st, frames_per_second=fps, ref_second=s0, ref_frame=f0)
        placed.append(buf)
        flags.append(ok)
        spfs.append(spf)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True, precision=precision,
                           fblock=-1)
    ingest_all(k, streams, fps, s0, f0)
    k.setDelays(poly, offs)
    k.process()
    v, c = k.visibilities(), k.row_counts()
    cfgm = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True, frame_ok=flags,
                             frame_samples=spfs)
    ref, rc = fxref.process(cfgm, placed, poly, offs, mode)
    np.testing.assert_array_equal(c, rc)
    assert c.min() < w.numffts
    assert eps(v, ref) < bound
    k.close()


def test_vdif_complete_stream_equals_plain_input(gpu):
    w, data, poly, offs, rng = case(3, 512, 80, seed=32)
    fps = 250
    streams = [vdif.frames_from_station(d, 2048, frames_per_second=fps, legacy=True).reshape(-1)
               for d in data]
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True)
    k.setInputData(data)
    k.setDelays(poly, offs)
    k.process()
    plain = k.visibilities()
    st = ingest_all(k, streams, fps, 0, 0)
    assert all(s["missing"] == 0 and s["placed"] == s["frames"] for s in st)
    k.process()
    np.testing.assert_array_equal(k.visibilities(), plain)
    # back to host input: the flags no longer apply
    k.setInputData(data)
    k.process()
    np.testing.assert_array_equal(k.visibilities(), plain)
    k.close()


def test_vdif_complex8(gpu):
    w, _, poly, offs, rng = case(3, 1024, 40, seed=33, is_complex=True)
    data = [rng.integers(-100, 101, w.station_bytes).astype(np.int8).view(np.uint8)
            for _ in range(w.nant)]
    fps = 1000
    streams = [damaged_stream(data[a], 4096, rng, 8, True, fps, 77, 0, [5] if a == 1 else [])
               for a in range(w.nant)]
    cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, nbit=8, is_complex=True,
                            autos=True)
    placed, flags, spfs = [], [], []
    for s in streams:
        st, ok, spf, _ = fxref.vdif_place(cfg, s, frames_per_second=fps, ref_second=77)
        placed.append(st)
        flags.append(ok)
        spfs.append(spf)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, 8, w.lo, w.bw, True, autos=True)
    ingest_all(k, streams, fps, 77, 0)
    k.setDelays(poly, offs)
    k.process()
    cfgm = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, nbit=8, is_complex=True,
                             autos=True, frame_ok=flags, frame_samples=spfs)
    ref, rc = fxref.process(cfgm, placed, poly, offs)
    np.testing.assert_array_equal(k.row_counts(), rc)
    assert eps(k.visibilities(), ref) < EPS
    k.close()


def test_vdif_errors(gpu):
    w, data, poly, offs, rng = case(3, 64, 16, seed=34)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw)
    good = vdif.frames_from_station(data[0], 64, frames_per_second=10).reshape(-1)
    dev = torch.from_numpy(good).to("cuda")
    with pytest.raises(gcorr_amd.GxcError):  # station out of range
        k.setInputVDIF(3, dev.data_ptr(), good.size, frames_per_second=10)
    with pytest.raises(gcorr_amd.GxcError):  # payload not a multiple of 8 bytes
        k.setInputVDIF(0, dev.data_ptr(), good.size, frames_per_second=10, frame_bytes=32 + 60)
    with pytest.raises(gcorr_amd.GxcError):  # misaligned
        k.setInputVDIF(0, dev.data_ptr() + 4, good.size - 8, frames_per_second=10)
    with pytest.raises(gcorr_amd.GxcError):
        k.setInputVDIF(0, dev.data_ptr(), good.size, frames_per_second=0)
    k.setInputVDIF(0, dev.data_ptr(), good.size, frames_per_second=10)
    k.setDelays(poly, offs)
    with pytest.raises(gcorr_amd.GxcError):  # stations 1 and 2 not ingested yet
        k.process()
    with pytest.raises(gcorr_amd.GxcError):
        k.frame_flags(1)
    torch.cuda.synchronize()
    k.close()


def test_vdif_header_size_given(gpu):
    """First frame with a stray legacy bit: header_bytes = 32 keeps the
    stream's geometry (one malformed frame, visibilities equal plain input);
    header_bytes = 0 follows the first frame, as the oracle does; a first
    header that does not match the configuration is rejected when the frame
    size is read from it."""
    w, data, poly, offs, rng = case(3, 512, 60, seed=35)
    fps, payload = 200, 2048
    cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True)
    streams = []
    for d in data:
        fr = vdif.frames_from_station(d, payload, frames_per_second=fps)
        stray = fr[0].copy()
        vdif.set_header(stray, legacy=1)
        streams.append(np.concatenate([stray[None], fr]).reshape(-1))
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True)
    k.setInputData(data)
    k.setDelays(poly, offs)
    k.process()
    plain = k.visibilities()
    devs = [torch.from_numpy(s).to("cuda") for s in streams]
    for hb in (32, 0):
        for a, dv in enumerate(devs):
            got = k.setInputVDIF(a, dv.data_ptr(), dv.numel(), frames_per_second=fps,
                                 frame_bytes=payload + 32, header_bytes=hb)
            _, ok, _, want = fxref.vdif_place(cfg, streams[a], frames_per_second=fps, header_bytes=hb)
            assert got == want, (hb, a)
            np.testing.assert_array_equal(k.frame_flags(a), ok)
        if hb == 32:
            k.process()
            np.testing.assert_array_equal(k.visibilities(), plain)
    with pytest.raises(gcorr_amd.GxcError):
        k.setInputVDIF(0, devs[0].data_ptr(), devs[0].numel(), frames_per_second=fps, header_bytes=24)
    bad = streams[0].copy()
    vdif.set_header(bad[:payload + 32], bits=8)
    dbad = torch.from_numpy(bad).to("cuda")
    with pytest.raises(gcorr_amd.GxcError):  # frame size and geometry from a malformed first header
        k.setInputVDIF(0, dbad.data_ptr(), dbad.numel(), frames_per_second=fps)
    got = k.setInputVDIF(0, dbad.data_ptr(), dbad.numel(), frames_per_second=fps, frame_bytes=payload + 32,
                         header_bytes=32)
    assert got["malformed"] == 1 and got["missing"] == 0
    torch.cuda.synchronize()
    k.close()
