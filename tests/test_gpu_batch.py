"""Multi-subband batch: every subband of a subintegration in one launch per
stage, against the oracle per subband and against per-subband handles.

Subbands of the batched workload (8 stations; 16..128 MHz, 256..4096
channels, FFT lengths 512..8192), each with its own seeded delays, at
1/250 of their 0.5 s FFT counts.
"""
import numpy as np
import pytest

import gcorr_amd
from gcorr_amd import metrics, synth
from oracle import fxref

pytestmark = pytest.mark.gpu

EPS = 1e-5


def subbands(div=250):
    return [w.scaled(max(8, w.numffts // div)) for w in synth.c4_subbands()]


def eps(v, ref):
    return metrics.eps(v, ref)


def random_inputs(ws, seed=3):
    rng = np.random.default_rng(seed)
    return [[rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(w.nant)] for w in ws]


def oracle(w, data, poly, offs, mode="f32"):
    cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, nbit=2, autos=True)
    return fxref.process(cfg, data, poly, offs, mode)


def run_batch(ws, datas, delays, **kw):
    b = gcorr_amd.Batch(ws[0].nant, [(w.nchan, w.numffts, w.lo, w.bw) for w in ws], autos=True, **kw)
    for s, (data, (poly, offs)) in enumerate(zip(datas, delays)):
        b.set_input(s, data)
        b.set_delays(s, poly, offs)
    b.process()
    out = [(b.visibilities(s), b.row_counts(s)) for s in range(len(ws))]
    b.close()
    return out


def test_batch_fp32_matches_oracle_and_single_handles(gpu):
    ws = subbands()
    datas = random_inputs(ws)
    delays = [synth.make_delays(w) for w in ws]
    got = run_batch(ws, datas, delays)
    for s, w in enumerate(ws):
        v, c = got[s]
        ref, rc = oracle(w, datas[s], *delays[s])
        np.testing.assert_array_equal(c, rc)
        assert eps(v, ref) < EPS, s
        k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, 2, w.lo, w.bw, autos=True)
        k.setInputData(datas[s])
        k.setDelays(*delays[s])
        k.process()
        assert eps(v, k.visibilities()) < 1e-6, s
        k.close()


def test_batch_fp16_vs_fp64_oracle(gpu):
    ws = subbands(div=125)
    gen = [synth.generate(w, device="cpu") for w in ws]
    datas = [[b.numpy() for b in g[0]] for g in gen]
    delays = [(g[1], g[2]) for g in gen]
    got = run_batch(ws, datas, delays, precision="fp16")
    for s, w in enumerate(ws):
        v, c = got[s]
        ref, rc = oracle(w, datas[s], *delays[s], mode="f64")
        np.testing.assert_array_equal(c, rc)
        e = eps(v, ref)
        print(w.name, "fp16 eps", e)
        assert e < 1e-3, s


@pytest.mark.parametrize("precision,bound,mode,own", [("fp32", EPS, "f32", 1), ("fp16", 1e-3, "f64", 1),
                                                      ("fp32", EPS, "f32", 2), ("fp16", 1e-3, "f64", 2)])
def test_batch_matrix_core_fstage(gpu, precision, bound, mode, own):
    """fstage_mfma 1: the 2048..8192-point subbands get a launch each, on the
    matrix-core F-stage (2048..4096 points for FP32) or else the register
    F-stage; 2: a launch each on the register F-stage; the others stay in
    the multi launch."""
    ws = subbands(div=125)
    gen = [synth.generate(w, device="cpu") for w in ws]
    datas = [[b.numpy() for b in g[0]] for g in gen]
    delays = [(g[1], g[2]) for g in gen]
    got = run_batch(ws, datas, delays, precision=precision, fstage_mfma=own, chunks=2)
    for s, w in enumerate(ws):
        v, c = got[s]
        ref, rc = oracle(w, datas[s], *delays[s], mode=mode)
        np.testing.assert_array_equal(c, rc)
        e = eps(v, ref)
        print(w.name, precision, "matrix-core batch eps", e)
        assert e < bound, s


def test_batch_matrix_core_modes_agree(gpu):
    ws = subbands(div=500)
    datas = random_inputs(ws, seed=7)
    delays = [synth.make_delays(w) for w in ws]
    base = None
    for overlap in (False, True):
        for graph in (False, True):
            got = run_batch(ws, datas, delays, precision="fp16", fstage_mfma=True, chunks=4, overlap=overlap,
                            graph=graph)
            if base is None:
                base = got
            for (v, c), (v0, c0) in zip(got, base):
                np.testing.assert_array_equal(c, c0)
                np.testing.assert_array_equal(v, v0)


@pytest.mark.parametrize("graph", [True, False])
def test_batch_pol_conversion(gpu, graph):
    """gxc_batch_set_pol_conversion on some subbands: those match the
    oracle's per-spectrum conversion (gpubench.cu:141-166) within the FP32
    bound, the others stay unconverted; turning it off restores them."""
    ws = subbands(div=500)
    datas = random_inputs(ws, seed=8)
    delays = [synth.make_delays(w) for w in ws]
    rng = np.random.default_rng(8)
    gains = {s: ((rng.standard_normal((w.nant, 2, w.nchan)) + 1j * rng.standard_normal((w.nant, 2, w.nchan)))
                 / np.sqrt(2)).astype(np.complex64) for s, w in enumerate(ws) if s % 3 == 1}
    b = gcorr_amd.Batch(8, [(w.nchan, w.numffts, w.lo, w.bw) for w in ws], autos=True, graph=graph)
    for s in range(len(ws)):
        b.set_input(s, datas[s])
        b.set_delays(s, *delays[s])
    b.process()
    plain = [b.visibilities(s) for s in range(len(ws))]
    for s, g in gains.items():
        b.set_pol_conversion(s, g)
    b.process()
    for s, w in enumerate(ws):
        v = b.visibilities(s)
        if s in gains:
            cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, nbit=2, autos=True, polgains=gains[s])
            ref, rc = fxref.process(cfg, datas[s], *delays[s])
            np.testing.assert_array_equal(b.row_counts(s), rc)
            assert eps(v, ref) < EPS, s
        else:
            np.testing.assert_array_equal(v, plain[s])
    for s in gains:
        b.set_pol_conversion(s, None)
    b.process()
    for s in range(len(ws)):
        np.testing.assert_array_equal(b.visibilities(s), plain[s])
    b.close()


@pytest.mark.parametrize("graph", [True, False])
def test_batch_vdif_input(gpu, graph):
    """gxc_batch_set_input_vdif on two subbands: their frame flags equal the
    oracle's placement and their visibilities the oracle's run on the placed
    buffers with the missing-frame flags (FP32 bound); the other subbands are
    unchanged; set_input leaves VDIF mode again."""
    import torch
    from gcorr_amd import vdif

    ws = subbands(div=500)
    datas = random_inputs(ws, seed=9)
    delays = [synth.make_delays(w) for w in ws]
    b = gcorr_amd.Batch(8, [(w.nchan, w.numffts, w.lo, w.bw) for w in ws], autos=True, graph=graph)
    for s in range(len(ws)):
        b.set_input(s, datas[s])
        b.set_delays(s, *delays[s])
    b.process()
    plain = [b.visibilities(s) for s in range(len(ws))]
    fps, s0, f0 = 1000, 7, 3
    for s in (2, 5):
        w = ws[s]
        payload = 1024
        cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, nbit=2, autos=True)
        placed, flags, spfs = [], [], []
        for a in range(w.nant):
            fr = vdif.frames_from_station(datas[s][a], payload, frames_per_second=fps, ref_second=s0,
                                          ref_frame=f0, thread=s)
            keep = [k for k in range(fr.shape[0]) if not (a == 1 and k in (2, 3))]
            stream = fr[keep].reshape(-1)
            buf, ok, spf, _ = fxref.vdif_place(cfg, stream, frames_per_second=fps, ref_second=s0, ref_frame=f0,
                                               thread_id=s)
            placed.append(buf)
            flags.append(ok)
            spfs.append(spf)
            dev = torch.from_numpy(stream).to("cuda")
            b.set_input_vdif(s, a, dev.data_ptr(), stream.size, frames_per_second=fps, ref_second=s0,
                             ref_frame=f0, thread_id=s, keepalive=dev)
        torch.cuda.synchronize()
        for a in range(w.nant):
            np.testing.assert_array_equal(b.frame_flags(s, a), flags[a])
        ws_ref = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, nbit=2, autos=True, frame_ok=flags,
                                   frame_samples=spfs)
        delays_s = delays[s]
        b.process()
        ref, rc = fxref.process(ws_ref, placed, *delays_s)
        np.testing.assert_array_equal(b.row_counts(s), rc)
        assert b.row_counts(s).min() < w.numffts
        assert eps(b.visibilities(s), ref) < EPS, s
    for s in range(len(ws)):
        if s not in (2, 5):
            np.testing.assert_array_equal(b.visibilities(s), plain[s])
    for s in (2, 5):
        b.set_input(s, datas[s])
    b.process()
    for s in range(len(ws)):
        np.testing.assert_array_equal(b.visibilities(s), plain[s])
    b.close()


def test_batch_modes_agree(gpu):
    """Overlap and graph replay change scheduling only: identical results."""
    ws = subbands(div=500)
    datas = random_inputs(ws, seed=4)
    delays = [synth.make_delays(w) for w in ws]
    base = None
    for overlap in (False, True):
        for graph in (False, True):
            got = run_batch(ws, datas, delays, chunks=4, overlap=overlap, graph=graph)
            if base is None:
                base = got
            for (v, c), (v0, c0) in zip(got, base):
                np.testing.assert_array_equal(c, c0)
                np.testing.assert_array_equal(v, v0)


def test_batch_graph_replays_with_new_delays_and_device_input(gpu):
    import torch
    ws = subbands(div=500)
    datas = random_inputs(ws, seed=5)
    delays = [synth.make_delays(w) for w in ws]
    dev = [[torch.from_numpy(a).to("cuda:0") for a in d] for d in datas]
    torch.cuda.synchronize()
    b = gcorr_amd.Batch(8, [(w.nchan, w.numffts, w.lo, w.bw) for w in ws], autos=True, graph=True)
    for s in range(len(ws)):
        if s % 2:
            b.set_input_device(s, [t.data_ptr() for t in dev[s]], keepalive=dev[s])
        else:
            b.set_input(s, datas[s])
    for rep in range(3):
        shifted = [(p, o + rep * 1.3 * w.sampletime) for (p, o), w in zip(delays, ws)]
        for s in range(len(ws)):
            b.set_delays(s, *shifted[s])
        b.process()
        eager = run_batch(ws, datas, shifted, graph=False)
        for s in range(len(ws)):
            np.testing.assert_array_equal(b.visibilities(s), eager[s][0])
            np.testing.assert_array_equal(b.row_counts(s), eager[s][1])
    b.close()


def test_batch_autotune_keeps_results(gpu):
    ws = subbands(div=500)
    datas = random_inputs(ws, seed=6)
    delays = [synth.make_delays(w) for w in ws]
    b = gcorr_amd.Batch(8, [(w.nchan, w.numffts, w.lo, w.bw) for w in ws], autos=True)
    for s in range(len(ws)):
        b.set_input(s, datas[s])
        b.set_delays(s, *delays[s])
    t = b.autotune()
    assert t["candidates"] >= 5 and t["fblock"] in (256, 512, -1) and t["process_ms"] > 0
    b.process()
    for s, w in enumerate(ws):
        ref, rc = oracle(w, datas[s], *delays[s])
        np.testing.assert_array_equal(b.row_counts(s), rc)
        assert eps(b.visibilities(s), ref) < EPS
    b.close()
