"""Parity of the HIP path (through the C-ABI) with the oracle.

FP32 visibilities must match the fxkernel restatement within 1e-5
normalised error: eps = max over (row, product) of
||V_gpu - V_ref||_2 / ||V_ref||_2 over channels.  Integer work (validity,
row counts, the integer X known-answer test) must match exactly.
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


def random_case(nant, nchan, numffts, seed=1, lo=1.65e9, bw=64e6, is_complex=False):
    w = synth.Workload("t", nant=nant, nchan=nchan, numffts=numffts, bw=bw, lo=lo,
                       is_complex=is_complex, seed=seed)
    rng = np.random.default_rng(seed)
    data = [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(nant)]
    poly, offs = synth.make_delays(w)
    return w, data, poly, offs


def run_gpu(w, data, poly, offs, autos=True, **kw):
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, w.is_complex,
                           autos=autos, **kw)
    k.setInputData(data)
    k.setDelays(poly, offs)
    k.process()
    v, c = k.visibilities(), k.row_counts()
    k.close()
    return v, c


def run_oracle(w, data, poly, offs, autos=True, mode="f32", quirk=True):
    cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, nbit=w.nbit,
                            is_complex=w.is_complex, autos=autos, quirk=quirk)
    return fxref.process(cfg, data, poly, offs, mode)


# ------------------------------------------------------------------ F-stage

@pytest.mark.parametrize("nchan,is_complex", [(32, False), (64, False), (128, False), (256, False),
                                              (512, False), (1024, False), (2048, False),
                                              (4096, False), (256, True), (1024, True),
                                              (8192, True), (8192, False), (16384, False),
                                              (16384, True), (32768, True)])
def test_fstage_matches_oracle(gpu, nchan, is_complex):
    check_fstage_vs_oracle(nchan, is_complex)


@pytest.mark.parametrize("nchan,is_complex", [(1024, False), (2048, False), (4096, False),
                                              (2048, True), (4096, True), (8192, True)])
def test_register_fstage_matches_oracle_with_invalid_ffts(gpu, nchan, is_complex):
    """The register kernel (fblock = L/32) at 2048..8192 points, where for
    real input one workgroup takes both polarisations of a station and FFT
    (complex input: one each): spectra per polarisation against the oracle,
    and invalid FFTs zero in both rows."""
    L = nchan if is_complex else 2 * nchan
    check_fstage_vs_oracle(nchan, is_complex, fblock=L // 32)


def check_fstage_vs_oracle(nchan, is_complex, **kw):
    w, data, poly, offs = random_case(3, nchan, 40, seed=nchan + is_complex, is_complex=is_complex,
                                      bw=128e6 if is_complex else 64e6)
    # station 1: large net delay so FFT 0 is out of the buffer; station 2:
    # negative net delay (odd sample offsets, last FFT out of the buffer).
    Ts = w.sampletime
    offs = offs.copy()
    offs[1] = poly[1][2] - 5.3 * Ts
    offs[2] = poly[2][2] + 3.6 * Ts
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, w.is_complex, **kw)
    k.setInputData(data)
    k.setDelays(poly, offs)
    spec, valid = k.fstage(0, w.numffts)
    k.close()
    cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, nbit=w.nbit,
                            is_complex=w.is_complex)
    for a in range(w.nant):
        for t in range(w.numffts):
            ref, ok = fxref.station_spectrum(cfg, data[a], poly[a], offs[a], t)
            assert bool(valid[a, t]) == ok, (a, t)
            got = spec[t, 2 * a:2 * a + 2]
            if not ok:
                assert not np.any(got)
                continue
            rms = np.sqrt(np.mean(np.abs(ref) ** 2))
            assert np.max(np.abs(got - ref)) / rms < 2e-5, (a, t)
    assert not valid[1, 0] and not valid[2, w.numffts - 1]


# ------------------------------------------------------------------ whole subintegration

def test_process_c1_full_size(gpu):
    """4 stations x 2 pols, 128 channels, 12 500 FFTs (0.1 s at 32 MS/s)."""
    w = synth.WORKLOADS["c1"]
    rng = np.random.default_rng(11)
    data = [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(w.nant)]
    poly, offs = synth.make_delays(w)
    v, c = run_gpu(w, data, poly, offs)
    ref, rc = run_oracle(w, data, poly, offs)
    np.testing.assert_array_equal(c, rc)
    assert eps(v, ref) < EPS


def test_process_c2_shape_reduced(gpu):
    """6 stations x 2 pols, 1024 channels, 2 000 FFTs of the 62 500."""
    w = synth.WORKLOADS["c2"].scaled(2000)
    rng = np.random.default_rng(12)
    data = [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(w.nant)]
    poly, offs = synth.make_delays(w)
    v, c = run_gpu(w, data, poly, offs)
    ref, rc = run_oracle(w, data, poly, offs)
    np.testing.assert_array_equal(c, rc)
    assert eps(v, ref) < EPS
    # the GPU is at least as close to the FP64 oracle as the FP32 restatement is
    r64, _ = run_oracle(w, data, poly, offs, mode="f64")
    assert eps(v, r64) < EPS
    assert eps(v, r64) <= 1.5 * eps(ref, r64) + 1e-7


@pytest.mark.parametrize("nant,nchan,numffts", [(1, 256, 50), (2, 64, 300), (5, 512, 97),
                                                (9, 128, 64), (16, 256, 24), (3, 1024, 1),
                                                (4, 32, 2), (32, 256, 12)])
def test_process_shapes(gpu, nant, nchan, numffts):
    w, data, poly, offs = random_case(nant, nchan, numffts, seed=nant * 100 + nchan)
    v, c = run_gpu(w, data, poly, offs)
    ref, rc = run_oracle(w, data, poly, offs)
    np.testing.assert_array_equal(c, rc)
    assert eps(v, ref) < EPS


def test_process_invalid_ffts_and_zero_count_rows(gpu):
    w, data, poly, offs = random_case(4, 256, 60, seed=3)
    Ts = w.sampletime
    offs = offs.copy()
    offs[0] = poly[0][2] - 7.4 * Ts     # first FFT invalid
    offs[1] = poly[1][2] + 9.6 * Ts     # last FFT invalid
    offs[3] = poly[3][2] + 1.0          # a second of offset: never valid
    v, c = run_gpu(w, data, poly, offs)
    ref, rc = run_oracle(w, data, poly, offs)
    np.testing.assert_array_equal(c, rc)
    assert rc[0] == w.numffts - 2          # (0,1): both edges lost
    assert rc[2] == 0 and not np.any(v[2])  # (0,3)
    assert eps(v, ref) < EPS


def test_process_exact_phase_law(gpu):
    w, data, poly, offs = random_case(3, 1024, 80, seed=8)
    v, c = run_gpu(w, data, poly, offs, exact_phase=True)
    ref, rc = run_oracle(w, data, poly, offs, quirk=False)
    np.testing.assert_array_equal(c, rc)
    assert eps(v, ref) < EPS
    stride_law, _ = run_oracle(w, data, poly, offs, quirk=True)
    assert eps(stride_law, ref) > 10 * EPS  # the two laws are distinguishable here


def test_process_fractional_lo(gpu):
    w, data, poly, offs = random_case(3, 512, 50, seed=9, lo=1.65e9 + 0.375)
    poly = poly.copy()
    offs = offs.copy()
    poly[:, 2] += 1.0   # whole seconds of delay exercise the (lo - int lo) term
    offs += 1.0
    v, c = run_gpu(w, data, poly, offs)
    ref, rc = run_oracle(w, data, poly, offs)
    np.testing.assert_array_equal(c, rc)
    assert eps(v, ref) < EPS


@pytest.mark.parametrize("nchan", [256, 1024])
def test_process_complex8(gpu, nchan):
    w, data, poly, offs = random_case(4, nchan, 60, seed=21, bw=128e6, is_complex=True)
    v, c = run_gpu(w, data, poly, offs)
    ref, rc = run_oracle(w, data, poly, offs)
    np.testing.assert_array_equal(c, rc)
    assert eps(v, ref) < EPS


def test_chunked_equals_single_pass(gpu):
    w, data, poly, offs = random_case(5, 256, 301, seed=4)
    v1, c1 = run_gpu(w, data, poly, offs)
    v2, c2 = run_gpu(w, data, poly, offs, chunk_ffts=37)
    np.testing.assert_array_equal(c1, c2)
    assert eps(v2, v1) < 1e-6
    ref, _ = run_oracle(w, data, poly, offs)
    assert eps(v2, ref) < EPS


def test_process_is_deterministic_and_reusable(gpu):
    w, data, poly, offs = random_case(4, 512, 120, seed=5)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True)
    k.setInputData(data)
    k.setDelays(poly, offs)
    k.process()
    a = k.visibilities()
    k.process()
    b = k.visibilities()
    np.testing.assert_array_equal(a, b)
    # new delays on the same handle
    offs2 = offs + 2.0 * w.sampletime
    k.setDelays(poly, offs2)
    k.process()
    ref, _ = run_oracle(w, data, poly, offs2)
    assert eps(k.visibilities(), ref) < EPS
    k.close()


# ------------------------------------------------------------------ X-engine KAT

@pytest.mark.parametrize("nant,nchan,nfft", [(3, 256, 4), (5, 1024, 100), (3, 512, 3000),
                                             (16, 128, 40), (9, 64, 50), (32, 64, 40),
                                             (10, 64, 33), (12, 32, 17), (13, 64, 50), (15, 64, 24)])
def test_xengine_integer_kat(gpu, nant, nchan, nfft):
    """gcorr/test/unit.cc:73-159: integer spectra in [-5, 5] make every sum
    exact, so the GPU result equals the exact sum / nfft bit for bit."""
    rng = np.random.default_rng(12335 + nfft)
    X = (rng.integers(-5, 6, (nfft, 2 * nant, nchan))
         + 1j * rng.integers(-5, 6, (nfft, 2 * nant, nchan))).astype(np.complex64)
    got = gcorr_amd.xengine(X, nant, autos=True)
    exact = fxref.xmac(X, nant, autos=True)  # sequential FP32, exact for these inputs
    Xd = X.astype(np.complex128)
    rows = [(j, k) for j in range(nant) for k in range(j + 1, nant)] + [(a, a) for a in range(nant)]
    for r, (j, k) in enumerate(rows):
        for l in range(2):
            for m in range(2):
                s = np.sum(Xd[:, 2 * j + l] * np.conj(Xd[:, 2 * k + m]), axis=0)
                np.testing.assert_array_equal(exact[r, 2 * l + m], s)
                expect = (s.real.astype(np.float32) / np.float32(nfft)) \
                    + 1j * (s.imag.astype(np.float32) / np.float32(nfft))
                np.testing.assert_array_equal(got[r, 2 * l + m], expect.astype(np.complex64))


@pytest.mark.parametrize("nant", [9, 10, 12, 13, 15, 16, 32])
def test_fp32_xtiles_agree_bitwise(gpu, nant):
    """At 17..32 inputs the FP32 X-stage stages single words in LDS under an
    XOR swizzle of input, channel pair and FFT row; at 64 inputs the
    8-channel tile loads the next stage during the MFMAs and the 4-channel
    tile does not.  A channel's sums run in the same order whatever the
    channel tile, so every tile must give the same visibilities bit for
    bit, and they match the oracle.  70 FFTs in 2 splits leave a partial
    last LDS stage in each split."""
    w, data, poly, offs = random_case(nant, 256, 70, seed=40 + nant)
    tiles = (4, 8) if nant > 16 else (4, 8, 16)
    out = {}
    for xt in tiles:
        out[xt], rc = run_gpu(w, data, poly, offs, xtile=xt, xsplits=2)
    ref, rc_ref = run_oracle(w, data, poly, offs)
    np.testing.assert_array_equal(rc, rc_ref)
    assert eps(out[tiles[-1]], ref) < EPS
    for xt in tiles[:-1]:
        np.testing.assert_array_equal(out[xt], out[tiles[-1]])


# ------------------------------------------------------------------ interface

def test_accumulate_and_save(gpu, tmp_path):
    w, data, poly, offs = random_case(3, 64, 30, seed=6)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw)
    k.setInputData(data)
    k.setDelays(poly, offs)
    k.process()
    v = k.visibilities()
    acc = np.ones((3, 4, 64), np.complex64)
    k.accumulate(acc)
    k.accumulate(acc)
    np.testing.assert_allclose(acc, 1 + 2 * v, rtol=1e-6, atol=1e-3)
    p = tmp_path / "vis.out"
    k.saveVisibilities(str(p))
    k.close()
    rows = [ln.split() for ln in p.read_text().splitlines()]
    assert len(rows) == 64 and all(len(r) == 2 + 3 * 4 * 4 for r in rows)
    for c in (0, 17, 63):
        r = rows[c]
        assert int(r[0]) == c
        assert float(r[1]) == pytest.approx((c + 0.5) / 64 * w.bw / 1e6, abs=1e-6)
        for b in range(3):
            for p_ in range(4):
                re, im, amp, ph = (float(x) for x in r[2 + 4 * (4 * b + p_):6 + 4 * (4 * b + p_)])
                z = v[b, p_, c]
                assert re == pytest.approx(z.real, abs=1e-5 + 1e-6 * abs(z))
                assert im == pytest.approx(z.imag, abs=1e-5 + 1e-6 * abs(z))
                assert amp == pytest.approx(abs(z), abs=1e-5 + 1e-6 * abs(z))


def test_errors_are_reported(gpu):
    import torch
    k = gcorr_amd.FxKernel(2, 64, 10, 2, 1.65e9, 16e6)
    with pytest.raises(gcorr_amd.GxcError):
        k.process()                    # no input, no delays
    with pytest.raises(ValueError):
        k.setInputData([np.zeros(10, np.uint8)] * 2)
    buf = torch.zeros(k.station_bytes + 8, dtype=torch.uint8, device="cuda:0")
    with pytest.raises(gcorr_amd.GxcError):  # device buffers must be 4-byte aligned
        k.setInputDevice([buf.data_ptr() + 1, buf.data_ptr()], keepalive=buf)
    with pytest.raises(gcorr_amd.GxcError):  # not an F-stage workgroup size
        gcorr_amd.FxKernel(2, 64, 10, 2, 1.65e9, 16e6, fblock=96)
    k.close()


# ------------------------------------------------------------------ streams and graphs

def test_overlap_and_graph_modes_agree(gpu):
    """Two-stream overlap and graph replay change scheduling only: with the
    same chunking the visibilities are identical bit for bit."""
    w, data, poly, offs = random_case(6, 512, 403, seed=8)
    ref, rc = run_oracle(w, data, poly, offs)
    out = {}
    for overlap in (False, True):
        for graph in (False, True):
            v, c = run_gpu(w, data, poly, offs, chunk_ffts=50, overlap=overlap, graph=graph)
            np.testing.assert_array_equal(c, rc)
            assert eps(v, ref) < EPS
            out[overlap, graph] = v
    for v in out.values():
        np.testing.assert_array_equal(v, out[False, False])


def test_graph_replays_match_eager(gpu):
    """A captured subintegration replayed with new delays and new input, with
    eager work on other handles between the replays, equals the eager run."""
    w, data, poly, offs = random_case(4, 256, 160, seed=9)
    rng = np.random.default_rng(10)
    g = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True,
                           chunk_ffts=40, graph=True)
    g.setInputData(data)
    for rep in range(4):
        offs_r = offs + rep * 0.7 * w.sampletime
        if rep == 2:
            data = [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(w.nant)]
            g.setInputData(data)
        g.setDelays(poly, offs_r)
        g.process()
        got = g.visibilities()
        eager, _ = run_gpu(w, data, poly, offs_r, chunk_ffts=40, graph=False)
        np.testing.assert_array_equal(got, eager)
        ref, _ = run_oracle(w, data, poly, offs_r)
        assert eps(got, ref) < EPS
    g.close()


def test_graph_with_device_input(gpu):
    """Station buffers handed over as device pointers, swapped between replays."""
    import torch
    w, data, poly, offs = random_case(3, 1024, 90, seed=11)
    rng = np.random.default_rng(12)
    data2 = [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(w.nant)]
    dev1 = [torch.from_numpy(a).to("cuda:0") for a in data]
    dev2 = [torch.from_numpy(a).to("cuda:0") for a in data2]
    torch.cuda.synchronize()
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, graph=True)
    k.setDelays(poly, offs)
    for dev, host in ((dev1, data), (dev2, data2), (dev1, data)):
        k.setInputDevice([t.data_ptr() for t in dev], keepalive=dev)
        k.process()
        ref, _ = run_oracle(w, host, poly, offs, autos=False)
        assert eps(k.visibilities(), ref) < EPS
    k.close()


# ------------------------------------------------------------------ FFT cross-check

@pytest.mark.parametrize("nchan,is_complex", [(256, False), (1024, False), (4096, False),
                                              (16384, False), (1024, True), (32768, True)])
def test_fstage_fft_matches_rocfft(gpu, nchan, is_complex):
    """With LO 0 and zero delays every rotator is exactly 1, so the F-stage
    output is the FFT of the unpacked samples: checked against rocFFT
    (torch.fft on the GPU), an FFT independent of both this build and the
    oracle."""
    import torch
    nant, nfft = 2, 6
    w = synth.Workload("t", nant=nant, nchan=nchan, numffts=nfft, bw=64e6, lo=0.0,
                       is_complex=is_complex)
    rng = np.random.default_rng(nchan)
    data = [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(nant)]
    k = gcorr_amd.FxKernel(nant, nchan, nfft, w.nbit, 0.0, w.bw, is_complex)
    k.setInputData(data)
    k.setDelays(np.zeros((nant, 3)), np.zeros(nant))
    spec, valid = k.fstage(0, nfft)
    k.close()
    assert valid.all()
    L = w.fftchannels
    for a in range(nant):
        b = torch.from_numpy(data[a]).to("cuda:0")
        if is_complex:
            v = b.view(torch.int8).float().reshape(-1, 2, 2)          # [t][pol][re, im]
            x = torch.complex(v[..., 0], v[..., 1]).T                 # [pol][t]
        else:
            lv = torch.tensor([-3.3359, -1.0, 1.0, 3.3359], device="cuda:0")
            codes = torch.stack([(b >> s) & 3 for s in (0, 2, 4, 6)], dim=1).long()  # A0 B0 A1 B1
            x = lv[codes].reshape(-1, 2, 2).permute(2, 0, 1).reshape(2, -1)   # [pol][t]
            x = x.to(torch.complex64)
        x = x[:, :nfft * L].reshape(2, nfft, L)
        ref = torch.fft.fft(x, dim=-1)[..., :nchan].permute(1, 0, 2).cpu().numpy()  # [t][pol][c]
        got = spec[:, 2 * a:2 * a + 2]
        scale = np.sqrt(np.mean(np.abs(ref) ** 2))
        assert np.max(np.abs(got - ref)) / scale < 2e-5, a


# ------------------------------------------------------------------ autotuner

@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_autotune_keeps_parity_and_reports_choice(gpu, precision):
    """The autotuner only changes launch shapes: after it, results still
    match the oracle (FP32), or the untuned handle within FP16 resolution
    (another F-stage kernel rounds the FP32 spectra differently, which moves
    some FP16 roundings by one unit: 2^-11 relative on single products).
    When it picks the matrix-core F-stage (fblock -1), whose FFT passes
    round to f16, the FP16 result is held to the FP16 bound of the FP64
    oracle instead."""
    w, data, poly, offs = random_case(6, 1024, 400, seed=13)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True,
                           precision=precision)
    k.setInputData(data)
    k.setDelays(poly, offs)
    k.process()
    before = k.visibilities()
    t = k.autotune()
    assert t["candidates"] >= 5 and (t["fblock"] > 0 or t["fblock"] == -1)
    assert t["xtile"] > 0 and t["xsplits"] >= 1
    assert t["process_ms"] > 0 and t["default_ms"] > 0
    k.process()
    after = k.visibilities()
    if precision == "fp32":
        ref, _ = run_oracle(w, data, poly, offs)
        assert eps(after, ref) < EPS
    if precision == "fp16" and t["fblock"] == -1:
        ref64, _ = run_oracle(w, data, poly, offs, mode="f64")
        assert eps(after, ref64) < 1e-3
    else:
        assert eps(after, before) < (1e-5 if precision == "fp32" else 1e-4)
    # the choice handed to a new handle through the config gives the same result
    k2 = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True,
                            precision=precision, fblock=t["fblock"], xtile=t["xtile"],
                            xsplits=t["xsplits"], graph=bool(t["graph"]))
    k2.setInputData(data)
    k2.setDelays(poly, offs)
    k2.process()
    np.testing.assert_array_equal(k2.visibilities(), after)
    k.close()
    k2.close()


@pytest.mark.parametrize("fblock_mult,xtile_mult", [(0.5, 1), (2, 1), (1, 0.5), (1, 2)])
def test_launch_shape_variants_match_oracle(gpu, fblock_mult, xtile_mult):
    """Every F workgroup size and X channel tile the autotuner can pick."""
    w, data, poly, offs = random_case(5, 512, 150, seed=14)
    fb = int(256 * fblock_mult)        # default for L = 1024 is 256 threads
    xt = int(16 * xtile_mult)          # default FP32 tile at M <= 16 is 16 channels
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True,
                           fblock=fb, xtile=xt)
    k.setInputData(data)
    k.setDelays(poly, offs)
    k.process()
    ref, rc = run_oracle(w, data, poly, offs)
    np.testing.assert_array_equal(k.row_counts(), rc)
    assert eps(k.visibilities(), ref) < EPS
    k.close()


@pytest.mark.parametrize("nchan,is_complex", [(1024, False), (2048, False), (4096, False),
                                              (2048, True), (4096, True), (8192, True)])
def test_long_fft_kernel_at_short_lengths(gpu, nchan, is_complex):
    """At FFT lengths 2048..8192 the register kernel (fblock = L/32; both
    polarisations of a station and FFT per workgroup for real input) is an
    alternative the autotuner can choose."""
    w, data, poly, offs = random_case(3, nchan, 40, seed=15 + nchan, is_complex=is_complex,
                                      bw=128e6 if is_complex else 64e6)
    L = w.fftchannels
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, w.is_complex,
                           autos=True, fblock=L // 32)
    k.setInputData(data)
    k.setDelays(poly, offs)
    k.process()
    ref, rc = run_oracle(w, data, poly, offs)
    np.testing.assert_array_equal(k.row_counts(), rc)
    assert eps(k.visibilities(), ref) < EPS
    k.close()
