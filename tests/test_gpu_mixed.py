"""Mixed-precision path: FP16 / BF16 / FP8 e4m3 spectra with the MFMA X-engine.

Exactness where the arithmetic allows it (integer spectra in [-5, 5] are
representable in all three formats and their products sum exactly in FP32),
and errors against the FP64 oracle on the generated signal otherwise.
"""
import numpy as np
import pytest

import gcorr_amd
from gcorr_amd import metrics, synth
from oracle import fxref

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp16", "bf16", "fp8"])
@pytest.mark.parametrize("nant,nchan,nfft", [(3, 64, 100), (8, 32, 300), (12, 32, 37),
                                             (16, 16, 64), (32, 8, 40)])
def test_mfma_xengine_integer_kat(gpu, precision, nant, nchan, nfft):
    """gcorr/test/unit.cc:73-159 pattern through the MFMA X-engine: 1, 2 and 4
    input blocks of 16, a tail of FFTs not a multiple of 32."""
    rng = np.random.default_rng(nant * 1000 + nfft)
    X = (rng.integers(-5, 6, (nfft, 2 * nant, nchan))
         + 1j * rng.integers(-5, 6, (nfft, 2 * nant, nchan))).astype(np.complex64)
    got = gcorr_amd.xengine(X, nant, autos=True, precision=precision)
    Xd = X.astype(np.complex128)
    rows = [(j, k) for j in range(nant) for k in range(j + 1, nant)] + [(a, a) for a in range(nant)]
    for r, (j, k) in enumerate(rows):
        for l in range(2):
            for m in range(2):
                s = np.sum(Xd[:, 2 * j + l] * np.conj(Xd[:, 2 * k + m]), axis=0)
                expect = (s.real.astype(np.float32) / np.float32(nfft)) \
                    + 1j * (s.imag.astype(np.float32) / np.float32(nfft))
                np.testing.assert_array_equal(got[r, 2 * l + m], expect.astype(np.complex64),
                                              err_msg="row %d prod %d" % (r, 2 * l + m))


_CACHE = {}


def _case(w):
    """Generated input and the FP64 oracle's visibilities, computed once per workload."""
    if w not in _CACHE:
        bufs, poly, offs = synth.generate(w, device="cpu")
        data = [b.numpy() for b in bufs]
        cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, nbit=w.nbit,
                                is_complex=w.is_complex, autos=True)
        ref, rc = fxref.process(cfg, data, poly, offs, "f64")
        _CACHE[w] = (data, poly, offs, ref, rc)
    return _CACHE[w]


def _mixed_vs_oracle(w, precision, **shape):
    data, poly, offs, ref, rc = _case(w)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, w.is_complex,
                           precision=precision, autos=True, **shape)
    k.setInputData(data)
    k.setDelays(poly, offs)
    k.process()
    v, c = k.visibilities(), k.row_counts()
    k.close()
    np.testing.assert_array_equal(c, rc)
    return v, ref


# bound on eps vs the FP64 oracle: FP16 is the 1e-3 target; BF16 has 3 fewer
# mantissa bits, e4m3 7 fewer.
BOUND = {"fp16": 1e-3, "bf16": 8e-3, "fp8": 0.12}


@pytest.mark.parametrize("precision", ["fp16", "bf16", "fp8"])
def test_mixed_process_vs_fp64_oracle(gpu, precision):
    """16 stations x 2 pols, 4096 channels with tones (the mixed-precision
    workload's shape), 600 of its 31 250 FFTs."""
    w = synth.WORKLOADS["c3"].scaled(600)
    v, ref = _mixed_vs_oracle(w, precision)
    e = metrics.eps(v, ref)
    errs = metrics.amp_phase_errors(v, ref)
    print(precision, "eps", e, errs)
    assert e < BOUND[precision]


@pytest.mark.parametrize("precision", ["fp16", "bf16", "fp8"])
@pytest.mark.parametrize("fblock", [-1, 256, 0])
def test_c3_shape_every_fstage_vs_fp64_oracle(gpu, precision, fblock):
    """The 16-station, 4096-channel tones workload's shape (600 of its 31 250
    FFTs) through each F-stage the autotuner chooses between at 8192 points:
    the matrix-core kernel (-1), the register kernel (256 threads)
    and the two-polarisation item kernel (default), against the FP64 oracle."""
    w = synth.WORKLOADS["c3"].scaled(600)
    v, ref = _mixed_vs_oracle(w, precision, fblock=fblock)
    e = metrics.eps(v, ref)
    print(precision, "fblock", fblock, "eps", e, metrics.amp_phase_errors(v, ref))
    assert e < BOUND[precision]


@pytest.mark.parametrize("precision", ["fp16", "fp8"])
def test_c3_shape_matrix_core_fstage_every_xtile(gpu, precision):
    """The matrix-core F-stage with each X-stage channel tile the library
    offers at 32 inputs (the autotuner's candidates), 600 FFTs at the
    16-station shape, against the FP64 oracle."""
    w = synth.WORKLOADS["c3"].scaled(600)
    ran = []
    for xtile in (8, 16, 32, 64):
        for xsplits in (0, 3):
            try:
                v, ref = _mixed_vs_oracle(w, precision, fblock=-1, xtile=xtile, xsplits=xsplits)
            except gcorr_amd.GxcError:
                continue  # not a tile of this shape
            e = metrics.eps(v, ref)
            print(precision, "xtile", xtile, "xsplits", xsplits, "eps", e)
            assert e < BOUND[precision], (xtile, xsplits)
            ran.append(xtile)
    assert len(set(ran)) >= 2, ran


def test_mixed_fp16_small_shapes(gpu):
    for key, n in (("c1", 2000), ("c2", 300)):
        w = synth.WORKLOADS[key].scaled(n)
        v, ref = _mixed_vs_oracle(w, "fp16")
        assert metrics.eps(v, ref) < 1e-3, key


def test_mixed_chunked_matches_single_pass(gpu):
    w = synth.WORKLOADS["c2"].scaled(400)
    bufs, poly, offs = synth.generate(w, device="cpu")
    data = [b.numpy() for b in bufs]
    out = []
    for chunk in (0, 97):
        k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, precision="fp16",
                               autos=True, chunk_ffts=chunk)
        k.setInputData(data)
        k.setDelays(poly, offs)
        k.process()
        out.append(k.visibilities())
        k.close()
    assert metrics.eps(out[1], out[0]) < 1e-5


@pytest.mark.parametrize("key,nfft,chunk", [("c5", 24, 3), ("c2", 2000, 16)])
def test_chunked_fp32_compensated_vs_fp64_oracle(gpu, key, nfft, chunk):
    """Subintegrations in many chunks (8 passes at the 32-station complex
    shape, 125 at the 6-station one): each later chunk adds into the split
    partials with Neumaier compensation, so the error against the FP64
    oracle stays that of one pass (fxkernel.cpp:327-367 sums every FFT into
    one FP32 accumulator; gcorr splits the FFT axis, gxkernel.cu:412-424)."""
    w = synth.WORKLOADS[key].scaled(nfft)
    v1, ref = _mixed_vs_oracle(w, "fp32")
    vc, _ = _mixed_vs_oracle(w, "fp32", chunk_ffts=chunk)
    e1, ec = metrics.eps(v1, ref), metrics.eps(vc, ref)
    print(key, "one pass eps", e1, "chunked eps", ec)
    assert ec < 1e-5
    assert ec <= 1.25 * e1 + 1e-8


def test_mixed_is_deterministic(gpu):
    w = synth.WORKLOADS["c1"].scaled(300)
    data, poly, offs, _, _ = _case(w)
    out = []
    for _ in range(2):
        k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, precision="fp16",
                               autos=True)
        k.setInputData(data)
        k.setDelays(poly, offs)
        k.process()
        out.append(k.visibilities())
        k.close()
    np.testing.assert_array_equal(out[0], out[1])


@pytest.mark.parametrize("precision", ["fp32", "fp16", "fp8"])
def test_c5_shape_vs_fp64_oracle(gpu, precision):
    """32 stations x 2 pols of complex int8, 32768-point FFTs (the largest
    workload's shape: 64 inputs, 4 blocks of 16 in the X-engine), 24 of its
    15 625 FFTs; FP8 e4m3 against the FP16 and FP32 forms and the FP64 oracle."""
    w = synth.WORKLOADS["c5"].scaled(24)
    v, ref = _mixed_vs_oracle(w, precision)
    e = metrics.eps(v, ref)
    print(precision, "eps", e, metrics.amp_phase_errors(v, ref))
    assert e < {"fp32": 1e-5, "fp16": 1e-3, "fp8": BOUND["fp8"]}[precision]


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_c5_shape_every_xtile(gpu, precision):
    """64 inputs (4 blocks of 16, 10 block tiles per channel): every X-stage
    channel tile the library offers there (FP32: 4 channels; FP16: 4 and 8),
    with one and with several FFT splits, against the FP64 oracle (24 FFTs
    of the 32-station shape)."""
    w = synth.WORKLOADS["c5"].scaled(24)
    ran = []
    for xtile in (2, 4, 8, 16):
        for xsplits in (1, 3):
            try:
                v, ref = _mixed_vs_oracle(w, precision, xtile=xtile, xsplits=xsplits)
            except gcorr_amd.GxcError:
                continue  # not a tile of this shape
            e = metrics.eps(v, ref)
            print(precision, "xtile", xtile, "xsplits", xsplits, "eps", e)
            assert e < {"fp32": 1e-5, "fp16": 1e-3}[precision], (xtile, xsplits)
            ran.append(xtile)
    assert len(set(ran)) >= (1 if precision == "fp32" else 2), ran
