"""The matrix-core F-stage (gxc_config.fblock = GXC_FBLOCK_MFMA).

Its FFT passes run as f16 MFMA products with FP32 accumulation.  For the
reduced-precision forms the data are rounded to f16 between passes, so the
visibilities must stay within the reduced-precision bound of the FP64 oracle
(1e-3 for FP16, as the FP32-F-stage form).  The FP32 form carries every
operand as an f16 pair plus its remainder (22 significant bits) and must meet
the FP32 bound, 1e-5.  Every supported length is covered, with the row
counts exact.
"""
import numpy as np
import pytest

import gcorr_amd
from gcorr_amd import synth
from oracle import fxref

pytestmark = pytest.mark.gpu

MFMA = -1


def eps(v, ref):
    num = np.linalg.norm(v - ref, axis=-1)
    den = np.linalg.norm(ref, axis=-1)
    mask = den > 0
    assert np.all(num[~mask] == 0)
    return float((num[mask] / den[mask]).max()) if mask.any() else 0.0


def case(nant, nchan, numffts, seed, rho=0.02):
    w = synth.Workload("m", nant=nant, nchan=nchan, numffts=numffts, bw=64e6, lo=1.65e9, rho=rho, seed=seed)
    data, poly, offs = synth.generate(w)
    return w, [np.asarray(d) for d in data], poly, offs


def run(w, data, poly, offs, **kw):
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True, **kw)
    k.setInputData(data)
    k.setDelays(poly, offs)
    k.process()
    v, c = k.visibilities(), k.row_counts()
    k.close()
    return v, c


@pytest.mark.parametrize("nchan,numffts", [(1024, 300), (2048, 200), (4096, 100)])
def test_mfma_fstage_fp16_within_bound(gpu, nchan, numffts):
    w, data, poly, offs = case(4, nchan, numffts, seed=61)
    ref, rc = fxref.process(fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True),
                            data, poly, offs, "f64")
    v, c = run(w, data, poly, offs, precision="fp16", fblock=MFMA)
    np.testing.assert_array_equal(c, rc)
    e = eps(v, ref)
    v0, _ = run(w, data, poly, offs, precision="fp16")
    e0 = eps(v0, ref)
    print("nchan %d: mfma F-stage eps %.3g, FP32 F-stage eps %.3g" % (nchan, e, e0))
    assert e < 1e-3


@pytest.mark.parametrize("precision,bound", [("bf16", 1e-2), ("fp8", 0.12)])
def test_mfma_fstage_other_precisions(gpu, precision, bound):
    w, data, poly, offs = case(3, 1024, 200, seed=62)
    ref, _ = fxref.process(fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True),
                           data, poly, offs, "f64")
    v, _ = run(w, data, poly, offs, precision=precision, fblock=MFMA)
    assert eps(v, ref) < bound


def test_mfma_fstage_random_bytes_and_invalid_ffts(gpu):
    # random packed bytes and delays that push the first FFTs out of range
    w = synth.Workload("r", nant=3, nchan=1024, numffts=120, bw=64e6, lo=1.65e9, seed=63)
    rng = np.random.default_rng(63)
    data = [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(3)]
    poly, offs = synth.make_delays(w)
    offs = offs.copy()
    offs[1] += 5 * w.fftchannels * w.sampletime
    ref, rc = fxref.process(fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True),
                            data, poly, offs, "f64")
    v, c = run(w, data, poly, offs, precision="fp16", fblock=MFMA)
    np.testing.assert_array_equal(c, rc)
    assert c.min() < w.numffts
    assert eps(v, ref) < 1e-3


@pytest.mark.parametrize("nchan,numffts", [(1024, 300), (2048, 200)])
def test_mfma_fstage_fp32_within_bound(gpu, nchan, numffts):
    w, data, poly, offs = case(4, nchan, numffts, seed=64)
    ref, rc = fxref.process(fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True),
                            data, poly, offs, "f64")
    v, c = run(w, data, poly, offs, precision="fp32", fblock=MFMA)
    np.testing.assert_array_equal(c, rc)
    e = eps(v, ref)
    v0, _ = run(w, data, poly, offs, precision="fp32")
    e0 = eps(v0, ref)
    print("nchan %d: FP32 mfma F-stage eps %.3g, FP32 F-stage eps %.3g" % (nchan, e, e0))
    assert e < 1e-5


def test_mfma_fstage_fp32_invalid_ffts(gpu):
    w = synth.Workload("r", nant=3, nchan=1024, numffts=120, bw=64e6, lo=1.65e9, seed=65)
    rng = np.random.default_rng(65)
    data = [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(3)]
    poly, offs = synth.make_delays(w)
    offs = offs.copy()
    offs[2] += 7 * w.fftchannels * w.sampletime
    ref, rc = fxref.process(fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True),
                            data, poly, offs, "f64")
    v, c = run(w, data, poly, offs, precision="fp32", fblock=MFMA)
    np.testing.assert_array_equal(c, rc)
    assert c.min() < w.numffts
    assert eps(v, ref) < 1e-5


@pytest.mark.parametrize("precision,bound", [("fp16", 1e-3), ("fp32", 1e-5)])
def test_mfma_fstage_few_units_and_chunks(gpu, precision, bound):
    """Two stations and 5 FFTs (fewer units than workgroups of the
    persistent grid), then the same subintegration in 3 uneven chunks
    (chunk_ffts = 2): every unit of every chunk is still correlated once."""
    w, data, poly, offs = case(2, 2048, 5, seed=66)
    ref, rc = fxref.process(fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True),
                            data, poly, offs, "f64")
    v, c = run(w, data, poly, offs, precision=precision, fblock=MFMA)
    np.testing.assert_array_equal(c, rc)
    assert eps(v, ref) < bound
    v2, c2 = run(w, data, poly, offs, precision=precision, fblock=MFMA, chunk_ffts=2)
    np.testing.assert_array_equal(c2, rc)
    assert eps(v2, ref) < bound


def test_mfma_fstage_rejected_where_unsupported(gpu):
    with pytest.raises(gcorr_amd.GxcError):  # FP32 form: up to 4096 points
        gcorr_amd.FxKernel(3, 4096, 10, 2, 1.65e9, 64e6, precision="fp32", fblock=MFMA)
    with pytest.raises(gcorr_amd.GxcError):  # 512-point FFT: not covered
        gcorr_amd.FxKernel(3, 256, 10, 2, 1.65e9, 64e6, precision="fp16", fblock=MFMA)
