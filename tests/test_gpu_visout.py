"""Visibility output and the GPU lag S/N (gxc_lag_snr, gxc_save_*).

The lag S/N of util/lagsnr.py:33-36 on the GPU must match the restatement
(oracle/lagref.py) on the reference's example visibilities
(examples/adam/vis.out, as a fixture), on closed-form delays up to 32768
channels and on a correlated subintegration; the binary and DiFX writers
must read back bit for bit.
"""
import os

import numpy as np
import pytest

import gcorr_amd
from gcorr_amd import synth, visio
from oracle.lagref import lag_snr

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_lag_snr_on_reference_example(gpu):
    vis = np.load(os.path.join(GOLDEN, "adam_vis.npz"))["vis"]
    snr, peak = gcorr_amd.lag_snr(vis.astype(np.complex64))
    want, wpeak = lag_snr(vis.astype(np.complex64))
    np.testing.assert_array_equal(peak, wpeak)
    np.testing.assert_allclose(snr, want, rtol=1e-4)


@pytest.mark.parametrize("n", [2, 64, 1024, 4096, 32768])
def test_lag_snr_closed_form(gpu, n):
    c = np.arange(n)
    ds = [0, 1, n // 3, n - 1]
    v = np.stack([np.exp(-2j * np.pi * c * d / n) for d in ds]).astype(np.complex64)
    snr, peak = gcorr_amd.lag_snr(v)
    assert list(peak) == ds
    np.testing.assert_allclose(snr, 1.0 / np.sqrt(1.0 / n - 1.0 / n ** 2), rtol=2e-4)


def test_lag_snr_random_batch(gpu):
    rng = np.random.default_rng(5)
    v = (rng.standard_normal((37, 512)) + 1j * rng.standard_normal((37, 512))).astype(np.complex64)
    v[:, :] += 3 * np.exp(-2j * np.pi * np.arange(512) * 77 / 512)
    snr, peak = gcorr_amd.lag_snr(v)
    want, wpeak = lag_snr(v)
    np.testing.assert_array_equal(peak, wpeak)
    np.testing.assert_allclose(snr, want, rtol=1e-4)


def correlated(nchan=256, numffts=400):
    w = synth.Workload("o", nant=4, nchan=nchan, numffts=numffts, bw=64e6, lo=1.65e9, rho=0.3, seed=41)
    data, poly, offs = synth.generate(w)
    k = gcorr_amd.FxKernel(w.nant, w.nchan, w.numffts, w.nbit, w.lo, w.bw, autos=True)
    k.setInputData([np.asarray(d) for d in data])
    k.setDelays(poly, offs)
    k.process()
    return w, k


def test_lag_snr_of_a_subintegration(gpu):
    w, k = correlated()
    vis = k.visibilities()
    snr, peak = k.lag_snr()
    want, wpeak = lag_snr(vis)
    np.testing.assert_allclose(snr, want, rtol=1e-4)
    lags = np.sort(np.abs(np.fft.ifft(vis.astype(np.complex128), axis=-1)), axis=-1)
    clear = lags[..., -1] > lags[..., -2] * (1 + 1e-3)  # peaks not decided by rounding
    np.testing.assert_array_equal(peak[clear], wpeak[clear])
    assert clear.sum() >= vis.shape[0]
    k.close()


def test_binary_and_difx_round_trip(gpu, tmp_path):
    w, k = correlated(nchan=128, numffts=100)
    vis, counts = k.visibilities(), k.row_counts()
    k.save_binary(tmp_path / "v.bin")
    b = visio.read_binary(tmp_path / "v.bin")
    np.testing.assert_array_equal(b["vis"], vis)
    np.testing.assert_array_equal(b["counts"], counts)
    rows = [(j, l) for j in range(4) for l in range(j + 1, 4)] + [(a, a) for a in range(4)]
    assert [tuple(x) for x in b["stations"]] == rows
    assert (b["nant"], b["nchan"], b["numffts"], b["autos"]) == (4, 128, 100, True)
    assert (b["bandwidth"], b["lo"]) == (w.bw, w.lo)
    uvw = np.arange(len(rows) * 3, dtype=np.float64).reshape(-1, 3) * 1.5
    k.save_difx(tmp_path / "v.difx", mjd=59123, seconds=4321.25, pols="RL", uvw=uvw, source_index=2,
                freq_index=3)
    recs = visio.read_difx(tmp_path / "v.difx", 128)
    assert len(recs) == len(rows) * 4
    for i, (h, v) in enumerate(recs):
        r, p = divmod(i, 4)
        j, l = rows[r]
        assert h["baseline"] == 256 * (j + 1) + (l + 1)
        assert h["polpair"] == "RL"[p >> 1] + "RL"[p & 1]
        assert (h["mjd"], h["seconds"], h["source_index"], h["freq_index"]) == (59123, 4321.25, 2, 3)
        assert h["weight"] == counts[r] / 100
        assert (h["u"], h["v"], h["w"]) == tuple(uvw[r])
        assert h["version"] == 1 and h["pulsar_bin"] == 0
        np.testing.assert_array_equal(v, vis[r, p])
    k.close()
