"""The lag S/N restatement (util/lagsnr.py:33-36) and the vis.out fixture.

Pinned by closed forms: a pure delay of d channels-worth of lag,
V_c = exp(-2 pi i c d / n), transforms under ifft to a single unit lag at d, so
S/N = 1 / sqrt(1/n - 1/n^2) and the peak is d; a constant series peaks at
lag 0.  The fixture is the reference's example output,
examples/adam/vis.out, parsed with lagsnr.py's column rule.
"""
import os

import numpy as np
import pytest

from oracle.lagref import lag_snr

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("n,d", [(2, 1), (64, 0), (64, 17), (1024, 1000), (4096, 3)])
def test_single_delay_closed_form(n, d):
    c = np.arange(n)
    v = np.exp(-2j * np.pi * c * d / n)
    snr, peak = lag_snr(v)
    assert peak == d
    assert snr == pytest.approx(1.0 / np.sqrt(1.0 / n - 1.0 / n ** 2), rel=1e-9)


def test_two_delays_and_batch_shape():
    n = 256
    c = np.arange(n)
    v = np.stack([np.exp(-2j * np.pi * c * 5 / n) + 0.5 * np.exp(-2j * np.pi * c * 40 / n),
                  np.ones(n)])
    snr, peak = lag_snr(v)
    assert list(peak) == [5, 0]
    lags = np.zeros(n)
    lags[5], lags[40] = 1.0, 0.5
    assert snr[0] == pytest.approx(1.0 / lags.std(), rel=1e-9)


def test_adam_fixture_loads():
    vis = np.load(os.path.join(GOLDEN, "adam_vis.npz"))["vis"]
    assert vis.shape == (4, 4096) and vis.dtype == np.complex128
    snr, peak = lag_snr(vis)
    assert np.all(np.isfinite(snr)) and np.all(snr > 3)
    assert np.all((peak >= 0) & (peak < 4096))
