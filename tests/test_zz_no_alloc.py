"""gxc_xengine_noalloc through the C-ABI: the allocation-free variant of
gxc_xengine, with its working memory in the library's fixed per-GPU
scratch array.

Checked against gxc_xengine on the same spectra (FP32 within the 1e-5
parity bar, the reduced precisions within the same bound since both pack
with unit scale and run the same kernels), on two calls in a row with the
second smaller (the scratch keeps the first call's contents), against the
exact integer known-answer pattern of gcorr/test/unit.cc:73-159, and with a
shape whose working memory exceeds the scratch rejected before any GPU work.
"""
import ctypes

import numpy as np
import pytest
import torch

import gcorr_amd
from gcorr_amd import _lib, metrics

pytestmark = pytest.mark.gpu


def noalloc(X, nant, autos, precision):
    """Run gxc_xengine_noalloc on device copies of X; returns (status, vis)."""
    L = _lib.load()
    nfft, ninp, nchan = X.shape
    d_x = torch.from_numpy(np.ascontiguousarray(X, np.complex64)).to("cuda")
    nrows = gcorr_amd.num_rows(nant, autos)
    d_v = torch.full((nrows, 4, nchan), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")
    stream = torch.cuda.current_stream()
    st = L.gxc_xengine_noalloc(ctypes.c_void_p(d_x.data_ptr()), int(nant), int(nchan), int(nfft),
                               int(bool(autos)), _lib.PRECISION[precision], ctypes.c_void_p(d_v.data_ptr()),
                               ctypes.c_void_p(stream.cuda_stream))
    stream.synchronize()
    return st, d_v.cpu().numpy()


def spectra(nfft, nant, nchan, seed):
    rng = np.random.default_rng(seed)
    return ((rng.normal(size=(nfft, 2 * nant, nchan)) + 1j * rng.normal(size=(nfft, 2 * nant, nchan)))
            * 3).astype(np.complex64)


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16", "fp8"])
def test_noalloc_matches_xengine_twice(gpu, precision):
    for nant, nchan, nfft, seed in ((16, 256, 300, 1), (4, 64, 50, 2)):
        X = spectra(nfft, nant, nchan, seed)
        print("noalloc", precision, nant, nchan, nfft, flush=True)
        st, got = noalloc(X, nant, True, precision)
        assert st == 0, _lib.load().gxc_last_error(None)
        want = gcorr_amd.xengine(X, nant, autos=True, precision=precision)
        assert np.isfinite(got).all()
        assert metrics.eps(got, want) < 1e-5, (precision, nant)


@pytest.mark.parametrize("precision", ["fp32", "fp16", "fp8"])
def test_noalloc_integer_kat(gpu, precision):
    """Integer spectra in [-5, 5]: every product and sum is exact."""
    nant, nchan, nfft = 8, 32, 300
    rng = np.random.default_rng(7)
    X = (rng.integers(-5, 6, (nfft, 2 * nant, nchan)) + 1j * rng.integers(-5, 6, (nfft, 2 * nant, nchan))
         ).astype(np.complex64)
    st, got = noalloc(X, nant, True, precision)
    assert st == 0
    Xd = X.astype(np.complex128)
    rows = [(j, k) for j in range(nant) for k in range(j + 1, nant)] + [(a, a) for a in range(nant)]
    for r, (j, k) in enumerate(rows):
        for l in range(2):
            for m in range(2):
                s = np.sum(Xd[:, 2 * j + l] * np.conj(Xd[:, 2 * k + m]), axis=0)
                expect = (s.real.astype(np.float32) / np.float32(nfft)) \
                    + 1j * (s.imag.astype(np.float32) / np.float32(nfft))
                np.testing.assert_array_equal(got[r, 2 * l + m], expect.astype(np.complex64))


def test_noalloc_rejects_shapes_beyond_the_scratch(gpu):
    L = _lib.load()
    # 64 inputs x 32768 channels: one split of partials is 1 GiB
    st = L.gxc_xengine_noalloc(ctypes.c_void_p(16), 32, 32768, 4, 1, _lib.PRECISION["fp32"],
                               ctypes.c_void_p(16), None)
    assert st == 1  # GXC_ERR_INVALID_ARG
    assert b"scratch" in L.gxc_last_error(None)
