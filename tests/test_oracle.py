"""Pins of the oracle (oracle/fxref.c, the CPU restatement of fxkernel).

Each check compares the oracle with something it was not derived from: an
independent encoder or decoder from another part of the reference, a closed
form, a library routine, or brute force on small inputs.
"""
import math

import numpy as np
import pytest

from oracle import fxref

HIMAG = np.float32(3.3359)
LEVELS = np.array([-HIMAG, -1.0, 1.0, HIMAG], np.float32)


def cfg_real(nant=2, nchan=128, numffts=10, lo=1.65e9, bw=16e6, autos=False, quirk=True):
    return fxref.make_config(nant, nchan, numffts, lo, bw, autos=autos, quirk=quirk)


# ---------------------------------------------------------------- unpack / LUT

def pack_like_generators(code_a, code_b):
    """generators/bitconversion.c:147-155 (packBit2, 2 channels): ch0=pol0 t, ch1=pol1 t,
    ch2=pol0 t+1, ch3=pol1 t+1, byte = ch0 | ch1<<2 | ch2<<4 | ch3<<6."""
    out = np.zeros(len(code_a) // 2, np.uint8)
    for i in range(len(out)):
        out[i] = (code_a[2 * i] | (code_b[2 * i] << 2) | (code_a[2 * i + 1] << 4)
                  | (code_b[2 * i + 1] << 6))
    return out


def unpack_like_gcorr(byte):
    """gcorr/src/gxkernel.cu:215-227 (unpack2bit_2chan): levels of src&3 -> pol0 t,
    (src>>2)&3 -> pol1 t, (src>>4)&3 -> pol0 t+1, (src>>6)&3 -> pol1 t+1."""
    return (LEVELS[byte & 3], LEVELS[(byte >> 2) & 3], LEVELS[(byte >> 4) & 3],
            LEVELS[(byte >> 6) & 3])


def test_lut_every_byte_matches_gcorr_decoder():
    lut = fxref.lut2bit()
    for b in range(256):
        a0, b0, a1, b1 = unpack_like_gcorr(b)
        assert tuple(lut[b]) == (a0, b0, a1, b1)


@pytest.mark.parametrize("offset", [0, 1, 2, 3, 101, 256, 257])
def test_unpack_roundtrip_with_generator_packer(offset):
    cfg = cfg_real(nchan=64, numffts=8)
    d = fxref.derive(cfg)
    n = 2 * d.bytes_per_station
    rng = np.random.default_rng(offset)
    ca = rng.integers(0, 4, n).astype(np.uint8)
    cb = rng.integers(0, 4, n).astype(np.uint8)
    data = pack_like_generators(ca, cb)
    out = fxref.unpack(cfg, data, offset)
    L = d.fftchannels
    np.testing.assert_array_equal(out[0].real, LEVELS[ca[offset:offset + L]])
    np.testing.assert_array_equal(out[1].real, LEVELS[cb[offset:offset + L]])
    assert not np.any(out.imag)


def test_unpack_complex8_layout():
    cfg = fxref.make_config(1, 64, 4, 1.65e9, 16e6, nbit=8, is_complex=True)
    d = fxref.derive(cfg)
    rng = np.random.default_rng(3)
    raw = rng.integers(-128, 128, d.bytes_per_station).astype(np.int8)
    out = fxref.unpack(cfg, raw.view(np.uint8), 5)
    t = raw.reshape(-1, 4)[5:5 + d.fftchannels].astype(np.float64)
    np.testing.assert_array_equal(out[0], t[:, 0] + 1j * t[:, 1])
    np.testing.assert_array_equal(out[1], t[:, 2] + 1j * t[:, 3])


# ---------------------------------------------------------------- delay model

@pytest.mark.parametrize("i", [0, 1, 17, 3199, 62499])
def test_station_delay_closed_forms(i):
    # linear: a = d1 slope, b = d(i).  a is a difference of two ~1e-3 s
    # values, so its error is a few ulp of 1e-3 s (2.2e-19 s), not of a.
    ulp = np.spacing(1.0e-3)
    m, a, b = fxref.station_delay([0.0, 4.8e-12, 1.0e-3], i)
    assert a == pytest.approx(4.8e-12, abs=4 * ulp)
    assert b == pytest.approx(4.8e-12 * i + 1.0e-3, rel=1e-13)
    assert m == pytest.approx(b + a / 2, rel=1e-13)
    # constant: no slope
    m, a, b = fxref.station_delay([0.0, 0.0, -2.5e-3], i)
    assert a == 0.0 and b == -2.5e-3 and m == -2.5e-3
    # quadratic: a = p0 (2i+1) + p1, b = d(i) - p0/12, mean = b + a/2
    p = [6.1e-15, 4.8e-12, 1.001e-3]
    m, a, b = fxref.station_delay(p, i)
    assert a == pytest.approx(p[0] * (2 * i + 1) + p[1], abs=8 * ulp)
    assert b == pytest.approx(p[0] * i * i + p[1] * i + p[2] - p[0] / 12, rel=1e-12)
    assert m == pytest.approx(b + a / 2, rel=1e-14)


def _const_delay_case(ndf, fftindex, numffts=10):
    cfg = cfg_real(nchan=128, numffts=numffts)
    Ts = 1.0 / (2.0 * 16e6)
    return cfg, fxref.delay_split(cfg, [0.0, 0.0, ndf * Ts], 0.0, fftindex), Ts


def test_delay_split_positive_rounds_half_up():
    cfg, (valid, off, frac, _, _), Ts = _const_delay_case(3.25, 1)
    assert valid and off == 256 - 3
    assert frac == pytest.approx(-0.25 * Ts, abs=1e-20)
    _, (valid, off, _, _, _), _ = _const_delay_case(3.75, 1)
    assert valid and off == 256 - 4


def test_delay_split_first_fft_invalid_and_slack():
    _, (valid, _, _, _, _), _ = _const_delay_case(3.25, 0)
    assert not valid                            # offset -3 leaves the buffer
    _, (valid, off, frac, _, _), Ts = _const_delay_case(0.75, 0)
    assert valid and off == 0                   # offset -1 -> 0, frac += Ts
    assert frac == pytest.approx((0.25 + 1.0) * Ts, abs=1e-20)


def test_delay_split_negative_truncates_toward_zero():
    # ndf = -1.75: int(-1.25) = -1 (truncation), so the last FFT sits at
    # maxoffset+1 and is pulled back by one sample; rounding to -2 would
    # make it invalid instead.
    _, (valid, off, frac, _, _), Ts = _const_delay_case(-1.75, 9)
    assert valid and off == 9 * 256
    assert frac == pytest.approx((0.75 - 1.0) * Ts, abs=1e-20)
    _, (valid, _, _, _, _), _ = _const_delay_case(-2.625, 9)
    assert not valid


# ---------------------------------------------------------------- fringe rotator

def _ideal_rotator(cfg, a, b, quirk):
    d = fxref.derive(cfg)
    L, S2, s = d.fftchannels, d.substride, d.stride
    n = np.arange(L)
    neff = (n % S2) + (n // S2) * (s if quirk else S2)
    ib = math.trunc(b)
    ph = cfg.lo * ((b - ib) + a * neff / L)
    if cfg.lo - math.trunc(cfg.lo) > 1e-11:
        ph = ph + (cfg.lo - math.trunc(cfg.lo)) * ib
    return np.exp(-2j * np.pi * (ph - np.trunc(ph)))


def test_rotator_identity_and_constant():
    cfg = cfg_real(nchan=1024)
    r = fxref.rotator(cfg, 0.0, 0.0)
    np.testing.assert_array_equal(r.real, 1.0)
    np.testing.assert_array_equal(r.imag, 0.0)
    b = 1.234567e-3
    r = fxref.rotator(cfg, 0.0, b, "f64")
    expect = np.exp(-2j * np.pi * ((1.65e9 * b) % 1.0))
    np.testing.assert_allclose(r, expect, atol=1e-9)


@pytest.mark.parametrize("nchan", [64, 128, 256, 1024, 2048])
@pytest.mark.parametrize("quirk", [True, False])
def test_rotator_matches_per_sample_law(nchan, quirk):
    cfg = cfg_real(nchan=nchan, quirk=quirk)
    a, b = 2.0e-10, -3.21e-3
    ideal = _ideal_rotator(cfg, a, b, quirk)
    np.testing.assert_allclose(fxref.rotator(cfg, a, b, "f64"), ideal, atol=2e-8)
    np.testing.assert_allclose(fxref.rotator(cfg, a, b, "f32"), ideal, atol=3e-6)
    if quirk:  # the stride law really differs from the exact one at this slope
        other = _ideal_rotator(cfg, a, b, False)
        assert np.max(np.abs(other - ideal)) > 1e-3


def test_rotator_fractional_lo_adds_integer_delay_term():
    cfg = cfg_real(nchan=256, lo=1.65e9 + 0.25)
    a, b = 1.0e-11, 1.0001  # one whole second of delay plus 100 us
    ideal = _ideal_rotator(cfg, a, b, True)
    np.testing.assert_allclose(fxref.rotator(cfg, a, b, "f64"), ideal, atol=1e-8)
    # without the (lo - int lo) * intdelay term the phase is off by 0.25 turn
    cfg0 = cfg_real(nchan=256, lo=1.65e9 + 0.25)
    no_term = np.exp(-2j * np.pi * ((cfg0.lo * (b - 1.0) + cfg0.lo * a * 0) % 1.0))
    assert abs(fxref.rotator(cfg0, a, b, "f64")[0] - no_term) > 0.5


# ---------------------------------------------------------------- fractional sample

@pytest.mark.parametrize("nchan", [128, 512, 1024])
def test_fracrot_closed_form(nchan):
    cfg = cfg_real(nchan=nchan)
    Ts = 1.0 / 32e6
    np.testing.assert_array_equal(fxref.fracrot(cfg, 0.0), 1.0 + 0j)
    # one sample: phase 2 pi c bw / nchan * Ts = pi c / nchan
    c = np.arange(nchan)
    np.testing.assert_allclose(fxref.fracrot(cfg, Ts, "f64"), np.exp(1j * np.pi * c / nchan),
                               atol=1e-12)
    frac = -0.3712 * Ts
    np.testing.assert_allclose(fxref.fracrot(cfg, frac, "f32"),
                               np.exp(1j * 2 * np.pi * c * 16e6 / nchan * frac), atol=2e-6)


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("L", [2, 4, 8, 32, 128, 256, 2048, 8192, 32768])
def test_fft_matches_numpy(L):
    rng = np.random.default_rng(L)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    ref = np.fft.fft(x)
    scale = np.sqrt(L)
    assert np.max(np.abs(fxref.fft(x, "f64") - ref)) / scale < 1e-13
    assert np.max(np.abs(fxref.fft(x, "f32") - ref)) / scale < 2e-6


@pytest.mark.parametrize("L", [4, 16, 64, 256])
def test_fft_matches_brute_force_dft(L):
    rng = np.random.default_rng(7 + L)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    n = np.arange(L)
    W = np.exp(-2j * np.pi * np.outer(n, n) / L)
    np.testing.assert_allclose(fxref.fft(x, "f64"), W @ x, atol=1e-11)


# ---------------------------------------------------------------- X-engine KAT

@pytest.mark.parametrize("nant,nchan,nfft", [(3, 256, 4), (5, 1024, 100), (3, 512, 300)])
def test_xmac_integer_kat(nant, nchan, nfft):
    """gcorr/test/unit.cc:42-159 pattern: integer data in [-5, 5] makes every
    FP32 sum exact, so the check is equality with the exact brute force."""
    rng = np.random.default_rng(12335 + nant)
    X = (rng.integers(-5, 6, (nfft, 2 * nant, nchan))
         + 1j * rng.integers(-5, 6, (nfft, 2 * nant, nchan)))
    got = fxref.xmac(X, nant, autos=True)
    rows = [(j, k) for j in range(nant) for k in range(j + 1, nant)] + [(a, a) for a in range(nant)]
    for r, (j, k) in enumerate(rows):
        for l in range(2):
            for m in range(2):
                expect = np.sum(X[:, 2 * j + l] * np.conj(X[:, 2 * k + m]), axis=0)
                np.testing.assert_array_equal(got[r, 2 * l + m], expect)


# ---------------------------------------------------------------- whole process

def _random_case(nant=3, nchan=128, numffts=200, seed=5, autos=True):
    cfg = cfg_real(nant=nant, nchan=nchan, numffts=numffts, autos=autos)
    d = fxref.derive(cfg)
    rng = np.random.default_rng(seed)
    data = [rng.integers(0, 256, d.bytes_per_station, dtype=np.uint8) for _ in range(nant)]
    Ts = d.sampletime
    delays = np.array([[0.0, 3.0e-12 * (a - 1), (1.0 + a) * 1e-3] for a in range(nant)])
    offsets = np.array([(1.0 + a) * 1e-3 + (0.3 + a) * Ts for a in range(nant)])
    return cfg, d, data, delays, offsets


def test_process_fp32_close_to_fp64():
    cfg, d, data, delays, offsets = _random_case()
    v32, c32 = fxref.process(cfg, data, delays, offsets, "f32")
    v64, c64 = fxref.process(cfg, data, delays, offsets, "f64")
    np.testing.assert_array_equal(c32, c64)
    eps = np.linalg.norm(v32 - v64, axis=-1) / np.linalg.norm(v64, axis=-1)
    assert eps.max() < 1e-5


def test_autos_closed_form_uniform_codes():
    # Uniform random bytes: every code equiprobable, E[x^2] = (1 + H^2)/2, so
    # E|X_c|^2 = L E[x^2] per FFT (white input).
    cfg, d, data, delays, offsets = _random_case(numffts=400)
    v, c = fxref.process(cfg, data, delays, offsets, "f64")
    nbl = d.nbaselines
    expect = d.fftchannels * (1 + float(HIMAG) ** 2) / 2
    for a in range(cfg.nant):
        for p in (0, 3):
            assert v[nbl + a, p].real.mean() == pytest.approx(expect, rel=0.01)
        # cross-hand autos are conjugates of each other
        np.testing.assert_allclose(v[nbl + a, 2], np.conj(v[nbl + a, 1]), rtol=1e-12, atol=1e-9)


def test_counts_follow_validity():
    cfg, d, data, delays, offsets = _random_case(numffts=50)
    # station 2 gets a net delay of +2.6 samples: FFT 0 leaves the buffer
    offsets = offsets.copy()
    delays = delays.copy()
    delays[2] = [0.0, 0.0, 5e-3]
    offsets[2] = 5e-3 - 2.6 * d.sampletime
    v, c = fxref.process(cfg, data, delays, offsets, "f32")
    # rows: (0,1),(0,2),(1,2), autos 0,1,2
    assert list(c) == [50, 49, 49, 50, 50, 49]


def test_accumulate_range_sums_to_process():
    cfg, d, data, delays, offsets = _random_case(numffts=60)
    v, c = fxref.process(cfg, data, delays, offsets, "f64")
    s1, c1 = fxref.accumulate_range(cfg, data, delays, offsets, 0, 25)
    s2, c2 = fxref.accumulate_range(cfg, data, delays, offsets, 25, 60)
    np.testing.assert_array_equal(c1 + c2, c)
    tot = (s1 + s2) / np.maximum(c, 1)[:, None, None]
    assert np.max(np.abs(tot - v)) / np.max(np.abs(v)) < 1e-5


def test_physics_cross_amplitude_on_generated_signal():
    """Generated common signal with delays and fringes: after fxkernel's
    delay, fringe and fractional-sample corrections the normalised cross
    amplitude is the 2-bit correlation eta*rho (eta = 0.8825 for these
    levels) with ~zero phase.  Fixes every sign convention together."""
    from gcorr_amd import synth
    w = synth.WORKLOADS["c1"].scaled(1500)
    bufs, poly, offs = synth.generate(w, device="cpu")
    data = [b.numpy() for b in bufs]
    cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw, autos=True)
    v, c = fxref.process(cfg, data, poly, offs, "f64")
    nbl = w.nant * (w.nant - 1) // 2
    eta = (2 * ((1 / math.sqrt(2 * math.pi)) * (1 - math.exp(-0.98 ** 2 / 2))
                + float(HIMAG) * math.exp(-0.98 ** 2 / 2) / math.sqrt(2 * math.pi))) ** 2 \
        / synth.quantised_power_2bit()
    assert eta == pytest.approx(0.8825, abs=5e-4)
    b = 0
    for j in range(w.nant - 1):
        for k in range(j + 1, w.nant):
            for p, (l, m) in ((0, (0, 0)), (3, (1, 1))):
                norm = np.sqrt(v[nbl + j, 3 * l].real * v[nbl + k, 3 * m].real)
                r = (v[b, p] / norm)[4:-4].mean()
                assert abs(r) == pytest.approx(eta * w.rho, rel=0.12)
                assert abs(np.angle(r)) < 0.15
            b += 1
    # the autocorrelation level is the quantiser's power, a closed form
    assert v[nbl:, 0].real.mean() / w.fftchannels == pytest.approx(synth.quantised_power_2bit(),
                                                                    rel=0.01)
