"""VDIF ingest on the CPU: the frame writer, the oracle's placement
(gcorr ingest, ingest/src/ingest.c:267-355) and the missing-frame rule.

The header layout is pinned by a hand-assembled header decoded the way
examples/mark/dump_vdif.py:36-48 decodes it; placement by round trips
through the writer with frames shuffled, dropped, duplicated, flagged
invalid, from other threads, outside the window and malformed; the
missing-frame rule by the closed-form count of FFTs that touch a missing
frame and by per-FFT sums of the unflagged oracle.
"""
import numpy as np
import pytest

from gcorr_amd import vdif
from oracle import fxref


def test_decode_hand_assembled_header():
    # w0: legacy, 12345 s; w1: epoch 42, frame 5; w2: version 1, 2 channels,
    # 36 x 8 bytes; w3: real, 2 bits, thread 64, station 1
    raw = np.array([0x40003039, 0x2A000005, 0x21000024, 0x04400001], "<u4").view(np.uint8)
    h = vdif.decode_header(raw)
    assert h == {"invalid": 0, "legacy": 1, "seconds": 12345, "frame": 5, "epoch": 42,
                 "frame_bytes": 288, "channels": 2, "complex": 0, "bits": 2, "thread": 64,
                 "station": 1}
    w = vdif.header_words(12345, 5, 288, 2, False, epoch=42, thread=64, station=1, legacy=True)
    # the writer reproduces every field (the version bits are not written)
    assert list(w) == [0x40003039, 0x2A000005, 0x01000024, 0x04400001]
    c = vdif.header_words(7, 1, 1056, 8, True, invalid=True)
    d = vdif.decode_header(c.astype("<u4").view(np.uint8))
    assert (d["invalid"], d["complex"], d["bits"], d["legacy"], d["frame_bytes"]) == (1, 1, 8, 0, 1056)


def small(nchan=64, numffts=20, nant=3, is_complex=False):
    cfg = fxref.make_config(nant, nchan, numffts, 1.65e9, 16e6, nbit=8 if is_complex else 2,
                            is_complex=is_complex)
    d = fxref.derive(cfg)
    return cfg, d


def test_round_trip_in_any_order():
    cfg, d = small()
    rng = np.random.default_rng(1)
    buf = rng.integers(0, 256, d.bytes_per_station, dtype=np.uint8)
    fr = vdif.frames_from_station(buf, 256, frames_per_second=100, ref_second=5, ref_frame=97)
    rng.shuffle(fr)
    st, ok, spf, stats = fxref.vdif_place(cfg, fr.reshape(-1), frames_per_second=100,
                                          ref_second=5, ref_frame=97)
    np.testing.assert_array_equal(st, buf)
    assert ok.all() and len(ok) == fr.shape[0]
    assert spf == 256 * 8 // 4  # 2 bits x 2 pols per time sample
    assert stats == {"frames": 6, "placed": 6, "missing": 0, "duplicate": 0, "invalid": 0,
                     "other_thread": 0, "out_of_window": 0, "malformed": 0}


def test_every_skip_rule():
    cfg, d = small()
    rng = np.random.default_rng(2)
    buf = rng.integers(0, 256, d.bytes_per_station, dtype=np.uint8)
    fr = vdif.frames_from_station(buf, 128, frames_per_second=4, ref_second=10, ref_frame=3)
    n = fr.shape[0]
    keep = [k for k in range(n) if k not in (2, 7)]  # frames 2 and 7 never arrive
    extra = []
    dup = fr[4].copy()
    dup[40:] ^= 0xFF                                  # later duplicate of slot 4: ignored
    extra.append(dup)
    bad = fr[5].copy()
    vdif.set_header(bad, invalid=1)                  # invalid copy of slot 5: skipped
    extra.append(bad)
    other = fr[6].copy()
    vdif.set_header(other, thread=3)                 # another thread
    extra.append(other)
    early = fr[0].copy()
    vdif.set_header(early, seconds=10, frame=2)      # one frame before the window
    extra.append(early)
    late = fr[0].copy()
    vdif.set_header(late, seconds=10 + (3 + n) // 4, frame=(3 + n) % 4)  # just past the end
    extra.append(late)
    ep = fr[1].copy()
    vdif.set_header(ep, epoch=1)                     # another epoch
    extra.append(ep)
    for f in (dict(bits=4), dict(channels=4), dict(complex=1), dict(frame_bytes=8 * 21),
              dict(frame=4), dict(legacy=1)):
        m = fr[3].copy()
        vdif.set_header(m, **f)
        extra.append(m)
    stream = np.concatenate([fr[keep]] + [e[None] for e in extra]).reshape(-1)
    st, ok, spf, stats = fxref.vdif_place(cfg, stream, frames_per_second=4, ref_second=10,
                                          ref_frame=3)
    want_ok = np.ones(n, np.uint8)
    want_ok[[2, 7]] = 0
    np.testing.assert_array_equal(ok, want_ok)
    for k in range(n):
        seg = slice(k * 128, min((k + 1) * 128, buf.size))
        np.testing.assert_array_equal(st[seg], buf[seg] if want_ok[k] else 0)
    assert stats == {"frames": len(keep) + len(extra), "placed": n - 2, "missing": 2,
                     "duplicate": 1, "invalid": 1, "other_thread": 1, "out_of_window": 3,
                     "malformed": 6}


def test_all_threads_and_frame_size_override():
    cfg, d = small()
    buf = np.arange(d.bytes_per_station, dtype=np.uint64).astype(np.uint8)
    fr = vdif.frames_from_station(buf, 512, frames_per_second=10, thread=9, legacy=True)
    st, ok, spf, stats = fxref.vdif_place(cfg, fr.reshape(-1), frames_per_second=10, thread_id=-1,
                                          frame_bytes=16 + 512)
    np.testing.assert_array_equal(st, buf)
    assert ok.all() and stats["other_thread"] == 0
    _, ok0, _, stats0 = fxref.vdif_place(cfg, fr.reshape(-1), frames_per_second=10, thread_id=0)
    assert not ok0.any() and stats0["other_thread"] == fr.shape[0]


def test_complex8_offset_binary():
    cfg, d = small(nchan=32, numffts=8, is_complex=True)
    rng = np.random.default_rng(3)
    buf = rng.integers(-127, 128, d.bytes_per_station).astype(np.int8).view(np.uint8)
    fr = vdif.frames_from_station(buf, 64, nbit=8, is_complex=True, frames_per_second=50)
    # on the wire: offset binary, so a two's-complement 0 is 0x80
    pay = fr[:, 32:].reshape(-1)[:buf.size]
    np.testing.assert_array_equal(pay, buf ^ 0x80)
    st, ok, spf, _ = fxref.vdif_place(cfg, fr.reshape(-1), frames_per_second=50)
    np.testing.assert_array_equal(st, buf)
    assert spf == 64 // 4
    raw = vdif.frames_from_station(buf, 64, nbit=8, is_complex=True, frames_per_second=50,
                                   offset_binary8=False)
    st2, _, _, _ = fxref.vdif_place(cfg, raw.reshape(-1), frames_per_second=50, offset_binary8=False)
    np.testing.assert_array_equal(st2, buf)


def test_unusable_streams_are_errors():
    cfg, d = small()
    buf = np.zeros(d.bytes_per_station, np.uint8)
    fr = vdif.frames_from_station(buf, 256, frames_per_second=10)
    with pytest.raises(ValueError):  # payload not a multiple of 8 bytes
        fxref.vdif_place(cfg, fr.reshape(-1), frames_per_second=10, frame_bytes=32 + 252)
    with pytest.raises(ValueError):
        fxref.vdif_place(cfg, fr.reshape(-1), frames_per_second=0)


def per_fft_sums(cfg, data, poly, offs):
    """Unnormalised FP32 sums and counts of every FFT alone."""
    out = []
    for t in range(cfg.numffts):
        out.append(fxref.accumulate_range(cfg, data, poly, offs, t, t + 1))
    return out


def test_missing_frame_rule_closed_form():
    nant, nchan, numffts = 3, 32, 24
    cfg, d = small(nchan, numffts, nant)
    L = d.fftchannels
    rng = np.random.default_rng(4)
    data = [rng.integers(0, 256, d.bytes_per_station, dtype=np.uint8) for _ in range(nant)]
    poly = np.zeros((nant, 3))
    offs = np.zeros(nant)  # zero delays: FFT t reads samples [tL, tL + L)
    payload = 24           # 48 time samples per frame: FFTs and frames do not line up
    spf = payload * 8 // 4
    nslots = -(-d.bytes_per_station // payload)
    flags = [np.ones(nslots, np.uint8) for _ in range(nant)]
    # all present: identical to no flags
    full = fxref.make_config(nant, nchan, numffts, 1.65e9, 16e6, frame_ok=flags,
                             frame_samples=[spf] * nant)
    v0, c0 = fxref.process(cfg, data, poly, offs)
    v1, c1 = fxref.process(full, data, poly, offs)
    np.testing.assert_array_equal(v0, v1)
    np.testing.assert_array_equal(c0, c1)
    # station 1 misses frame 9 (samples 432..479), station 2 misses frame 30
    flags[1][9] = 0
    flags[2][30] = 0
    cfgm = fxref.make_config(nant, nchan, numffts, 1.65e9, 16e6, frame_ok=flags,
                             frame_samples=[spf] * nant)
    v2, c2 = fxref.process(cfgm, data, poly, offs)

    def bad(f):
        return {t for t in range(numffts) if t * L < (f + 1) * spf and t * L + L > f * spf}

    b1, b2 = bad(9), bad(30)
    assert b1 == {6, 7} and b2 == {22, 23}  # L = 64: frame 9 spans FFTs 6-7, frame 30 FFTs 22-23
    valid = [set(range(numffts)), set(range(numffts)) - b1, set(range(numffts)) - b2]
    pairs = [(0, 1), (0, 2), (1, 2)]
    np.testing.assert_array_equal(c2, [len(valid[j] & valid[k]) for j, k in pairs])
    sums = per_fft_sums(cfg, data, poly, offs)
    for r, (j, k) in enumerate(pairs):
        acc = np.zeros((4, nchan), np.complex128)
        for t in sorted(valid[j] & valid[k]):
            acc += sums[t][0][r]
        np.testing.assert_allclose(v2[r], acc / c2[r], rtol=2e-5, atol=2e-5 * np.abs(acc).max() / c2[r])


def test_header_size_given_or_from_first_frame():
    """A stream whose first frame claims a legacy (16-byte) header: by
    default the geometry follows that frame and every well-formed frame is
    malformed against it; with header_bytes = 32 the stray first frame is the
    only malformed one and the station buffer is whole."""
    cfg, d = small()
    buf = np.arange(d.bytes_per_station, dtype=np.uint64).astype(np.uint8)
    fr = vdif.frames_from_station(buf, 256, frames_per_second=10)
    stray = fr[0].copy()
    vdif.set_header(stray, legacy=1)
    stream = np.concatenate([stray[None], fr]).reshape(-1)
    n = fr.shape[0]
    st, ok, _, stats = fxref.vdif_place(cfg, stream, frames_per_second=10, header_bytes=32)
    np.testing.assert_array_equal(st, buf)
    assert ok.all() and stats["malformed"] == 1 and stats["placed"] == n
    _, ok0, _, stats0 = fxref.vdif_place(cfg, stream, frames_per_second=10)
    assert stats0["malformed"] == n and stats0["placed"] == 1 and ok0.sum() == 1
