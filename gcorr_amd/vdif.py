"""VDIF frames for the ingest path (gxc_set_input_vdif).

Writes and reads the frame layout that gcorr's ingest consumes
(ingest/src/ingest.c:267-355) and examples/mark/dump_vdif.py:36-55 decodes:
four (legacy) or eight little-endian 32-bit header words, then the payload.

    w0 = invalid:1 | legacy:1 | seconds from epoch:30
    w1 = unassigned:2 | epoch:6 | frame number within the second:24
    w2 = version:3 | log2(channels):5 | frame length in 8-byte units:24
    w3 = complex:1 | bits per sample - 1:5 | thread:10 | station:16

The payload of a two-channel (two-polarisation) frame is the fxkernel
station layout itself: real 2-bit samples packed LSB first as
[A_t, B_t, A_t+1, B_t+1], complex 8-bit as [Are, Aim, Bre, Bim] in offset
binary (the station buffers hold two's complement, so 8-bit bytes are
XORed with 0x80 on the way in and out).
"""
import numpy as np


def header_words(seconds, frame, frame_bytes, nbit, is_complex, *, epoch=0, thread=0, station=0,
                 invalid=False, legacy=False, log2_channels=1):
    """The header words of one frame (uint32 array of 4 or 8)."""
    w = np.zeros(4 if legacy else 8, np.uint32)
    w[0] = (int(bool(invalid)) << 31) | (int(bool(legacy)) << 30) | (int(seconds) & 0x3FFFFFFF)
    w[1] = ((int(epoch) & 0x3F) << 24) | (int(frame) & 0xFFFFFF)
    w[2] = ((int(log2_channels) & 0x1F) << 24) | ((int(frame_bytes) // 8) & 0xFFFFFF)
    w[3] = (int(bool(is_complex)) << 31) | (((int(nbit) - 1) & 0x1F) << 26) | \
        ((int(thread) & 0x3FF) << 16) | (int(station) & 0xFFFF)
    return w


def decode_header(buf):
    """Fields of the header at the start of buf (as dump_vdif.py:36-48 reads them)."""
    w = np.frombuffer(bytes(buf[:16]), "<u4")
    return {
        "invalid": int(w[0] >> 31), "legacy": int((w[0] >> 30) & 1),
        "seconds": int(w[0] & 0x3FFFFFFF), "frame": int(w[1] & 0xFFFFFF),
        "epoch": int((w[1] >> 24) & 0x3F), "frame_bytes": int(w[2] & 0xFFFFFF) * 8,
        "channels": 1 << int((w[2] >> 24) & 0x1F), "complex": int(w[3] >> 31),
        "bits": int((w[3] >> 26) & 0x1F) + 1, "thread": int((w[3] >> 16) & 0x3FF),
        "station": int(w[3] & 0xFFFF),
    }


def frames_from_station(buf, payload_bytes, nbit=2, is_complex=False, *, frames_per_second,
                        ref_second=0, ref_frame=0, epoch=0, thread=0, station=0, legacy=False,
                        offset_binary8=True):
    """Cut a station buffer into consecutive VDIF frames.

    Returns uint8 [nframes, frame_bytes]; frame k carries bytes
    [k*payload_bytes, (k+1)*payload_bytes) of buf (the last one zero-padded)
    and the time stamp of slot k after (ref_second, ref_frame)."""
    buf = np.ascontiguousarray(buf, np.uint8).reshape(-1)
    if payload_bytes % 8:
        raise ValueError("VDIF payloads are a multiple of 8 bytes")
    hdr = 16 if legacy else 32
    fb = hdr + payload_bytes
    n = -(-buf.size // payload_bytes)
    out = np.zeros((n, fb), np.uint8)
    pay = np.zeros(n * payload_bytes, np.uint8)
    pay[:buf.size] = buf ^ np.uint8(0x80) if (nbit == 8 and offset_binary8) else buf
    out[:, hdr:] = pay.reshape(n, payload_bytes)
    k = np.arange(n, dtype=np.int64) + int(ref_frame)
    sec = int(ref_second) + k // int(frames_per_second)
    frm = k % int(frames_per_second)
    words = np.zeros((n, hdr // 4), np.uint32)
    base = header_words(0, 0, fb, nbit, is_complex, epoch=epoch, thread=thread, station=station,
                        legacy=legacy)
    words[:] = base
    words[:, 0] = (words[:, 0] & np.uint32(0xC0000000)) | (sec & 0x3FFFFFFF).astype(np.uint32)
    words[:, 1] = (words[:, 1] & np.uint32(0xFF000000)) | (frm & 0xFFFFFF).astype(np.uint32)
    out[:, :hdr] = words.astype("<u4").view(np.uint8).reshape(n, hdr)
    return out


def set_header(buf, **fields):
    """Rewrite fields of the header at the start of buf in place (keys of
    decode_header)."""
    frame = buf
    h = decode_header(frame)
    h.update(fields)
    w = header_words(h["seconds"], h["frame"], h["frame_bytes"], h["bits"], h["complex"],
                     epoch=h["epoch"], thread=h["thread"], station=h["station"],
                     invalid=h["invalid"], legacy=h["legacy"],
                     log2_channels=int(h["channels"]).bit_length() - 1)
    frame[:16] = w[:4].astype("<u4").view(np.uint8)


def read_frames(path):
    """A VDIF file as a read-only uint8 memory map (frames back to back)."""
    return np.memmap(path, np.uint8, mode="r")
