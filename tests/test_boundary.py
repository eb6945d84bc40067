"""The C-ABI boundary on the CPU: the library loads, exports every function
include/gxcorr.h declares, validates configurations, and reads the
reference's .conf fixtures exactly as fxkernel/src/common.cpp:54-114 does.
No kernel runs here."""
import ctypes
import glob
import os
import subprocess

import numpy as np
import pytest

import gcorr_amd
from gcorr_amd import _lib
from conftest import GOLDEN, ROOT

CONF = os.path.join(GOLDEN, "conf")


def test_library_exports_every_declared_function():
    L = _lib.load()
    names = _lib.declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n


def test_library_is_gfx950_only_code():
    # the shared object carries a gfx950 code object and no other target
    with open(_lib.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"gfx950" in blob
    for other in (b"gfx942", b"gfx90a", b"gfx1100"):
        assert other not in blob


def test_create_without_device_or_bad_config_fails_cleanly():
    def create(**kw):
        c = _lib.Config()
        _lib.load().gxc_config_init(ctypes.byref(c))
        c.nant, c.nchan, c.numffts, c.bw, c.lo = 4, 128, 10, 16e6, 1.65e9
        for k, v in kw.items():
            setattr(c, k, v)
        h = ctypes.c_void_p()
        return _lib.load().gxc_create(ctypes.byref(c), ctypes.byref(h))

    assert create(nchan=96) == 1          # not a power of two
    assert create(numffts=0) == 1
    assert create(nbit=4) == 2            # fxkernel supports 2-bit real only
    assert create(npol=1) == 2
    assert create(is_complex=1) == 2      # complex must be 8-bit
    assert create(nchan=32768) == 2       # real input: FFT length 65536 is beyond 32768
    assert create(precision=7) == 1


def test_station_bytes_closed_form():
    # common.cpp:16: nchan * cfactor * (numffts+1) * nbit/8 * npol
    assert gcorr_amd.station_bytes(4, 1024, 3200) == 1024 * 2 * 3201 * 2 // 8 * 2
    assert gcorr_amd.station_bytes(6, 1024, 62500) == 64001024
    # complex 8-bit: 4 bytes per time sample (2 pols x re, im)
    assert gcorr_amd.station_bytes(8, 1024, 1024, nbit=8, is_complex=True) == 4 * 1024 * 1025
    assert gcorr_amd.num_rows(6) == 15 and gcorr_amd.num_rows(6, autos=True) == 21


def test_parse_reference_test_conf():
    c = gcorr_amd.parse_config(os.path.join(CONF, "fxkernel_test.conf"))
    assert (c.nbit, c.npol, c.is_complex, c.nchan, c.nant) == (2, 2, False, 1024, 4)
    assert (c.lo, c.bw, c.numffts) == (1650000000.0, 32000000.0, 3200)
    assert c.antennas == ["Aa", "Bb", "Cc", "Dd"]
    assert c.files == ["aa.bin", "bb.bin", "cc.bin", "dd.bin"]
    np.testing.assert_array_equal(c.delays[0], [6.1e-19, 4.8e-15, 100.1e-9])
    np.testing.assert_array_equal(c.offsets, [100e-9, 110e-9, 120e-9, 130e-9])


def test_parse_defaults_and_complex():
    c = gcorr_amd.parse_config(os.path.join(CONF, "gcorr_bench", "test6.conf"))
    assert c.npol == 2 and c.nant == 6 and c.nchan == 1024  # no NPOL line: default 2
    c = gcorr_amd.parse_config(os.path.join(CONF, "gcorr_bench", "test8-8.conf"))
    assert c.is_complex and c.nbit == 8 and c.nant == 8 and c.bw == 128e6
    c = gcorr_amd.parse_config(os.path.join(CONF, "gv023.conf"))
    assert (c.lo, c.bw, c.nchan, c.numffts, c.nant) == (8.24e9, 64e6, 4096, 4096, 2)
    np.testing.assert_array_equal(c.delays[1], [-2.76004e-19, 1.40150e-10, 1.04602069e-02])


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(CONF, "*", "*.conf"))))
def test_parse_every_bench_conf(path):
    c = gcorr_amd.parse_config(path)
    assert c.nant == len(c.antennas) and c.nant >= 4
    assert c.nchan > 0 and (c.nchan & (c.nchan - 1)) == 0 and c.numffts > 0
    assert c.delays.shape == (c.nant, 3)
    # the same file read as plain whitespace-separated text
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    ant = lines[-c.nant:]
    for i, tok in enumerate(ant):
        assert tok[0] == c.antennas[i]
        np.testing.assert_array_equal(c.delays[i], [float(x) for x in tok[2:5]])


def test_parse_unparsable_line_is_an_error(tmp_path):
    p = tmp_path / "bad.conf"
    p.write_text("NBIT 2\n\nNCHAN 64\n")
    with pytest.raises(gcorr_amd.GxcError):
        gcorr_amd.parse_config(str(p))
    with pytest.raises(gcorr_amd.GxcError):
        gcorr_amd.parse_config(str(tmp_path / "missing.conf"))


def test_parse_unknown_keyword_is_skipped(tmp_path):
    p = tmp_path / "x.conf"
    p.write_text("FOO 3\nnchan 256\nNANT 1\nA a.bin 0 0 1e-3 1e-3\n")
    c = gcorr_amd.parse_config(str(p))
    assert c.nchan == 256 and c.nant == 1 and c.delays[0][2] == 1e-3


def test_write_config_roundtrip(tmp_path):
    c = gcorr_amd.parse_config(os.path.join(CONF, "fxkernel_test.conf"))
    p = tmp_path / "rt.conf"
    gcorr_amd.write_config(str(p), c)
    d = gcorr_amd.parse_config(str(p))
    np.testing.assert_array_equal(c.delays, d.delays)
    np.testing.assert_array_equal(c.offsets, d.offsets)
    assert (c.nchan, c.numffts, c.lo, c.bw) == (d.nchan, d.numffts, d.lo, d.bw)


def test_cli_usage():
    exe = os.path.join(ROOT, "gcorr_amd", "gxcorr_run")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "Usage" in r.stdout


def test_product_does_not_reference_oracle():
    """The product path never loads the oracle: no import, no link."""
    import re
    for f in glob.glob(os.path.join(ROOT, "gcorr_amd", "*.py")):
        with open(f) as fh:
            assert not re.search(r"^\s*(from|import)\s+oracle", fh.read(), re.M), f
    with open(_lib.LIB_PATH, "rb") as fh:
        assert b"fxref" not in fh.read()


def test_batch_create_checks_config_before_device():
    L = _lib.load()

    def create(subs, **kw):
        c = _lib.BatchConfig()
        L.gxc_batch_config_init(ctypes.byref(c))
        c.nant = 8
        arr = (_lib.Subband * max(1, len(subs)))(*[_lib.Subband(*s) for s in subs])
        c.nsub = len(subs)
        c.subbands = arr
        for k, v in kw.items():
            setattr(c, k, v)
        h = ctypes.c_void_p()
        return L.gxc_batch_create(ctypes.byref(c), ctypes.byref(h))

    ok = [(256, 100, 1.65e9, 16e6), (4096, 50, 1.65e9, 128e6)]
    assert create([]) == 1                                     # no subbands
    assert create(ok, nbit=8, is_complex=1) == 2               # real 2-bit only
    assert create([(64, 100, 1.65e9, 16e6)]) == 2              # FFT length 128 < 256
    assert create([(8192, 10, 1.65e9, 16e6)]) == 2             # FFT length 16384 > 8192
    assert create([(100, 10, 1.65e9, 16e6)]) == 1              # not a power of two
    assert create(ok, fstage_mfma=3) == 1                      # own launches: 0, 1 or 2
    assert create(ok, fstage_mfma=-1) == 1
    assert create(ok) == 3                                     # valid, but no GPU here
    assert create(ok, fstage_mfma=2) == 3
    msg = L.gxc_batch_last_error(None).decode()
    assert "device" in msg


def test_autotune_null_handle_is_an_error():
    L = _lib.load()
    t = _lib.Tuning()
    assert L.gxc_autotune(None, ctypes.byref(t)) == 1
    assert L.gxc_batch_autotune(None, ctypes.byref(t)) == 1
