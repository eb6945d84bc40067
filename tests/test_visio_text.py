"""vis.out reader and the verifyfxkernel-style comparator (CPU).

The files are written here in the text layout of
fxkernel/src/fxkernel.cpp:389-399 (%5d channel, %11.6f MHz, then per row and
product %11.5f re, %11.5f im, %11.5f amplitude, %10.5f phase).  The
comparator is pinned by known injected errors: the fraction it reports for
the damaged series is the injected size over that series' median amplitude,
every other series stays at the text format's rounding, and the cases the
reference script's signed arithmetic gets wrong are caught.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from gcorr_amd import visio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def write_vis_out(path, vis, bw):
    nrows, nprod, nchan = vis.shape
    with open(path, "w") as f:
        for c in range(nchan):
            f.write("%5d %11.6f" % (c, (c + 0.5) / nchan * bw / 1e6))
            for r in range(nrows):
                for p in range(nprod):
                    x = vis[r, p, c]
                    f.write(" %11.5f %11.5f %11.5f %10.5f" % (x.real, x.imag, abs(x), np.angle(x)))
            f.write("\n")


def _vis(nrows=6, nchan=64, seed=3):
    rng = np.random.default_rng(seed)
    v = (rng.normal(size=(nrows, 4, nchan)) + 1j * rng.normal(size=(nrows, 4, nchan))) * 10 + 30
    return v.astype(np.complex64)


def test_read_text_round_trip(tmp_path):
    v = _vis()
    p = tmp_path / "vis.out"
    write_vis_out(p, v, 64e6)
    r = visio.read_text(str(p))
    assert r["vis"].shape == v.shape
    np.testing.assert_array_equal(r["channel"], np.arange(v.shape[2]))
    np.testing.assert_allclose(r["freq_mhz"], (np.arange(64) + 0.5) / 64 * 64, atol=5e-7)
    # %11.5f: each component within half a unit of the fifth decimal, plus
    # the complex64 rounding of the parsed value
    for comp in ("real", "imag"):
        x, y = getattr(r["vis"], comp), getattr(v, comp)
        assert np.all(np.abs(x.astype(np.float64) - y) <= 5.0001e-6 + np.spacing(np.abs(y)))
    np.testing.assert_allclose(r["amp"], np.abs(v), atol=1e-5)


def test_read_text_rejects_other_layouts(tmp_path):
    p = tmp_path / "bad.out"
    p.write_text("0 1.0 2.0 3.0\n1 1.0 2.0 3.0\n")
    with pytest.raises(ValueError):
        visio.read_text(str(p))


def test_compare_identical_files_pass(tmp_path):
    v = _vis()
    a, b = tmp_path / "a.out", tmp_path / "b.out"
    write_vis_out(a, v, 64e6)
    write_vis_out(b, v, 64e6)
    r = visio.compare_text(str(a), str(b), 6)
    assert r["pass"] and r["max_abs_diff"] == 0.0 and r["fail"] == []


@pytest.mark.parametrize("row,prod,chan,delta", [(2, 1, 7, 0.01 + 0.0j), (5, 3, 63, -0.02j), (0, 0, 0, 0.5)])
def test_compare_finds_injected_error(tmp_path, row, prod, chan, delta):
    v = _vis()
    w = v.copy()
    w[row, prod, chan] += delta
    a, b = tmp_path / "a.out", tmp_path / "b.out"
    write_vis_out(a, v, 64e6)
    write_vis_out(b, w, 64e6)
    r = visio.compare_text(str(a), str(b), 6, maxepsilon=1e-4)
    expect = abs(delta) / np.median(np.abs(v[row, prod].astype(np.complex128)))
    assert [(x[0], x[1]) for x in r["fail"]] == [(row, prod)]
    assert r["worst"][:2] == (row, prod)
    # the text format rounds each component to 1e-5
    assert abs(r["worst"][2] - expect) <= 2e-5 / np.median(np.abs(v[row, prod]))
    others = np.delete(r["fraction"].reshape(-1), row * 4 + prod)
    assert others.max() < 1e-6


def test_compare_catches_what_signed_arithmetic_misses(tmp_path):
    """A series with a negative median real part and a negative difference:
    verifyfxkernel.py:65-69 divides signed maxima by a signed median and lets
    it through; |delta| over the median amplitude does not."""
    v = _vis()
    v[1, 2] = -np.abs(v[1, 2].real) - 1j * np.abs(v[1, 2].imag)
    w = v.copy()
    w[1, 2, 10] -= 3.0
    a, b = tmp_path / "a.out", tmp_path / "b.out"
    write_vis_out(a, v, 64e6)
    write_vis_out(b, w, 64e6)
    r = visio.compare_text(str(a), str(b), 6, maxepsilon=1e-3)
    assert not r["pass"] and (1, 2) in [(x[0], x[1]) for x in r["fail"]]


def test_compare_zero_reference_series(tmp_path):
    """A row with no valid FFT is zero: any difference there fails."""
    v = _vis()
    v[4] = 0
    w = v.copy()
    w[4, 0, 3] = 1e-3
    a, b = tmp_path / "a.out", tmp_path / "b.out"
    write_vis_out(a, v, 64e6)
    write_vis_out(b, w, 64e6)
    r = visio.compare_text(str(a), str(b), 6)
    assert [(x[0], x[1]) for x in r["fail"]] == [(4, 0)] and r["fail"][0][2] == np.inf


def test_verify_cli(tmp_path):
    v = _vis(nrows=3)
    a, b, c = tmp_path / "a.out", tmp_path / "b.out", tmp_path / "c.out"
    write_vis_out(a, v, 16e6)
    write_vis_out(b, v, 16e6)
    w = v.copy()
    w[2, 1, 5] += 1.0
    write_vis_out(c, w, 16e6)
    conf = tmp_path / "x.conf"
    conf.write_text("NCHAN 64\nNANT 3\nNPOL 2\n")
    run = lambda *f: subprocess.run([sys.executable, "-m", "gcorr_amd.verify", "--config", str(conf),
                                     "--maxepsilon", "1e-4"] + [str(x) for x in f],
                                    cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    ok = run(a, b)
    assert ok.returncode == 0 and ok.stdout.startswith("PASS")
    bad = run(a, c)
    assert bad.returncode == 1 and "FAIL: Baseline 2 pol product 1" in bad.stdout
    missing = run(a, tmp_path / "none.out")
    assert missing.returncode == 2
