"""gxcorr_run end to end: the testfxkernel program (fxkernel/src/testfxkernel.cpp:21-124)
built against the C++ facade, on the GPU.

A .conf and one packed file per station are written, gxcorr_run correlates
them, and its vis.out (text) and binary output are checked against the
oracle: FP32 visibilities within 1e-5, row counts exact, the text within its
printed precision of the binary values, and the verifyfxkernel-style
comparator passing against the oracle's own vis.out.
"""
import os
import subprocess

import numpy as np
import pytest

import gcorr_amd
from gcorr_amd import metrics, synth, visio
from oracle import fxref

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gcorr_amd", "gxcorr_run")


def _write_vis_out(path, vis, bw):
    nrows, nprod, nchan = vis.shape
    with open(path, "w") as f:
        for c in range(nchan):
            f.write("%5d %11.6f" % (c, (c + 0.5) / nchan * bw / 1e6))
            for r in range(nrows):
                for p in range(nprod):
                    x = complex(vis[r, p, c])
                    f.write(" %11.5f %11.5f %11.5f %10.5f" % (x.real, x.imag, abs(x), np.angle(x)))
            f.write("\n")


@pytest.mark.parametrize("key,nfft", [("c2", 200), ("c1", 700)])
def test_gxcorr_run_matches_oracle(gpu, tmp_path, key, nfft):
    w = synth.WORKLOADS[key].scaled(nfft)
    bufs, poly, offs = synth.generate(w, device="cpu")
    data = [b.numpy() for b in bufs]
    files = []
    for a, d in enumerate(data):
        f = tmp_path / ("st%d.bin" % a)
        d.tofile(f)
        files.append(str(f))
    conf = tmp_path / "corr.conf"
    gcorr_amd.write_config(str(conf), gcorr_amd.CorrConfig(
        nbit=2, npol=2, is_complex=False, nchan=w.nchan, nant=w.nant, lo=w.lo, bw=w.bw,
        numffts=w.numffts, antennas=["S%d" % a for a in range(w.nant)], files=files,
        delays=poly, offsets=offs))
    print("gxcorr_run", key, nfft, flush=True)
    r = subprocess.run([EXE, str(conf), str(tmp_path / "vis.bin")], cwd=tmp_path, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Got NANT %d" % w.nant in r.stdout
    assert (tmp_path / "runtime.fxkernel.log").read_text().startswith("Run time was")

    cfg = fxref.make_config(w.nant, w.nchan, w.numffts, w.lo, w.bw)
    ref, rc = fxref.process(cfg, data, poly, offs, "f32")
    b = visio.read_binary(str(tmp_path / "vis.bin"))
    np.testing.assert_array_equal(b["counts"], rc)
    assert b["vis"].shape == ref.shape
    assert metrics.eps(b["vis"], ref) < 1e-5

    t = visio.read_text(str(tmp_path / "vis.out"))
    assert t["vis"].shape == ref.shape
    for comp in ("real", "imag"):
        x, y = getattr(t["vis"], comp).astype(np.float64), getattr(b["vis"], comp).astype(np.float64)
        assert np.all(np.abs(x - y) <= 5.0001e-6 + 2 * np.spacing(np.abs(y).astype(np.float32)))

    _write_vis_out(tmp_path / "oracle.out", ref, w.bw)
    cmp = visio.compare_text(str(tmp_path / "oracle.out"), str(tmp_path / "vis.out"), len(rc), maxepsilon=1e-4)
    assert cmp["pass"], cmp["worst"]
