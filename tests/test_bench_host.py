"""Host logic of bench.py that runs without a GPU: the one-GPU guard and the
selection of the committed HBM-traffic figures."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("name", ["traffic_c2_fp32.json", "traffic_c3_fp32.json", "traffic_c3_fp16.json"])
def test_traffic_files_name_their_shape(name):
    with open(os.path.join(ROOT, "profiles", name)) as f:
        t = json.load(f)
    for key in ("fstage", "fblock", "fstage_launches_per_step", "fetch_scale", "write_scale", "source"):
        assert key in t
    assert os.path.exists(os.path.join(ROOT, t["source"]))


def test_traffic_used_only_for_the_profiled_shape():
    with open(os.path.join(ROOT, "profiles", "traffic_c2_fp32.json")) as f:
        t = json.load(f)
    fb, lps = t["fblock"], t["fstage_launches_per_step"]
    assert bench.traffic_from_profiles("c2", "fp32", fb, float(lps))["fstage"] == t["fstage"]
    # another F-stage shape, another chunking or a workload without a profile
    assert bench.traffic_from_profiles("c2", "fp32", fb + 1, float(lps)) is None
    assert bench.traffic_from_profiles("c2", "fp32", fb, 2.0 * lps) is None
    assert bench.traffic_from_profiles("c1", "fp32", fb, float(lps)) is None


def test_given_shapes():
    got = bench.given_shapes(["fp32:256,16,8", "fp16:-1,32,8,0"])
    assert got["fp32"] == {"fblock": 256, "xtile": 16, "xsplits": 8, "graph": True}
    assert got["fp16"] == {"fblock": -1, "xtile": 32, "xsplits": 8, "graph": False}
    for bad in ("fp32:256,16", "fp32:1,2,3,4,5", "fp64:1,2,3"):
        with pytest.raises(SystemExit):
            bench.given_shapes([bad])


@pytest.mark.parametrize("name", ["traffic_c3_fp32.json", "traffic_c3_fp16.json"])
def test_traffic_limit_cites_committed_counters(name):
    """The roofline's 'limit' note comes with the traffic figures and names
    the committed counter file it is read from."""
    with open(os.path.join(ROOT, "profiles", name)) as f:
        t = json.load(f)
    note = t["limit"]["fstage"]
    cited = [w.strip("()") for w in note.split() if w.strip("()").startswith("profiles/")]
    assert cited and all(os.path.exists(os.path.join(ROOT, c)) for c in cited), note
