"""Compile the gfx950 HIP library (gcorr_amd/libgxcorr.so) and the CLI.

Run as ``python -m gcorr_amd.build``.  Every call recompiles.  Objects go to
gcorr_amd/_build/ (not committed); the shared library is written in-tree so it
travels with the source tree to the machine that runs it.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgxcorr.so")
CLI = os.path.join(PKG, "gxcorr_run")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for every HIP translation unit; the
# compiler still emits scalar FP32 FMA/ADD/MUL.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

SOURCES = ["fstage.hip", "fstage_half.hip", "fstage_double.hip", "fstage_big.hip", "fstage_mfma.hip", "fstage_multi.hip", "xengine.hip", "counts.hip", "vdif.hip", "lags.hip", "polconv.hip", "requant.hip", "gxc_noalloc.hip", "gxc_api.cpp", "gxc_vdif_api.cpp", "gxc_vdif_state.cpp", "gxc_convert.cpp", "gxc_stream.cpp", "gxc_output.cpp", "gxc_tune.cpp", "gxc_host.cpp", "gxc_batch.cpp", "config.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), r.stdout))
    return r.stdout


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    _run([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", out])
    return out


def build_library(jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_cli():
    src = os.path.join(CSRC, "gxcorr_run.cpp")
    tmp = CLI + ".tmp"
    _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", tmp,
          "-L", PKG, "-lgxcorr", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CLI)
    return CLI


def build_oracle():
    _run(["make", "-B", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_library()
    build_cli()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIB, CLI)
    sys.exit(0)
