"""Fixture of the reference's example visibilities, examples/adam/vis.out
(2 stations, 4096 channels, NANT 2 in examples/adam/gv023.conf).

Reads that data file (run in a checkout that has the reference at
/root/reference) and stores its four (baseline, product) series of channel
visibilities as complex128 in adam_vis.npz, so the lag S/N check of
util/lagsnr.py:33-36 can be run on them without the reference.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.lagref import read_vis_out  # noqa: E402

SRC = "/root/reference/examples/adam/vis.out"

if __name__ == "__main__":
    vis = read_vis_out(SRC)
    np.savez_compressed(os.path.join(HERE, "adam_vis.npz"), vis=vis)
    print(vis.shape)
