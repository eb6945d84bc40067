"""The fxkernel .conf format (fxkernel/src/common.cpp:54-114) via gxc_parse_config."""
import ctypes
import os
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _lib


@dataclass
class CorrConfig:
    nbit: int
    npol: int
    is_complex: bool
    nchan: int
    nant: int
    lo: float
    bw: float
    numffts: int
    antennas: List[str] = field(default_factory=list)
    files: List[str] = field(default_factory=list)
    delays: np.ndarray = None   # [nant][3], seconds, t in FFT intervals
    offsets: np.ndarray = None  # [nant], seconds

    @property
    def fftchannels(self):
        return self.nchan if self.is_complex else 2 * self.nchan

    @property
    def sampletime(self):
        st = 1.0 / (2.0 * self.bw)
        return st * 2 if self.is_complex else st


def parse_config(path):
    """parseConfig: returns a CorrConfig; raises GxcError for an unparsable file."""
    L = _lib.load()
    cf = _lib.ConfFile()
    _lib.check(L.gxc_parse_config(os.fsencode(path), ctypes.byref(cf)))
    try:
        n = cf.cfg.nant
        names = [cf.names[i].decode() for i in range(cf.nant_read)]
        files = [cf.files[i].decode() for i in range(cf.nant_read)]
        delays = np.array([[cf.delays[3 * i + k] for k in range(3)] for i in range(n)],
                          np.float64).reshape(n, 3)
        offsets = np.array([cf.offsets[i] for i in range(n)], np.float64)
        c = cf.cfg
        return CorrConfig(nbit=c.nbit, npol=c.npol, is_complex=bool(c.is_complex), nchan=c.nchan,
                          nant=n, lo=c.lo, bw=c.bw, numffts=c.numffts, antennas=names, files=files,
                          delays=delays, offsets=offsets)
    finally:
        L.gxc_conffile_free(ctypes.byref(cf))


def write_config(path, cfg):
    """Write a CorrConfig in the same keyword format."""
    with open(path, "w") as f:
        f.write("NBIT %d\nNPOL %d\nCOMPLEX %d\nNCHAN %d\nLO %.17g\nBANDWIDTH %.17g\nNUMFFTS %d\nNANT %d\n"
                % (cfg.nbit, cfg.npol, int(cfg.is_complex), cfg.nchan, cfg.lo, cfg.bw, cfg.numffts,
                   cfg.nant))
        for i in range(cfg.nant):
            d = cfg.delays[i]
            f.write("%s %s %.17g %.17g %.17g %.17g\n" % (cfg.antennas[i], cfg.files[i], d[0], d[1], d[2],
                                                          cfg.offsets[i]))
