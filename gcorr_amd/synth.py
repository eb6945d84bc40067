"""Seeded synthetic station voltages for the FX-correlation workloads.

Signal model (per station i, polarisation p):

    x_i(n) = Re{ z_p(n + d_i(n)) exp(+2 pi i lo tau_i(n)) } * sqrt(rho)
             + tones + sqrt(1 - rho) * noise_ip(n)

z_p is the analytic form of a common unit-variance white Gaussian signal
(one independent signal per polarisation), tau_i(t) the station's quadratic
delay model and d_i(n) = (tau_i - fileoffset_i)/Ts the net delay in samples.
Sample n of station i then holds the common signal at n*Ts + tau_i - f_i,
which is the alignment that fxkernel's integer shift, fringe rotation and
fractional-sample correction undo (fxkernel/src/fxkernel.cpp:288-324,
498-567, 595-634).  The fractional delay is applied per block of samples as
a phase ramp on the block's spectrum (blocks overlap by a guard band that is
discarded); the fringe phase is applied per sample in FP64.

Real data are quantised to 2 bits at +-0.98 sigma with levels
{-3.3359, -1, +1, +3.3359} and packed LSB first as [A_t, B_t, A_t+1, B_t+1]
(fxkernel.cpp:443-476; generators/src/bitconversion.c:119-163 packs the same
way).  Complex data are int8 (sigma 20, clipped to +-127) laid out
[Are, Aim, Bre, Bim] per time sample.

Runs on any torch device; the bench generates its input directly in HBM.
"""
import math
from dataclasses import dataclass, field, replace
from typing import List, Tuple

import numpy as np
import torch

HIMAG = 3.3359
THRESH = 0.98


@dataclass(frozen=True)
class Workload:
    """One subintegration's shape and signal parameters."""
    name: str
    nant: int
    nchan: int
    numffts: int
    bw: float
    lo: float = 1.65e9
    rho: float = 0.02
    is_complex: bool = False
    ntones: int = 0
    tone_amp: float = 0.1
    sigma8: float = 20.0
    seed: int = 20261202

    @property
    def nbit(self):
        return 8 if self.is_complex else 2

    @property
    def fftchannels(self):
        return self.nchan if self.is_complex else 2 * self.nchan

    @property
    def sampletime(self):
        st = 1.0 / (2.0 * self.bw)
        return 2 * st if self.is_complex else st

    @property
    def nsamples(self):
        """Time samples per station buffer: numffts+1 FFTs (fxkernel/src/common.cpp:16)."""
        return (self.numffts + 1) * self.fftchannels

    @property
    def station_bytes(self):
        return 4 * self.nsamples if self.is_complex else self.nsamples // 2

    @property
    def input_samples(self):
        """Samples (all stations x pols) processed per subintegration."""
        return self.nant * 2 * self.numffts * self.fftchannels

    @property
    def subint_seconds(self):
        return self.numffts * self.fftchannels * self.sampletime

    def scaled(self, numffts):
        return replace(self, numffts=int(numffts))


# The five workloads: 2-bit real at 4, 6 and 16 stations; complex int8 at 32.
WORKLOADS = {
    "c1": Workload("4st_2pol_128ch_32MSps_0.1s", nant=4, nchan=128, numffts=12500, bw=16e6,
                   rho=0.05, seed=20261202 + 0),
    "c2": Workload("6st_2pol_1024ch_128MSps_1s", nant=6, nchan=1024, numffts=62500, bw=64e6,
                   rho=0.02, seed=20261202 + 1),
    "c3": Workload("16st_2pol_4096ch_256MSps_1s_tones", nant=16, nchan=4096, numffts=31250,
                   bw=128e6, rho=0.01, ntones=3, seed=20261202 + 2),
    "c5": Workload("32st_2pol_32768ch_cplx8_512MSps_1s", nant=32, nchan=32768, numffts=15625,
                   bw=512e6, rho=0.005, is_complex=True, seed=20261202 + 4),
}

# Uneven subbands of the batched workload: (bandwidth Hz, nchan); 8 stations, 0.5 s.
SUBBANDS_C4 = [(16e6, 256), (16e6, 512), (32e6, 512), (32e6, 1024), (64e6, 1024), (64e6, 2048),
               (128e6, 2048), (128e6, 4096)]


def c4_subbands(seed=20261202 + 3):
    """The 8 subbands of the batched workload as Workloads (0.5 s each)."""
    out = []
    for k, (bw, nchan) in enumerate(SUBBANDS_C4):
        L = 2 * nchan
        numffts = int(round(0.5 * 2 * bw / L))
        out.append(Workload("c4_sb%d_%dMHz_%dch" % (k, int(bw / 1e6), nchan), nant=8, nchan=nchan,
                            numffts=numffts, bw=bw, rho=0.03, seed=seed * 16 + k))
    return out


def make_delays(w: Workload):
    """Seeded quadratic delays: |tau| <= 5 ms, |dtau/dt| <= 1e-6 over the subint.

    Returns (poly[nant][3] in fxkernel's convention d(t) = d0 t^2 + d1 t + d2,
    seconds with t in FFT intervals; fileoffsets[nant] seconds) chosen so the
    net delay tau - offset is within a few samples at mid-subintegration.
    """
    rng = np.random.default_rng(w.seed + 7)
    Ts, Tf = w.sampletime, w.fftchannels * w.sampletime
    T = w.subint_seconds
    c0 = rng.uniform(-5e-3, 5e-3, w.nant)
    c1 = rng.uniform(-5e-7, 5e-7, w.nant)
    c2 = rng.uniform(-2.5e-7, 2.5e-7, w.nant) / T
    poly = np.stack([c2 * Tf * Tf, c1 * Tf, c0], axis=1)
    tmid = T / 2
    tau_mid = c2 * tmid * tmid + c1 * tmid + c0
    offsets = tau_mid + rng.uniform(-3.0, 3.0, w.nant) * Ts
    return poly.astype(np.float64), offsets.astype(np.float64)


def _net_delay_samples(poly, offset, Ts, L, n):
    u = n.double() / L
    tau = (poly[0] * u + poly[1]) * u + poly[2]
    return (tau - offset) / Ts, tau


def _common_signal(w, pol, nlen, device):
    g = torch.Generator(device=device)
    g.manual_seed(int(w.seed) * 4 + pol)
    if w.is_complex:
        s = torch.complex(torch.randn(nlen, generator=g, device=device),
                          torch.randn(nlen, generator=g, device=device)) * math.sqrt(0.5)
    else:
        s = torch.randn(nlen, generator=g, device=device)
    return s * math.sqrt(w.rho)


def _tones(w, margin, nlen, device):
    if w.ntones == 0:
        return None
    rng = np.random.default_rng(w.seed + 11)
    freqs = rng.uniform(0.05, 0.95, w.ntones) * w.bw
    phases = rng.uniform(0, 2 * math.pi, w.ntones)
    m = torch.arange(nlen, device=device, dtype=torch.float64) - margin
    tot = torch.zeros(nlen, device=device, dtype=torch.float64)
    for f, ph in zip(freqs, phases):
        arg = torch.remainder(f * w.sampletime * m, 1.0) * (2 * math.pi) + ph
        tot += w.tone_amp * torch.cos(arg)
    return tot.float()


def _delayed_station(w, common, margin, poly, offset, device, block=3840, guard=128):
    """Analytic common signal at n + d_i(n), times the fringe phase; complex [nsamples]."""
    n_total = w.nsamples
    L = w.fftchannels
    N = block + 2 * guard
    nblk = (n_total + block - 1) // block
    centres = torch.arange(nblk, device=device, dtype=torch.float64) * block + block / 2
    d, _ = _net_delay_samples(poly, offset, w.sampletime, L, centres)
    D = torch.round(d)
    delta = (d - D).float()
    start = (torch.arange(nblk, device=device) * block + D.long() - guard + margin)
    out = torch.empty(nblk * block, dtype=torch.complex64, device=device)
    k = torch.arange(N, device=device, dtype=torch.float32)
    if not w.is_complex:
        weight = torch.zeros(N, device=device)
        weight[0] = 1.0
        weight[1:N // 2] = 2.0
        weight[N // 2] = 1.0
    step = 2048
    for b0 in range(0, nblk, step):
        b1 = min(nblk, b0 + step)
        idx = start[b0:b1, None] + torch.arange(N, device=device)[None, :]
        seg = common[idx]
        S = torch.fft.fft(seg, dim=1)
        ramp = torch.polar(torch.ones_like(delta[b0:b1, None]) * 1.0,
                           2 * math.pi * k[None, :] * delta[b0:b1, None] / N)
        if w.is_complex:
            Z = S * ramp
        else:
            Z = S * ramp * weight[None, :]
        z = torch.fft.ifft(Z, dim=1)[:, guard:guard + block]
        out[b0 * block:b1 * block] = z.reshape(-1)
    z = out[:n_total]
    # fringe phase exp(+2 pi i lo tau(n)), FP64 argument reduced to turns
    n = torch.arange(n_total, device=device, dtype=torch.float64)
    _, tau = _net_delay_samples(poly, offset, w.sampletime, L, n)
    turns = torch.remainder(w.lo * tau, 1.0)
    phase = (2 * math.pi * turns).float()
    return z * torch.polar(torch.ones_like(phase), phase)


def _quantise2(x, sigma):
    t = THRESH * sigma
    code = torch.full(x.shape, 1, dtype=torch.uint8, device=x.device)
    code[x >= 0] = 2
    code[x >= t] = 3
    code[x < -t] = 0
    return code


def pack2bit(code_a, code_b):
    """Pack 2-bit codes of two pols LSB first: [A_t, B_t, A_t+1, B_t+1] per byte."""
    a0, a1 = code_a[0::2], code_a[1::2]
    b0, b1 = code_b[0::2], code_b[1::2]
    return (a0 | (b0 << 2) | (a1 << 4) | (b1 << 6)).to(torch.uint8)


def generate(w: Workload, device="cpu", delays=None):
    """Station buffers (list of uint8 tensors on `device`), poly[nant][3], offsets[nant]."""
    poly, offsets = make_delays(w) if delays is None else delays
    L = w.fftchannels
    n_total = w.nsamples
    # net delay range over the buffer bounds the margin of the common signal
    ends = torch.tensor([0.0, n_total / 2, float(n_total)], dtype=torch.float64)
    dmax = 0.0
    for a in range(w.nant):
        d, _ = _net_delay_samples(poly[a], offsets[a], w.sampletime, L, ends)
        dmax = max(dmax, float(d.abs().max()))
    margin = int(dmax) + 3840 + 512
    nlen = n_total + 2 * margin
    tone = _tones(w, margin, nlen, device)
    sigma = math.sqrt(1.0 + (w.ntones * w.tone_amp ** 2 / 2.0))
    commons = []
    for pol in range(2):
        c = _common_signal(w, pol, nlen, device)
        if tone is not None:
            c = c + tone
        commons.append(c)
    out = []
    for a in range(w.nant):
        codes = []
        for pol in range(2):
            z = _delayed_station(w, commons[pol], margin, poly[a], offsets[a], device)
            g = torch.Generator(device=device)
            g.manual_seed(int(w.seed) * 1000 + 10 * a + pol + 17)
            if w.is_complex:
                noise = torch.complex(torch.randn(n_total, generator=g, device=device),
                                      torch.randn(n_total, generator=g, device=device))
                x = z + noise * math.sqrt((1.0 - w.rho) * 0.5)
                q = torch.clamp(torch.round(torch.view_as_real(x) * w.sigma8), -127, 127)
                codes.append(q.to(torch.int8))
            else:
                x = z.real + torch.randn(n_total, generator=g, device=device) * math.sqrt(1.0 - w.rho)
                codes.append(_quantise2(x, sigma))
            del z
        if w.is_complex:
            buf = torch.stack([codes[0], codes[1]], dim=1).reshape(-1).view(torch.uint8)
        else:
            buf = pack2bit(codes[0], codes[1])
        out.append(buf.contiguous())
    return out, poly, offsets


def random_stations(w: Workload, seed=1):
    """Uniformly random packed bytes per station (numpy), for parity tests."""
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 256, w.station_bytes, dtype=np.uint8) for _ in range(w.nant)]


def quantised_power_2bit():
    """E[x^2] of the 2-bit quantiser on unit Gaussian input (levels +-1, +-HIMAG, cut +-0.98)."""
    from math import erf, sqrt
    p_hi = 1.0 - erf(THRESH / sqrt(2.0))  # P(|x| > 0.98)
    return p_hi * HIMAG ** 2 + (1.0 - p_hi) * 1.0
