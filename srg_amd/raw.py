"""Decoded Sentinel-1 echo lines for the range-compression path.

A decoded line (the input of sentinel/processsub.f90:108) is 30000 complex
float samples: the first 10 (80 bytes) carry the packet header, the echo
samples follow.  `header_bytes` writes the fields the range compression and
the burst driver read, at the byte offsets they read them
(processsub.f90:75-86, sentinel/backprojectgpusub.f90:262-300 and :406-429): big-endian
integers, sign-and-magnitude for the chirp ramp rate and start frequency,
the start time as a native double at byte 68.

`IW1` holds published Sentinel-1 IW1 chirp values (annotation of IW SLC
products: pulse length 52.40481 us, ramp rate 1.078230e12 Hz/s, start
frequency -28.25216 MHz, range decimation code 8 -> 64.345238 MHz); the
header codes are those values quantised the way the instrument encodes them.

`synth_lines` (numpy, small) and `device_lines` (GPU kernel, full bursts)
make lines whose echoes are delayed copies of the transmitted chirp from
point targets of a BurstScene, plus white Gaussian receiver noise; a target
at slant range R appears after range compression at backprojection index
(R - rngstart) / dmrg, counting the 10 header samples as the reference does.
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass

import numpy as np

from . import tdbp

FREF = 37.53472224          # MHz, processsub.f90:34
RAW_STRIDE = 30000          # complex samples per decoded line
HDR = 10                    # header samples


@dataclass
class HeaderCodes:
    """Integer header fields (the instrument's codes)."""

    rangedecimation: int = 8
    txprr: int = 0x8000 | 1605      # ramp rate: sign bit (1 = positive) + 15 bits
    txpsf: int = 12335              # start frequency: sign bit 0 = negative
    txpl: int = 1967                # pulse length in 1/fref us
    nsamps: int = 24400             # echo samples (written as nsamps / 2)
    pri: int = 21859                # pulse repetition interval in 1/fref us
    rank: int = 9
    swst: int = 0
    starttime: float = 0.0
    azbeam: int = 512               # azimuth beam address (mid)


IW1 = HeaderCodes()


def header_bytes(c: HeaderCodes) -> bytes:
    h = bytearray(80)
    h[40] = c.rangedecimation & 255
    h[42:44] = struct.pack(">H", c.txprr)
    h[44:46] = struct.pack(">H", c.txpsf)
    h[46:49] = struct.pack(">I", c.txpl)[1:]
    h[49] = c.rank & 31
    h[50:53] = struct.pack(">I", c.pri)[1:]
    h[53:56] = struct.pack(">I", c.swst)[1:]
    h[60:62] = struct.pack(">H", c.azbeam & 1023)
    h[65:67] = struct.pack(">H", c.nsamps // 2)
    h[68:76] = struct.pack("<d", c.starttime)
    return bytes(h)


def header_samples(h: bytes) -> np.ndarray:
    """The 80 header bytes as the 10 complex64 samples that carry them."""
    return np.frombuffer(h, np.complex64).copy()


def chirp_of(codes: HeaderCodes = IW1) -> tdbp.Chirp:
    return tdbp.chirp_decode(header_bytes(codes))


@dataclass
class EchoModel:
    """Per pulse and target: delay (raw samples after the header) and the
    carrier phase exp(-j 4 pi R / lambda)."""

    delay: np.ndarray      # [L][T] float64
    carrier: np.ndarray    # [L][T] float64 (cycles, fractional part)
    amp: np.ndarray        # [T] complex


def echo_model(sc, c: tdbp.Chirp) -> EchoModel:
    """Delays from the scene's per-pulse satellite positions: compressed index
    (R - rngstart) / dmrg = delay + 10."""
    g = sc.geom
    R = np.linalg.norm(sc.satloc[:, None, :] - sc.targets_xyz[None, :, :], axis=2)
    delay = (R - g.rngstart) / g.dmrg - HDR
    cyc = R * (2.0 / g.wvl)
    return EchoModel(delay=delay, carrier=cyc - np.floor(cyc), amp=np.asarray(sc.targets_amp))


def chirp_wave(c: tdbp.Chirp, u: np.ndarray) -> np.ndarray:
    """Transmitted chirp at fractional sample offsets u (0 <= u < npts)."""
    t = u / float(c.samplefreq)
    ph = 2.0 * math.pi * float(c.startfreq) * t + math.pi * float(c.ramprate) * t * t
    return np.exp(1j * ph)


def synth_lines(sc, c: tdbp.Chirp, header: bytes, noise: float = 0.0, seed: int = 0,
                stride: int = RAW_STRIDE) -> np.ndarray:
    """Decoded lines (complex64 [L][stride]) in numpy: small bursts only."""
    m = echo_model(sc, c)
    L = sc.L
    out = np.zeros((L, stride), np.complex128)
    n = np.arange(c.nsamps, dtype=np.float64)
    for p in range(L):
        for t in range(m.delay.shape[1]):
            u = n - m.delay[p, t]
            ok = (u >= 0) & (u < c.npts)
            out[p, HDR + np.nonzero(ok)[0]] += (m.amp[t] *
                                                np.exp(-2j * math.pi * m.carrier[p, t]) *
                                                chirp_wave(c, u[ok]))
    if noise > 0:
        rng = np.random.default_rng(seed)
        out[:, HDR:HDR + c.nsamps] += (rng.standard_normal((L, c.nsamps)) +
                                       1j * rng.standard_normal((L, c.nsamps))) * (
            noise / math.sqrt(2.0))
    out = out.astype(np.complex64)
    out[:, :HDR] = header_samples(header)
    return out


def device_lines(sc, c: tdbp.Chirp, header: bytes, d_raw: "tdbp.DeviceArray",
                 noise: float = 0.0, seed: int = 0, stream=None):
    """The same model on the GPU into d_raw (cf32 [L][stride])."""
    m = echo_model(sc, c)
    tdbp.synth_raw(d_raw, c, header, m.delay, m.carrier, m.amp, noise, seed, stream)
