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


@dataclass
class RawScene:
    """A burst of decoded lines made self-consistent with the driver's own
    derivation: per-line headers (times, chirp, range window codes), the orbit
    state vectors, the satellite positions the pipeline will compute, a DEM
    grid around the footprint, and point targets on grid pixels inside the
    burst's focusable zone."""

    headers: np.ndarray        # uint8 [L][80]
    chirp: "tdbp.Chirp"
    timeorbit: np.ndarray
    xx: np.ndarray
    vv: np.ndarray
    times: np.ndarray
    satloc: np.ndarray         # [L][3], quantised as the position file
    dem: np.ndarray            # int16 [length][width]
    firstlat: float
    firstlon: float
    deltalat: float
    deltalon: float
    geom: "tdbp.BurstGeometry"
    targets_xyz: np.ndarray
    targets_amp: np.ndarray
    targets_px: np.ndarray     # grid (row, col)

    @property
    def L(self):
        return self.headers.shape[0]


def _quantize(a, s):
    ai = np.trunc(a)
    return (ai * s + np.rint((a - ai) * s)) / s


def raw_burst_scene(L: int = 600, nsamps: int = 6000, centre_range: float = 850e3,
                    posting: float = 0.0004, ntargets: int = 4, seed: int = 0,
                    t0: float = 43200.0, heading_deg: float = 0.0) -> RawScene:
    from . import synth
    rng = np.random.default_rng(seed)
    c = 299792458.0
    orbit = synth.make_orbit(35.0, -118.0, t0, heading_deg=heading_deg, seed=seed)
    base = HeaderCodes(nsamps=nsamps)
    bp = tdbp.burst_decode(header_bytes(base))
    chirp = tdbp.chirp_decode(header_bytes(base))
    S = chirp.nvalid + HDR
    dmrg = c / 2.0 / bp.samplefreq
    range0 = centre_range - 0.5 * S * dmrg
    swst = (range0 - 60.0) * 2.0 / c - base.rank * bp.pri
    code = int(round(swst / float(np.float32(1e-6)) * FREF))
    if not 0 <= code < (1 << 24):
        raise ValueError(f"range window code {code}")
    prf = bp.prf
    tstart = t0 - 0.5 * (L - 1) / prf
    times = tstart + np.arange(L) / prf
    headers = np.zeros((L, 80), np.uint8)
    for i in range(L):
        headers[i] = np.frombuffer(header_bytes(HeaderCodes(nsamps=nsamps, swst=code,
                                                            starttime=float(times[i]))), np.uint8)
    times = np.array([struct.unpack("<d", bytes(h[68:76]))[0] for h in headers])
    tmid = 0.5 * (times[0] + times[0] + (L - 1) / prf)
    timeorbit, xx, vv = orbit.state_vectors(tmid)
    sat, vel = tdbp.orbit_interp(timeorbit, xx, vv, times)
    sat, vel = _quantize(sat, 1e7), _quantize(vel, 1e8)
    bp = tdbp.burst_decode(headers[0])
    rngstart = (2 * bp.range0 / c) * c / 2.0
    rngend = rngstart + (S - 1) * dmrg
    ll = tdbp.bounds(times[0] - 2.0, times[0] + (L - 1) / prf + 2.0, rngstart, rngend,
                     timeorbit, xx, vv)
    m = 4 * posting
    firstlat = ll[1] + m
    firstlon = ll[2] - m
    length = int(math.ceil((ll[1] - ll[0] + 2 * m) / posting))
    width = int(math.ceil((ll[3] - ll[2] + 2 * m) / posting))
    dem = np.zeros((length, width), np.int16)
    geom, ll2, ins = tdbp.burst_setup(headers, S, vel[L // 2 - 1], timeorbit, xx, vv, width,
                                      length, firstlat, firstlon, -posting, posting)
    # targets on pixels whose zero-Doppler pulse and range sit well inside
    A = geom.A
    rows, cols = [], []
    for _ in range(5000):
        if len(rows) == ntargets:
            break
        r = int(rng.integers(geom.firstline, geom.firstline + geom.nlines))
        col = int(rng.integers(geom.firstpix, geom.lastpix + 1))
        x = synth.llh_to_ecef(firstlat - r * posting, firstlon + col * posting, 0.0)
        tz = synth.zero_doppler_time(orbit, x, tmid)
        p = (tz - times[0]) * prf
        R = float(np.linalg.norm(x - orbit.pos(tz)))
        i = (R - rngstart) / dmrg
        if A // 2 + 8 <= p <= L - 1 - A // 2 - 8 and 200 <= i <= S - 200:
            rows.append(r)
            cols.append(col)
    rows = np.array(rows)
    cols = np.array(cols)
    txyz = synth.llh_to_ecef(firstlat - rows * posting, firstlon + cols * posting,
                             np.zeros(len(rows))).reshape(-1, 3)
    amp = np.exp(1j * rng.uniform(-np.pi, np.pi, len(rows)))
    return RawScene(headers=headers, chirp=chirp, timeorbit=timeorbit, xx=xx, vv=vv,
                    times=times, satloc=sat, dem=dem, firstlat=firstlat, firstlon=firstlon,
                    deltalat=-posting, deltalon=posting, geom=geom, targets_xyz=txyz,
                    targets_amp=amp, targets_px=np.stack([rows, cols], 1))


def device_raw_scene(rs: RawScene, d_lines: "tdbp.DeviceArray", noise: float = 0.0,
                     seed: int = 0):
    """The scene's decoded lines into d_lines (cf32 [L][stride]): chirp echoes
    of the targets from the quantised positions, per-line headers."""
    g = rs.geom
    R = np.linalg.norm(rs.satloc[:, None, :] - rs.targets_xyz[None, :, :], axis=2)
    delay = (R - g.rngstart) / g.dmrg - HDR
    cyc = R * (2.0 / g.wvl)
    tdbp.synth_raw(d_lines, rs.chirp, bytes(rs.headers[0]), delay, cyc - np.floor(cyc),
                   rs.targets_amp, noise, seed)
    tdbp.headers_htod(d_lines, rs.headers)
