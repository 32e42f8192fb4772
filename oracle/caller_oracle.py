"""CPU restatement of the burst driver's parameter derivation (test
infrastructure: only tests/ use it), in the reference's storage types:
numpy float32 where the Fortran variables or literals are default reals.

* burstparams, sentinel/backprojectgpusub.f90:406-439 (d8 :441-451): the
  literals 1.e-6 and 0.05546576 are default reals; fref is 37.53472224d0.
* fdopcoefs :262-300 and linfit :355-383: fref is assigned from the
  default-real literal 37.53472224 (FP32-rounded); ang, prf, t are default
  reals; linfit sums FP32 values and FP32 products in FP64 and returns FP32.
* the integration length :193-209.
* the line times and state-vector index of sentineltimingsub.f90:55-66, and
  the f16.7 / f14.8 text round trip of the position file the driver reads
  back (:176-179 of backprojectgpusub.f90).

The footprint bounds.f90 is not restated here: the reference's own Fortran
is compiled into oracle/_ref/libref_bounds.so (oracle/Makefile) and is the
checker.
"""
from __future__ import annotations

import math
import struct

import numpy as np

from . import range_oracle as RO

F32 = np.float32


def in3(b, i):
    return RO.in3(b, i)


def burstparams(header) -> dict:
    b = np.frombuffer(bytes(header), np.uint8)
    fref = 37.53472224
    c = 299792458.0
    e6 = float(F32(1.e-6))
    starttime = struct.unpack("<d", bytes(b[68:76]))[0]
    prf = fref / in3(b, 50) * 1.e6
    code = int(b[40])
    samplefreq = RO.samplefrequency(code) * 1.e6
    swst = in3(b, 53) / fref * e6
    irank = int(b[49]) & 31
    pri = in3(b, 50) / fref * e6
    range0 = c / 2.0 * (irank * pri + swst) + 60.0
    return {"starttime": starttime, "prf": prf, "samplefreq": samplefreq, "pri": pri,
            "swst": swst, "range0": range0, "wvl": float(F32(0.05546576)),
            "rangedecimation": code, "rank": irank}


def steering_fit(headers: np.ndarray):
    """fdopcoefs + linfit over lines 21..lines-20 (1-based)."""
    h = np.asarray(headers, np.uint8)
    lines = h.shape[0]
    fref = float(F32(37.53472224))
    sumx = sumx2 = sumxy = sumy = 0.0
    n = lines - 40
    for i in range(21, lines - 19):
        b = h[i - 1]
        word = RO.in2(b, 60)
        azb = word & 1023
        ang = F32(F32(F32(azb) / F32(1024.0)) * F32(1.8)) - F32(0.9)
        prf = F32(fref / in3(b, 50) * 1.e6)
        t = F32(F32(F32(i) - F32(F32(lines) / F32(2.0))) / prf)
        sumx = sumx + float(t)
        sumx2 = sumx2 + float(F32(t * t))
        sumxy = sumxy + float(F32(t * ang))
        sumy = sumy + float(ang)
    den = n * sumx2 - sumx ** 2
    c1 = F32((n * sumxy - sumx * sumy) / den)
    c0 = F32((sumy * sumx2 - sumx * sumxy) / den)
    return float(c0), float(c1)


def aperture(vel, range0, wvl, prf, angc1) -> int:
    pi = 4.0 * math.atan2(1.0, 1.0)
    re, ht = 6378137.0, 700000.0
    v = np.asarray(vel, np.float64)
    vmag = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * (re / (re + ht))
    at = (range0 * wvl / 2 / vmag / vmag * prf) / (1.0 + range0 * angc1 * pi / 180.0 / vmag)
    return int(at * prf * 0.5)


def line_time(header) -> float:
    return struct.unpack("<d", bytes(bytearray(header)[68:76]))[0]


def sv_index(t, timeorbit) -> int:
    """ilocation of sentineltimingsub.f90:62 (1-based, truncated)."""
    return int((t - timeorbit[0]) / (timeorbit[1] - timeorbit[0]))


def text_round_trip(x, fmt: str) -> float:
    """Value written with a Fortran F edit descriptor and read back."""
    return float(fmt % x)
