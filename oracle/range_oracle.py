"""CPU restatement of the range compression of sentinel/processsub.f90:75-126
(test infrastructure: only tests/, smoke() and bench.py's CPU baseline use it).

Header decoding follows processsub.f90:75-86 with the readers in2/in3 of
:163-173 and the sampling-rate table of :181-212, keeping the reference's
storage types: samplefreq, ramprate, startfreq and pulselength are real*4
(FP32) variables, the table ratios are single-precision literals (3./7.*4.),
npts = int(pulselength * samplefreq) is an FP32 product truncated.  The
reference function follows :96-102 (t = (j-1)/samplefreq in FP32, phase in
FP64, cos/sin rounded to FP32, divided by npts in FP32).  The per-line
transform follows :108-123 but in FP64: numpy's FFT of the zero-padded echo,
times conj of the reference spectrum, inverse transform scaled by N (FFTW's
backward transform is unnormalised).  The reference runs FFTW in FP32; this
oracle is exact to FP64 rounding, so device results are compared to it with
a tolerance.
"""
from __future__ import annotations

import math

import numpy as np

FREF = 37.53472224
NFFT = 32768
HDR = 10

_RATIO = {0: (3, 4), 1: (2, 3), 3: (5, 9), 4: (4, 9), 5: (3, 8), 6: (1, 3), 7: (1, 6),
          8: (3, 7), 9: (5, 16), 10: (3, 26), 11: (4, 11)}


def in2(b, i):
    """processsub.f90:163-167 on bytes b[i], b[i+1] (0-based)."""
    return int(b[i + 1]) + 256 * int(b[i])


def in3(b, i):
    """processsub.f90:169-173."""
    return int(b[i + 2]) + 256 * int(b[i + 1]) + 65536 * int(b[i])


def samplefrequency(code: int) -> float:
    """processsub.f90:181-212: FP32 quotient times 4 (FP32), times fref (FP64)."""
    p, q = _RATIO[code]
    r4 = np.float32(np.float32(p) / np.float32(q)) * np.float32(4.0)
    return float(np.float64(np.float32(r4)) * FREF)


def chirp_params(header) -> dict:
    """processsub.f90:75-86."""
    b = np.frombuffer(bytes(header), np.uint8)
    code = int(b[40])
    sf = np.float32(samplefrequency(code))
    w42, w44 = in2(b, 42), in2(b, 44)
    ip42 = (w42 & 32768) // 32768
    rr = np.float32(float((-1) ** (1 - ip42) * (w42 & 32767)) * FREF * FREF / 2 ** 21)
    ip44 = (w44 & 32768) // 32768
    sfreq = np.float32(float(np.float32(rr / np.float32(4.0))) / FREF +
                       float((-1) ** (1 - ip44) * (w44 & 32767)) * FREF / 2 ** 14)
    pl = np.float32(in3(b, 46) / FREF)
    nsamps = in2(b, 65) * 2
    npts = int(np.float32(pl * sf))
    return {"rangedecimation": code, "samplefreq": sf, "ramprate": rr, "startfreq": sfreq,
            "pulselength": pl, "nsamps": nsamps, "npts": npts, "nvalid": nsamps - npts}


def reference_function(c: dict, n: int = NFFT) -> np.ndarray:
    """reftime of processsub.f90:96-101 (complex64, zero beyond npts)."""
    npts = c["npts"]
    j = np.arange(npts)
    t = (j.astype(np.float32) / np.float32(c["samplefreq"])).astype(np.float64)
    pi = 4.0 * math.atan2(1.0, 1.0)
    phase = 2.0 * pi * float(c["startfreq"]) * t + pi * float(c["ramprate"]) * t * t
    ref = np.zeros(n, np.complex64)
    fn = np.float32(npts)
    ref[:npts] = (np.cos(phase).astype(np.float32) / fn) + 1j * (np.sin(phase).astype(np.float32)
                                                                / fn)
    return ref


def compress(lines: np.ndarray, c: dict, n: int = NFFT) -> np.ndarray:
    """processsub.f90:108-123 in FP64: complex128 [nlines][HDR + nvalid], the
    header samples copied, then the nvalid compressed samples."""
    lines = np.asarray(lines)
    nl = lines.shape[0]
    nsamps, nvalid = c["nsamps"], c["nvalid"]
    R = np.conj(np.fft.fft(reference_function(c, n).astype(np.complex128)))
    out = np.empty((nl, HDR + nvalid), np.complex128)
    for i in range(nl):
        x = np.zeros(n, np.complex128)
        x[:nsamps] = lines[i, HDR:HDR + nsamps]
        y = np.fft.ifft(np.fft.fft(x) * R) * n
        out[i, :HDR] = lines[i, :HDR]
        out[i, HDR:] = y[:nvalid]
    return out
