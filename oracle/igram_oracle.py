"""CPU restatement of interferogram formation, int/crossmul.f90 and
int/makecc.f90 (test infrastructure: only tests/ and bench.py's CPU baseline
use it).

crossmul (crossmul.f90:86-199) per output line of looksdn rows: masks from
the unscaled rows (complex sum over the looksdn rows per column, then over
looksac columns, |sum| <= 1e-6 = no data, :110-122); rows times scale,
zero-padded to nnn = next power of two >= na (:72-76), forward FFT, the
spectrum's first nnn/2 bins at the start and last nnn/2 at the end of a
2 nnn array, inverse FFT (unnormalised), / nnn / 2 (:127-148); igram =
up1 conj(up2), amp = (|up1|^2, |up2|^2) on the first 2 na samples, adjacent
pairs summed, then rows, then looksac columns, amp square-rooted per
component, masked pixels zero (:150-184).  The transforms run in FP64
(numpy) where the reference runs FFTW in FP32; device results are compared
with a tolerance.

makecc (makecc.f90:46-60): amp = re(b) im(b); cc = |a| / amp where amp > 0,
else 0; output line [amp | cc].
"""
from __future__ import annotations

import numpy as np


def fft_length(na: int) -> int:
    """crossmul.f90:72-76: the first 2**i (i >= 1) not below na."""
    for i in range(1, 25):
        if 2 ** i >= na:
            return 2 ** i
    raise ValueError(na)


def upsample2(rows: np.ndarray, nnn: int) -> np.ndarray:
    """crossmul.f90:127-148 on rows [n][na] -> [n][2 nnn] (complex128)."""
    n, na = rows.shape
    x = np.zeros((n, nnn), np.complex128)
    x[:, :na] = rows
    X = np.fft.fft(x, axis=1)
    up = np.zeros((n, 2 * nnn), np.complex128)
    up[:, :nnn // 2] = X[:, :nnn // 2]
    up[:, 2 * nnn - nnn // 2:] = X[:, nnn // 2:]
    return np.fft.ifft(up, axis=1) * (2 * nnn) / nnn / 2.0


def crossmul(img1: np.ndarray, img2: np.ndarray, nvalid: int = None, scale: float = 1.0,
             looksac: int = 1, looksdn: int = 1, mask: bool = True):
    """Returns (igram, amp), complex128 [nvalid/looksdn][na/looksac]."""
    img1 = np.asarray(img1)
    img2 = np.asarray(img2)
    nl, na = img1.shape
    nvalid = nl if nvalid is None else nvalid
    nout, nw = nvalid // looksdn, na // looksac
    nnn = fft_length(na)
    rows = nout * looksdn
    u1 = upsample2(img1[:rows].astype(np.complex128) * np.float32(scale), nnn)[:, :2 * na]
    u2 = upsample2(img2[:rows].astype(np.complex128) * np.float32(scale), nnn)[:, :2 * na]
    ig = u1 * np.conj(u2)
    am = np.abs(u1) ** 2 + 1j * np.abs(u2) ** 2
    ig = ig[:, 0::2] + ig[:, 1::2]                                 # [rows][na]
    am = am[:, 0::2] + am[:, 1::2]
    ig = ig.reshape(nout, looksdn, na).sum(1)[:, :nw * looksac]
    am = am.reshape(nout, looksdn, na).sum(1)[:, :nw * looksac]
    ig = ig.reshape(nout, nw, looksac).sum(2)
    am = am.reshape(nout, nw, looksac).sum(2)
    am = np.sqrt(am.real) + 1j * np.sqrt(am.imag)
    if mask:
        def flags(img):
            s = img[:rows].astype(np.complex64).reshape(nout, looksdn, na).sum(1, dtype=np.complex64)
            s = s[:, :nw * looksac].reshape(nout, nw, looksac).sum(2, dtype=np.complex64)
            return np.abs(s) > 1e-6
        keep = flags(img1) & flags(img2)
        ig = np.where(keep, ig, 0)
        am = np.where(keep, am, 0)
    return ig, am


def makecc(igram: np.ndarray, amp: np.ndarray) -> np.ndarray:
    """[nlines][2 npix]: amplitude rows then correlation rows, per line."""
    a = np.asarray(amp.real, np.float64) * np.asarray(amp.imag, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        cc = np.where(a > 0, np.abs(igram) / np.where(a > 0, a, 1), 0.0)
    return np.concatenate([a, cc], axis=1)
