"""Host-side tests of the range-compression path: the oracle's header decoding
against published Sentinel-1 IW1 chirp values and against the library's own
decoder bit for bit, the sampling-rate table, the oracle's FFT correlation
against a direct sum, and the compressed point target's closed form."""
import math

import numpy as np
import pytest

from oracle import range_oracle as RO
from srg_amd import raw
from srg_amd import tdbp as T


def test_decode_iw1_matches_published_values():
    c = RO.chirp_params(raw.header_bytes(raw.IW1))
    assert c["rangedecimation"] == 8
    assert abs(float(c["samplefreq"]) - 64.345238) < 1e-5            # MHz
    assert abs(float(c["ramprate"]) - 1.078230) <= RO.FREF ** 2 / 2 ** 21   # one code step
    assert abs(float(c["startfreq"]) - (-28.25216)) <= RO.FREF / 2 ** 14
    assert abs(float(c["pulselength"]) - 52.40481) <= 1.0 / RO.FREF
    # pulse length x sampling rate = 3371.99 samples; the reference's FP32
    # product rounds up to 3372.0 before the truncation
    assert abs(c["npts"] - 52.40481 * 64.345238) < 1 and c["npts"] == 3372
    assert c["nsamps"] == 24400 and c["nvalid"] == 24400 - c["npts"]
    # the chirp sweeps the band symmetrically: f0 + K T ~ -f0
    f_end = float(c["startfreq"]) + float(c["ramprate"]) * float(c["pulselength"])
    assert abs(f_end + float(c["startfreq"])) < 0.05


def test_samplefrequency_table():
    exact = {0: 3 / 4, 5: 3 / 8, 9: 5 / 16}
    for code, (p, q) in RO._RATIO.items():
        v = RO.samplefrequency(code)
        assert abs(v / (4 * RO.FREF * p / q) - 1) < 1e-7
        if code in exact:
            assert v == 4 * exact[code] * RO.FREF
    # the FP32 quotient, not the exact one, is what the reference computes
    assert RO.samplefrequency(8) != 4 * RO.FREF * 3 / 7


def random_header(rng):
    codes = raw.HeaderCodes(rangedecimation=int(rng.choice(list(RO._RATIO))),
                            txprr=int(rng.integers(0, 65536)), txpsf=int(rng.integers(0, 65536)),
                            txpl=int(rng.integers(1, 1 << 24)),
                            nsamps=2 * int(rng.integers(1, 32768)))
    return raw.header_bytes(codes)


def test_library_decoder_matches_oracle_bitwise():
    rng = np.random.default_rng(5)
    for _ in range(400):
        h = random_header(rng)
        o = RO.chirp_params(h)
        c = T.chirp_decode(h)
        for k in ("rangedecimation", "nsamps", "npts", "nvalid"):
            assert getattr(c, k) == o[k], k
        for k in ("samplefreq", "ramprate", "startfreq", "pulselength"):
            assert np.float32(getattr(c, k)).view(np.uint32) == np.float32(o[k]).view(np.uint32), k


def test_decoder_rejects_unknown_decimation():
    h = bytearray(raw.header_bytes(raw.IW1))
    h[40] = 2                                   # no such code in the table
    with pytest.raises(T.TdbpError):
        T.chirp_decode(bytes(h))


def small_chirp(npts=17, nsamps=100):
    return {"rangedecimation": 8, "samplefreq": np.float32(64.345238),
            "ramprate": np.float32(40.0), "startfreq": np.float32(-3.0),
            "pulselength": np.float32(npts / 64.345238), "nsamps": nsamps, "npts": npts,
            "nvalid": nsamps - npts}


def test_compress_is_circular_correlation():
    n = 128
    c = small_chirp()
    rng = np.random.default_rng(2)
    lines = (rng.standard_normal((3, 120)) + 1j * rng.standard_normal((3, 120))).astype(
        np.complex64)
    out = RO.compress(lines, c, n)
    r = RO.reference_function(c, n).astype(np.complex128)
    for i in range(3):
        x = np.zeros(n, np.complex128)
        x[:c["nsamps"]] = lines[i, RO.HDR:RO.HDR + c["nsamps"]]
        direct = [n * np.sum(x[(np.arange(n) + k) % n] * np.conj(r)) for k in range(c["nvalid"])]
        np.testing.assert_allclose(out[i, RO.HDR:], direct, rtol=0, atol=1e-9 * n)
        assert np.array_equal(out[i, :RO.HDR], lines[i, :RO.HDR])


def test_point_target_compresses_to_n():
    c = RO.chirp_params(raw.header_bytes(raw.IW1))
    ref = RO.reference_function(c)
    d = 5000
    line = np.zeros((1, RO.HDR + c["nsamps"]), np.complex64)
    line[0, RO.HDR + d:RO.HDR + d + c["npts"]] = ref[:c["npts"]] * np.float32(c["npts"])
    out = RO.compress(line, c)[0, RO.HDR:]
    assert int(np.argmax(np.abs(out))) == d
    assert abs(out[d] / RO.NFFT - 1) < 1e-6            # N * sum |chirp|^2 / npts = N
    side = np.abs(np.delete(out, np.arange(d - 2, d + 3)))
    assert side.max() < 0.3 * RO.NFFT


def test_shorter_transform_gives_the_reference_samples():
    """The kept samples are a linear correlation of the echo window: any
    transform length >= nsamps gives the 32768-point values up to the length
    factor of the unnormalised inverse (the device uses 24576 for IW1)."""
    c = RO.chirp_params(raw.header_bytes(raw.IW1))
    rng = np.random.default_rng(9)
    lines = (rng.standard_normal((2, 10 + c["nsamps"])) +
             1j * rng.standard_normal((2, 10 + c["nsamps"]))).astype(np.complex64)
    full = RO.compress(lines, c)[:, RO.HDR:]
    short = RO.compress(lines, c, n=24576)[:, RO.HDR:] * (32768 / 24576)
    np.testing.assert_allclose(short, full, rtol=0, atol=1e-9 * np.abs(full).max())
