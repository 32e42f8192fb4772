"""Host-side pins of the interferogram oracle (int/crossmul.f90,
int/makecc.f90): the FFT upsampling against a direct DFT and against the
closed form of a band-limited row, an image with itself (real interferogram,
correlation 1), a constant phase screen, looks as block sums, masks, and the
makecc formula."""
import math

import numpy as np

from oracle import igram_oracle as IO


def direct_upsample(row, nnn):
    na = len(row)
    x = np.zeros(nnn, np.complex128)
    x[:na] = row
    k = np.arange(nnn)
    X = np.array([np.sum(x * np.exp(-2j * math.pi * k * m / nnn)) for m in range(nnn)])
    U = np.zeros(2 * nnn, np.complex128)
    U[:nnn // 2] = X[:nnn // 2]
    U[2 * nnn - nnn // 2:] = X[nnn // 2:]
    m = np.arange(2 * nnn)
    y = np.array([np.sum(U * np.exp(2j * math.pi * m * n / (2 * nnn))) for n in range(2 * nnn)])
    return y / nnn / 2.0


def test_fft_length():
    assert [IO.fft_length(n) for n in (1, 2, 3, 4000, 4096, 4097)] == [2, 2, 4, 4096, 4096, 8192]


def test_upsample_matches_direct_dft():
    rng = np.random.default_rng(1)
    row = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    np.testing.assert_allclose(IO.upsample2(row[None, :], 8)[0], direct_upsample(row, 8),
                               atol=1e-12)


def test_upsample_band_limited_closed_form():
    nnn = 64
    n = np.arange(nnn)
    rng = np.random.default_rng(2)
    ks = [-31, -7, 0, 3, 30]
    amps = rng.standard_normal(len(ks)) + 1j * rng.standard_normal(len(ks))
    f = lambda t: sum(a * np.exp(2j * math.pi * k * t / nnn) for a, k in zip(amps, ks))  # noqa
    up = IO.upsample2(f(n)[None, :], nnn)[0]
    np.testing.assert_allclose(up[0::2], f(n) / 2, atol=1e-12)          # the reference's 1/2
    np.testing.assert_allclose(up[1::2], f(n + 0.5) / 2, atol=1e-12)


def test_image_with_itself_is_real_and_fully_correlated():
    rng = np.random.default_rng(3)
    img = (rng.standard_normal((12, 50)) + 1j * rng.standard_normal((12, 50))).astype(np.complex64)
    ig, am = IO.crossmul(img, img, looksac=2, looksdn=3)
    assert ig.shape == (4, 25)
    assert np.all(np.abs(ig.imag) <= 1e-12 * ig.real) and np.all(ig.real > 0)
    np.testing.assert_array_equal(am.real, am.imag)
    cc = IO.makecc(ig, am)
    np.testing.assert_allclose(cc[:, 25:], 1.0, rtol=1e-12)
    np.testing.assert_allclose(cc[:, :25], am.real * am.imag)


def test_constant_phase_screen():
    rng = np.random.default_rng(4)
    img = (rng.standard_normal((8, 40)) + 1j * rng.standard_normal((8, 40))).astype(np.complex64)
    phi = 0.7
    img2 = (img * np.exp(-1j * phi)).astype(np.complex64)
    ig, am = IO.crossmul(img, img2, looksac=4, looksdn=2)
    np.testing.assert_allclose(np.angle(ig), phi, atol=1e-6)


def test_looks_are_block_sums():
    rng = np.random.default_rng(5)
    a = (rng.standard_normal((12, 30)) + 1j * rng.standard_normal((12, 30))).astype(np.complex64)
    b = (rng.standard_normal((12, 30)) + 1j * rng.standard_normal((12, 30))).astype(np.complex64)
    ig1, am1 = IO.crossmul(a, b, mask=False)
    ig, am = IO.crossmul(a, b, looksac=3, looksdn=4, mask=False)
    np.testing.assert_allclose(ig, ig1.reshape(3, 4, 10, 3).sum((1, 3)), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(am.real ** 2, (am1.real ** 2).reshape(3, 4, 10, 3).sum((1, 3)),
                               rtol=1e-12)


def test_mask_zeroes_no_data_blocks():
    rng = np.random.default_rng(6)
    a = (rng.standard_normal((8, 32)) + 1j * rng.standard_normal((8, 32))).astype(np.complex64)
    b = a.copy()
    a[:4, :8] = 0                        # no data in image 1 there
    b[4:, 24:] = 0                       # and in image 2 here
    ig, am = IO.crossmul(a, b, looksac=4, looksdn=4)
    assert np.all(ig[0, :2] == 0) and np.all(am[0, :2] == 0)
    assert np.all(ig[1, 6:] == 0) and np.all(am[1, 6:] == 0)
    assert np.all(ig[0, 2:] != 0) and np.all(ig[1, :6] != 0)
    igu, _ = IO.crossmul(a, b, looksac=4, looksdn=4, mask=False)
    assert np.all(igu[0, :2] != 0)       # unmasked: the upsampling spreads energy in


def test_makecc_formula():
    ig = np.array([[3 + 4j, 1 + 0j, 0j]])
    am = np.array([[2 + 5j, 0 + 3j, 1 + 1j]])
    out = IO.makecc(ig, am)
    np.testing.assert_allclose(out, [[10.0, 0.0, 1.0, 0.5, 0.0, 0.0]])
